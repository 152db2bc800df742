"""BF16 MFMA CNN kernels against the FP64 oracle (oracle/cnn_ref.py).

With bf16=True the oracle rounds weights and stored activations / gradients to
bfloat16 where the kernels store them; what remains is FP32-vs-FP64
accumulation order, so forward and gradients agree closely.  Against the
unrounded FP64 network the BF16 path is within bfloat16 error."""
import numpy as np
import pytest

import cnn_ref as CR
from conftest import rel_l2
from ctlab_amd import cnn as CN
from ctlab_amd import phantom as PH

pytestmark = pytest.mark.gpu


def _data(shape, seed):
    YN, XN, ZN = shape
    ells = PH.random_ellipsoids(seed, 4, XN, YN, ZN, 1.0, 1.0, semi_xy=(2, 6), semi_z=(1, 4), mu=(0.2, 1.0))
    t = PH.voxelize(ells, XN, YN, ZN, 1.0, 1.0)
    x = t + np.random.default_rng(seed).normal(0, 0.1, t.shape).astype(np.float32)
    return x, t


@pytest.mark.parametrize("shape", [(6, 7, 16), (5, 9, 24), (5, 17, 40), (9, 10, 72)])
def test_forward_matches_oracle(cuda, shape):
    import torch
    x, _ = _data(shape, 1)
    p = CN.init_params(3).astype(np.float64)
    p[-1] = 0.05
    ws, bs = CR.unflatten(p)
    ref_bf, _ = CR.forward(ws, bs, x, shape, bf16=True)
    ref_64, _ = CR.forward(ws, bs, x, shape, bf16=False)
    for training in (False, True):
        net = CN.CNN(shape, training=training, params=p)
        out = net(torch.from_numpy(x).cuda())
        torch.cuda.synchronize()
        got = out.cpu().numpy().reshape(-1).astype(np.float64)
        xf = x.reshape(-1).astype(np.float64)
        e_bf = rel_l2(got - xf, ref_bf - xf)
        e_64 = rel_l2(got - xf, ref_64 - xf)
        print(f"cnn forward training={training}: vs bf16-emulating oracle {e_bf:.2e}, vs FP64 {e_64:.2e}")
        assert e_bf < 2e-3
        assert e_64 < 3e-2


# ZN % 32 == 0 takes the x-sweep weight-gradient kernel (one or several runs
# of 32 z per column); the other shapes take the per-(dy, dx) kernel
@pytest.mark.parametrize("shape", [(6, 5, 16), (5, 17, 40), (6, 5, 32), (4, 9, 64), (3, 11, 96)])
def test_gradients_match_oracle(cuda, shape):
    import torch
    x, t = _data(shape, 2)
    p = CN.init_params(4).astype(np.float64)
    ws, bs = CR.unflatten(p)
    L_ref, dws, dbs = CR.loss_and_grads(ws, bs, x, t, shape, bf16=True)
    g_ref = CR.flatten(dws, dbs)
    net = CN.CNN(shape, training=True, params=p)
    L, g = net.loss_grad(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda(), want_grad=True)
    assert abs(L - L_ref) < 1e-3 * L_ref
    wr, br = CR.unflatten(g_ref)
    wg, bg = CR.unflatten(g.astype(np.float64))
    for l in range(5):
        ew, eb = rel_l2(wg[l], wr[l]), rel_l2(bg[l], br[l])
        print(f"layer {l + 1}: dW rel L2 {ew:.2e}, db rel L2 {eb:.2e}")
        assert ew < 2e-2 and eb < 2e-2


def test_forward_at_evaluation_size_matches_oracle_on_crops(cuda):
    """512x512x64 forward (the evaluation size, 65536 tiles of each layer's
    kernel): the network output over a 22^3 crop in the interior and a 16^3
    crop at the volume's corner, against the oracle run on the crop alone.  A
    margin of 5 voxels (the 5 layers' receptive-field radius) on every side that
    is not the volume's own boundary is left out of the comparison."""
    import torch
    shape = (512, 512, 64)
    rng = np.random.default_rng(8)
    x = rng.random(shape, dtype=np.float32)
    p = CN.init_params(6).astype(np.float64)
    p[-1] = 0.05
    ws, bs = CR.unflatten(p)
    net = CN.CNN(shape, training=False, params=p)
    out = net(torch.from_numpy(x.reshape(-1)).cuda()).cpu().numpy().reshape(shape).astype(np.float64)
    for (y0, x0, z0), c, keep in (((300, 123, 21), 22, (slice(5, 17),) * 3),
                                  ((0, 0, 0), 16, (slice(0, 11),) * 3)):
        crop = x[y0:y0 + c, x0:x0 + c, z0:z0 + c].astype(np.float64)
        ref, _ = CR.forward(ws, bs, crop.reshape(-1), (c, c, c), bf16=True)
        ref = ref.reshape(c, c, c)[keep]
        got = out[y0:y0 + c, x0:x0 + c, z0:z0 + c][keep]
        base = crop[keep]
        e = rel_l2((got - base).reshape(-1), (ref - base).reshape(-1))
        print(f"512x512x64 forward, crop at {(y0, x0, z0)}: vs bf16-emulating oracle {e:.2e}")
        assert e < 2e-3


def test_training_reduces_denoising_loss(cuda):
    import torch
    shape = (12, 12, 16)
    net = CN.CNN(shape, training=True, params=CN.init_params(5))
    losses = []
    for it in range(60):
        x, t = _data(shape, 100 + it % 6)
        losses.append(net.loss_grad(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()))
        net.adam_step(2e-3)
    print("loss", losses[0], "->", np.mean(losses[-6:]))
    assert np.mean(losses[-6:]) < 0.7 * np.mean(losses[:6])
