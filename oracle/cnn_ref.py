"""FP64 reference of the learned proximal CNN (5 layers of 3x3x3 conv, 1 -> 32
-> 32 -> 32 -> 32 -> 1, ReLU after layers 1-4, residual output x + net(x)).
TEST INFRASTRUCTURE ONLY: the checker of the HIP training / inference kernels.
Weights: w[l] of shape (27, cout, cin), b[l] of shape (cout,).  Volumes are
(YN, XN, ZN) in the (Z,X,Y) order; activations (nvox, C)."""
import numpy as np

import oracle as O

CH = [1, 32, 32, 32, 32, 1]


def bf16_round(a):
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(ws, bs, x, shape, bf16=False):
    """returns output (nvox,) and the list of post-activation tensors a[0..5].
    bf16=True rounds weights and the layer 1-4 activations to bfloat16 where the
    GPU kernels store them (accumulation stays exact here)."""
    rnd = bf16_round if bf16 else (lambda v: v)
    a = [np.asarray(x, np.float64).reshape(-1, 1)]
    for l in range(5):
        z = O.conv3d(a[-1], rnd(ws[l]), bs[l], shape)
        a.append(rnd(np.maximum(z, 0.0)) if l < 4 else z)
    return a[0][:, 0] + a[5][:, 0], a


def loss_and_grads(ws, bs, x, target, shape, bf16=False):
    """L = 0.5 * mean((x + net(x) - target)^2); returns L, dws, dbs"""
    rnd = bf16_round if bf16 else (lambda v: v)
    out, a = forward(ws, bs, x, shape, bf16)
    n = out.size
    r = out - np.asarray(target, np.float64).reshape(-1)
    L = 0.5 * float(r @ r) / n
    g = (r / n).reshape(-1, 1)                    # dL/dz5 (FP32 on the GPU)
    dws, dbs = [None] * 5, [None] * 5
    for l in range(4, -1, -1):
        dws[l], dbs[l] = O.conv3d_wgrad(a[l], g, CH[l], CH[l + 1], shape)
        if l > 0:
            g = rnd(O.conv3d_dgrad(g, rnd(ws[l]), shape) * (a[l] > 0))
    return L, dws, dbs


def flatten(ws, bs):
    """flat parameter vector in the library order W1 b1 W2 b2 ... W5 b5"""
    return np.concatenate([np.concatenate([w.reshape(-1), b.reshape(-1)]) for w, b in zip(ws, bs)])


def unflatten(p):
    ws, bs, o = [], [], 0
    for l in range(5):
        n = 27 * CH[l + 1] * CH[l]
        ws.append(np.asarray(p[o:o + n], np.float64).reshape(27, CH[l + 1], CH[l]))
        o += n
        bs.append(np.asarray(p[o:o + CH[l + 1]], np.float64))
        o += CH[l + 1]
    return ws, bs


def init_weights(seed):
    """He-normal weights, zero biases; the last layer starts near zero so the
    network starts close to the identity map."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ws, bs = [], []
    for l in range(5):
        cin, cout = CH[l], CH[l + 1]
        std = np.sqrt(2.0 / (27 * cin)) * (0.1 if l == 4 else 1.0)
        ws.append(rng.standard_normal((27, cout, cin)) * std)
        bs.append(np.zeros(cout))
    return ws, bs
