"""Every benchmark configuration at its full size, on the GPU.

cfg2  512x512x64 volume, 984 views, 888x64 arc detector: both operators at full
      size, checked against the FP64 oracle on 48 views spread over the rotation
      (relative L2 <= 1e-5), and the full-size adjoint identity (<= 1e-6).
cfg3  256x256 fan beam, 360 views, 888 cells, Poisson noise I0 = 1e5: OS-SART
      (10 subsets x 20 iterations) and FISTA-TV (100 iterations, lambda 1e-4)
      against the FP64 oracle drivers (relative L2 <= 1e-4).
cfg4  the 32 uneven OS-SART problems (seeds 100-131, up to 512x512x64 and 984
      views): one batched iteration (persistent launches, HIP graph) equals the
      per-problem iterations bit for bit.
cfg5  512x512x64, 120 views, I0 = 1e4: the learned step's safeguard (an
      accepted learned output never has a larger ||Ax - b||^2 than the plain
      iterate it replaces) and a zero network giving plain FISTA-TV's outputs
      bit for bit.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, rel_l2

pytestmark = pytest.mark.gpu

sys.path.insert(0, ROOT)


def test_cfg2_full_size(cuda):
    import torch
    import bench
    from ctlab_amd import dd
    g = bench.cfg2_geometry()
    vol = bench.cfg2_phantom(g)
    plan = dd.Plan(g)
    p = plan.proj(torch.from_numpy(vol).cuda())
    b = plan.back(p)
    torch.cuda.synchronize()
    views = bench.sample_views(g.PN, 48)
    sub = g.subset(views)
    p_np = p.cpu().numpy()
    e_p = rel_l2(p_np[views], O.proj(sub, vol))
    bs = dd.Plan(sub).back(torch.from_numpy(np.ascontiguousarray(p_np[views])).cuda())
    torch.cuda.synchronize()
    e_b = rel_l2(bs.cpu().numpy(), O.back(sub, p_np[views].reshape(-1)))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(3)
    x = torch.rand(g.vol_shape, generator=gen, device="cuda")
    gen.manual_seed(4)
    y = torch.rand(g.prj_shape, generator=gen, device="cuda")
    lhs = float((plan.proj(x).double() * y.double()).sum())
    rhs = float((x.double() * plan.back(y).double()).sum())
    adj = abs(lhs - rhs) / abs(lhs)
    print(f"cfg2 full size: proj {e_p:.2e}, back {e_b:.2e} (48 views vs FP64), adjoint {adj:.2e}, "
          f"backprojection norm {float(b.norm()):.4e}")
    assert e_p <= 1e-5 and e_b <= 1e-5 and adj <= 1e-6
    assert torch.isfinite(b).all()


def test_cfg3_os_sart_and_fista_tv(cuda):
    import torch
    import recon_ref as RR
    from ctlab_amd import dd
    from ctlab_amd import geometry as G
    from ctlab_amd import phantom as PH
    from ctlab_amd import recon as RC
    g = G.fan_multislice(XN=256, YN=256, SLN=1, dx=1.0, PN=360, DNU=888, col_size=1.0239, sid=541.0, sdd=949.0)
    ph = PH.voxelize(PH.random_ellipses(1, 10, 256, 256, g.dx), 256, 256, 1, g.dx, 1.0)
    plan = dd.Plan(g)
    clean = plan.proj(torch.from_numpy(np.ascontiguousarray(ph)).cuda()).cpu().numpy()
    b = PH.add_poisson_noise(clean, 1e5, 3).reshape(-1)
    bt = torch.from_numpy(b).cuda()
    rc = RC.Recon(plan, 10)
    rc.set_data(bt)
    xs = torch.zeros(g.vol_shape, device="cuda")
    rc.os_sart(xs, 20)
    rf = RC.Recon(plan, 1)
    rf.set_data(bt)
    L = float(np.float32(1.01 * rf.lipschitz(30)))
    xf = torch.zeros(g.vol_shape, device="cuda")
    rf.fista_tv(xf, 100, 1e-4, 10, L=L)
    torch.cuda.synchronize()
    e_s = rel_l2(xs.cpu().numpy(), RR.os_sart(g, b, 10, 20))
    xf_ref, _ = RR.fista_tv(g, b, 100, 1e-4, 10, L)
    e_f = rel_l2(xf.cpu().numpy(), xf_ref)
    print(f"cfg3: OS-SART 10x20 {e_s:.2e}, FISTA-TV 100 {e_f:.2e} vs the FP64 drivers")
    assert e_s <= 1e-4 and e_f <= 1e-4


def test_cfg4_all_problems_batched_equals_per_problem(cuda):
    import torch
    import bench
    from ctlab_amd import dd
    from ctlab_amd import learned as LE
    from ctlab_amd import phantom as PH
    from ctlab_amd import recon as RC
    recs, shapes = [], []
    for sp in bench.cfg4_specs(32):
        g = bench.cfg4_geometry(sp)
        plan = dd.Plan(g)
        ph = LE.voxelize_torch(PH.random_ellipsoids(sp["seed"], 20, g.XN, g.YN, g.ZN, g.dx, g.dz), g.vol_shape,
                               g.dx, g.dz, "cuda")
        rc = RC.Recon(plan, 10)
        rc.set_data(LE.noisy_projection(plan, ph, sp["I0"], sp["seed"]))
        recs.append(rc)
        shapes.append(g.vol_shape)
        del ph
    print(f"cfg4: 32 problems set up, {torch.cuda.memory_allocated() / 2**30:.1f} GiB (torch)", flush=True)
    bt = RC.Batch(recs)
    xb = [torch.zeros(s, device="cuda") for s in shapes]
    bt.os_sart(xb, 1)
    torch.cuda.synchronize()
    nu, nm = bt.tiles()
    for rc, s, x in zip(recs, shapes, xb):
        rc.set_tiles(nu, nm)
        xs = torch.zeros(s, device="cuda")
        rc.os_sart(xs, 1)
        torch.cuda.synchronize()
        assert torch.equal(xs, x)
        assert float(x.abs().sum()) > 0


def test_cfg5_full_size_safeguard_and_zero_network(cuda):
    import torch
    from ctlab_amd import cnn as CN
    from ctlab_amd import dd
    from ctlab_amd import learned as LE
    from ctlab_amd import recon as RC
    ge = LE.cone_geometry(512, 512, 64, 120)
    pe = dd.Plan(ge)
    re_ = RC.Recon(pe, 1)
    ph = LE.make_problem(ge, pe, re_, 7, 1e4, "cuda")
    L = re_.lipschitz(30) * 1.01
    ze = LE.covered_slices(ge)
    iters, lam = 20, 0.1
    _, h_plain, _ = LE.fista_tv_run(re_, L, ge.vol_shape, iters, lam, 10, ph, ze, "cuda")
    weights = os.path.join(ROOT, "ctlab_amd", "weights", "cnn_prox_cfg5.npy")
    params = (np.load(weights, allow_pickle=False).astype(np.float32) if os.path.exists(weights)
              else CN.init_params(0))
    dca = LE.DataConsistent(LE.ScaledCNN(ge.vol_shape, params), re_, to_fidelity=True, m_max=10)
    _, h_l, _ = LE.fista_tv_run(re_, L, ge.vol_shape, iters, lam, 10, ph, ze, "cuda", learned=dca, learned_every=5,
                                feedback=False)
    accepted = [h for h in h_l if h["learned"]]
    for h in accepted:
        assert h["fidelity"] <= h_plain[h["iter"] - 1]["fidelity"]
    for h in h_l:
        if not h["learned"]:
            assert h["psnr"] == h_plain[h["iter"] - 1]["psnr"]
    print(f"cfg5: {len(accepted)} of {iters // 5} learned outputs accepted, each at a residual no higher than "
          f"plain's; PSNR {h_plain[-1]['psnr']:.2f} -> {h_l[-1]['psnr']:.2f} dB at k = {iters}")
    # a network that outputs 0 leaves every output equal to plain FISTA-TV's
    zero = LE.ScaledCNN(ge.vol_shape, np.zeros_like(params))
    x_out = torch.empty(ge.vol_shape, device="cuda")
    xk = torch.rand(ge.vol_shape, device="cuda")
    zero(xk, x_out)
    assert torch.equal(x_out, xk)
    _, h_z, _ = LE.fista_tv_run(re_, L, ge.vol_shape, iters, lam, 10, ph, ze, "cuda", learned=zero,
                                learned_every=5, feedback=False)
    assert [h["psnr"] for h in h_z] == [h["psnr"] for h in h_plain]
    assert [h["fidelity"] for h in h_z] == [h["fidelity"] for h in h_plain]
