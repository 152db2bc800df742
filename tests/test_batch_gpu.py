"""Batched OS-SART over problems of uneven volume, view, detector and subset
sizes: the persistent batched launches must reproduce the per-problem
OS-SART (itself checked against the FP64 reference driver) bit for bit."""
import numpy as np
import pytest

import recon_ref as RR
from conftest import rel_l2
from ctlab_amd import dd
from ctlab_amd import geometry as G
from ctlab_amd import phantom as PH
from ctlab_amd import recon as RC

pytestmark = pytest.mark.gpu

# (XN, YN, ZN, PN, DNU, DNV, osNum, ang0, masked)
SHAPES = [
    (40, 36, 8, 48, 72, 10, 4, 0.10, False),
    (32, 48, 12, 30, 60, 14, 3, 0.40, True),
    (24, 24, 6, 20, 40, 8, 1, 0.00, False),
    (56, 40, 16, 36, 90, 18, 5, 0.77, False),
    (20, 28, 4, 25, 33, 5, 2, 1.30, False),
]


def _problems(tiles=(16, 16)):
    import torch
    out = []
    for i, (XN, YN, ZN, PN, DNU, DNV, os_, a0, masked) in enumerate(SHAPES):
        mask = None
        if masked:
            mask = (np.random.default_rng(40 + i).random(XN * YN) > 0.15).astype(np.uint8)
        g = G.circular_cone(XN=XN, YN=YN, ZN=ZN, dx=1.0, dz=1.2, PN=PN, DNU=DNU, DNV=DNV, col_size=1.1,
                            row_size=1.3, sid=110.0, sdd=190.0, ang0=a0, mask=mask)
        ells = PH.random_ellipsoids(30 + i, 5, XN, YN, ZN, g.dx, g.dz, semi_xy=(2, 10), semi_z=(1, 4))
        ph = PH.voxelize(ells, XN, YN, ZN, g.dx, g.dz)
        plan = dd.Plan(g)
        plan.set_tiles(*tiles)
        b = plan.proj(torch.from_numpy(ph).cuda())
        rc = RC.Recon(plan, os_)
        rc.set_data(b)
        out.append((g, plan, rc, b))
    return out


def _per_problem(probs, iters, lam=1.0, nonneg=False):
    import torch
    xs = []
    for g, _, rc, _ in probs:
        x = torch.zeros(g.vol_shape, device="cuda")
        rc.os_sart(x, iters, lam, nonneg, use_graph=False)
        xs.append(x)
    torch.cuda.synchronize()
    return xs


@pytest.mark.parametrize("nonneg", [False, True])
def test_batch_equals_per_problem_bit_exact(cuda, nonneg):
    import torch
    probs = _problems()
    ref = _per_problem(probs, 2, 0.9, nonneg)
    bt = RC.Batch([p[2] for p in probs])
    assert bt.tiles() == (16, 16)
    xg = [torch.zeros(p[0].vol_shape, device=cuda) for p in probs]
    bt.os_sart(xg, 2, 0.9, nonneg, use_graph=True)
    xe = [torch.zeros(p[0].vol_shape, device=cuda) for p in probs]
    bt.os_sart(xe, 1, 0.9, nonneg, use_graph=False)
    bt.os_sart(xe, 1, 0.9, nonneg, use_graph=False)
    torch.cuda.synchronize()
    for i, (a, b, c) in enumerate(zip(ref, xg, xe)):
        assert torch.equal(a, b), f"problem {i}: graph replay differs"
        assert torch.equal(a, c), f"problem {i}: eager batch differs"
        assert float(a.abs().max()) > 0
    # one more replay on volumes that already hold an iterate (graph reused)
    ref3 = [x.clone() for x in ref]
    for (g, _, rc, _), x in zip(probs, ref3):
        rc.os_sart(x, 1, 0.9, nonneg, use_graph=False)
    bt.os_sart(xg, 1, 0.9, nonneg, use_graph=True)
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(ref3, xg)):
        assert torch.equal(a, b), f"problem {i}: third iteration differs"


def test_batch_matches_fp64_driver(cuda):
    import torch
    probs = _problems()
    bt = RC.Batch([p[2] for p in probs])
    xs = [torch.zeros(p[0].vol_shape, device=cuda) for p in probs]
    bt.os_sart(xs, 2)
    torch.cuda.synchronize()
    for (g, _, rc, b), x in zip(probs, xs):
        ref = RR.os_sart(g, b.cpu().numpy(), rc.osNum, 2)
        e = rel_l2(x.cpu().numpy(), ref)
        print(f"batched OS-SART {g.vol_shape} PN={g.PN} osNum={rc.osNum}: rel L2 vs FP64 driver {e:.2e}")
        assert e < 1e-4


@pytest.mark.parametrize("tiles", [(8, 12), (12, 8)])
def test_batch_tiles_and_autotune(cuda, tiles):
    import torch
    probs = _problems(tiles)
    ref = _per_problem(probs, 1)
    bt = RC.Batch([p[2] for p in probs])
    bt.set_tiles(*tiles)
    xs = [torch.zeros(p[0].vol_shape, device=cuda) for p in probs]
    bt.os_sart(xs, 1)
    torch.cuda.synchronize()
    for a, b in zip(ref, xs):
        assert torch.equal(a, b)
    scratch = [torch.zeros(p[0].vol_shape, device=cuda) for p in probs]
    ms = bt.autotune(scratch)
    assert set(ms["proj"]) == {8, 12, 16} and min(ms["back"].values()) > 0
    nu, nm = bt.tiles()
    assert nu in (8, 12, 16) and nm in (8, 12, 16)
    xs2 = [torch.zeros(p[0].vol_shape, device=cuda) for p in probs]
    bt.os_sart(xs2, 1)
    torch.cuda.synchronize()
    for a, b in zip(ref, xs2):
        assert rel_l2(b.cpu().numpy(), a.cpu().numpy()) < 1e-5
    # per-problem runs on the tuned tiles (graph path, recaptured) match bit for bit
    for g, _, rc, _ in probs:
        rc.set_tiles(nu, nm)
    for (g, _, rc, _), b in zip(probs, xs2):
        x = torch.zeros(g.vol_shape, device=cuda)
        rc.os_sart(x, 1, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(x, b)
