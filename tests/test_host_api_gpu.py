"""OS_SART / CG / evalProjBack with the reference's host-level argument lists
(src/DD3_GPU_recon.h:21-103) and raw device file output, on the GPU: the
drivers equal the device-resident ctlab_os_sart / ctlab_cg bit for bit, and
each intermediate file holds that iteration's volume."""
import numpy as np
import pytest

from ctlab_amd import _lib, dd, rawio
from ctlab_amd import geometry as G
from ctlab_amd import phantom as PH
from ctlab_amd import recon as RC

pytestmark = pytest.mark.gpu

SCAN = dict(sid=150.0, sdd=260.0, DNU=72, DNV=12, col_size=1.3, row_size=1.6, col_offset=0.25, row_offset=-0.5)
VOL = dict(XN=40, YN=36, ZN=12, dx=1.0, dz=1.5)
PN = 48


def _problem():
    angs = (0.1 + np.arange(PN) * (2 * np.pi / PN)).astype(np.float32)
    zpos = np.linspace(-3.0, 3.0, PN).astype(np.float32)          # a short helix
    # the drivers' own detector (built from float32 scan arguments)
    xds, yds, zds = RC.driver_detector(SCAN["sid"], SCAN["sdd"], SCAN["DNU"], SCAN["DNV"], SCAN["col_size"],
                                       SCAN["row_size"], SCAN["col_offset"], SCAN["row_offset"])
    g = G.ConeGeometry(x0=0.0, y0=SCAN["sid"], z0=0.0, DNU=SCAN["DNU"], DNV=SCAN["DNV"], xds=xds, yds=yds, zds=zds,
                       imgXCenter=0.0, imgYCenter=0.0, imgZCenter=0.0, angs=angs, zpos=zpos, **VOL)
    ells = PH.random_ellipsoids(7, 6, g.XN, g.YN, g.ZN, g.dx, g.dz, semi_xy=(3.0, 15.0), semi_z=(2.0, 6.0))
    vol = PH.voxelize(ells, g.XN, g.YN, g.ZN, g.dx, g.dz)
    return g, angs, zpos, vol


def _args(angs, zpos, name, osNum, iters, med):
    return dict(hangs=angs, hzPos=zpos, mask=None, VolumeName=name, osNum=osNum, iterNum=iters, outputMedRes=med,
                sid=SCAN["sid"], sdd=SCAN["sdd"], DNU=SCAN["DNU"], DNV=SCAN["DNV"], PN=PN, imgXCenter=0.0,
                imgYCenter=0.0, imgZCenter=0.0, XN=VOL["XN"], YN=VOL["YN"], ZN=VOL["ZN"], dx=VOL["dx"],
                dz=VOL["dz"], col_size=SCAN["col_size"], row_size=SCAN["row_size"], col_offset=SCAN["col_offset"],
                row_offset=SCAN["row_offset"])


def test_os_sart_driver_equals_device_driver(cuda, tmp_path):
    import torch
    g, angs, zpos, vol = _problem()
    plan = dd.Plan(g)
    b = plan.proj(torch.from_numpy(vol).cuda())
    hprj = b.cpu().numpy()
    name = str(tmp_path / "sart.raw")
    out = RC.OS_SART(None, hprj, **_args(angs, zpos, name, 4, 3, True))
    rc = RC.Recon(plan, 4)
    rc.set_data(b)
    x = torch.zeros(g.vol_shape, device=cuda)
    per_iter = []
    for _ in range(3):
        rc.os_sart(x, 1, lam=1.0)
        per_iter.append(x.cpu().numpy())
    assert np.array_equal(out, per_iter[-1])
    assert np.array_equal(rawio.read_raw(name, shape=g.vol_shape), out)
    for k in range(3):
        assert np.array_equal(rawio.read_raw(f"{name}_{k + 1}", shape=g.vol_shape), per_iter[k])
    err = np.linalg.norm(out - vol) / np.linalg.norm(vol)
    print(f"OS_SART host driver, 3 iterations: rel error vs phantom {err:.3f}")
    assert err < 0.6
    # warm start from the first iterate continues the same iteration
    out2 = RC.OS_SART(per_iter[0], hprj, **_args(angs, zpos, "", 4, 2, False))
    assert np.array_equal(out2, per_iter[-1])


def test_cg_driver_equals_device_driver(cuda, tmp_path):
    import torch
    g, angs, zpos, vol = _problem()
    plan = dd.Plan(g)
    b = plan.proj(torch.from_numpy(vol).cuda())
    hprj = b.cpu().numpy()
    name = str(tmp_path / "cg.raw")
    for cgm in (RC.CG_FR, RC.CG_DY):
        out = RC.CG(None, hprj, **_args(angs, zpos, name, 1, 4, True), cgm=cgm)
        rc = RC.Recon(plan, 1)
        rc.set_data(b)
        x = torch.zeros(g.vol_shape, device=cuda)
        rc.cg(x, 4, cgm=cgm)
        assert np.array_equal(out, x.cpu().numpy())
        mids = [rawio.read_raw(f"{name}_{k}", shape=g.vol_shape) for k in range(1, 5)]
        assert np.array_equal(mids[-1], out)
        res = [np.linalg.norm(plan.proj(torch.from_numpy(m).cuda()).cpu().numpy() - hprj) for m in mids]
        print(f"CG cgm={cgm}: ||Ax-b|| per iteration {[f'{r:.3g}' for r in res]}")
        assert res[-1] < res[0]


def test_eval_proj_back_and_errors(cuda, tmp_path):
    import torch
    g, angs, zpos, vol = _problem()
    plan = dd.Plan(g)
    rows = RC.evalProjBack(plan, True)
    cols = RC.evalProjBack(plan, False)
    assert torch.equal(rows, plan.proj(torch.ones(g.vol_shape, device=cuda)))
    assert torch.equal(cols, plan.back(torch.ones(g.prj_shape, device=cuda)))
    with pytest.raises(_lib.CtlabError) as e:
        RC.evalProjBack(plan, True, forwardMethod=2)
    assert e.value.code == 2
    with pytest.raises(_lib.CtlabError) as e:
        RC.OS_SART(None, np.zeros(PN * SCAN["DNU"] * SCAN["DNV"], np.float32),
                   **{**_args(angs, zpos, "", 2, 1, False), "forwardMethod": 2})
    assert e.value.code == 2


def test_raw_device_round_trip(cuda, tmp_path):
    import torch
    t = torch.randn(3, 5, 7, device=cuda, dtype=torch.float64)
    p = str(tmp_path / "d.raw")
    rawio.write_raw_device(p, t)
    assert np.array_equal(rawio.read_raw(p, dtype=np.float64, shape=(3, 5, 7)), t.cpu().numpy())
    u = torch.empty_like(t)
    rawio.read_raw_device(p, u)
    assert torch.equal(u, t)
