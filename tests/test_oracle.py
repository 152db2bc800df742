"""Pins of the FP64 CPU oracle (CPU only).

The reference's DD bodies are not in its tree, so the oracle is pinned by
(a) brute force: the dense matrix from direct interval overlaps, (b) the
textbook identity branchless DD == conventional overlap-merge DD in 2-D,
(c) closed forms (uniform box -> exact chord length), (d) exact symmetries
(90-degree rotation == view shift, mirror == reversed detector), (e) the
adjoint identity, and (f) analytic ellipsoid line integrals.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from ctlab_amd import geometry as G
from ctlab_amd import phantom as PH


def small_cone(PN=12, ang0=0.1, **kw):
    args = dict(XN=6, YN=5, ZN=4, dx=2.0, dz=2.5, PN=PN, DNU=9, DNV=5, col_size=3.0,
                row_size=2.0, sid=40.0, sdd=70.0, ang0=ang0)
    args.update(kw)
    return G.circular_cone(**args)


def test_dense_equals_sat_projector_and_transpose():
    g = small_cone()
    rng = np.random.default_rng(1)
    x = rng.random(g.nvox())
    y = rng.random(g.nprj())
    A = O.dense(g)
    assert rel_l2(O.proj(g, x), A @ x) < 1e-13
    assert rel_l2(O.back(g, y), A.T @ y) < 1e-13
    assert np.count_nonzero(A) > 0.05 * A.size


def test_dense_flat_detector_and_mask():
    mask = np.ones(6 * 5, np.uint8)
    mask[[0, 7, 13, 29]] = 0
    g = small_cone(arc=False, mask=mask, PN=7, ang0=0.3)
    rng = np.random.default_rng(2)
    x = rng.random(g.nvox())
    A = O.dense(g)
    assert rel_l2(O.proj(g, x), A @ x) < 1e-13
    cols = np.repeat(mask == 0, g.ZN)
    assert np.all(A[:, cols] == 0)


def small_helix(**kw):
    args = dict(XN=6, YN=5, ZN=6, dx=2.0, dz=2.5, PN=16, DNU=9, DNV=4, col_size=3.0, row_size=2.0,
                pitch=9.0, turns=1.5, sid=40.0, sdd=70.0, ang0=0.2)
    args.update(kw)
    return G.helical_cone(**args)


def test_dense_equals_sat_projector_helical():
    """Helical scan (per-view source z): the SAT projector and its transpose
    equal the direct interval-overlap matrix."""
    g = small_helix()
    assert np.ptp(g.zpos) > 10.0
    rng = np.random.default_rng(21)
    x = rng.random(g.nvox())
    y = rng.random(g.nprj())
    A = O.dense(g)
    assert rel_l2(O.proj(g, x), A @ x) < 1e-13
    assert rel_l2(O.back(g, y), A.T @ y) < 1e-13
    # the views see different slabs of z: some (view, slice) pairs are disjoint
    seen = (np.abs(A).reshape(g.PN, -1, g.YN * g.XN, g.ZN).sum(axis=(1, 2)) > 0)
    assert seen.any(axis=1).all() and not seen.all()


def test_helical_rigid_translation():
    """Moving the source path (hzPos) and the volume (imgZCenter) by the same
    distance along z leaves every projection unchanged; moving only the source
    path by one slice equals moving the volume content down by one slice."""
    g = small_helix()
    rng = np.random.default_rng(22)
    x = rng.random(g.nvox())
    ref = O.proj(g, x)
    d = 3.75    # hzPos are multiples of 27/32 mm: every shifted value stays exact in float32
    g2 = small_helix()
    g2.zpos = (g.zpos.astype(np.float64) + d).astype(np.float32)
    g2.imgZCenter = g.imgZCenter + d
    assert rel_l2(O.proj(g2, x), ref) < 1e-13
    # source path up by one slice (dz = 2.5 mm) = content down by one slice,
    # for content that stays inside the volume (bottom slice empty)
    v = x.reshape(g.YN, g.XN, g.ZN).copy()
    v[:, :, 0] = 0.0
    down = np.zeros_like(v)
    down[:, :, :-1] = v[:, :, 1:]
    g3 = small_helix()
    g3.zpos = (g.zpos.astype(np.float64) + g.dz).astype(np.float32)
    assert rel_l2(O.proj(g3, v.reshape(-1)), O.proj(g, down.reshape(-1))) < 1e-13


@pytest.mark.parametrize("kind", ["cone", "fan1", "fan3"])
def test_adjoint_identity(kind):
    if kind == "cone":
        g = small_cone(PN=16, ang0=0.05)
    else:
        g = G.fan_multislice(XN=20, YN=17, SLN=1 if kind == "fan1" else 3, dx=1.5, PN=30, DNU=41,
                             col_size=1.2, sid=50.0, sdd=90.0, ang0=0.2)
    rng = np.random.default_rng(3)
    x = rng.random(g.nvox())
    y = rng.random(g.nprj())
    lhs = O.proj(g, x) @ y
    rhs = x @ O.back(g, y)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def _fan_pts(g):
    xd = g.xds.astype(np.float64)
    yd = g.yds.astype(np.float64)
    return O.boundaries(xd), O.boundaries(yd), xd, yd


def test_branchless_equals_conventional_2d():
    g = G.fan_multislice(XN=32, YN=32, SLN=1, dx=1.0, PN=24, DNU=50, col_size=1.0, sid=60.0,
                         sdd=100.0, ang0=0.05)
    rng = np.random.default_rng(4)
    img = rng.random(g.nvox())
    bx, by, cx, cy = _fan_pts(g)
    angs = g.angs.astype(np.float64)
    pc = O.ddm_pts(False, 60.0, 1.0, 1.0, 32, 32, angs, bx, by, cx, cy, img)
    assert rel_l2(O.proj(g, img), pc) < 1e-13
    y = rng.random(g.nprj())
    bc = O.ddm_pts(True, 60.0, 1.0, 1.0, 32, 32, angs, bx, by, cx, cy, y)
    assert rel_l2(O.back(g, y), bc) < 1e-13
    # float build of the conventional path stays within FP32 rounding
    pf = O.ddm_pts(False, 60.0, 1.0, 1.0, 32, 32, angs, bx, by, cx, cy, img, dtype=np.float32)
    assert rel_l2(pf, pc) < 2e-6


def test_branchless_equals_conventional_cfg1_size():
    """The 2-D case at its full size (256x256 image of 10 seeded random
    ellipses, 360 views over 360 degrees, 888 arc cells of 1.0239 mm, SID 541,
    SDD 949): the branchless FP64 oracle equals the conventional overlap-merge
    DD (the reference's DDM_*_proj / DDM_*_bproj restated) for both operators."""
    g = G.fan_multislice(XN=256, YN=256, SLN=1, dx=1.0, PN=360, DNU=888, col_size=1.0239, sid=541.0,
                         sdd=949.0)
    img = PH.voxelize(PH.random_ellipses(1, 10, 256, 256, g.dx), 256, 256, 1, g.dx, 1.0).reshape(-1)
    img = img.astype(np.float64)
    assert img.max() > 0
    bx, by, cx, cy = _fan_pts(g)
    angs = g.angs.astype(np.float64)
    pc = O.ddm_pts(False, 541.0, 1.0, 1.0, 256, 256, angs, bx, by, cx, cy, img)
    assert rel_l2(O.proj(g, img), pc) < 1e-13
    y = np.random.default_rng(3).random(g.nprj())
    bc = O.ddm_pts(True, 541.0, 1.0, 1.0, 256, 256, angs, bx, by, cx, cy, y)
    assert rel_l2(O.back(g, y), bc) < 1e-13


def test_ddm_ea_api_matches_branchless_arc():
    """DDM_EA_proj's exact-arc boundaries equal DD3Boundaries of arc centres in
    angle (the arc is centred on the source), so only the two edge cells differ."""
    S2O, S2D, DN, pitch = 541.0, 949.0, 888, 1.0239
    g = G.fan_multislice(XN=64, YN=64, SLN=1, dx=4.0, PN=36, DNU=DN, col_size=pitch, sid=S2O,
                         sdd=S2D, ang0=0.01)
    rng = np.random.default_rng(5)
    img = rng.random(g.nvox())
    pe = O.ddm("ea_proj", S2O, S2D - S2O, 256.0, 256.0, pitch / S2D, (DN - 1) / 2.0, 64, 64, DN,
               g.angs.astype(np.float64), img).reshape(g.PN, DN)
    pb = O.proj(g, img).reshape(g.PN, DN)
    assert rel_l2(pb[:, 1:-1], pe[:, 1:-1]) < 1e-6
    ed = O.ddm("ed_proj", S2O, S2D - S2O, 256.0, 256.0, 0.6, (DN - 1) / 2.0, 64, 64, DN,
               g.angs.astype(np.float64), img)
    y = rng.random(ed.size)
    lhs = ed @ y
    rhs = img @ O.ddm("ed_bproj", S2O, S2D - S2O, 256.0, 256.0, 0.6, (DN - 1) / 2.0, 64, 64, DN,
                      g.angs.astype(np.float64), y)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_uniform_box_equals_chord_length():
    """Footprints inside a uniform box: each slab contributes mu*L exactly, so
    the DD value is mu times the exact segment length between the outer slab
    planes along the cell's centre ray."""
    mu = 0.7
    g = G.circular_cone(XN=20, YN=20, ZN=10, dx=1.0, dz=1.0, PN=10, DNU=4, DNV=2, col_size=0.5,
                        row_size=0.5, sid=100.0, sdd=150.0, ang0=0.3)
    vol = np.full(g.nvox(), mu)
    p = O.proj(g, vol).reshape(g.PN, g.DNU, g.DNV)
    orient = O.orientation(g)
    for k in range(g.PN):
        th = float(g.angs[k])
        c, s = np.cos(th), np.sin(th)
        S = np.array([-g.y0 * s, g.y0 * c, g.z0])
        for u in range(g.DNU):
            for v in range(g.DNV):
                C = np.array([g.xds[u] * c - g.yds[u] * s, g.xds[u] * s + g.yds[u] * c, g.zds[v]],
                             dtype=np.float64)
                d = C - S
                ax = 1 if orient[k] else 0
                length = 20.0 * np.linalg.norm(d) / abs(d[ax])
                assert abs(p[k, u, v] - mu * length) < 1e-12 * length


def test_rotation_by_90_degrees_is_a_view_shift():
    """Rotating the volume by +90 deg about z equals shifting the views by PN/4
    (no view on a 45-degree tie), which pins the x-slab / y-slab handling."""
    PN = 12
    g = G.circular_cone(XN=8, YN=8, ZN=3, dx=1.0, dz=1.0, PN=PN, DNU=15, DNV=3, col_size=1.0,
                        row_size=1.5, sid=30.0, sdd=50.0, ang0=0.0)
    orient = O.orientation(g)
    assert orient.sum() not in (0, PN)
    rng = np.random.default_rng(6)
    vol = rng.random(g.vol_shape)
    # rotate by +90 deg: new(x, y) = old(y', x') with (x,y) = R90 (x', y')
    rot = np.transpose(vol, (1, 0, 2))[:, ::-1, :]
    p0 = O.proj(g, vol).reshape(PN, g.DNU, g.DNV)
    p1 = O.proj(g, np.ascontiguousarray(rot)).reshape(PN, g.DNU, g.DNV)
    # float32 view angles (the reference ABI type) make theta + pi/2 inexact
    assert rel_l2(np.roll(p0, PN // 4, axis=0), p1) < 1e-6


def test_mirror_is_reversed_detector():
    """Mirroring x -> -x equals negating the view angles and reversing u."""
    PN = 10
    g = G.circular_cone(XN=9, YN=7, ZN=3, dx=1.0, dz=1.0, PN=PN, DNU=13, DNV=3, col_size=1.0,
                        row_size=1.5, sid=30.0, sdd=50.0, ang0=0.2)
    gm = G.circular_cone(XN=9, YN=7, ZN=3, dx=1.0, dz=1.0, PN=PN, DNU=13, DNV=3, col_size=1.0,
                         row_size=1.5, sid=30.0, sdd=50.0, ang0=0.2)
    gm.angs = (2 * np.pi - g.angs.astype(np.float64)).astype(np.float32)
    rng = np.random.default_rng(7)
    vol = rng.random(g.vol_shape)
    p0 = O.proj(g, vol).reshape(PN, g.DNU, g.DNV)
    p1 = O.proj(gm, np.ascontiguousarray(vol[:, ::-1, :])).reshape(PN, g.DNU, g.DNV)
    assert rel_l2(p0[:, ::-1, :], p1) < 1e-6   # float32 angles: 2*pi - a rounds


def test_ellipsoid_line_integrals():
    """DD of a voxelised smooth ellipsoid approaches the exact line integrals."""
    g = G.circular_cone(XN=96, YN=96, ZN=24, dx=1.0, dz=1.0, PN=6, DNU=80, DNV=16, col_size=2.0,
                        row_size=2.0, sid=200.0, sdd=350.0, ang0=0.1)
    ells = np.array([[3.0, -2.0, 0.5, 30.0, 22.0, 8.0, 0.4, 0.02]])
    vol = PH.voxelize(ells, g.XN, g.YN, g.ZN, g.dx, g.dz)
    p = O.proj(g, vol).reshape(g.PN, g.DNU, g.DNV)
    err, ref = [], []
    for k in range(g.PN):
        th = float(g.angs[k])
        c, s = np.cos(th), np.sin(th)
        S = np.array([-g.y0 * s, g.y0 * c, 0.0])
        U, V = np.meshgrid(np.arange(g.DNU), np.arange(g.DNV), indexing="ij")
        C = np.stack([g.xds[U] * c - g.yds[U] * s, g.xds[U] * s + g.yds[U] * c, g.zds[V]], -1)
        li = PH.line_integral(ells, np.broadcast_to(S, C.shape), C.astype(np.float64))
        err.append(p[k] - li)
        ref.append(li)
    assert rel_l2(np.array(err) + np.array(ref), np.array(ref)) < 0.05


def test_boundaries_rule():
    # src/utilities.cuh:413-433
    assert np.allclose(O.boundaries(np.array([0.0, 1.0, 3.0])), [-0.5, 0.5, 2.0, 4.0])
    assert np.allclose(O.boundaries(np.array([2.0])), [1.5, 2.5])


def test_block_decomposition_sums_to_full_operator():
    """The identity blocked plans rest on: the projection of a volume of more
    than 64 slices onto a detector of more than 64 rows is the sum, over
    (slice block, row block) pairs, of the projection of the slice block alone
    (a volume of its own, centred on the block) onto the row block alone.  The
    row centres are exact binary fractions, so the blocks' boundary rule gives
    the full detector's boundaries.  Also the backprojection, block by block."""
    g = G.circular_cone(XN=12, YN=10, ZN=150, dx=1.0, dz=0.5, PN=10, DNU=20, DNV=140, col_size=1.5,
                        row_size=0.5, sid=60.0, sdd=100.0, ang0=0.3)
    assert np.all(np.diff(g.zds.astype(np.float64)) == 0.5)
    rng = np.random.default_rng(30)
    vol = rng.random(g.vol_shape)
    y = rng.random(g.nprj())
    full_p = O.proj(g, vol).reshape(g.PN, g.DNU, g.DNV)
    full_b = O.back(g, y).reshape(g.vol_shape)
    yb = y.reshape(g.PN, g.DNU, g.DNV)
    sum_p = np.zeros_like(full_p)
    sum_b = np.zeros_like(full_b)
    zmin = g.imgZCenter - g.ZN * g.dz / 2
    for z0 in range(0, g.ZN, 64):
        nz = min(64, g.ZN - z0)
        for v0 in range(0, g.DNV, 64):
            nv = min(64, g.DNV - v0)
            sub = G.ConeGeometry(**{k: getattr(g, k) for k in g.__dataclass_fields__})
            sub.ZN, sub.DNV = nz, nv
            sub.imgZCenter = zmin + (z0 + 0.5 * nz) * g.dz
            sub.zds = g.zds[v0:v0 + nv].copy()
            part = O.proj(sub, np.ascontiguousarray(vol[:, :, z0:z0 + nz]))
            sum_p[:, :, v0:v0 + nv] += part.reshape(g.PN, g.DNU, nv)
            back = O.back(sub, np.ascontiguousarray(yb[:, :, v0:v0 + nv]))
            sum_b[:, :, z0:z0 + nz] += back.reshape(g.YN, g.XN, nz)
    assert rel_l2(sum_p, full_p) < 1e-12
    assert rel_l2(sum_b, full_b) < 1e-12


def test_ddm_ed_proj_matches_branchless_flat():
    """DDM_ED_proj (flat detector at y = -O2D, cell pitch detSize; src/DDM_proj.h:
    42-55) equals the branchless oracle on the same flat-detector fan geometry.
    The cell centres are exact binary fractions, so DD3Boundaries of the centres
    gives the flat detector's own boundaries and every cell agrees (a detector
    placed at +O2D, or a shifted centre index, fails)."""
    S2O, S2D, DN, pitch = 541.0, 949.0, 600, 0.5
    g = G.fan_multislice(XN=64, YN=64, SLN=1, dx=4.0, PN=36, DNU=DN, col_size=pitch, sid=S2O, sdd=S2D,
                         arc=False, ang0=0.01)
    assert np.all(np.diff(g.xds.astype(np.float64)) == pitch)
    rng = np.random.default_rng(8)
    img = rng.random(g.nvox())
    for dt in (np.float64, np.float32):
        ed = O.ddm("ed_proj", S2O, S2D - S2O, 256.0, 256.0, pitch, (DN - 1) / 2.0, 64, 64, DN,
                   g.angs.astype(np.float64), img, dtype=dt)
        tol = 1e-12 if dt == np.float64 else 2e-6
        assert rel_l2(ed, O.proj(g, img)) < tol
