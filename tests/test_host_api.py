"""Host-level functions that need no GPU: raw file input / output
(src/readwrite.hpp) and the helical view / slice range helpers
(src/utilities.cuh:478-532), checked through the C ABI.

The range helpers are pinned by what they are for: every slice that the
oracle's dense system matrix shows a view range touching lies inside the slice
range returned for it, and every view touching a slice range lies inside the
view range returned for it."""
import math

import numpy as np
import pytest

import oracle as O
from ctlab_amd import _lib, dd, rawio
from ctlab_amd import geometry as G


def test_raw_round_trip_float_and_double(tmp_path):
    rng = np.random.default_rng(5)
    for dt in (np.float32, np.float64):
        a = rng.standard_normal((3, 4, 5)).astype(dt)
        p = tmp_path / f"a_{np.dtype(dt).name}.raw"
        rawio.write_raw(p, a)
        assert p.stat().st_size == a.nbytes
        assert np.array_equal(np.fromfile(p, dtype=dt).reshape(a.shape), a)   # plain native-endian array
        assert np.array_equal(rawio.read_raw(p, dtype=dt, shape=a.shape), a)
        assert np.array_equal(rawio.read_raw(p, n=7, dtype=dt), a.reshape(-1)[:7])


def test_raw_errors_leave_buffer(tmp_path):
    p = tmp_path / "short.raw"
    np.arange(5, dtype=np.float32).tofile(p)
    L = _lib.lib()
    buf = np.full(8, -1.0, np.float32)
    rc = L.ctlab_read_raw(str(p).encode(), buf.ctypes.data, 8, 4)
    assert rc == 6 and L.ctlab_last_error() == 6          # CTLAB_ERR_IO
    assert np.all(buf == -1.0)
    assert L.ctlab_read_raw(str(tmp_path / "missing.raw").encode(), buf.ctypes.data, 1, 4) == 6
    assert L.ctlab_read_raw(str(p).encode(), buf.ctypes.data, 1, 2) == 1     # element size
    assert L.ctlab_write_raw(str(tmp_path / "nodir" / "x.raw").encode(), buf.ctypes.data, 1, 4) == 6
    with pytest.raises(_lib.CtlabError):
        rawio.read_raw(p, n=6)


def test_rawio_cli(tmp_path, capsys):
    v = np.arange(2 * 3 * 4, dtype=np.float64).reshape(2, 3, 4)
    src = tmp_path / "v64.raw"
    v.tofile(src)
    dst = tmp_path / "v32.raw"
    assert rawio.main(["convert", str(src), str(dst), "--from", "f64", "--to", "f32"]) == 0
    assert np.array_equal(np.fromfile(dst, np.float32), v.reshape(-1).astype(np.float32))
    assert rawio.main(["info", str(dst), "--shape", "2,3,4"]) == 0
    assert "24 x f32" in capsys.readouterr().out
    pgm = tmp_path / "s.pgm"
    assert rawio.main(["slice", str(dst), str(pgm), "--shape", "2,3,4", "--z", "1"]) == 0
    data = pgm.read_bytes()
    assert data.startswith(b"P5\n3 2\n255\n") and len(data) == len(b"P5\n3 2\n255\n") + 6


def _helix():
    return G.helical_cone(XN=6, YN=5, ZN=12, dx=2.0, dz=2.5, PN=24, DNU=9, DNV=4, col_size=3.0, row_size=2.0,
                          pitch=12.0, turns=2.0, sid=40.0, sdd=70.0, ang0=0.2, row_offset=0.25)


def _touch(g):
    """touch[p, k]: view p has a non-zero system-matrix entry in slice k (dense oracle)."""
    A = O.dense(g).reshape(g.PN, g.DNU * g.DNV, g.YN * g.XN, g.ZN)
    return np.abs(A).sum(axis=(1, 2)) > 0


def _helper_args(g):
    # row centre index and pitch at the detector; centre slice index of the volume
    return dict(detCntIdxV=(g.DNV - 1) / 2.0 + 0.25, detStpZ=2.0, DNV=g.DNV,
                objCntIdxZ=(g.ZN - 1) / 2.0 - g.imgZCenter / g.dz, dz=g.dz, ZN=g.ZN)


def test_vol_z_range_contains_every_touched_slice():
    g = _helix()
    touch = _touch(g)
    a = _helper_args(g)
    informative = 0
    for ps in range(0, g.PN, 3):
        for pe in (ps + 1, ps + 4, g.PN):
            if pe > g.PN:
                continue
            s, e = dd.getVolZIdxPair(g.zpos, ps, pe, a["detCntIdxV"], a["detStpZ"], a["DNV"], a["objCntIdxZ"],
                                     a["dz"], a["ZN"])
            ks = np.nonzero(touch[ps:pe].any(axis=0))[0]
            assert 0 <= s <= e <= g.ZN
            assert ks.size == 0 or (s <= ks.min() and ks.max() < e), (ps, pe, s, e, ks)
            informative += (e - s) < g.ZN
    assert informative > 5          # ranges narrower than the volume for short view ranges


def test_prj_range_contains_every_touching_view():
    g = _helix()
    touch = _touch(g)
    a = _helper_args(g)
    informative = 0
    for ks in range(0, g.ZN, 2):
        for ke in (ks + 1, ks + 3):
            if ke > g.ZN:
                continue
            s, e = dd.getPrjIdxPair(g.zpos, ks, ke, a["objCntIdxZ"], a["dz"], a["ZN"], a["detCntIdxV"],
                                    a["detStpZ"], a["DNV"])
            ps = np.nonzero(touch[:, ks:ke].any(axis=1))[0]
            assert 0 <= s <= e <= g.PN
            assert ps.size == 0 or (s <= ps.min() and ps.max() < e), (ks, ke, s, e, ps)
            informative += (e - s) < g.PN
    assert informative > 3


def test_range_helper_argument_checks():
    L = _lib.lib()
    z = np.array([0.0, 1.0, 0.5], np.float32)          # not ascending
    with pytest.raises(_lib.CtlabError):
        dd.getPrjIdxPair(z, 0, 2, 1.0, 1.0, 4, 1.5, 1.0, 4)
    with pytest.raises(_lib.CtlabError):
        dd.getVolZIdxPair(np.zeros(4, np.float32), 3, 2, 1.5, 1.0, 4, 1.5, 1.0, 4)     # empty view range
    assert L.ctlab_last_error() == 1


def test_driver_detector_is_the_arc_detector():
    """The OS_SART / CG drivers' detector equals geometry.arc_detector built
    from the same float32 arguments, to float32 rounding of the trigonometry."""
    from ctlab_amd import recon as RC
    args = (150.0, 260.0, 72, 12, 1.3, 1.6, 0.25, -0.5)
    got = RC.driver_detector(*args)
    f = [float(np.float32(a)) if isinstance(a, float) else a for a in args]
    ref = G.arc_detector(*f)
    for a, b in zip(got, ref):
        assert np.allclose(a, b, rtol=0, atol=2e-5 * max(1.0, float(np.abs(b).max())))
    assert np.array_equal(got[2], ref[2])


def test_cxx_header_reference_names(tmp_path):
    """A C++ program written against the reference's names (readData_host,
    getVolZIdxPair, OS_SART, ...) compiles with include/ctlab_mi355x.hpp and
    runs against the library."""
    import os
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "ctlab_amd")
    exe = tmp_path / "hpp_api"
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "cxx", "hpp_api.cpp"), "-L", libdir, "-lctlab_mi355x",
                    f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), str(tmp_path)], check=True, capture_output=True, text=True).stdout.split()
    assert [float(v) for v in out[:5]] == [1.5, -2.25, 3.0, 0.1, 0.2]
    s, e, ps, pe, threw = (int(v) for v in out[5:])
    z = np.array([-6.0, -2.0, 2.0, 6.0], np.float32)
    assert (s, e) == dd.getVolZIdxPair(z, 1, 3, 1.5, 2.0, 4, 7.5, 1.0, 16)
    assert (ps, pe) == dd.getPrjIdxPair(z, 4, 8, 7.5, 1.0, 16, 1.5, 2.0, 4)
    assert threw == 1


def test_c_call_site_compiles_and_links(tmp_path):
    """The reference's C call site (DD3Proj_gpu with its 24 arguments) compiles
    against include/ctlab_mi355x.h with a C99 compiler, links against the
    library, and an unsupported prjMode is reported without touching the output."""
    import os
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "ctlab_amd")
    exe = tmp_path / "c_abi"
    subprocess.run([gcc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "cxx", "c_abi.c"), "-L", libdir, "-lctlab_mi355x",
                    f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split()[:2] == ["2", "1"]


def test_learning_rate_schedules():
    """The learned prox's Adam step sizes: the half-cosine decay starts at lr,
    passes the midpoint of lr and its 0.05 floor halfway and ends on the floor,
    never rising; the step schedule keeps lr and drops the last epoch to 0.3 lr."""
    from ctlab_amd import learned as LE
    lr, T = 1e-3, 600
    cos = [LE.learning_rate(lr, "cosine", s, T, 0, 1) for s in range(T + 1)]
    assert cos[0] == pytest.approx(lr) and cos[T] == pytest.approx(0.05 * lr)
    assert cos[T // 2] == pytest.approx(0.525 * lr)
    assert all(a >= b for a, b in zip(cos, cos[1:]))
    assert [LE.learning_rate(lr, "step", 0, T, e, 3) for e in range(3)] == [lr, lr, pytest.approx(0.3 * lr)]
    assert LE.learning_rate(lr, "step", 0, T, 0, 1) == lr
    with pytest.raises(ValueError):
        LE.learning_rate(lr, "linear", 0, T, 0, 1)


def _vol_pair_ref(z_lo, z_hi, cnt, stp, DNV, ocnt, dz, ZN, promote=True):
    """getVolZIdxPair (src/utilities.cuh:478-499) with T = float: the terms with
    0.5 are double expressions rounded to float once when stored."""
    f32, f64 = np.float32, np.float64
    lower = f32((f64(cnt) + 0.5) * f64(stp)) if promote else f32((cnt + f32(0.5)) * stp)
    upper = f32(f32(f32(DNV) * stp) - lower)
    sp, ep = f32(z_lo - lower), f32(z_hi + upper)
    s = math.floor(f32(f32(f32(sp / dz) + ocnt) - f32(1)))
    e = math.ceil(f32(f32(f32(ep / dz) + ocnt) + f32(1))) + 1
    return min(max(s, 0), ZN), min(max(e, 0), ZN)


def _prj_pair_ref(zpos, o0, o1, ocnt, dz, cnt, stp, DNV, promote=True):
    """getPrjIdxPair (src/utilities.cuh:505-532) with T = float."""
    f32, f64 = np.float32, np.float64
    d = f32(f32(o0) - ocnt)
    lowv = f32((f64(d) - 0.5) * f64(dz)) if promote else f32((d - f32(0.5)) * dz)
    highv = f32(lowv + f32(f32(o1 - o0) * dz))
    lower = f32((f64(cnt) + 0.5) * f64(stp)) if promote else f32((cnt + f32(0.5)) * stp)
    upper = f32(f32(f32(DNV) * stp) - lower)
    lo, hi = f32(lowv - upper), f32(highv + lower)
    s = int(np.searchsorted(zpos, lo, side="right")) - 1
    e = int(np.searchsorted(zpos, hi, side="right")) + 2
    PN = zpos.size
    return min(max(s, 0), PN), min(max(e, 0), PN)


def test_range_helpers_round_like_the_reference():
    """The range helpers evaluate (detCntIdxV + 0.5) * detStpZ and
    (ObjZIdx_Start - objCntIdxZ - 0.5) * dz in double and store them as float,
    as the reference's float instantiation does; the cases are built so that the
    all-float evaluation lands on the other side of a floor / ceil / sorted
    position for some of them."""
    f32 = np.float32
    rng = np.random.default_rng(40)
    n_vol = n_prj = differ_vol = differ_prj = 0
    while (n_vol < 300 or differ_vol < 3) and n_vol < 20000:
        cnt, stp = f32(rng.uniform(0, 128)), f32(rng.uniform(0.2, 2.0))
        lp = f32((np.float64(cnt) + 0.5) * np.float64(stp))
        dz = f32(rng.choice([0.25, 0.5, 1.0]))
        zp = f32(lp + f32(int(rng.integers(-6, 6))) * dz)
        ocnt = f32(rng.choice([10.0, 31.5, 63.5]))
        zpos = np.array([zp, zp + f32(3.0)], dtype=np.float32)
        got = dd.getVolZIdxPair(zpos, 0, 2, cnt, stp, 64, ocnt, dz, 4096)
        assert got == _vol_pair_ref(zpos[0], zpos[1], cnt, stp, 64, ocnt, dz, 4096)
        differ_vol += got != _vol_pair_ref(zpos[0], zpos[1], cnt, stp, 64, ocnt, dz, 4096, promote=False)
        n_vol += 1
    while (n_prj < 300 or differ_prj < 3) and n_prj < 20000:
        ocnt, dz = f32(rng.uniform(0, 100)), f32(rng.uniform(0.3, 2.0))
        o0 = int(rng.integers(0, 100))
        cnt, stp = f32(rng.uniform(0, 64)), f32(rng.uniform(0.2, 2.0))
        lowv = f32((np.float64(f32(f32(o0) - ocnt)) - 0.5) * np.float64(dz))
        upper = f32(f32(f32(64) * stp) - f32((np.float64(cnt) + 0.5) * np.float64(stp)))
        base = f32(lowv - upper)
        zpos = np.sort(np.concatenate([base + f32(0.5) * np.arange(-8, 8, dtype=np.float32),
                                       np.array([base], dtype=np.float32)])).astype(np.float32)
        got = dd.getPrjIdxPair(zpos, o0, o0 + 5, ocnt, dz, 200, cnt, stp, 64)
        assert got == _prj_pair_ref(zpos, o0, o0 + 5, ocnt, dz, cnt, stp, 64)
        differ_prj += got != _prj_pair_ref(zpos, o0, o0 + 5, ocnt, dz, cnt, stp, 64, promote=False)
        n_prj += 1
    assert differ_vol >= 3 and differ_prj >= 3, (differ_vol, differ_prj)
