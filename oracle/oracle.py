"""ctypes front end of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker / CPU baseline; the
product package ``ctlab_amd`` never imports it.

Conventions are those of the GPU library (volume ``(y*XN + x)*ZN + z``,
projection ``(p*DNU + u)*DNV + v``); see the header of ``dd_oracle.c``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


class Geom(C.Structure):
    _fields_ = [
        ("XN", C.c_int), ("YN", C.c_int), ("ZN", C.c_int),
        ("dx", C.c_double), ("dz", C.c_double),
        ("cx", C.c_double), ("cy", C.c_double), ("cz", C.c_double),
        ("DNU", C.c_int), ("DNV", C.c_int),
        ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
        ("xds", _dp), ("yds", _dp), ("zds", _dp),
        ("PN", C.c_int),
        ("angs", _dp), ("zpos", _dp),
        ("mask", C.POINTER(C.c_ubyte)),
        ("parallel", C.c_int),
    ]


def _declare(L):
    gp = C.POINTER(Geom)
    L.ddo_proj.argtypes = [gp, _dp, _dp]
    L.ddo_proj_range.argtypes = [gp, _dp, _dp, C.c_int, C.c_int]
    L.ddo_back.argtypes = [gp, _dp, _dp]
    L.ddo_dense.argtypes = [gp, _dp]
    L.ddo_orientation.argtypes = [gp, C.POINTER(C.c_int)]
    L.ddo_boundaries.argtypes = [C.c_int, _dp, _dp]
    L.ddo_set_threads.argtypes = [C.c_int]
    L.ddo_set_threads.restype = None
    L.ddo_max_threads.restype = C.c_int
    for t, P in (("d", _dp), ("f", _fp)):
        T = C.c_double if t == "d" else C.c_float
        for nm in ("ddm_proj_pts_", "ddm_back_pts_"):
            getattr(L, nm + t).argtypes = [P, P, T, T, T, C.c_int, C.c_int, C.c_int, C.c_int,
                                           P, P, P, P, P]
        for nm in ("ddm_ed_proj_", "ddm_ea_proj_", "ddm_ed_bproj_", "ddm_ea_bproj_"):
            getattr(L, nm + t).argtypes = [P, P, T, T, T, T, T, T, C.c_int, C.c_int, C.c_int,
                                           C.c_int, P]
    L.ro_sart_prj_weight.argtypes = [_dp, _dp, _dp, C.c_long]
    L.ro_sart_bak_weight.argtypes = [_dp, _dp, _dp, C.c_long]
    L.ro_ossart_update.argtypes = [_dp, _dp, _dp, _dp, C.c_double, C.c_long]
    L.ro_fista.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_long]
    L.ro_soft_threshold.argtypes = [_dp, C.c_double, C.c_long]
    L.ro_cg_beta.argtypes = [_dp, _dp, _dp, C.c_long, C.c_int]
    L.ro_cg_beta.restype = C.c_double
    L.ro_tv_prox.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int]
    L.ro_tv_value.argtypes = [_dp, C.c_int, C.c_int, C.c_int]
    L.ro_tv_value.restype = C.c_double
    if hasattr(L, "co_conv3d_fwd"):
        _declare_conv(L)


def _declare_conv(L):
    L.co_conv3d_fwd.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.co_conv3d_dgrad.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.co_conv3d_wgrad.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


class OracleGeom:
    """FP64 geometry holder mirroring the DD3Proj_gpu argument list."""

    def __init__(self, g):
        # g: ctlab_amd.geometry.ConeGeometry (or anything with the same attributes)
        self.keep = []
        s = Geom()
        s.XN, s.YN, s.ZN = int(g.XN), int(g.YN), int(g.ZN)
        s.dx, s.dz = float(g.dx), float(g.dz)
        s.cx, s.cy, s.cz = float(g.imgXCenter), float(g.imgYCenter), float(g.imgZCenter)
        s.DNU, s.DNV = int(g.DNU), int(g.DNV)
        s.x0, s.y0, s.z0 = float(g.x0), float(g.y0), float(g.z0)
        for nm in ("xds", "yds", "zds", "angs", "zpos"):
            arr, ptr = _d(getattr(g, nm))
            self.keep.append(arr)
            setattr(s, nm, ptr)
        s.PN = int(g.PN)
        if getattr(g, "mask", None) is not None:
            m = np.ascontiguousarray(g.mask, dtype=np.uint8)
            self.keep.append(m)
            s.mask = m.ctypes.data_as(C.POINTER(C.c_ubyte))
        s.parallel = int(getattr(g, "parallel", 0))
        self.s = s
        self.g = g

    def nvox(self):
        return self.s.XN * self.s.YN * self.s.ZN

    def nprj(self):
        return self.s.PN * self.s.DNU * self.s.DNV


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def proj(g, vol, views=None):
    og = OracleGeom(g)
    vol, vp = _d(vol.reshape(-1))
    out = np.zeros(og.nprj(), dtype=np.float64)
    if views is None:
        _chk(lib().ddo_proj(C.byref(og.s), vp, out.ctypes.data_as(_dp)), "proj")
    else:
        _chk(lib().ddo_proj_range(C.byref(og.s), vp, out.ctypes.data_as(_dp), views[0], views[1]), "proj")
    return out


def back(g, prj):
    og = OracleGeom(g)
    prj, pp = _d(prj.reshape(-1))
    out = np.zeros(og.nvox(), dtype=np.float64)
    _chk(lib().ddo_back(C.byref(og.s), pp, out.ctypes.data_as(_dp)), "back")
    return out


def dense(g):
    og = OracleGeom(g)
    A = np.zeros((og.nprj(), og.nvox()), dtype=np.float64)
    _chk(lib().ddo_dense(C.byref(og.s), A.ctypes.data_as(_dp)), "dense")
    return A


def orientation(g):
    og = OracleGeom(g)
    out = np.zeros(og.s.PN, dtype=np.int32)
    lib().ddo_orientation(C.byref(og.s), out.ctypes.data_as(C.POINTER(C.c_int)))
    return out


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops."""
    lib().ddo_set_threads(int(n))


def max_threads():
    return int(lib().ddo_max_threads())


def boundaries(c):
    c, cp = _d(c)
    b = np.zeros(c.size + 1, dtype=np.float64)
    lib().ddo_boundaries(c.size + 1, cp, b.ctypes.data_as(_dp))
    return b


def ddm(kind, S2O, O2D, objSizeX, objSizeY, pitch, detCntIdx, XN, YN, DN, angs, data, dtype=np.float64):
    """Conventional DD: kind in {'ed_proj','ea_proj','ed_bproj','ea_bproj'}."""
    cv = _d if dtype == np.float64 else _f
    fn = getattr(lib(), f"ddm_{kind}_{'d' if dtype == np.float64 else 'f'}")
    angs, ap = cv(angs)
    PN = angs.size
    if kind.endswith("bproj"):
        prj, pp = cv(data.reshape(-1))
        img, ip = cv(np.zeros(XN * YN))
    else:
        img, ip = cv(data.reshape(-1))
        prj, pp = cv(np.zeros(PN * DN))
    fn(pp, ip, S2O, O2D, objSizeX, objSizeY, pitch, detCntIdx, XN, YN, DN, PN, ap)
    return img if kind.endswith("bproj") else prj


def ddm_pts(back, S2O, dx, dy, XN, YN, angs, bx, by, cx, cy, data, dtype=np.float64):
    cv = _d if dtype == np.float64 else _f
    fn = getattr(lib(), f"ddm_{'back' if back else 'proj'}_pts_{'d' if dtype == np.float64 else 'f'}")
    angs, ap = cv(angs)
    bx, bxp = cv(bx); by, byp = cv(by); cx, cxp = cv(cx); cy, cyp = cv(cy)
    DN = cx.size
    PN = angs.size
    if back:
        prj, pp = cv(data.reshape(-1))
        img, ip = cv(np.zeros(XN * YN))
    else:
        img, ip = cv(data.reshape(-1))
        prj, pp = cv(np.zeros(PN * DN))
    fn(pp, ip, S2O, dx, dy, XN, YN, DN, PN, ap, bxp, byp, cxp, cyp)
    return img if back else prj


# ---- recon helpers ---------------------------------------------------------
def sart_prj_weight(prj, real, rowsum):
    prj, p = _d(np.array(prj, dtype=np.float64))
    real, r = _d(real); rowsum, w = _d(rowsum)
    lib().ro_sart_prj_weight(p, r, w, prj.size)
    return prj


def sart_bak_weight(vol, recon, colsum):
    vol, v = _d(vol); colsum, w = _d(colsum)
    recon, r = _d(np.array(recon, dtype=np.float64))
    lib().ro_sart_bak_weight(v, r, w, vol.size)
    return recon


def ossart_update(img, cor, weg, msk, lam):
    img, i = _d(np.array(img, dtype=np.float64))
    cor, c = _d(cor); weg, w = _d(weg); msk, m = _d(msk)
    lib().ro_ossart_update(i, c, w, m, lam, img.size)
    return img


def fista(cur, las, t1, t2, minV=-2000.0, maxV=6000.0):
    cur, c = _d(np.array(cur, dtype=np.float64))
    las, l = _d(las)
    lib().ro_fista(c, l, t1, t2, minV, maxV, cur.size)
    return cur


def soft_threshold(x, g):
    x, p = _d(np.array(x, dtype=np.float64))
    lib().ro_soft_threshold(p, g, x.size)
    return x


def cg_beta(gk, gk1, d, case):
    gk, a = _d(gk); gk1, b = _d(gk1); d, e = _d(d)
    return lib().ro_cg_beta(a, b, e, gk.size, case)


def tv_prox(b, XN, YN, ZN, lam, iters):
    b, bp = _d(b.reshape(-1))
    x = np.zeros_like(b)
    _chk(lib().ro_tv_prox(bp, x.ctypes.data_as(_dp), XN, YN, ZN, lam, iters), "tv_prox")
    return x


def tv_value(x, XN, YN, ZN):
    x, p = _d(x.reshape(-1))
    return lib().ro_tv_value(p, XN, YN, ZN)


# ---- 3x3x3 convolution (learned prox CNN checker) ---------------------------
def conv3d(x, w, b, shape):
    """x: (nvox, cin), w: (27, cout, cin), b: (cout,) or None; shape = (YN, XN, ZN)."""
    YN, XN, ZN = shape
    x, xp = _d(x); w, wp = _d(w)
    cout, cin = w.shape[1], w.shape[2]
    out = np.zeros((x.shape[0], cout))
    bp = None
    if b is not None:
        b, bp = _d(b)
    lib().co_conv3d_fwd(xp, wp, bp, out.ctypes.data_as(_dp), XN, YN, ZN, cin, cout, 0)
    return out


def conv3d_dgrad(gout, w, shape):
    YN, XN, ZN = shape
    gout, gp = _d(gout); w, wp = _d(w)
    cout, cin = w.shape[1], w.shape[2]
    gin = np.zeros((gout.shape[0], cin))
    lib().co_conv3d_dgrad(gp, wp, gin.ctypes.data_as(_dp), XN, YN, ZN, cin, cout)
    return gin


def conv3d_wgrad(x, gout, cin, cout, shape):
    YN, XN, ZN = shape
    x, xp = _d(x); gout, gp = _d(gout)
    gw = np.zeros((27, cout, cin))
    gb = np.zeros(cout)
    lib().co_conv3d_wgrad(xp, gp, gw.ctypes.data_as(_dp), gb.ctypes.data_as(_dp), XN, YN, ZN, cin, cout)
    return gw, gb
