"""Device-resident reconstruction (OS-SART, FISTA-TV, CG) over a projector plan.

Mirrors CTLAB's drivers (src/DD3_GPU_recon.h:21-103) and update functors
(src/SARTWeighting.h, src/FastIterativeShrinkageAlgorithm.h,
src/utilities.hpp:1062-1268) through libctlab_mi355x.so; every vector is a
float32 torch tensor in HBM and every kernel runs on the GPU.
"""
import ctypes as C

import numpy as np

from . import _lib
from .dd import Plan, _stream_handle

CG_FR, CG_PRP, CG_CW, CG_DI, CG_DY = range(5)      # CGMethod, src/DD3_GPU_recon.h:14-19


def _p(t):
    return C.c_void_p(t.data_ptr())


class Recon:
    def __init__(self, plan: Plan, osNum=1):
        self.plan = plan
        h = _lib.lib().ctlab_recon_create(plan.h, osNum)
        if not h:
            _lib.check()
            raise _lib.CtlabError(-1, "recon creation failed")
        self.h = C.c_void_p(h)
        self.osNum = osNum

    def set_tiles(self, nu, nm):
        _lib.check(_lib.lib().ctlab_recon_set_tiles(self.h, nu, nm))

    def set_data(self, b, stream=None):
        _lib.check(_lib.lib().ctlab_recon_set_data(self.h, _p(b), _stream_handle(stream)))

    def os_sart(self, x, iters, lam=1.0, nonneg=False, use_graph=True, stream=None):
        _lib.check(_lib.lib().ctlab_os_sart(self.h, _p(x), iters, lam, int(nonneg), int(use_graph),
                                            _stream_handle(stream)))
        return x

    def fista_tv(self, x, iters, tv_lambda, tv_iters=10, L=0.0, nonneg=False, stream=None):
        obj = np.zeros(max(iters, 1), dtype=np.float64)
        _lib.check(_lib.lib().ctlab_fista_tv(self.h, _p(x), iters, tv_lambda, tv_iters, L, int(nonneg),
                                             obj.ctypes.data_as(C.POINTER(C.c_double)),
                                             _stream_handle(stream)))
        return obj[:iters]

    def cg(self, x, iters, cgm=CG_FR, stream=None):
        res = np.zeros(max(iters, 1), dtype=np.float64)
        _lib.check(_lib.lib().ctlab_cg(self.h, _p(x), iters, cgm, res.ctypes.data_as(C.POINTER(C.c_double)),
                                       _stream_handle(stream)))
        return res[:iters]

    def cg_to_fidelity(self, x, f_target, max_iters, stream=None):
        """Fletcher-Reeves CG from x until ||Ax - b||^2 <= f_target or max_iters
        steps; returns (steps taken, ||Ax - b||^2 after each step)."""
        fid = np.zeros(max(max_iters, 1), dtype=np.float64)
        done = C.c_int(0)
        _lib.check(_lib.lib().ctlab_cg_to_fidelity(self.h, _p(x), max_iters, float(f_target),
                                                    fid.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done),
                                                    _stream_handle(stream)))
        return done.value, fid[:done.value]

    def data_fidelity(self, x, stream=None):
        v = _lib.lib().ctlab_recon_data_fidelity(self.h, _p(x), _stream_handle(stream))
        _lib.check()
        return v

    def gradient(self, y, g, stream=None):
        v = _lib.lib().ctlab_recon_gradient(self.h, _p(y), _p(g), _stream_handle(stream))
        _lib.check()
        return v

    def lipschitz(self, iters=20, stream=None):
        v = _lib.lib().ctlab_recon_lipschitz(self.h, iters, _stream_handle(stream))
        _lib.check()
        return v

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().ctlab_recon_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many OS-SART reconstructions advanced together (ctlab_batch_*): every
    phase of a subset step is one persistent launch over all problems and one
    iteration is a HIP graph.  Results equal the per-problem Recon.os_sart."""

    def __init__(self, recons):
        self.recons = list(recons)
        arr = (C.c_void_p * len(self.recons))(*[r.h.value for r in self.recons])
        h = _lib.lib().ctlab_batch_create(arr, len(self.recons))
        if not h:
            _lib.check()
            raise _lib.CtlabError(-1, "batch creation failed")
        self.h = C.c_void_p(h)

    def _vols(self, xs):
        if len(xs) != len(self.recons):
            raise ValueError(f"need {len(self.recons)} volumes, got {len(xs)}")
        return (C.c_void_p * len(xs))(*[x.data_ptr() for x in xs])

    def set_tiles(self, nu, nm):
        _lib.check(_lib.lib().ctlab_batch_set_tiles(self.h, nu, nm))

    def tiles(self):
        a, b = C.c_int(), C.c_int()
        _lib.check(_lib.lib().ctlab_batch_get_tiles(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def os_sart(self, xs, iters, lam=1.0, nonneg=False, use_graph=True, stream=None):
        _lib.check(_lib.lib().ctlab_batch_os_sart(self.h, self._vols(xs), iters, lam, int(nonneg),
                                                  int(use_graph), _stream_handle(stream)))
        return xs

    def autotune(self, xs, lam=1.0):
        """Times one iteration per tile size on xs (updated); keeps the fastest.
        Returns {'proj': {8: ms, 12: ms, 16: ms}, 'back': {...}}."""
        ms = np.zeros(6, dtype=np.float32)
        _lib.check(_lib.lib().ctlab_batch_autotune(self.h, self._vols(xs), lam,
                                                   ms.ctypes.data_as(C.POINTER(C.c_float))))
        return {"proj": dict(zip((8, 12, 16), ms[:3].tolist())), "back": dict(zip((8, 12, 16), ms[3:].tolist()))}

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().ctlab_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- elementwise kernels (reference functors) ------------------------------
def sart_prj_weight(prj, real, rowsum, stream=None):
    _lib.check(_lib.lib().ctlab_sart_prj_weight(_p(prj), _p(real), _p(rowsum), prj.numel(), _stream_handle(stream)))
    return prj


def sart_bak_weight(vol, recon, colsum, stream=None):
    _lib.check(_lib.lib().ctlab_sart_bak_weight(_p(vol), _p(recon), _p(colsum), vol.numel(), _stream_handle(stream)))
    return recon


def ossart_update(img, cor, weg, msk, lam, stream=None):
    _lib.check(_lib.lib().ctlab_ossart_update(_p(img), _p(cor), _p(weg), _p(msk), lam, img.numel(),
                                              _stream_handle(stream)))
    return img


def fista_momentum(cur, las, t1, t2, minV=-2000.0, maxV=6000.0, stream=None):
    _lib.check(_lib.lib().ctlab_fista_momentum(_p(cur), _p(las), t1, t2, minV, maxV, cur.numel(),
                                               _stream_handle(stream)))
    return cur


def axpby(a, x, b, y, stream=None):
    """y = a*x + b*y"""
    _lib.check(_lib.lib().ctlab_axpby(a, _p(x), b, _p(y), y.numel(), _stream_handle(stream)))
    return y


def soft_threshold(x, g, stream=None):
    _lib.check(_lib.lib().ctlab_soft_threshold(_p(x), g, x.numel(), _stream_handle(stream)))
    return x


def cg_beta(gk, gk1, d, cases, stream=None):
    out = np.zeros(2, dtype=np.float64)
    _lib.check(_lib.lib().ctlab_cg_beta(_p(gk), _p(gk1), _p(d), cases, gk.numel(),
                                        out.ctypes.data_as(C.POINTER(C.c_double)), _stream_handle(stream)))
    return out[0] / out[1]


def tv_prox_into(b, x, lam, iters, shape, work, stream=None):
    """x = TV prox of b with a caller-owned work buffer of 9*numel floats."""
    YN, XN, ZN = shape
    _lib.check(_lib.lib().ctlab_tv_prox(_p(b), _p(x), lam, iters, XN, YN, ZN, _p(work), _stream_handle(stream)))
    return x


def tv_prox(b, lam, iters, shape, stream=None):
    import torch
    YN, XN, ZN = shape
    x = torch.empty_like(b)
    work = torch.empty(9 * b.numel(), dtype=torch.float32, device=b.device)
    _lib.check(_lib.lib().ctlab_tv_prox(_p(b), _p(x), lam, iters, XN, YN, ZN, _p(work), _stream_handle(stream)))
    return x


# ---- host-level drivers with the reference's argument lists ---------------

def evalProjBack(plan: Plan, forwardProjection, forwardMethod=0, backwardMethod=0, stream=None):
    """A*1 (forwardProjection) or A^T*1 over the plan's geometry, as a new
    tensor (src/DD3_GPU_recon.h:21-28)."""
    import torch
    g = plan.g
    shape = g.prj_shape if forwardProjection else g.vol_shape
    out = torch.empty(shape, dtype=torch.float32, device=f"cuda:{plan.device}")
    rc = _lib.lib().ctlab_evalProjBack(plan.h, 1 if forwardProjection else 0, int(forwardMethod),
                                       int(backwardMethod), out.data_ptr(), _stream_handle(stream))
    _lib.check(rc)
    return out


def driver_detector(sid, sdd, DNU, DNV, col_size, row_size, col_offset=0.0, row_offset=0.0):
    """(xds, yds, zds) that OS_SART / CG build from their scan arguments (the
    library's own construction, from float32 arguments)."""
    xds, yds, zds = (np.zeros(DNU, np.float32), np.zeros(DNU, np.float32), np.zeros(DNV, np.float32))
    rc = _lib.lib().ctlab_arc_detector(float(sid), float(sdd), int(DNU), int(DNV), float(col_size), float(row_size),
                                       float(col_offset), float(row_offset), _lib.fptr(xds), _lib.fptr(yds),
                                       _lib.fptr(zds))
    _lib.check(rc)
    return xds, yds, zds


def _host_driver(name, extra, initImg, hprj, hangs, hzPos, mask, VolumeName, osNum, iterNum, outputMedRes, sid,
                 sdd, DNU, DNV, PN, imgXCenter, imgYCenter, imgZCenter, XN, YN, ZN, dx, dz, col_size, row_size,
                 col_offset, row_offset, forwardMethod, backwardMethod):
    f32 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    hprj, hangs, hzPos, initImg = f32(hprj), f32(hangs), f32(hzPos), f32(initImg)
    if hprj.size != PN * DNU * DNV or hangs.size != PN:
        raise ValueError("projection or view-angle array does not match PN, DNU, DNV")
    if initImg is not None and initImg.size != XN * YN * ZN:
        raise ValueError("initial volume does not match XN, YN, ZN")
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
    out = np.zeros(XN * YN * ZN, dtype=np.float32)
    fp = lambda a: None if a is None else _lib.fptr(a)
    name_b = VolumeName.encode() if VolumeName else None
    rc = getattr(_lib.lib(), name)(
        _lib.fptr(out), fp(initImg), _lib.fptr(hprj), _lib.fptr(hangs), fp(hzPos), _lib.bptr(m), name_b, int(osNum),
        int(iterNum), 1 if outputMedRes else 0, float(sid), float(sdd), int(DNU), int(DNV), int(PN),
        float(imgXCenter), float(imgYCenter), float(imgZCenter), int(XN), int(YN), int(ZN), float(dx), float(dz),
        float(col_size), float(row_size), float(col_offset), float(row_offset), int(forwardMethod),
        int(backwardMethod), *extra)
    _lib.check(rc)
    return out.reshape(YN, XN, ZN)


def OS_SART(initImg, hprj, hangs, hzPos, mask, VolumeName, osNum, iterNum, outputMedRes, sid, sdd, DNU, DNV, PN,
            imgXCenter, imgYCenter, imgZCenter, XN, YN, ZN, dx, dz, col_size, row_size, col_offset, row_offset,
            forwardMethod=0, backwardMethod=0):
    """OS-SART with the argument list of src/DD3_GPU_recon.h:60-65 (host
    arrays; returns reconImg as (YN, XN, ZN)).  Arc detector built from sid,
    sdd, cell sizes and offsets; the result also goes to VolumeName (raw
    float, (Z,X,Y)) and, with outputMedRes, iteration k to VolumeName_<k>."""
    return _host_driver("ctlab_OS_SART", (), initImg, hprj, hangs, hzPos, mask, VolumeName, osNum, iterNum,
                        outputMedRes, sid, sdd, DNU, DNV, PN, imgXCenter, imgYCenter, imgZCenter, XN, YN, ZN, dx, dz,
                        col_size, row_size, col_offset, row_offset, forwardMethod, backwardMethod)


def CG(initImg, hprj, hangs, hzPos, mask, VolumeName, osNum, iterNum, outputMedRes, sid, sdd, DNU, DNV, PN,
       imgXCenter, imgYCenter, imgZCenter, XN, YN, ZN, dx, dz, col_size, row_size, col_offset, row_offset,
       forwardMethod=0, backwardMethod=0, cgm=CG_FR):
    """CG with the argument list of src/DD3_GPU_recon.h:98-103 (osNum is
    accepted and unused: CG runs over all views)."""
    return _host_driver("ctlab_CG", (int(cgm),), initImg, hprj, hangs, hzPos, mask, VolumeName, osNum, iterNum,
                        outputMedRes, sid, sdd, DNU, DNV, PN, imgXCenter, imgYCenter, imgZCenter, XN, YN, ZN, dx, dz,
                        col_size, row_size, col_offset, row_offset, forwardMethod, backwardMethod)
