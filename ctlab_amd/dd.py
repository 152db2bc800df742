"""Python mirror of CTLAB's projector / backprojector entry points.

* ``DD3Proj_gpu`` / ``DD3Back_gpu`` / ``DD2_multiGPU`` call the library's
  C ABI with the reference's argument list (src/DD3_GPU_Proj.h:41-44,
  src/DD3_GPU_Back.h:40-44, src/multiSlices_ker.cuh:35-38) on host arrays.
* ``Plan`` wraps the device-resident API: geometry tables live on the GPU and
  ``proj`` / ``back`` take and return torch tensors already in HBM, launched on
  torch's current stream.

Both raise ``CtlabError`` when the library records an error; nothing here
computes on the CPU.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import bptr, fptr
from .geometry import ConeGeometry


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def DD3Proj_gpu(g: ConeGeometry, vol, gpunum=0, prjMode=0):
    """Host-pointer projection with the DD3Proj_gpu argument list; returns prj (PN, DNU, DNV)."""
    L = _lib.lib()
    vol = _f32(vol).reshape(-1)
    if vol.size != g.nvox():
        raise ValueError("volume size does not match the geometry")
    prj = np.zeros(g.nprj(), dtype=np.float32)
    L.DD3Proj_gpu(g.x0, g.y0, g.z0, g.DNU, g.DNV, fptr(g.xds), fptr(g.yds), fptr(g.zds),
                  g.imgXCenter, g.imgYCenter, g.imgZCenter, fptr(g.angs), fptr(g.zpos), g.PN,
                  g.XN, g.YN, g.ZN, fptr(vol), fptr(prj), g.dx, g.dz, bptr(g.mask), gpunum, prjMode)
    _lib.check()
    return prj.reshape(g.prj_shape)


def DD3Back_gpu(g: ConeGeometry, prj, gpunum=0, squared=0, prjMode=0):
    """Host-pointer backprojection (exact adjoint); returns vol (YN, XN, ZN)."""
    L = _lib.lib()
    prj = _f32(prj).reshape(-1)
    if prj.size != g.nprj():
        raise ValueError("projection size does not match the geometry")
    vol = np.zeros(g.nvox(), dtype=np.float32)
    L.DD3Back_gpu(g.x0, g.y0, g.z0, g.DNU, g.DNV, fptr(g.xds), fptr(g.yds), fptr(g.zds),
                  g.imgXCenter, g.imgYCenter, g.imgZCenter, fptr(g.angs), fptr(g.zpos), g.PN,
                  g.XN, g.YN, g.ZN, fptr(vol), fptr(prj), g.dx, g.dz, bptr(g.mask), gpunum,
                  squared, prjMode)
    _lib.check()
    return vol.reshape(g.vol_shape)


def _split(start, n):
    s = np.asarray(start if start is not None else [0] * n, dtype=np.int32)
    return s, s.ctypes.data_as(C.POINTER(C.c_int))


def DD3Proj_multiGPU(g: ConeGeometry, vol, startPN, gpuNum, prjMode=0, helical=False):
    """DD3Proj_multiGPU (or DD3ProjHelical_3GPU / _4GPU when helical, gpuNum 3 or 4):
    views split at startPN over gpuNum GPUs in the reference; one GPU here."""
    L = _lib.lib()
    vol = _f32(vol).reshape(-1)
    if vol.size != g.nvox():
        raise ValueError("volume size does not match the geometry")
    prj = np.zeros(g.nprj(), dtype=np.float32)
    s, sp = _split(startPN, gpuNum)
    args = (g.x0, g.y0, g.z0, g.DNU, g.DNV, fptr(g.xds), fptr(g.yds), fptr(g.zds), g.imgXCenter, g.imgYCenter,
            g.imgZCenter, fptr(g.angs), fptr(g.zpos), g.PN, g.XN, g.YN, g.ZN, fptr(vol), fptr(prj), g.dx, g.dz,
            bptr(g.mask), prjMode, sp)
    if helical:
        getattr(L, f"DD3ProjHelical_{gpuNum}GPU")(*args)
    else:
        L.DD3Proj_multiGPU(*args, gpuNum)
    _lib.check()
    return prj.reshape(g.prj_shape)


def DD3Back_multiGPU(g: ConeGeometry, prj, startVOL, gpuNum, methodId=0, helical=False):
    """DD3Back_multiGPU (or DD3BackHelical_3GPU / _4GPU): slices split at startVOL."""
    L = _lib.lib()
    prj = _f32(prj).reshape(-1)
    if prj.size != g.nprj():
        raise ValueError("projection size does not match the geometry")
    vol = np.zeros(g.nvox(), dtype=np.float32)
    s, sp = _split(startVOL, gpuNum)
    args = (g.x0, g.y0, g.z0, g.DNU, g.DNV, fptr(g.xds), fptr(g.yds), fptr(g.zds), g.imgXCenter, g.imgYCenter,
            g.imgZCenter, fptr(g.angs), fptr(g.zpos), g.PN, g.XN, g.YN, g.ZN, fptr(vol), fptr(prj), g.dx, g.dz,
            bptr(g.mask), methodId, sp)
    if helical:
        getattr(L, f"DD3BackHelical_{gpuNum}GPU")(*args)
    else:
        L.DD3Back_multiGPU(*args, gpuNum)
    _lib.check()
    return vol.reshape(g.vol_shape)


def DD2_multiGPU(g: ConeGeometry, data, method, startIdx=None, gpuNum=1):
    """Multi-slice fan beam: method 0 projects ``data`` (image), 2 backprojects it."""
    if not g.parallel:
        raise ValueError("DD2_multiGPU needs a multi-slice (parallel) geometry")
    L = _lib.lib()
    if method == 0:
        vol = _f32(data).reshape(-1)
        prj = np.zeros(g.nprj(), dtype=np.float32)
    else:
        prj = _f32(data).reshape(-1)
        vol = np.zeros(g.nvox(), dtype=np.float32)
    sidx = np.zeros(max(gpuNum, 1), dtype=np.int32) if startIdx is None else np.asarray(startIdx, np.int32)
    L.DD2_multiGPU(fptr(vol), fptr(prj), method, g.x0, g.y0, fptr(g.xds), fptr(g.yds), g.DNU,
                   g.ZN, g.imgXCenter, g.imgYCenter, g.XN, g.YN, g.dx, fptr(g.angs), g.PN,
                   bptr(g.mask), sidx.ctypes.data_as(C.POINTER(C.c_int)), gpuNum)
    _lib.check()
    return prj.reshape(g.prj_shape) if method == 0 else vol.reshape(g.vol_shape)


def _stream_handle(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class Plan:
    """Device-resident projector / backprojector for one geometry."""

    def __init__(self, g: ConeGeometry, gpunum=0, sat_bits=64):
        L = _lib.lib()
        self.g = g
        if g.parallel:
            h = L.ctlab_plan_create_fan(g.x0, g.y0, fptr(g.xds), fptr(g.yds), g.DNU, g.ZN,
                                        g.imgXCenter, g.imgYCenter, g.XN, g.YN, g.dx,
                                        fptr(g.angs), g.PN, bptr(g.mask), gpunum)
        else:
            h = L.ctlab_plan_create(g.x0, g.y0, g.z0, g.DNU, g.DNV, fptr(g.xds), fptr(g.yds),
                                    fptr(g.zds), g.imgXCenter, g.imgYCenter, g.imgZCenter,
                                    fptr(g.angs), fptr(g.zpos), g.PN, g.XN, g.YN, g.ZN, g.dx, g.dz,
                                    bptr(g.mask), gpunum)
        if not h:
            _lib.check()
            raise _lib.CtlabError(-1, "plan creation failed")
        self.h = C.c_void_p(h)
        self.device = gpunum
        if sat_bits != 64:
            self.set_sat_bits(sat_bits)

    def set_sat_bits(self, bits):
        rc = _lib.lib().ctlab_plan_set_sat_bits(self.h, bits)
        _lib.check(rc)

    def set_tiles(self, nu, nm):
        _lib.check(_lib.lib().ctlab_plan_set_tiles(self.h, nu, nm))

    def tiles(self):
        a, b = C.c_int(), C.c_int()
        _lib.lib().ctlab_plan_get_tiles(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def autotune(self, vol, prj):
        """Times the tile variants on (vol, prj) -- both overwritten -- and keeps
        the fastest; returns {'proj': {8: ms, 12: ms, 16: ms}, 'back': {...}}."""
        ms = np.zeros(6, dtype=np.float32)
        self._check_tensor(vol, self.g.nvox(), "volume")
        self._check_tensor(prj, self.g.nprj(), "projection")
        rc = _lib.lib().ctlab_plan_autotune(self.h, C.c_void_p(vol.data_ptr()), C.c_void_p(prj.data_ptr()),
                                            ms.ctypes.data_as(C.POINTER(C.c_float)))
        _lib.check(rc)
        return {"proj": dict(zip((8, 12, 16), ms[:3].tolist())),
                "back": dict(zip((8, 12, 16), ms[3:].tolist()))}

    def views(self):
        a, b = C.c_int(), C.c_int()
        _lib.lib().ctlab_plan_views(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def _check_tensor(self, t, n, what):
        import torch
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.numel() == n):
            raise ValueError(f"{what}: need a contiguous float32 CUDA tensor with {n} elements")

    def proj(self, vol, out=None, stream=None):
        import torch
        g = self.g
        self._check_tensor(vol, g.nvox(), "volume")
        if out is None:
            out = torch.empty(g.prj_shape, dtype=torch.float32, device=vol.device)
        self._check_tensor(out, g.nprj(), "projection")
        rc = _lib.lib().ctlab_proj(self.h, C.c_void_p(vol.data_ptr()), C.c_void_p(out.data_ptr()),
                                   _stream_handle(stream))
        _lib.check(rc)
        return out

    def back(self, prj, out=None, stream=None):
        import torch
        g = self.g
        self._check_tensor(prj, g.nprj(), "projection")
        if out is None:
            out = torch.empty(g.vol_shape, dtype=torch.float32, device=prj.device)
        self._check_tensor(out, g.nvox(), "volume")
        rc = _lib.lib().ctlab_back(self.h, C.c_void_p(prj.data_ptr()), C.c_void_p(out.data_ptr()),
                                   _stream_handle(stream))
        _lib.check(rc)
        return out

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().ctlab_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def getVolZIdxPair(zPos, PrjIdx_Start, PrjIdx_End, detCntIdxV, detStpZ, DNV, objCntIdxZ, dz, ZN):
    """Slices [start, end) that views [PrjIdx_Start, PrjIdx_End) of a helical
    scan can reach (src/utilities.cuh:478-499; zPos ascending; detCntIdxV and
    detStpZ: row centre index and row pitch at the detector; objCntIdxZ: the
    volume's centre slice index)."""
    z = _f32(zPos)
    s, e = C.c_int(), C.c_int()
    rc = _lib.lib().ctlab_getVolZIdxPair(fptr(z), z.size, int(PrjIdx_Start), int(PrjIdx_End), float(detCntIdxV),
                                         float(detStpZ), int(DNV), float(objCntIdxZ), float(dz), int(ZN),
                                         C.byref(s), C.byref(e))
    _lib.check(rc)
    return s.value, e.value


def getPrjIdxPair(zPos, ObjZIdx_Start, ObjZIdx_End, objCntIdxZ, dz, ZN, detCntIdxV, detStpZ, DNV):
    """Views [start, end) whose detector can reach slices [ObjZIdx_Start,
    ObjZIdx_End) (src/utilities.cuh:505-532)."""
    z = _f32(zPos)
    s, e = C.c_int(), C.c_int()
    rc = _lib.lib().ctlab_getPrjIdxPair(fptr(z), z.size, int(ObjZIdx_Start), int(ObjZIdx_End), float(objCntIdxZ),
                                        float(dz), int(ZN), float(detCntIdxV), float(detStpZ), int(DNV),
                                        C.byref(s), C.byref(e))
    _lib.check(rc)
    return s.value, e.value
