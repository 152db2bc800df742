"""ctypes binding of libctlab_mi355x.so (built in-tree by ctlab_amd.build).

There is no CPU fallback: if the library is missing, every entry point raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libctlab_mi355x.so")
_lib = None

_fp = C.POINTER(C.c_float)
_bp = C.POINTER(C.c_ubyte)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class CtlabError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ctlab error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m ctlab_amd.build` "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    f, i = C.c_float, C.c_int
    L.DD3Proj_gpu.argtypes = [f, f, f, i, i, _fp, _fp, _fp, f, f, f, _fp, _fp, i, i, i, i, _fp, _fp,
                              f, f, _bp, i, i]
    L.DD3Proj_gpu.restype = None
    L.DD3Back_gpu.argtypes = [f, f, f, i, i, _fp, _fp, _fp, f, f, f, _fp, _fp, i, i, i, i, _fp, _fp,
                              f, f, _bp, i, i, i]
    L.DD3Back_gpu.restype = None
    L.DD2_multiGPU.argtypes = [_fp, _fp, i, f, f, _fp, _fp, i, i, f, f, i, i, f, _fp, i, _bp, _ip, i]
    L.DD2_multiGPU.restype = None
    geo = [f, f, f, i, i, _fp, _fp, _fp, f, f, f, _fp, _fp, i, i, i, i, _fp, _fp, f, f, _bp, i, _ip]
    for name in ("DD3Proj_multiGPU", "DD3Back_multiGPU"):
        getattr(L, name).argtypes = geo + [i]
        getattr(L, name).restype = None
    for name in ("DD3ProjHelical_3GPU", "DD3ProjHelical_4GPU", "DD3BackHelical_3GPU", "DD3BackHelical_4GPU"):
        getattr(L, name).argtypes = geo
        getattr(L, name).restype = None
    L.ctlab_last_error.restype = i
    L.ctlab_error_string.argtypes = [i]
    L.ctlab_error_string.restype = C.c_char_p
    L.ctlab_last_error_message.restype = C.c_char_p
    L.ctlab_plan_create.argtypes = [f, f, f, i, i, _fp, _fp, _fp, f, f, f, _fp, _fp, i, i, i, i, f, f,
                                    _bp, i]
    L.ctlab_plan_create.restype = _vp
    L.ctlab_plan_create_fan.argtypes = [f, f, _fp, _fp, i, i, f, f, i, i, f, _fp, i, _bp, i]
    L.ctlab_plan_create_fan.restype = _vp
    L.ctlab_plan_destroy.argtypes = [_vp]
    L.ctlab_plan_destroy.restype = None
    L.ctlab_plan_set_sat_bits.argtypes = [_vp, i]
    L.ctlab_plan_set_sat_bits.restype = i
    L.ctlab_proj.argtypes = [_vp, _vp, _vp, _vp]
    L.ctlab_proj.restype = i
    L.ctlab_back.argtypes = [_vp, _vp, _vp, _vp]
    L.ctlab_back.restype = i
    L.ctlab_plan_set_tiles.argtypes = [_vp, i, i]
    L.ctlab_plan_set_tiles.restype = i
    L.ctlab_plan_get_tiles.argtypes = [_vp, _ip, _ip]
    L.ctlab_plan_get_tiles.restype = i
    L.ctlab_plan_autotune.argtypes = [_vp, _vp, _vp, C.POINTER(C.c_float)]
    L.ctlab_plan_autotune.restype = i
    L.ctlab_plan_views.argtypes = [_vp, _ip, _ip]
    L.ctlab_plan_views.restype = i
    L.ctlab_plan_workspace_bytes.argtypes = [_vp]
    L.ctlab_plan_workspace_bytes.restype = C.c_longlong
    L.ctlab_plan_cache_set_limit.argtypes = [i]
    L.ctlab_plan_cache_set_limit.restype = i
    L.ctlab_plan_cache_clear.argtypes = []
    L.ctlab_plan_cache_clear.restype = None
    L.ctlab_plan_cache_size.argtypes = []
    L.ctlab_plan_cache_size.restype = i
    _declare_optional(L)


def _declare_optional(L):
    """Entry points of the other translation units (recon, conv, batch)."""
    from . import _lib_ext
    _lib_ext.declare(L)


def last_error():
    L = lib()
    return L.ctlab_last_error(), L.ctlab_last_error_message().decode()


def check(rc=None):
    code, msg = last_error()
    if rc not in (None, 0) and code == 0:
        code = rc
    if code != 0:
        raise CtlabError(code, msg)


def fptr(a):
    return a.ctypes.data_as(_fp)


def bptr(a):
    return None if a is None else a.ctypes.data_as(_bp)
