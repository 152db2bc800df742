"""C-ABI checks that need no GPU: the library loads, exports every function
declared in include/*.h, and rejects bad arguments before touching the device
(the reference would exit(1) under DEBUG or ignore the error)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from ctlab_amd import _lib, dd
from ctlab_amd import geometry as G


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"//[^\n]*", "", txt)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt):
            nm = m.group(1)
            if nm in ("if", "sizeof", "defined"):
                continue
            pre = txt[:m.start()].rstrip()
            if pre.endswith("(") or pre.endswith(","):
                continue
            names.add(nm)
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared_functions()
    assert {"DD3Proj_gpu", "DD3Back_gpu", "DD2_multiGPU", "ctlab_last_error"} <= names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def _tiny():
    return G.circular_cone(XN=8, YN=8, ZN=4, dx=1.0, dz=1.0, PN=4, DNU=12, DNV=4, col_size=1.0,
                           row_size=1.0, sid=40.0, sdd=70.0)


def test_unsupported_mode_sets_error_and_leaves_output():
    L = _lib.lib()
    g = _tiny()
    vol = np.ones(g.nvox(), np.float32)
    prj = np.full(g.nprj(), 7.0, np.float32)
    L.DD3Proj_gpu(g.x0, g.y0, g.z0, g.DNU, g.DNV, _lib.fptr(g.xds), _lib.fptr(g.yds), _lib.fptr(g.zds),
                  0.0, 0.0, 0.0, _lib.fptr(g.angs), _lib.fptr(g.zpos), g.PN, g.XN, g.YN, g.ZN,
                  _lib.fptr(vol), _lib.fptr(prj), g.dx, g.dz, None, 0, 4)
    code, msg = _lib.last_error()
    assert code == 2 and "prjMode" in msg
    assert np.all(prj == 7.0)
    L.DD3Back_gpu(g.x0, g.y0, g.z0, g.DNU, g.DNV, _lib.fptr(g.xds), _lib.fptr(g.yds), _lib.fptr(g.zds),
                  0.0, 0.0, 0.0, _lib.fptr(g.angs), _lib.fptr(g.zpos), g.PN, g.XN, g.YN, g.ZN,
                  _lib.fptr(vol), _lib.fptr(prj), g.dx, g.dz, None, 0, 0, 3)
    assert _lib.last_error()[0] == 2
    assert np.all(vol == 1.0)
    with pytest.raises(_lib.CtlabError):
        dd.DD2_multiGPU(G.fan_multislice(8, 8, 1, 1.0, 4, 12, 1.0), np.ones(64), method=1)


def test_invalid_sizes_and_null_pointers():
    L = _lib.lib()
    g = _tiny()
    vol = np.ones(g.nvox(), np.float32)
    prj = np.zeros(g.nprj(), np.float32)
    L.DD3Proj_gpu(g.x0, g.y0, g.z0, g.DNU, g.DNV, _lib.fptr(g.xds), _lib.fptr(g.yds), _lib.fptr(g.zds),
                  0.0, 0.0, 0.0, _lib.fptr(g.angs), _lib.fptr(g.zpos), 0, g.XN, g.YN, g.ZN,
                  _lib.fptr(vol), _lib.fptr(prj), g.dx, g.dz, None, 0, 0)
    assert _lib.last_error()[0] == 1
    L.DD3Proj_gpu(g.x0, g.y0, g.z0, g.DNU, g.DNV, None, _lib.fptr(g.yds), _lib.fptr(g.zds),
                  0.0, 0.0, 0.0, _lib.fptr(g.angs), _lib.fptr(g.zpos), g.PN, g.XN, g.YN, g.ZN,
                  _lib.fptr(vol), _lib.fptr(prj), g.dx, g.dz, None, 0, 0)
    assert _lib.last_error()[0] == 1
    L.DD3Proj_gpu(g.x0, g.y0, g.z0, g.DNU, g.DNV, _lib.fptr(g.xds), _lib.fptr(g.yds), _lib.fptr(g.zds),
                  0.0, 0.0, 0.0, _lib.fptr(g.angs), _lib.fptr(g.zpos), g.PN, g.XN, g.YN, g.ZN,
                  None, _lib.fptr(prj), g.dx, g.dz, None, 0, 0)
    assert _lib.last_error()[0] == 1
    assert np.all(prj == 0)


def test_error_strings():
    L = _lib.lib()
    assert L.ctlab_error_string(0) == b"success"
    assert L.ctlab_error_string(2) == b"unsupported projection mode"
    assert L.ctlab_error_string(99) == b"unknown error"


def test_plan_recon_batch_argument_checks():
    """Entry points beyond the reference ABI reject bad arguments before any device call."""
    L = _lib.lib()
    assert L.ctlab_plan_set_tiles(None, 16, 16) == 1
    assert not L.ctlab_recon_create(None, 1)
    assert _lib.last_error()[0] == 1
    assert not L.ctlab_batch_create(None, 0)
    assert _lib.last_error()[0] == 1
    arr = (C.c_void_p * 2)(None, None)
    assert not L.ctlab_batch_create(arr, 2)
    assert _lib.last_error()[0] == 1 and "reconstruction 0" in _lib.last_error()[1]
    assert L.ctlab_batch_set_tiles(None, 24, 8) == 1
    assert L.ctlab_batch_os_sart(None, None, 1, 1.0, 0, 1, None) == 1
    assert L.ctlab_batch_autotune(None, None, 1.0, None) == 1
    assert L.ctlab_batch_size(None) == 0
    assert L.ctlab_recon_set_tiles(None, 8, 8) == 1
    done = C.c_int(-1)
    assert L.ctlab_cg_to_fidelity(None, None, 3, 0.0, None, None, None) == 1
    assert L.ctlab_cg_to_fidelity(None, None, 3, float("nan"), None, C.byref(done), None) == 1
    assert L.ctlab_cg_to_fidelity(None, None, 3, 0.0, None, C.byref(done), None) == 1
    assert _lib.last_error()[0] == 1
    # CNN objects: sizes checked before any device call
    assert not L.ctlab_cnn_create(0, 8, 8, 0, 0)
    assert not L.ctlab_cnn_create(8, 8, 12, 1, 0)                 # training needs ZN % 8 == 0
    assert not L.ctlab_cnn_create(1 << 20, 1 << 20, 1 << 10, 0, 0)  # beyond the kernels' grid
    assert _lib.last_error()[0] == 1 and "grid" in _lib.last_error()[1]


def test_multi_gpu_split_validation():
    """The split entry points check mode and split before any device work."""
    g = _tiny()
    vol = np.ones(g.nvox(), np.float32)
    with pytest.raises(_lib.CtlabError) as e:
        dd.DD3Proj_multiGPU(g, vol, [0, 3, 2], 3)          # decreasing
    assert e.value.code == 1
    with pytest.raises(_lib.CtlabError):
        dd.DD3Proj_multiGPU(g, vol, [1, 2], 2)             # must start at 0
    with pytest.raises(_lib.CtlabError):
        dd.DD3Proj_multiGPU(g, vol, [0, 9], 2)             # beyond PN = 4
    with pytest.raises(_lib.CtlabError) as e:
        dd.DD3Proj_multiGPU(g, vol, [0, 2], 2, prjMode=1)
    assert e.value.code == 2
    prj = np.ones(g.nprj(), np.float32)
    with pytest.raises(_lib.CtlabError):
        dd.DD3Back_multiGPU(g, prj, [0, 5, 6], 3, helical=True)   # beyond ZN = 4
    with pytest.raises(_lib.CtlabError) as e:
        dd.DD3Back_multiGPU(g, prj, [0, 1, 2, 3], 4, methodId=4, helical=True)
    assert e.value.code == 2
