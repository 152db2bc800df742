"""Raw volume / projection files, as CTLAB reads and writes them
(readData_host / writeData_host / writeFileFromDeviceToDisk,
src/readwrite.hpp:15-113): n native-endian float32 or float64 elements, no
header.  Volumes are stored in (Z,X,Y) order (z fastest, then x, then y;
src/DD3_GPU_recon.h:37), projections as [view][column][row] (row fastest).

The functions go through the library's C ABI (ctlab_read_raw /
ctlab_write_raw for numpy arrays, the _device forms for torch tensors on the
GPU).  Command line:

    python -m ctlab_amd.rawio info FILE [--dtype f32|f64] [--shape D0,D1,...]
    python -m ctlab_amd.rawio convert IN OUT --from f64 --to f32
    python -m ctlab_amd.rawio slice FILE --shape YN,XN,ZN --z K OUT.pgm [--dtype f32]

``--shape`` is the array shape with the fastest index last: (YN, XN, ZN) for a
volume, (PN, DNU, DNV) for projections.
"""
import argparse
import os
import sys

import numpy as np

from . import _lib

_DT = {"f32": np.float32, "f64": np.float64}


def _elem(dtype):
    dt = np.dtype(dtype)
    if dt not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise ValueError(f"raw files hold float32 or float64, not {dt}")
    return dt


def read_raw(path, n=None, dtype=np.float32, shape=None):
    """n elements (default: the whole file, or prod(shape)) of dtype from path."""
    dt = _elem(dtype)
    if n is None:
        n = int(np.prod(shape)) if shape is not None else os.path.getsize(path) // dt.itemsize
    out = np.empty(int(n), dtype=dt)
    L = _lib.lib()
    rc = L.ctlab_read_raw(os.fsencode(path), out.ctypes.data if n else None, int(n), dt.itemsize)
    _lib.check(rc)
    return out.reshape(shape) if shape is not None else out


def write_raw(path, a):
    """Writes a float32 / float64 array (C order) to path."""
    a = np.ascontiguousarray(a)
    dt = _elem(a.dtype)
    L = _lib.lib()
    rc = L.ctlab_write_raw(os.fsencode(path), a.ctypes.data if a.size else None, int(a.size), dt.itemsize)
    _lib.check(rc)


def read_raw_device(path, out):
    """Fills the contiguous float32 / float64 CUDA tensor `out` from path."""
    size = {4: 4, 8: 8}[out.element_size()]
    assert out.is_cuda and out.is_contiguous()
    rc = _lib.lib().ctlab_read_raw_device(os.fsencode(path), out.data_ptr(), out.numel(), size)
    _lib.check(rc)
    return out


def write_raw_device(path, t):
    """Writes the contiguous float32 / float64 CUDA tensor t to path."""
    assert t.is_cuda and t.is_contiguous()
    rc = _lib.lib().ctlab_write_raw_device(os.fsencode(path), t.data_ptr(), t.numel(), t.element_size())
    _lib.check(rc)


def _shape(s):
    return tuple(int(v) for v in s.split(",")) if s else None


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m ctlab_amd.rawio", description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("info", help="element count and statistics")
    a.add_argument("file")
    a.add_argument("--dtype", default="f32", choices=sorted(_DT))
    a.add_argument("--shape", default="")
    c = sub.add_parser("convert", help="float32 <-> float64")
    c.add_argument("src")
    c.add_argument("dst")
    c.add_argument("--from", dest="src_dtype", default="f64", choices=sorted(_DT))
    c.add_argument("--to", dest="dst_dtype", default="f32", choices=sorted(_DT))
    s = sub.add_parser("slice", help="one z slice of a (Z,X,Y) volume as an 8-bit PGM image")
    s.add_argument("file")
    s.add_argument("out")
    s.add_argument("--shape", required=True, help="YN,XN,ZN")
    s.add_argument("--z", type=int, required=True)
    s.add_argument("--dtype", default="f32", choices=sorted(_DT))
    args = ap.parse_args(argv)
    if args.cmd == "info":
        shape = _shape(args.shape)
        x = read_raw(args.file, dtype=_DT[args.dtype], shape=shape)
        size = os.path.getsize(args.file)
        print(f"{args.file}: {x.size} x {args.dtype} ({size} bytes"
              f"{', shape ' + str(x.shape) if shape else ''}) min {x.min():.6g} max {x.max():.6g} "
              f"mean {x.mean(dtype=np.float64):.6g}" if x.size else f"{args.file}: empty")
    elif args.cmd == "convert":
        write_raw(args.dst, read_raw(args.src, dtype=_DT[args.src_dtype]).astype(_DT[args.dst_dtype]))
    else:
        YN, XN, ZN = _shape(args.shape)
        if not 0 <= args.z < ZN:
            ap.error(f"--z must be in [0, {ZN})")
        sl = read_raw(args.file, dtype=_DT[args.dtype], shape=(YN, XN, ZN))[:, :, args.z].astype(np.float64)
        lo, hi = float(sl.min()), float(sl.max())
        img = np.zeros_like(sl) if hi <= lo else (sl - lo) / (hi - lo)
        with open(args.out, "wb") as f:
            f.write(f"P5\n{XN} {YN}\n255\n".encode())
            f.write(np.round(img * 255.0).astype(np.uint8).tobytes())
    return 0


if __name__ == "__main__":
    sys.exit(main())
