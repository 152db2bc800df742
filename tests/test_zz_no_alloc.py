"""ctlab_cg_beta_static: the no-allocation form of ctlab_cg_beta, called through
the C ABI, gives the same two FP64 sums bit for bit, on two calls in a row (the
second with a smaller input, so stale partial sums of the first would show)."""
import ctypes as C

import numpy as np
import pytest

from ctlab_amd import _lib

pytestmark = pytest.mark.gpu


def _sums_alloc(L, gk, gk1, d, cases):
    h = np.zeros(2, dtype=np.float64)
    rc = L.ctlab_cg_beta(C.c_void_p(gk.data_ptr()), C.c_void_p(gk1.data_ptr()), C.c_void_p(d.data_ptr()), cases,
                         gk.numel(), h.ctypes.data_as(C.POINTER(C.c_double)), None)
    assert rc == 0, _lib.last_error()
    return h


def _sums_static(L, gk, gk1, d, cases, out):
    import torch
    rc = L.ctlab_cg_beta_static(C.c_void_p(gk.data_ptr()), C.c_void_p(gk1.data_ptr()), C.c_void_p(d.data_ptr()),
                                cases, gk.numel(), C.c_void_p(out.data_ptr()), None)
    assert rc == 0, _lib.last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy().copy()


def test_static_beta_equals_allocating_beta(cuda):
    import torch
    L = _lib.lib()
    out = torch.full((2,), float("nan"), dtype=torch.float64, device=cuda)
    gen = torch.Generator(device="cpu").manual_seed(5)
    for n, cases in ((3_000_001, 1), (777, 4), (1_000_000, 2), (5, 0), (64, 3)):
        v = torch.randn(3, n, generator=gen).to(cuda)
        gk, gk1, d = v[0].contiguous(), v[1].contiguous(), v[2].contiguous()
        print(f"n={n} case {cases}", flush=True)
        ref = _sums_alloc(L, gk, gk1, d, cases)
        got = _sums_static(L, gk, gk1, d, cases, out)
        assert np.array_equal(got, ref), (n, cases, got, ref)
    # invalid arguments are refused with an error, no launch
    rc = L.ctlab_cg_beta_static(C.c_void_p(gk.data_ptr()), C.c_void_p(gk1.data_ptr()), C.c_void_p(d.data_ptr()),
                                7, gk.numel(), C.c_void_p(out.data_ptr()), None)
    assert rc == 1
    rc = L.ctlab_cg_beta_static(C.c_void_p(gk.data_ptr()), C.c_void_p(gk1.data_ptr()), C.c_void_p(d.data_ptr()),
                                0, gk.numel(), None, None)
    assert rc == 1


def test_static_beta_under_graph_capture(cuda):
    """Captured in a HIP graph and replayed three times, with eager work and
    synchronisation between the replays, each replay equals the eager call."""
    import torch
    L = _lib.lib()
    n = 200_003
    gen = torch.Generator(device="cpu").manual_seed(6)
    v = torch.randn(3, n, generator=gen).to(cuda)
    gk, gk1, d = v[0].contiguous(), v[1].contiguous(), v[2].contiguous()
    ref = _sums_alloc(L, gk, gk1, d, 1)
    out = torch.zeros(2, dtype=torch.float64, device=cuda)
    # an eager call first: it looks the scratch address up outside the capture
    assert np.array_equal(_sums_static(L, gk, gk1, d, 1, out), ref)
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        rc = L.ctlab_cg_beta_static(C.c_void_p(gk.data_ptr()), C.c_void_p(gk1.data_ptr()),
                                    C.c_void_p(d.data_ptr()), 1, n, C.c_void_p(out.data_ptr()),
                                    C.c_void_p(s.cuda_stream))
    assert rc == 0
    for k in range(3):
        out.zero_()
        small = torch.randn(1000, device=cuda)
        _sums_alloc(L, small, small, small, 0)      # other work on the scratch's neighbours
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), ref), k
