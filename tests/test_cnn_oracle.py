"""CPU pins of the conv3d / CNN oracle (checker of the BF16 MFMA kernels).

Forward conv against scipy.ndimage.correlate (a library routine with the same
zero-padded 3x3x3 cross-correlation); input and weight gradients against
central finite differences; the whole 5-layer network gradient likewise."""
import numpy as np
import pytest
from scipy import ndimage

import cnn_ref as CR
import oracle as O


def test_conv3d_forward_matches_scipy():
    rng = np.random.default_rng(0)
    shape = (5, 6, 7)        # (YN, XN, ZN)
    cin, cout = 3, 2
    x = rng.standard_normal((np.prod(shape), cin))
    w = rng.standard_normal((27, cout, cin))
    b = rng.standard_normal(cout)
    out = O.conv3d(x, w, b, shape)
    for co in range(cout):
        ref = np.full(shape, b[co])
        for ci in range(cin):
            k = w[:, co, ci].reshape(3, 3, 3)           # taps ordered (dy, dx, dz)
            ref += ndimage.correlate(x[:, ci].reshape(shape), k, mode="constant", cval=0.0)
        assert np.allclose(out[:, co], ref.reshape(-1), atol=1e-12)


def test_conv3d_gradients_finite_differences():
    rng = np.random.default_rng(1)
    shape = (3, 4, 5)
    cin, cout = 2, 3
    x = rng.standard_normal((np.prod(shape), cin))
    w = rng.standard_normal((27, cout, cin))
    gout = rng.standard_normal((np.prod(shape), cout))
    f = lambda xx, ww: float(np.sum(gout * O.conv3d(xx, ww, None, shape)))
    gin = O.conv3d_dgrad(gout, w, shape)
    gw, gb = O.conv3d_wgrad(x, gout, cin, cout, shape)
    eps = 1e-6
    for idx in [(0, 0), (7, 1), (59, 0)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += eps; xm[idx] -= eps
        assert abs((f(xp, w) - f(xm, w)) / (2 * eps) - gin[idx]) < 1e-6
    for idx in [(0, 0, 0), (13, 2, 1), (26, 1, 0)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += eps; wm[idx] -= eps
        assert abs((f(x, wp) - f(x, wm)) / (2 * eps) - gw[idx]) < 1e-6
    assert np.allclose(gb, gout.sum(0))


def test_network_gradient_finite_differences():
    rng = np.random.default_rng(2)
    shape = (4, 4, 5)
    ws, bs = CR.init_weights(3)
    ws = [w[:, :min(w.shape[1], 4), :min(w.shape[2], 4)] for w in ws]   # narrow net for speed
    CR_CH = CR.CH
    try:
        CR.CH = [1, 4, 4, 4, 4, 1]
        bs = [rng.standard_normal(w.shape[1]) * 0.1 for w in ws]
        x = rng.standard_normal(np.prod(shape))
        t = rng.standard_normal(np.prod(shape))
        L, dws, dbs = CR.loss_and_grads(ws, bs, x, t, shape)
        eps = 1e-6
        for l, idx in [(0, (4, 1, 0)), (2, (13, 3, 2)), (4, (20, 0, 3))]:
            wp = [w.copy() for w in ws]; wm = [w.copy() for w in ws]
            wp[l][idx] += eps; wm[l][idx] -= eps
            num = (CR.loss_and_grads(wp, bs, x, t, shape)[0] - CR.loss_and_grads(wm, bs, x, t, shape)[0]) / (2 * eps)
            assert abs(num - dws[l][idx]) < 1e-6 * max(1.0, abs(num))
        for l in (1, 3):
            bp = [b.copy() for b in bs]; bm = [b.copy() for b in bs]
            bp[l][0] += eps; bm[l][0] -= eps
            num = (CR.loss_and_grads(ws, bp, x, t, shape)[0] - CR.loss_and_grads(ws, bm, x, t, shape)[0]) / (2 * eps)
            assert abs(num - dbs[l][0]) < 1e-6 * max(1.0, abs(num))
    finally:
        CR.CH = CR_CH
