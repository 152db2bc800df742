"""FP64 CPU reference drivers of the reconstruction loops, on the oracle
projector.  TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker).

They follow the same steps as the library's device drivers:
  OS-SART  src/DD3_GPU_recon.h:60-65 with SART_Weighting (src/SARTWeighting.h:22-91),
           interleaved subsets (view p in subset p % osNum), rowSum = A_s 1, colSum = A_s^T 1;
  FISTA-TV gradient step, FGP TV prox (recon_oracle.c), FISTA_functor momentum;
  CG       normal equations, updates src/utilities.hpp:1121-1156, beta
           src/utilities.hpp:1228-1268 evaluated on gradients g = -r.
"""
import numpy as np

import oracle as O


def subsets(PN, osNum):
    return [np.arange(k, PN, osNum) for k in range(osNum)]


def os_sart(g, b, osNum, iters, lam=1.0, nonneg=False, x0=None):
    b = np.asarray(b, np.float64).reshape(g.PN, -1)
    x = np.zeros(g.nvox()) if x0 is None else np.array(x0, np.float64).reshape(-1)
    subs = []
    for views in subsets(g.PN, osNum):
        gs = g.subset(views)
        rowsum = O.proj(gs, np.ones(g.nvox()))
        colsum = O.back(gs, np.ones(gs.nprj()))
        subs.append((gs, b[views].reshape(-1), rowsum, colsum))
    for _ in range(iters):
        for gs, bs, rowsum, colsum in subs:
            p = O.proj(gs, x)
            r = O.sart_prj_weight(p, bs, rowsum)
            v = O.back(gs, r)
            x = x + lam * (O.sart_bak_weight(v, np.zeros_like(v), colsum))
            if nonneg:
                x = np.maximum(x, 0.0)
    return x


def fista_tv(g, b, iters, tv_lambda, tv_iters, L, nonneg=False, x0=None):
    b = np.asarray(b, np.float64).reshape(-1)
    x = np.zeros(g.nvox()) if x0 is None else np.array(x0, np.float64).reshape(-1)
    y, xold, t = x.copy(), x.copy(), 1.0
    obj = []
    for _ in range(iters):
        r = O.proj(g, y) - b
        obj.append(float(r @ r))
        z = y - O.back(g, r) / L
        x = O.tv_prox(z, g.XN, g.YN, g.ZN, tv_lambda / L, tv_iters)
        tn = 0.5 * (1.0 + np.sqrt(1.0 + 4.0 * t * t))
        if nonneg:
            x = np.clip(x, 0.0, 6000.0)
        y = O.fista(x, xold, t, tn, 0.0 if nonneg else -2000.0, 6000.0)
        xold = x.copy()
        t = tn
    return x, np.array(obj)


def cg(g, b, iters, cgm, x0=None):
    b = np.asarray(b, np.float64).reshape(-1)
    x = np.zeros(g.nvox()) if x0 is None else np.array(x0, np.float64).reshape(-1)
    r = O.back(g, b - O.proj(g, x))
    d = r.copy()
    res = []
    for _ in range(iters):
        q = O.proj(g, d)
        qq = q @ q
        if not qq > 0:
            break
        alpha = (r @ d) / qq
        x = x + alpha * d
        rn = r - alpha * O.back(g, q)
        beta = O.cg_beta(-r, -rn, d, cgm)
        d = rn + beta * d
        r = rn
        res.append(float(r @ r))
    return x, np.array(res)


def lipschitz(g, iters=20):
    v = np.ones(g.nvox())
    lam = 0.0
    for _ in range(iters):
        v /= np.linalg.norm(v)
        w = O.back(g, O.proj(g, v))
        lam = v @ w
        v = w
    return lam
