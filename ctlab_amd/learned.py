"""Learned proximal step inside a safeguarded FISTA (no counterpart in CTLAB).

    z_k     = y_k - A^T(A y_k - b) / L                      (gradient step)
    x_cnn   = z_k + net(z_k)                                 (learned prox, BF16 MFMA)
    x_{k+1} = x_cnn   if ||A x_cnn - b||^2 < ||A x_k - b||^2 (safeguard on the exact data fidelity)
              prox_{lambda/L TV}(z_k)  otherwise             (the plain FISTA-TV step)
    y_{k+1} = FISTA momentum (src/FastIterativeShrinkageAlgorithm.h:17-42)

Training pairs come from the projector itself: seeded random-ellipsoid
phantoms, their sparse-view projections with Poisson noise, and the
gradient-step iterate z_K of K plain FISTA iterations (K random), with the
phantom as target.  The network is trained by the library's own forward /
backward / Adam kernels (ctlab_amd.cnn).  Phantom voxelisation and Poisson
sampling use torch on the device (data synthesis, not the reconstruction).
"""
import math
import time

import numpy as np

from . import cnn as CN
from . import dd
from . import geometry as G
from . import phantom as PH
from . import recon as RC

INPUT_SCALE = 50.0     # attenuation [1/mm] * 50 ~ O(1) for the network


def cone_geometry(XN, YN, ZN, PN, dx=0.5, dz=1.0):
    """cfg2 scanner (888 x 64 arc detector, 1.0239 x 1.0963 mm, SID 541, SDD 949)."""
    return G.circular_cone(XN=XN, YN=YN, ZN=ZN, dx=dx, dz=dz, PN=PN, DNU=888, DNV=64, col_size=1.0239,
                           row_size=1.0963, sid=541.0, sdd=949.0)


def voxelize_torch(ells, shape, dx, dz, dev):
    """Same sampling as phantom.voxelize (centre-sampled indicators, summed), on the device."""
    import torch
    YN, XN, ZN = shape
    xs = (torch.arange(XN, device=dev, dtype=torch.float64) - (XN - 1) / 2.0) * dx
    ys = (torch.arange(YN, device=dev, dtype=torch.float64) - (YN - 1) / 2.0) * dx
    zs = (torch.arange(ZN, device=dev, dtype=torch.float64) - (ZN - 1) / 2.0) * dz
    Y, X = torch.meshgrid(ys, xs, indexing="ij")
    vol = torch.zeros(shape, device=dev, dtype=torch.float64)
    for (ex, ey, ez, a, b, c, ang, m) in ells:
        ca, sa = math.cos(ang), math.sin(ang)
        u = (X - ex) * ca + (Y - ey) * sa
        w = -(X - ex) * sa + (Y - ey) * ca
        q = (u / a) ** 2 + (w / b) ** 2
        if math.isinf(c):
            inside = (q <= 1.0)[:, :, None].expand(shape)
        else:
            inside = (q[:, :, None] + ((zs[None, None, :] - ez) / c) ** 2) <= 1.0
        vol += m * inside
    return vol.float().contiguous()


def noisy_projection(plan, x, I0, seed):
    """-log(max(Poisson(I0 exp(-Ax)), 1) / I0)"""
    import torch
    p = plan.proj(x)
    gen = torch.Generator(device=p.device)
    gen.manual_seed(int(seed))
    counts = torch.poisson(I0 * torch.exp(-p.double()), generator=gen)
    return (-torch.log(torch.clamp(counts, min=1.0) / I0)).float().contiguous()


def fista_iterate(rc, L, iters, shape, dev):
    """K plain FISTA gradient steps from 0; returns the last z = y - A^T(Ay-b)/L."""
    import torch
    y = torch.zeros(shape, device=dev)
    x_old = torch.zeros_like(y)
    g = torch.empty_like(y)
    z = torch.empty_like(y)
    t = 1.0
    for k in range(iters):
        rc.gradient(y, g)
        z.copy_(y)
        RC.axpby(-1.0 / L, g, 1.0, z)
        if k == iters - 1:
            break
        tn = 0.5 * (1.0 + math.sqrt(1.0 + 4.0 * t * t))
        y.copy_(z)
        RC.fista_momentum(y, x_old, t, tn)
        x_old.copy_(z)
        t = tn
    return z


def psnr(x, ref, zsl=None):
    """10 log10(max(ref)^2 / MSE) over the slices zsl (default: all)."""
    if zsl is not None:
        x, ref = x[:, :, zsl], ref[:, :, zsl]
    mse = float(((x.double() - ref.double()) ** 2).mean())
    return 10.0 * math.log10(float(ref.max()) ** 2 / max(mse, 1e-30))


def covered_slices(g):
    """z slices seen by every view: |z| < half detector height * (nearest slab distance) / SDD."""
    rmax = math.hypot(g.XN * g.dx / 2, g.YN * g.dx / 2)
    half = 0.5 * g.DNV * 1.0963 * (541.0 - rmax) / 949.0
    zc = (np.arange(g.ZN) - (g.ZN - 1) / 2.0) * g.dz
    idx = np.nonzero(np.abs(zc) + 0.5 * g.dz <= half)[0]
    return slice(int(idx[0]), int(idx[-1]) + 1)


class ScaledCNN:
    """x -> x + net(s x)/s, the network seeing attenuation scaled by INPUT_SCALE."""

    def __init__(self, shape, params, training=False):
        self.net = CN.CNN(shape, training=training, params=params)
        self.s = INPUT_SCALE

    def __call__(self, x, out):
        import torch
        xs = x * self.s
        self.net(xs, out=out)
        RC.axpby(0.0, out, 1.0 / self.s, out)
        return out


def generate_pairs(n, shape, views, I0, seed0, dev, kmin=2, kmax=20, log=None):
    """n (input z_K, phantom) pairs on the device, plus the geometry's L."""
    import torch
    YN, XN, ZN = shape
    g = cone_geometry(XN, YN, ZN, views)
    plan = dd.Plan(g)
    rc = RC.Recon(plan, 1)
    L = rc.lipschitz(30) * 1.01
    rng = np.random.default_rng(seed0)
    xs, ts = [], []
    t0 = time.perf_counter()
    for i in range(n):
        ells = PH.random_ellipsoids(seed0 + i, 20, XN, YN, ZN, g.dx, g.dz)
        ph = voxelize_torch(ells, g.vol_shape, g.dx, g.dz, dev)
        b = noisy_projection(plan, ph, I0, seed0 + i)
        rc.set_data(b)
        K = int(rng.integers(kmin, kmax + 1))
        xs.append(fista_iterate(rc, L, K, g.vol_shape, dev))
        ts.append(ph)
        if log and (i + 1) % 250 == 0:
            torch.cuda.synchronize()
            log(f"pairs {i + 1}/{n} ({time.perf_counter() - t0:.1f} s)")
    torch.cuda.synchronize()
    return xs, ts, L, time.perf_counter() - t0


def train(xs, ts, shape, epochs, lr=1e-3, seed=0, log=None):
    """Adam on 0.5*mean((z + net(z) - phantom)^2) in the scaled units; returns params and losses."""
    import torch
    net = CN.CNN(shape, training=True, params=CN.init_params(seed))
    s = INPUT_SCALE
    losses = []
    order_rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    steps = 0
    for ep in range(epochs):
        order = order_rng.permutation(len(xs))
        ep_loss = 0.0
        for k, i in enumerate(order):
            cur_lr = lr * (0.3 if ep == epochs - 1 and epochs > 1 else 1.0)
            loss = net.loss_grad(xs[i] * s, ts[i] * s)
            net.adam_step(cur_lr)
            ep_loss += loss
            steps += 1
        losses.append(ep_loss / len(order))
        if log:
            log(f"epoch {ep + 1}/{epochs}: mean loss {losses[-1]:.4e} ({time.perf_counter() - t0:.1f} s)")
    torch.cuda.synchronize()
    return net.params(), losses, time.perf_counter() - t0, steps


def _timed(fn):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = fn()
    torch.cuda.synchronize()
    return r, time.perf_counter() - t0


def fista_tv_run(rc, L, b_shape_vol, iters, tv_lambda, tv_iters, truth, zsl, dev, learned=None,
                 target_db=40.0):
    """Plain FISTA-TV (learned=None) or the safeguarded learned FISTA, step by
    step with the library's kernels.  Returns per-iteration history: wall time
    (GPU work only, synchronised), PSNR, data fidelity, accepted learned step."""
    import torch
    shape = b_shape_vol
    n = int(np.prod(shape))
    x = torch.zeros(shape, device=dev)
    y = torch.zeros_like(x)
    x_old = torch.zeros_like(x)
    g = torch.empty_like(x)
    z = torch.empty_like(x)
    xc = torch.empty_like(x)
    work = torch.empty(7 * n, device=dev)
    t = 1.0
    f_x = rc.data_fidelity(x)
    hist = []
    elapsed = 0.0
    for k in range(iters):
        def step():
            nonlocal t, f_x
            rc.gradient(y, g)
            z.copy_(y)
            RC.axpby(-1.0 / L, g, 1.0, z)
            accepted = False
            if learned is not None:
                learned(z, xc)
                f_c = rc.data_fidelity(xc)
                if f_c < f_x:
                    x.copy_(xc)
                    f_x = f_c
                    accepted = True
            if not accepted:
                RC.tv_prox_into(z, x, tv_lambda / L, tv_iters, shape, work)
                f_x = rc.data_fidelity(x)
            tn = 0.5 * (1.0 + math.sqrt(1.0 + 4.0 * t * t))
            y.copy_(x)
            RC.fista_momentum(y, x_old, t, tn)
            x_old.copy_(x)
            t = tn
            return accepted
        acc, dt = _timed(step)
        elapsed += dt
        hist.append({"iter": k + 1, "t": elapsed, "psnr": psnr(x, truth, zsl), "fidelity": f_x, "learned": acc})
    return x, hist


def time_to_psnr(hist, target):
    for h in hist:
        if h["psnr"] >= target:
            return h["t"], h["iter"], h["fidelity"]
    return None, None, None


def cnn_forward_rate(shape, params, dev, reps=3):
    """FLOP/s of one network forward at this volume shape (169,344 FLOP per voxel)."""
    import torch
    net = CN.CNN(shape, training=False, params=params)
    x = torch.rand(shape, device=dev)
    out = torch.empty_like(x)
    net(x, out=out)
    best = 1e30
    for _ in range(reps):
        _, dt = _timed(lambda: net(x, out=out))
        best = min(best, dt)
    flops = 169344.0 * x.numel()
    return flops / best, best
