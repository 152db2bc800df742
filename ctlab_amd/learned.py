"""Learned proximal step inside a safeguarded FISTA (no counterpart in CTLAB).

    z_k     = y_k - A^T(A y_k - b) / L                      (gradient step)
    x_tv    = prox_{lambda/L TV}(z_k)                        (the plain FISTA-TV step)
    x_cnn   = x_tv + net(x_tv)                               (learned correction, BF16 MFMA)
    x_{k+1} = x_cnn   if ||A x_cnn - b||^2 < ||A x_k - b||^2 (safeguard on the exact data fidelity)
              x_tv    otherwise
    y_{k+1} = FISTA momentum (src/FastIterativeShrinkageAlgorithm.h:17-42); after an
              accepted learned step the momentum restarts (t = 1, x_{k} := x_{k+1}), so
              the extrapolation does not amplify the network's correction

The learned step may include a data-consistency stage (DataConsistent: a
few CG iterations on ||Ax - b||^2 after the network), so that its output keeps
the iterate's fit to the data while the network's correction removes
artefacts; the safeguard is unchanged.

With learned_every = N > 1 the learned step is tried every N-th iteration
only (the others are plain FISTA-TV steps), so the network sees iterates close
to the plain FISTA-TV ones it was trained on.  Output mode (feedback = False):
the iterates are exactly those of plain FISTA-TV, and every N iterations the
output is x_cnn = x_k + net(x_k) when ||A x_cnn - b||^2 <= ||A x_k - b||^2, else
x_k; the correction never feeds back into the iteration.  With noise_level
given (the discrepancy principle, a variant of the safeguard), a learned output
is also accepted when ||A x_cnn - b||^2 <= noise_level: it fits the data to
within the noise.  Near convergence the plain iterates fit the noise, so a
strict decrease rejects every step towards the truth.

The learned part is a residual on top of the TV prox: a network that has
learned nothing (net = 0) gives back plain FISTA-TV exactly.  Training pairs
come from the projector itself: seeded random-ellipsoid phantoms, their
sparse-view projections with Poisson noise, and snapshots of the iterates x_k
of plain FISTA-TV on them (iterations k spread over the run), with the phantom
as target: the network learns to correct the iterates it will see.  The network is trained by the library's own forward /
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
    """x -> x + net(s x)/s, the network seeing attenuation scaled by INPUT_SCALE.
    The network's correction is taken back out of its output exactly
    ((s x + net) - s x), so a network that outputs 0 returns x bit for bit."""

    def __init__(self, shape, params, training=False):
        self.net = CN.CNN(shape, training=training, params=params)
        self.s = INPUT_SCALE

    def __call__(self, x, out, f_ref=None):
        xs = x * self.s
        self.net(xs, out=out)                     # out = s x + net(s x)
        RC.axpby(-1.0, xs, 1.0, out)              # net(s x)
        RC.axpby(1.0, x, 1.0 / self.s, out)       # x + net(s x) / s
        return out


EVAL_POPULATION = (512, 512, 64)   # ellipsoid population box (XN, YN, ZN) of the evaluation phantom


class DataConsistent:
    """Learned step with a data-consistency stage: x_c = x + net(x), then m
    CG iterations on ||Ax - b||^2 started from x_c (the reconstruction's own
    CG on the normal equations).  The network moves the estimate towards the
    phantom class, which also undoes the part of the iterate's fit that
    follows the noise; the CG steps restore the fit along the directions the
    data constrain most, while the correction stays in the directions they
    constrain little (the artefacts of sparse views).  The safeguard then
    applies to the result as to any learned step.

    With to_fidelity, the number of CG steps is not fixed: CG runs until
    ||Ax_c - b||^2 is at the reference value f_ref the caller passes (the plain
    iterate's), at most m_max steps, so that the candidate is compared with the
    plain iterate at equal data residual.  The stop aims 1e-5 (relative) below
    f_ref, so that the exact re-evaluation of the safeguard does not reject a
    candidate on the last digits of the value CG carries along."""

    def __init__(self, net, rc, m=3, to_fidelity=False, m_max=10):
        self.net, self.rc, self.m = net, rc, int(m)
        self.to_fidelity, self.m_max = bool(to_fidelity), int(m_max)
        self.last_steps = None

    def __call__(self, x, out, f_ref=None):
        self.net(x, out)
        if self.to_fidelity and f_ref is not None:
            self.last_steps, _ = self.rc.cg_to_fidelity(out, f_ref * (1.0 - 1e-5), self.m_max)
        elif self.m > 0:
            self.rc.cg(out, self.m)
            self.last_steps = self.m


def noise_energy(b, I0):
    """Expected ||noise||^2 of log-transformed Poisson data, from the data alone:
    var(-log(N/I0)) ~ 1/E[N] = exp(b)/I0 per ray."""
    import torch
    return float((torch.exp(b.double()) / I0).sum())


def make_problem(g, plan, rc, seed, I0, dev, population=EVAL_POPULATION, with_noise=False):
    """Seeded phantom on the device and its noisy data set on rc; returns the
    phantom (and the data's expected noise energy when with_noise).

    The phantom is a window of g's size cut at a seeded random position from 20
    random ellipsoids spread over a population box (voxel counts, same voxel
    size), so that volumes smaller than the box see the same object sizes and
    density as a volume of the box's size."""
    PX, PY, PZ = population
    ells = PH.random_ellipsoids(seed, 20, PX, PY, PZ, g.dx, g.dz)
    rng = np.random.default_rng(seed)
    off = [rng.uniform(-0.5, 0.5) * max(P - N, 0) * d for P, N, d in
           ((PX, g.XN, g.dx), (PY, g.YN, g.dx), (PZ, g.ZN, g.dz))]
    ells = ells.copy()
    ells[:, 0] -= off[0]
    ells[:, 1] -= off[1]
    ells[:, 2] -= off[2]
    ph = voxelize_torch(ells, g.vol_shape, g.dx, g.dz, dev)
    b = noisy_projection(plan, ph, I0, seed)
    rc.set_data(b)
    return (ph, noise_energy(b, I0)) if with_noise else ph


def snapshot_iters(rng, kmin, kmax, n):
    """n distinct FISTA iterations in [kmin, kmax], log-uniform (denser early)."""
    ks = set()
    while len(ks) < n:
        ks.add(int(np.clip(round(np.exp(rng.uniform(np.log(kmin), np.log(kmax)))), kmin, kmax)))
    return sorted(ks)


def generate_pairs(n_phantoms, per_phantom, shape, views, I0, seed0, dev, tv_lambda, tv_iters=10, kmin=5,
                   kmax=100, log=None, population=EVAL_POPULATION):
    """n_phantoms * per_phantom (x_k, phantom) pairs on the device: x_k is the
    FISTA-TV (tv_lambda) iterate at snapshot iterations k.  Returns pairs, L, seconds."""
    import torch
    YN, XN, ZN = shape
    g = cone_geometry(XN, YN, ZN, views)
    plan = dd.Plan(g)
    rc = RC.Recon(plan, 1)
    L = rc.lipschitz(30) * 1.01
    rng = np.random.default_rng(seed0)
    xs, ts = [], []
    t0 = time.perf_counter()
    for i in range(n_phantoms):
        ph = make_problem(g, plan, rc, seed0 + i, I0, dev, population)
        ks = snapshot_iters(rng, kmin, kmax, per_phantom)
        _, _, snaps = fista_tv_run(rc, L, g.vol_shape, ks[-1], tv_lambda, tv_iters, None, None, dev, snaps=ks)
        for k in ks:
            xs.append(snaps[k])
            ts.append(ph)
        if log and (i + 1) % 100 == 0:
            torch.cuda.synchronize()
            log(f"phantoms {i + 1}/{n_phantoms}, pairs {len(xs)} ({time.perf_counter() - t0:.1f} s)")
    torch.cuda.synchronize()
    return xs, ts, L, time.perf_counter() - t0


def identity_loss(xs, ts):
    """The training loss of net = 0: 0.5*mean((x - phantom)^2) in the scaled units."""
    s2 = INPUT_SCALE * INPUT_SCALE
    return float(np.mean([0.5 * s2 * float(((x - t) ** 2).mean()) for x, t in zip(xs, ts)]))


def network_loss(xs, ts, shape, params):
    """The same loss with the trained correction applied (forward only)."""
    import torch
    net = ScaledCNN(shape, params)
    out = torch.empty_like(xs[0])
    s2 = INPUT_SCALE * INPUT_SCALE
    vals = []
    for x, t in zip(xs, ts):
        net(x, out)
        vals.append(0.5 * s2 * float(((out - t) ** 2).mean()))
    return float(np.mean(vals))


def learning_rate(lr, schedule, step, total, epoch, epochs):
    """Step size of Adam step `step` (0-based) of `total`: "step" keeps lr and
    runs the last epoch at 0.3*lr; "cosine" decays from lr to 0.05*lr along a
    half cosine over all steps."""
    if schedule == "cosine":
        return lr * (0.05 + 0.95 * 0.5 * (1.0 + math.cos(math.pi * step / max(total, 1))))
    if schedule == "step":
        return lr * (0.3 if epoch == epochs - 1 and epochs > 1 else 1.0)
    raise ValueError(f"unknown learning-rate schedule {schedule!r}")


def train(xs, ts, shape, epochs, lr=1e-3, seed=0, log=None, schedule="step"):
    """Adam on 0.5*mean((x + net(x) - phantom)^2) in the scaled units, step
    sizes by learning_rate(schedule).  Returns params, per-epoch mean losses,
    seconds, steps."""
    import torch
    net = CN.CNN(shape, training=True, params=CN.init_params(seed))
    s = INPUT_SCALE
    losses = []
    order_rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    steps = 0
    total = epochs * len(xs)
    for ep in range(epochs):
        order = order_rng.permutation(len(xs))
        ep_loss = 0.0
        for k, i in enumerate(order):
            cur_lr = learning_rate(lr, schedule, steps, total, ep, epochs)
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


def fista_tv_run(rc, L, b_shape_vol, iters, tv_lambda, tv_iters, truth, zsl, dev, learned=None, snaps=None,
                 learned_every=1, feedback=True, noise_level=None):
    """Plain FISTA-TV (learned=None), the safeguarded learned FISTA (learned
    given; tried every learned_every iterations) or its output mode
    (feedback=False), step by step with the library's kernels.  Returns x, the
    per-iteration history (wall time of the GPU work, synchronised; PSNR of the
    output and of a rejected learned candidate when truth is given; data
    fidelity; accepted learned step) and {k: copy of the iterate x_k}."""
    import torch
    shape = b_shape_vol
    n = int(np.prod(shape))
    x = torch.zeros(shape, device=dev)
    y = torch.zeros_like(x)
    x_old = torch.zeros_like(x)
    g = torch.empty_like(x)
    z = torch.empty_like(x)
    xc = torch.empty_like(x)
    xt = torch.empty_like(x) if learned is not None else None
    work = torch.empty(9 * n, device=dev)
    t = 1.0
    f_x = rc.data_fidelity(x)
    in_loop = learned is not None and feedback
    out = x                               # the output of the current iteration
    f_out = f_x
    cand_psnr = cand_fid = None
    hist = []
    taken = {}
    elapsed = 0.0
    for k in range(iters):
        tried = learned is not None and ((k + 1) % learned_every == 0 or k + 1 == iters)
        # in-loop mode compares with f(x_k) at the next tried iteration: keep it current
        nxt = learned is not None and ((k + 2) % learned_every == 0 or k + 2 == iters)

        def step():
            nonlocal t, f_x
            rc.gradient(y, g)
            z.copy_(y)
            RC.axpby(-1.0 / L, g, 1.0, z)
            nonlocal out, f_out, cand_psnr, cand_fid
            accepted = False
            cand_psnr = cand_fid = None
            if not in_loop or not tried:
                RC.tv_prox_into(z, x, tv_lambda / L, tv_iters, shape, work)
                f_x = None
                out, f_out = x, None
                if tried:
                    f_x = rc.data_fidelity(x)
                    learned(x, xc, f_ref=f_x)
                    f_c = rc.data_fidelity(xc)
                    f_out = f_x
                    if f_c <= f_x or (noise_level is not None and f_c <= noise_level):
                        out, f_out = xc, f_c
                        accepted = True
                    else:
                        cand_fid = f_c
                elif in_loop and nxt:
                    f_x = rc.data_fidelity(x)
                    f_out = f_x
            else:
                RC.tv_prox_into(z, xt, tv_lambda / L, tv_iters, shape, work)
                learned(xt, xc, f_ref=f_x)
                f_c = rc.data_fidelity(xc)
                if f_c < f_x:
                    x.copy_(xc)
                    f_x = f_c
                    accepted = True
                else:
                    cand_fid = f_c
                    x.copy_(xt)
                    f_x = rc.data_fidelity(x) if nxt else None
                out, f_out = x, f_x
            if accepted and in_loop:   # momentum restart
                t = 1.0
                x_old.copy_(x)
            tn = 0.5 * (1.0 + math.sqrt(1.0 + 4.0 * t * t))
            y.copy_(x)
            RC.fista_momentum(y, x_old, t, tn)
            x_old.copy_(x)
            t = tn
            return accepted
        acc, dt = _timed(step)
        elapsed += dt
        # monitoring outside the timed region: the output's fidelity where the
        # algorithm did not need it, PSNRs, snapshots
        if f_out is None:
            f_out = rc.data_fidelity(out)
        if out is x:
            f_x = f_out
        if cand_fid is not None and truth is not None:
            cand_psnr = psnr(xc, truth, zsl)
        if snaps is not None and (k + 1) in snaps:
            taken[k + 1] = x.clone()
        hist.append({"iter": k + 1, "t": elapsed, "psnr": psnr(out, truth, zsl) if truth is not None else None,
                     "fidelity": f_out, "learned": acc, "rejected_candidate_psnr": cand_psnr,
                     "rejected_candidate_fidelity": cand_fid,
                     "dc_steps": getattr(learned, "last_steps", None) if tried else None})
    return out, hist, taken


def at_matched_fidelity(h_ref, h, level):
    """First iteration of each run whose output fidelity is <= level:
    (time, iteration, PSNR) for h_ref and for h."""
    def first(hh):
        for x in hh:
            if x["fidelity"] <= level:
                return x["t"], x["iter"], x["psnr"]
        return None, None, None
    return first(h_ref), first(h)


def time_to_psnr(hist, target):
    for h in hist:
        if h["psnr"] >= target:
            return h["t"], h["iter"], h["fidelity"]
    return None, None, None


def cnn_train_step_rate(shape, params, dev, reps=3):
    """FLOP/s of one training step (forward + input and weight gradients + Adam;
    3 x 169,344 FLOP per voxel for the convolutions)."""
    import torch
    net = CN.CNN(shape, training=True, params=params)
    x = torch.rand(shape, device=dev)
    t = torch.rand(shape, device=dev)
    net.loss_grad(x, t)
    net.adam_step(1e-4)
    best = 1e30
    for _ in range(reps):
        _, dt = _timed(lambda: (net.loss_grad(x, t), net.adam_step(1e-4)))
        best = min(best, dt)
    return 3 * 169344.0 * x.numel() / best, best


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
