"""Learned proximal step: a network trained by the library's own forward /
backward / Adam kernels on pairs the projector generates lowers the prox loss
below the identity map (the method without it) on held-out pairs, and the
safeguarded FISTA accepts a learned step only when ||Ax - b||^2 decreases."""
import numpy as np
import pytest

from ctlab_amd import dd
from ctlab_amd import learned as LE
from ctlab_amd import recon as RC

pytestmark = pytest.mark.gpu

SHAPE = (48, 48, 16)          # (YN, XN, ZN); training needs ZN % 8 == 0
POP = (96, 96, 32)            # ellipsoid population box the windows are cut from
VIEWS = 30


def test_trained_prox_gains_over_identity_and_safeguard_holds(cuda):
    import torch
    pairs = dict(shape=SHAPE, views=VIEWS, I0=1e4, dev=cuda, tv_lambda=0.3, kmin=2, kmax=20, population=POP)
    xs, ts, L, _ = LE.generate_pairs(30, 3, seed0=100, **pairs)
    hx, ht, _, _ = LE.generate_pairs(8, 3, seed0=900, **pairs)
    params, losses, _, steps = LE.train(xs, ts, SHAPE, 4, lr=1e-3)
    assert steps == 4 * len(xs) and np.all(np.isfinite(losses))
    id_loss = LE.identity_loss(hx, ht)
    net_loss = LE.network_loss(hx, ht, SHAPE, params)
    print(f"held-out prox loss: identity {id_loss:.4e}, trained network {net_loss:.4e} "
          f"({100 * (1 - net_loss / id_loss):.1f} % lower)")
    assert net_loss < id_loss

    g = LE.cone_geometry(SHAPE[1], SHAPE[0], SHAPE[2], VIEWS)
    plan = dd.Plan(g)
    rc = RC.Recon(plan, 1)
    ph = LE.make_problem(g, plan, rc, 4242, 1e4, cuda, POP)
    net = LE.ScaledCNN(g.vol_shape, params)
    f0 = rc.data_fidelity(torch.zeros(g.vol_shape, device=cuda))
    _, hist, _ = LE.fista_tv_run(rc, L, g.vol_shape, 15, 0.3, 10, ph, None, cuda, learned=net)
    prev = f0
    for h in hist:
        if h["learned"]:
            assert h["fidelity"] < prev, h
        prev = h["fidelity"]
    _, plain, _ = LE.fista_tv_run(rc, L, g.vol_shape, 15, 0.3, 10, ph, None, cuda)
    print("PSNR after 15 iterations: plain %.2f dB, learned %.2f dB (%d learned steps accepted)"
          % (plain[-1]["psnr"], hist[-1]["psnr"], sum(h["learned"] for h in hist)))

    # data-consistency stage: CG steps after the network lower ||Ax - b||^2 of
    # the network's output, and the safeguard still holds with it
    _, _, snaps = LE.fista_tv_run(rc, L, g.vol_shape, 8, 0.3, 10, ph, None, cuda, snaps=[8])
    x8 = snaps[8]
    raw = torch.empty_like(x8)
    net(x8, raw)
    dcs = LE.DataConsistent(net, rc, 3)
    out = torch.empty_like(x8)
    dcs(x8, out)
    f_raw, f_dc = rc.data_fidelity(raw), rc.data_fidelity(out)
    print(f"data consistency: ||Ax-b||^2 network output {f_raw:.4g}, after 3 CG steps {f_dc:.4g}")
    assert f_dc < f_raw
    _, hd, _ = LE.fista_tv_run(rc, L, g.vol_shape, 15, 0.3, 10, ph, None, cuda, learned=dcs, learned_every=3,
                               feedback=False)
    for a, b in zip(hd, plain):
        if a["learned"]:
            assert a["fidelity"] <= b["fidelity"], (a, b)     # output mode: against the plain iterate

    # adaptive stage: CG until the candidate's ||Ax - b||^2 is at the plain
    # iterate's.  It stops at the first step that gets there (no step when the
    # network's output already is), so the safeguard accepts its outputs, and
    # the single-shot form (one learned step after N plain iterations) gives the
    # plain iterate back with the same timing path when the step is rejected
    dca = LE.DataConsistent(net, rc, to_fidelity=True, m_max=10)
    f8 = rc.data_fidelity(x8)
    out2 = torch.empty_like(x8)
    dca(x8, out2, f_ref=f8)
    f_ad = rc.data_fidelity(out2)
    print(f"adaptive data consistency: {dca.last_steps} CG steps, ||Ax-b||^2 {f_ad:.6g} (plain iterate {f8:.6g})")
    assert 0 <= dca.last_steps <= 10
    if dca.last_steps < 10:
        assert f_ad <= f8 * (1 + 1e-6)
    if dca.last_steps > 0:
        f_less = torch.empty_like(x8)
        net(x8, f_less)
        rc.cg(f_less, dca.last_steps - 1)
        assert rc.data_fidelity(f_less) > f8 * (1 - 2e-5)     # one step fewer does not get there
    _, ha, _ = LE.fista_tv_run(rc, L, g.vol_shape, 15, 0.3, 10, ph, None, cuda, learned=dca, learned_every=5,
                               feedback=False)
    for a, b in zip(ha, plain):
        assert a["fidelity"] <= b["fidelity"] * (1 + 1e-6), (a, b)
        if a["dc_steps"] is not None:
            assert a["learned"] or a["rejected_candidate_fidelity"] is not None
    _, hs, _ = LE.fista_tv_run(rc, L, g.vol_shape, 8, 0.3, 10, ph, None, cuda, learned=dca, learned_every=8,
                               feedback=False)
    assert len(hs) == 8 and hs[-1]["dc_steps"] is not None and all(h["dc_steps"] is None for h in hs[:-1])
