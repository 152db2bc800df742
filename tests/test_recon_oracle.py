"""CPU pins of the reconstruction helpers.

* SART weighting and FISTA momentum against CTLAB's own headers, run on seeded
  inputs by oracle/_ref/ref_probe (fixture tests/golden/functors_ref.json;
  tests/golden/make_functor_golden.py rebuilds it from the reference tree).
* CG beta rules against their closed forms and the exact-line-search identity
  (FR == PRP == HS on a quadratic with exact steps).
* The FGP TV prox against a brute-force minimisation of the prox objective.
* Convergence invariants of the reference drivers (oracle/recon_ref.py).
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import recon_ref as RR
from conftest import ROOT, rel_l2
from ctlab_amd import geometry as G

GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "functors_ref.json")))


def f32(a):
    return np.asarray(a, dtype=np.float32)


def test_sart_weighting_matches_reference_headers():
    got = O.sart_prj_weight(GOLD["prj_in"], GOLD["real"], GOLD["rowSum"])
    ref = f32(GOLD["prj_out"])
    assert np.array_equal(got == 0, ref == 0)          # epsilon guard 1e-7 on the same cells
    assert np.allclose(got, ref, rtol=2e-6, atol=1e-7)
    got = O.sart_bak_weight(GOLD["vol"], GOLD["recon_in"], GOLD["colSum"])
    assert np.allclose(got, f32(GOLD["recon_out"]), rtol=2e-6, atol=2e-7)


def test_fista_matches_reference_header():
    t1, t2 = GOLD["t1"], GOLD["t2"]
    got = O.fista(GOLD["cur_in"], GOLD["las"], t1, t2)
    assert np.allclose(got, f32(GOLD["cur_out"]), rtol=1e-6, atol=1e-3)
    got = O.fista(GOLD["cur_in"], GOLD["las"], t1, t2, -500.0, 500.0)
    assert np.allclose(got, f32(GOLD["cur_out_clamped"]), rtol=1e-6, atol=1e-3)
    assert np.all(np.abs(np.asarray(GOLD["cur_out_clamped"])) <= 500.0)


@pytest.mark.skipif(not os.path.exists("/root/reference/src/SARTWeighting.h"), reason="reference tree absent")
def test_golden_fixture_is_current():
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ref_probe")], check=True,
                         capture_output=True, text=True).stdout
    live = json.loads(out)
    for k in ("prj_out", "recon_out", "cur_out", "cur_out_clamped"):
        assert live[k] == GOLD[k]


def test_cg_beta_closed_forms():
    rng = np.random.default_rng(0)
    gk, gk1, d = rng.standard_normal((3, 200))
    exp = [gk1 @ gk1 / (gk @ gk), gk1 @ (gk1 - gk) / (gk @ gk), gk1 @ (gk1 - gk) / (d @ (gk1 - gk)),
           gk1 @ gk1 / (-d @ gk), gk1 @ gk1 / (d @ (gk1 - gk))]
    for c in range(5):
        assert abs(O.cg_beta(gk, gk1, d, c) - exp[c]) < 1e-12 * abs(exp[c])


def test_cg_beta_rules_agree_under_exact_line_search():
    """On a quadratic with exact line search, FR, PRP, HS, CD and DY coincide."""
    rng = np.random.default_rng(1)
    M = rng.standard_normal((30, 12))
    Q = M.T @ M
    bb = rng.standard_normal(12)
    x = np.zeros(12)
    g0 = Q @ x - bb
    d = -g0
    a = -(g0 @ d) / (d @ Q @ d)
    g1 = Q @ (x + a * d) - bb
    betas = [O.cg_beta(g0, g1, d, c) for c in range(5)]
    assert np.allclose(betas, betas[0], rtol=1e-9)


def test_tv_prox_solves_the_prox_problem():
    from scipy.optimize import minimize
    rng = np.random.default_rng(2)
    XN, YN, ZN = 5, 4, 3
    b = rng.random(XN * YN * ZN)
    lam = 0.05
    x = O.tv_prox(b, XN, YN, ZN, lam, 3000)

    def obj(v):
        return 0.5 * np.sum((v - b) ** 2) + lam * O.tv_value(v, XN, YN, ZN)
    best = minimize(obj, x, method="Powell", options={"maxiter": 20000, "xtol": 1e-10, "ftol": 1e-14})
    assert obj(x) <= best.fun + 1e-6
    assert obj(x) < obj(b) and obj(x) < obj(np.full_like(b, b.mean()))
    # zero lambda is the identity, and TV of a constant is zero
    assert np.allclose(O.tv_prox(b, XN, YN, ZN, 0.0, 5), b)
    assert O.tv_value(np.full(60, 3.0), XN, YN, ZN) == 0.0


def _small():
    g = G.circular_cone(XN=24, YN=24, ZN=6, dx=1.0, dz=1.0, PN=36, DNU=48, DNV=8, col_size=1.2,
                        row_size=1.2, sid=80.0, sdd=140.0, ang0=0.1)
    from ctlab_amd import phantom as PH
    ells = PH.random_ellipsoids(3, 4, g.XN, g.YN, g.ZN, g.dx, g.dz, semi_xy=(2, 8), semi_z=(1, 3))
    x = PH.voxelize(ells, g.XN, g.YN, g.ZN, g.dx, g.dz).reshape(-1).astype(np.float64)
    return g, x, O.proj(g, x)


def test_os_sart_decreases_residual_on_consistent_data():
    g, x, b = _small()
    prev = np.linalg.norm(b)
    xk = None
    for _ in range(3):
        xk = RR.os_sart(g, b, 4, 1, x0=xk)
        r = np.linalg.norm(O.proj(g, xk) - b)
        assert r < prev
        prev = r


def test_fista_tv_and_cg_reduce_the_objective():
    g, x, b = _small()
    L = RR.lipschitz(g) * 1.01
    _, obj = RR.fista_tv(g, b, 8, 1e-4, 10, L)
    assert obj[-1] < 0.05 * obj[0]
    _, res = RR.cg(g, b, 6, 0)
    assert res[-1] < res[0]
