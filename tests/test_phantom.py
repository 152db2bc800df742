"""Phantoms and their exact line integrals (CPU).

* The Shepp-Logan table and the chord length through a rotated ellipse against
  CTLAB's own Ellipsoid.h (src/Ellipsoid.h:209-317), compiled by
  `make -C oracle ref` into oracle/_ref/ellipse_probe; fixture
  tests/golden/ellipse_ref.json (tests/golden/make_ellipse_golden.py).
* Closed forms for spheres: a line through the centre gives 2r, a line at
  distance d gives 2*sqrt(r^2 - d^2), a tangent or missing line gives 0 (the
  tangent up to the square root of rounding).
"""
import json
import os

import numpy as np
import pytest

from conftest import ROOT
from ctlab_amd import phantom as PH

GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "ellipse_ref.json")))


def test_shepp_logan_matches_reference_table():
    ref = np.array(GOLD["shepp_logan"], dtype=np.float64)
    got = PH.shepp_logan(1.0)
    assert got.shape == (10, 8)
    np.testing.assert_allclose(got[:, [0, 1, 3, 4, 6, 7]], ref, rtol=1e-6, atol=1e-7)
    assert np.all(np.isinf(got[:, 5])) and np.all(got[:, 2] == 0.0)
    e = PH.shepp_logan(200.0, mu=0.02, height_mm=50.0)
    assert e[0, 3] == pytest.approx(69.0) and e[0, 5] == 25.0 and e[1, 7] == pytest.approx(-0.016)


def test_line_integral_matches_reference_intersection_length():
    ch = np.array(GOLD["chords"], dtype=np.float64)
    cx, cy, a, b, rot, sx, sy, ex, ey, ref = ch.T
    assert np.count_nonzero(ref) > 10 and np.count_nonzero(ref == 0) > 10
    for k in range(ch.shape[0]):
        ell = np.array([[cx[k], cy[k], 0.0, a[k], b[k], np.inf, rot[k], 1.0]])
        got = PH.line_integral(ell, np.array([sx[k], sy[k], 0.0]), np.array([ex[k], ey[k], 0.0]))
        # the reference computes in float
        assert abs(got - ref[k]) <= 5e-5 * max(1.0, ref[k]), (k, got, ref[k])


def test_line_integral_sphere_closed_forms():
    rng = np.random.default_rng(50)
    for _ in range(50):
        c = rng.uniform(-5, 5, 3)
        r = rng.uniform(0.5, 3.0)
        ell = np.array([[c[0], c[1], c[2], r, r, r, rng.uniform(0, np.pi), 1.0]])
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        n = np.cross(d, rng.standard_normal(3))
        n /= np.linalg.norm(n)                          # unit normal to the line
        for off, exp in ((0.0, 2 * r), (0.6 * r, 2 * np.sqrt(r * r - (0.6 * r) ** 2)), (r, 0.0), (1.5 * r, 0.0)):
            p0 = c + off * n - 7.0 * d
            got = PH.line_integral(ell, p0, p0 + 3.0 * d)
            # at a tangent the discriminant is 0 up to rounding, and the chord
            # its square root: about sqrt(machine epsilon) * r
            tol = 1e-6 * r if off == r else 1e-9
            assert abs(got - exp) < tol, (off, got, exp)


@pytest.mark.skipif(not os.path.exists("/root/reference/src/Ellipsoid.h"), reason="reference tree absent")
def test_ellipse_fixture_is_current():
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ellipse_probe")], check=True,
                         capture_output=True, text=True).stdout
    live = json.loads(out)
    assert live["shepp_logan"] == GOLD["shepp_logan"] and live["chords"] == GOLD["chords"]
