"""Host logic of bench.py that the reported figures depend on (CPU only): the
CPU-baseline view sample and the learned single-shot report (earliest N whose
output reaches a PSNR; the same with a data residual no higher than plain
FISTA-TV's best iterate; ratios against plain's own times)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def test_sample_views_spread_over_the_rotation():
    v = bench.sample_views(984, 240)
    assert v.size == 240 and v[0] == 0 and np.all(np.diff(v) > 0) and v[-1] < 984
    assert np.all(np.isin(np.diff(v), (4, 5)))          # evenly spread
    assert np.array_equal(bench.sample_views(10, 240), np.arange(10))   # capped at PN


def _hist(psnrs, fids, dt=0.01):
    return [{"iter": k + 1, "t": dt * (k + 1), "psnr": p, "fidelity": f} for k, (p, f) in
            enumerate(zip(psnrs, fids))]


def test_single_shot_report_picks_earliest_qualifying_N():
    # plain FISTA-TV: PSNR rises to 39 dB at iteration 10 (t = 0.10 s), best fidelity 100 there
    h_plain = _hist([30, 32, 34, 35, 36, 37, 38, 38.5, 38.8, 39.0], [900, 500, 300, 200, 150, 130, 120, 110, 105, 100])
    single = {2: {"t_s": 0.03, "psnr": 37.0, "fidelity": 400},
              4: {"t_s": 0.05, "psnr": 39.5, "fidelity": 180},      # passes plain's best PSNR, not its residual
              6: {"t_s": 0.07, "psnr": 40.2, "fidelity": 95},       # 40 dB and residual <= 100
              8: {"t_s": 0.09, "psnr": 40.5, "fidelity": 90}}
    r = bench.single_shot_report(single, h_plain, 40.0)
    assert r["N_to_plain_best"] == 4 and r["t_to_plain_best_s"] == 0.05
    assert r["speedup_to_plain_best"] == pytest.approx(0.10 / 0.05)
    assert r["N_to_40dB"] == 6 and r["speedup_40dB_vs_plain"] is None      # plain never reaches 40 dB
    assert r["N_to_plain_best_at_its_residual"] == 6
    assert r["speedup_to_plain_best_at_its_residual"] == pytest.approx(0.10 / 0.07)


def test_single_shot_report_when_nothing_qualifies():
    h_plain = _hist([30, 41], [200, 100])
    r = bench.single_shot_report({2: {"t_s": 0.05, "psnr": 35.0, "fidelity": 150}}, h_plain, 40.0)
    assert r["t_to_40dB_s"] is None and r["speedup_to_plain_best"] is None
    assert r["t_to_plain_best_at_its_residual_s"] is None
