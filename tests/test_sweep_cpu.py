"""Host logic of the sweep driver: the grid, the logistic fit and the
bisection (GAGA/CGIHTcode/shared/tests_delta_fixed.m, calc_logit.m), with a
planted success model in place of the GPU."""
import math

import numpy as np

from gaga_amd import sweep


def test_default_grid_is_the_4096_instance_sweep():
    specs = sweep.grid_specs()
    assert len(specs) == 4096
    assert len({s["instance"] for s in specs}) == 4096
    for s in specs[::97]:
        assert s["m"] == math.ceil(s["delta"] * s["n"] - 1e-9)
        assert s["k"] == max(1, math.ceil(s["rho"] * s["m"] - 1e-9))
        assert 1 <= s["k"] <= s["m"] <= s["n"]
        assert (s["ens"] == "smv") == ("p" in s)
    gen = [s for s in specs if s["ens"] == "gen"]
    # dense footprint of the gen half: sum 4 m n
    total = sum(4.0 * s["m"] * s["n"] for s in gen)
    assert 85e9 < total < 95e9


def test_logit_model_closed_form():
    b = (30.0, 4.0)
    assert abs(sweep.logit_model(0.25, b) - 0.5) < 1e-12  # rho = 1/b1
    assert sweep.logit_model(0.0, b) > 0.99
    assert sweep.logit_model(0.5, b) < 0.01


def test_calc_logit_recovers_planted_transition():
    m = 400
    k = np.arange(20, 200, 12)
    b_true = (25.0, 1 / 0.22)
    frac = sweep.logit_model(k / m, b_true)
    b = sweep.calc_logit(k, np.full(len(k), m), frac, np.full(len(k), 10.0))
    assert abs(1 / b[1] - 0.22) < 0.01
    # a noisy step is fitted near its edge
    step = (k / m < 0.3).astype(float)
    b = sweep.calc_logit(k, np.full(len(k), m), step, np.full(len(k), 10.0))
    assert 0.26 < 1 / b[1] < 0.34


def test_bisection_with_planted_model():
    rng = np.random.default_rng(0)
    rho_star = {200: 0.3, 500: 0.4}
    calls = []

    def solve(specs):
        calls.append(len(specs))
        out = []
        for s in specs:
            p = sweep.logit_model(s["k"] / s["m"], (60.0, 1 / rho_star[s["m"]]))
            out.append(rng.random() < p)
        return np.array(out)

    res = sweep.transition_bisect("CGIHT", "gen", [200, 500], 1000, None, tests=10, solve=solve)
    for m, (kmin, kmax) in res.items():
        assert kmin < rho_star[m] * m < kmax
        assert kmax - kmin < 0.4 * m
    # both bisections run in the same batches
    assert calls[0] == 20
