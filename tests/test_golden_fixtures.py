"""The oracle against the reference's own pickled solutions.

`tests/golden/processes/solution_{dist}.npz` and `tests/golden/regression/solution.npz`
hold the reference's value functions (float64) for its processes test model
(`tests/test_processes.py:33-70`) and its regression model
(`tests/test_regression_test.py:29-66`), decoded from the pickles by
`tests/golden/extract_reference_pickles.py`. The reference checks its own
solve against them at decimal 5; the oracle reproduces them to rounding, which
pins its Tauchen, Gauss-Hermite normal / log-normal, uniform and Rouwenhorst
tables, the Markov transition of a discrete state, off-grid extrapolation and
the age-dependent wage.
"""

from pathlib import Path

import numpy as np
import pytest

from oracle import gridsearch as og
from tests import models as tm

GOLDEN = Path(__file__).resolve().parent / "golden"
DISTS = ["uniform", "normal", "lognormal", "tauchen", "rouwenhorst"]


def load_golden(path):
    z = np.load(path)
    out = {}
    for k in z.files:
        t, r = k.split("/")
        out.setdefault(int(t), {})[r] = z[k]
    return out


def reference_decimal5(expected, got):
    """`numpy.testing.assert_array_almost_equal(..., decimal=5)`, the
    reference's bar: |expected - got| < 1.5e-5 elementwise."""
    for t in expected:
        for r in expected[t]:
            np.testing.assert_array_almost_equal(expected[t][r], got[t][r], decimal=5)


def max_abs_diff(expected, got):
    return max(float(np.max(np.abs(np.asarray(expected[t][r]) - np.asarray(got[t][r]))))
               for t in expected for r in expected[t])


@pytest.mark.parametrize("dist", DISTS)
def test_oracle_reproduces_reference_process_solutions(dist):
    exp = load_golden(GOLDEN / "processes" / f"solution_{dist}.npz")
    V, _ = og.solve(tm.processes_model(4, dist), tm.reference_processes_params(dist))
    reference_decimal5(exp, V)
    assert max_abs_diff(exp, V) < 1e-12


def test_oracle_reproduces_reference_regression_solution():
    exp = load_golden(GOLDEN / "regression" / "solution.npz")
    V, _ = og.solve(tm.regression_model(4), tm.regression_params(4))
    reference_decimal5(exp, V)
    assert max_abs_diff(exp, V) < 1e-12
    # the last decision period reduces to max over c <= w of log(c) (dead next)
    assert np.isclose(exp[2]["working_life"].max(), np.log(400.0))


def test_fixture_shapes_follow_the_value_function_layout():
    """V axes: discrete / process states in declaration order, then wealth."""
    exp = load_golden(GOLDEN / "processes" / "solution_tauchen.npz")
    m = tm.processes_model(4, "tauchen")
    assert exp[0]["alive"].shape == m.layouts["alive"].V_shape == (5, 2, 5)
    assert exp[3]["dead"].shape == ()


def test_fixture_sensitivity_to_the_process_table():
    """A plausible slip in a process table (Tauchen's matrix replaced by
    Rouwenhorst's with the same parameters) moves V far outside the bar."""
    exp = load_golden(GOLDEN / "processes" / "solution_tauchen.npz")
    V, _ = og.solve(tm.processes_model(4, "rouwenhorst"), tm.reference_processes_params("rouwenhorst"))
    assert max_abs_diff(exp, V) > 1e-2


# ------------------------------------------------ the reference's simulations


def _oracle_records(model, params, ic, prec):
    """The oracle's simulation of the visited states: records of each agent's
    first period and the one after (as `tests.reference_frames.compare` reads them)."""
    from oracle import simulate as osim
    from pylcm_amd import DiscreteGrid

    dt = np.float32 if prec == "f32" else np.float64
    V, _ = og.solve(model, params, grid_dtype=dt)
    out = osim.simulate(model, params, V, ic, seed=3, grid_dtype=dt)
    names = list(model.regime_names)
    n = len(ic["regime_id"])
    start = np.array([model.ages.age_to_period(a) for a in ic["age"]])
    first = {"value": np.empty(n), "consumption": np.empty(n), "labor_supply": np.empty(n, dtype=object)}
    second = {"wealth": np.full(n, np.nan)}
    for k in range(n):
        t = start[k]
        lay = model.layouts[names[out["regime"][t, k]]]
        a = out["action"][t, k]
        cpts = og.grid_points(lay.c_grid, dt)
        first["value"][k] = out["value"][t, k]
        first["consumption"][k] = cpts[a % lay.n_c]
        labels = [g.categories for g in lay.da_grids if isinstance(g, DiscreteGrid)]
        first["labor_supply"][k] = labels[0][a // lay.n_c] if labels else ""
        if t + 1 < model.n_periods and not model.layouts[names[out["regime"][t + 1, k]]].terminal:
            second["wealth"][k] = out["w"][t + 1, k]
    return first, second


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("name", ["mortality_simulation", "precautionary_savings_simulation"])
def test_oracle_simulation_reproduces_reference_frames(name, prec):
    """The oracle's value, choice and next wealth at every state the
    reference's simulation visited (its f64 and f32 runs)."""
    from tests import reference_frames as rf

    frame = rf.load(prec, name)
    model, params = rf.model_and_params(name)
    rows, ic = rf.visited_states(frame, model)
    first, second = _oracle_records(model, params, ic, prec)
    n_rows, n_next = rf.compare(frame, model, rows, first, second)
    assert n_rows >= 11 and n_next >= 4


# ------------------------------- stochastic regime transitions: the mass rule


def _finite_periods(V):
    return {t: {r: bool(np.all(np.isfinite(v))) for r, v in V[t].items()} for t in V}


def test_oracle_divides_by_the_represented_regime_mass():
    """Probabilities scaled by a mass within 1e-3 of one give the unscaled
    solution (the blend divides by the mass); beyond 1e-3, or with a negative
    probability, every value that reads the blend is NaN."""
    T = 4
    m = tm.mortality_mass_model(T, n_wealth=12, n_consumption=15)
    V1, P1 = og.solve(m, tm.mortality_mass_params(T))
    Vs, Ps = og.solve(m, tm.mortality_mass_params(T, mass_scale=1.0008))
    for t in V1:
        for r in V1[t]:
            np.testing.assert_allclose(Vs[t][r], V1[t][r], rtol=1e-13, atol=0)
            assert np.array_equal(Ps[t][r], P1[t][r])
    for params in (tm.mortality_mass_params(T, mass_scale=1.0012),
                   tm.mortality_mass_params(T, mass_scale=0.998),
                   tm.mortality_mass_params(T, shift=0.05)):  # P(retired) = -0.05 after work
        Vb, _ = og.solve(m, params)
        shifted = params["working_life"]["next_regime"]["shift"] != 0
        for t in range(T - 1):
            for r in ("working_life", "retirement"):
                if shifted and (r == "retirement" or t == T - 2):
                    # the negative weight sits on the working regime's
                    # transition, in periods with a survival chance only
                    if r == "retirement":
                        np.testing.assert_allclose(Vb[t][r], V1[t][r], rtol=1e-13)
                    continue
                assert np.all(np.isnan(Vb[t][r])), (t, r)


def test_mortality_blend_matches_a_hand_computation():
    """One period of the mortality model by hand with plain loops: V at the
    last decision period is max log c (death next), and one period earlier
    each action's value is u + beta * (sp * V_next(w')) (the dead regime is
    worth zero), with V_next interpolated linearly on the wealth grid."""
    from pylcm_amd import examples as ex

    T = 3
    m = ex.mortality_model(T, n_wealth=9, n_consumption=11)
    p = ex.mortality_params(T)
    V, _ = og.solve(m, p)
    w = np.linspace(1.0, 400.0, 9)
    c = np.linspace(1.0, 400.0, 11)
    beta, r, wage, sp = 0.95, 0.05, 10.0, 0.98
    # last decision period (survival 0): retire-or-work, dead next
    V1 = np.array([max(np.log(ci) for ci in c if ci <= wi) for wi in w])
    np.testing.assert_allclose(V[1]["retirement"], V1, rtol=1e-14)
    np.testing.assert_allclose(V[1]["working_life"], V1, rtol=1e-14)

    def interp(x):
        k = min(max(int(np.floor((x - 1.0) / (w[1] - w[0]))), 0), len(w) - 2)
        t = (x - w[k]) / (w[1] - w[0])
        return (1 - t) * V1[k] + t * V1[k + 1]

    exp_work, exp_ret = [], []
    for wi in w:
        best_w, best_r = -np.inf, -np.inf
        for ci in c:
            if ci > wi:
                continue
            for work in (True, False):
                nxt = (1 + r) * (wi - ci) + (wage if work else 0.0)
                q = np.log(ci) - (0.5 if work else 0.0) + beta * sp * interp(nxt)
                best_w = max(best_w, q)
            best_r = max(best_r, np.log(ci) + beta * sp * interp((1 + r) * (wi - ci)))
        exp_work.append(best_w)
        exp_ret.append(best_r)
    np.testing.assert_allclose(V[0]["working_life"], exp_work, rtol=1e-13)
    np.testing.assert_allclose(V[0]["retirement"], exp_ret, rtol=1e-13)
