"""The HIP path against the reference's own golden vectors and against the
oracle on models with stochastic regime transitions.

* solution fixtures (`tests/golden/processes/*.npz`, `tests/golden/regression/solution.npz`,
  decoded from the reference's pickles): float64 solves within 1e-10 relative,
  float32 solves at the reference's decimal-5 bar;
* simulation fixtures (`tests/golden/regression/{f32,f64}/*_simulation.npz`): the
  value, choice and next wealth at every state the reference's simulation
  visited, in both precisions (`tests/reference_frames.py`);
* the mortality model (survival probabilities in (0, 1)) against the oracle at
  full size, the regime-mass rule (division by the mass, NaN off 1e-3 or with a
  negative probability), and simulated survival frequencies.
"""

import numpy as np
import pytest

from oracle import gridsearch as og
from pylcm_amd import examples as ex
from tests import models as tm
from tests import reference_frames as rf
from tests.parity import compare_solution, rel_err
from tests.test_golden_fixtures import DISTS, GOLDEN, load_golden, reference_decimal5

pytestmark = pytest.mark.gpu


def _solve(model, params, dtype):
    return model.solve(params=params, log_level="off", dtype=dtype)


@pytest.mark.parametrize("dist", DISTS)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_process_solutions(native, dist, dtype):
    exp = load_golden(GOLDEN / "processes" / f"solution_{dist}.npz")
    V = _solve(tm.processes_model(4, dist), tm.reference_processes_params(dist), dtype)
    reference_decimal5(exp, V)
    if dtype == np.float64:
        assert max(rel_err(V[t][r], exp[t][r]) for t in exp for r in exp[t]) <= 1e-10


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_regression_solution(native, dtype):
    exp = load_golden(GOLDEN / "regression" / "solution.npz")
    V = _solve(tm.regression_model(4), tm.regression_params(4), dtype)
    reference_decimal5(exp, V)
    if dtype == np.float64:
        assert max(rel_err(V[t][r], exp[t][r]) for t in exp for r in exp[t]) <= 1e-10


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("name", rf.FRAMES)
def test_reference_simulation_frames(native, name, prec):
    dtype = np.float32 if prec == "f32" else np.float64
    frame = rf.load(prec, name)
    model, params = rf.model_and_params(name)
    rows, ic = rf.visited_states(frame, model)
    V = _solve(model, params, dtype)
    df = model.simulate(params=params, initial_conditions=ic, period_to_regime_to_V_arr=V, log_level="off",
                        seed=5, dtype=dtype).to_dataframe(terminal_rows="all")
    start = np.array([model.ages.age_to_period(a) for a in ic["age"]])
    n = len(rows)
    first = {"value": np.empty(n), "consumption": np.empty(n), "labor_supply": np.empty(n, dtype=object)}
    second = {"wealth": np.full(n, np.nan)}
    term = rf.terminal_names(model)
    for k in range(n):
        mine = df[df["subject_id"] == k].set_index("period")
        row = mine.loc[start[k]]
        first["value"][k] = row["value"]
        first["consumption"][k] = row["consumption"]
        first["labor_supply"][k] = str(row["labor_supply"]) if "labor_supply" in mine else ""
        if start[k] + 1 in mine.index and str(mine.loc[start[k] + 1, "regime_name"]) not in term:
            second["wealth"][k] = mine.loc[start[k] + 1, "wealth"]
    n_rows, n_next = rf.compare(frame, model, rows, first, second)
    print(f"{name} {prec}: {n_rows} visited states, {n_next} next-wealth checks")
    assert n_rows >= 11 and n_next >= 4


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_mortality_model_full_size(native, dtype):
    """Survival probabilities in (0, 1): the continuation blends the alive
    regimes with the dead one (wealth 100 x consumption 500, T=4 and T=8)."""
    for T in (4, 8):
        m, p = ex.mortality_model(T), ex.mortality_params(T)
        V, P = m.solve(params=p, log_level="off", dtype=dtype, return_policy=True)
        Vo, Po, Qo = og.solve(m, p, grid_dtype=dtype, return_Q=True)
        worst, mism, ties = compare_solution(V, P, Vo, Po, Qo, dtype, model=m, params=p)
        print(f"T={T} {np.dtype(dtype)}: max rel V err {worst:.3e}, mismatches {mism}, near-ties {ties}")
        assert worst <= (1e-10 if dtype == np.float64 else 1e-5)
        assert mism == 0


@pytest.mark.parametrize("mass_scale,shift", [(1.0008, 0.0), (1.0012, 0.0), (0.998, 0.0), (1.0, 0.05)])
def test_regime_mass_rule(native, mass_scale, shift):
    """Within 1e-3 of one the blend divides by the mass; beyond it, or with a
    negative probability, the values that read the blend are NaN — the same
    NaN pattern as the oracle's, and the same values elsewhere."""
    T = 4
    m = tm.mortality_mass_model(T)
    p = tm.mortality_mass_params(T, mass_scale=mass_scale, shift=shift)
    V, P = m.solve(params=p, log_level="off", dtype=np.float64, return_policy=True)
    Vo, Po = og.solve(m, p)
    n_nan = 0
    for t in Vo:
        for r in Vo[t]:
            a, b = np.asarray(V[t][r]), np.asarray(Vo[t][r])
            assert np.array_equal(np.isnan(a), np.isnan(b)), (t, r)
            n_nan += int(np.isnan(b).sum())
            assert rel_err(a, b) <= 1e-10
    if mass_scale == 1.0008:
        assert n_nan == 0
        V1 = m.solve(params=tm.mortality_mass_params(T), log_level="off", dtype=np.float64)
        assert max(rel_err(V[t][r], V1[t][r]) for t in V1 for r in V1[t]) <= 1e-13
    else:
        assert n_nan > 0


def test_simulated_survival_frequencies(native):
    """Regime draws from probability rows: the share of agents alive after
    one and two periods matches the survival probabilities (5 sigma)."""
    T = 5
    m, p = ex.mortality_model(T, n_wealth=60, n_consumption=120), ex.mortality_params(T)
    n = 200_000
    ic = {"age": np.full(n, 40.0), "wealth": np.full(n, 100.0), "regime_id": np.zeros(n, dtype=np.int32)}
    df = m.simulate(params=p, initial_conditions=ic, period_to_regime_to_V_arr=None, log_level="off", seed=11,
                    dtype=np.float64).to_dataframe(terminal_rows="all")
    sp = ex.default_survival_probs(T)
    alive = ~df["regime_name"].astype(str).isin(["dead"])
    expected = 1.0
    for t in (1, 2, 3):
        expected *= sp[t - 1]
        share = float(alive[df["period"] == t].mean())
        sigma = np.sqrt(expected * (1 - expected) / n)
        print(f"period {t}: alive share {share:.5f}, expected {expected:.5f} (sigma {sigma:.1e})")
        assert abs(share - expected) <= 5 * sigma + 1e-12
