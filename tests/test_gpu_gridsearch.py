"""GPU parity of the fused grid-search kernel against the FP64 oracle.

Bars: float64 V within 1e-10 relative, float32 V within 1e-5 relative of the
oracle (which solves the same float32-rounded grids in float64); policies
identical except flagged near-ties: oracle Q within one ulp of the dtype, or
up to 8 ulp where Q in the reference's own arithmetic at that precision
(from the GPU's V_{t+1}) also departs from the float64 argmax or agrees with
the GPU (`tests/parity.py`).
"""

import numpy as np
import pytest

from oracle import gridsearch as og
from pylcm_amd import examples as ex
from tests.parity import compare_solution, rel_err

pytestmark = pytest.mark.gpu

RTOL = {np.float64: 1e-10, np.float32: 1e-5}


def _check(model, params, dtype, taste_regimes=()):
    V, P = model.solve(params=params, log_level="off", dtype=dtype, return_policy=True)
    Vo, Po, Qo = og.solve(model, params, grid_dtype=dtype, return_Q=True)
    worst, mism, ties, split = compare_solution(V, P, Vo, Po, Qo, dtype, taste_regimes, model=model,
                                                params=params, detail=True)
    print(f"max rel V err {worst:.3e}, policy mismatches {mism}, flagged near-ties {ties} {split}")
    assert worst <= RTOL[dtype]
    assert mism == 0
    return V, P


def test_deterministic_workload_f64(native):
    """T=10, wealth 100 x consumption 200, float64 (full size)."""
    m, p = ex.deterministic_workload()
    _check(m, p, np.float64)


def test_stochastic_small_f32(native):
    m = ex.retirement_model(8, n_wealth=60, n_consumption=80, income_nodes=7)
    p = ex.retirement_params(8, income={"rho": 0.95, "sigma": 0.1, "mu": 0.0})
    _check(m, p, np.float32)


def test_stochastic_small_f64(native):
    m = ex.retirement_model(6, n_wealth=50, n_consumption=70, income_nodes=5)
    p = ex.retirement_params(6, income={"rho": 0.9, "sigma": 0.2, "mu": 0.1})
    _check(m, p, np.float64)


def test_log_utility_ragged_sizes_f32(native):
    """Grid sizes that are not multiples of the tile (ragged last tile)."""
    m = ex.retirement_model(5, n_wealth=67, n_consumption=129, log_utility=True)
    p = ex.retirement_params(5, log_utility=True, wage=20.0, interest_rate=0.0,
                             discount_factor=0.98, disutility_of_work=1.0)
    _check(m, p, np.float32)


def test_batch_matches_single_solves(native):
    """Different params in one launch give the same answers as one-at-a-time."""
    m = ex.retirement_model(6, n_wealth=40, n_consumption=50, income_nodes=5)
    plist = [ex.retirement_params(6, discount_factor=b, crra=c,
                                  income={"rho": r, "sigma": 0.1, "mu": 0.0})
             for b, c, r in [(0.9, 1.2, 0.8), (0.95, 2.0, 0.9), (0.97, 2.9, 0.95)]]
    batch = m.solve_batch(plist, dtype=np.float32)
    for p, Vb in zip(plist, batch):
        Vs = m.solve(params=p, log_level="off", dtype=np.float32)
        for t in Vs:
            for r in Vs[t]:
                assert np.array_equal(Vs[t][r], Vb[t][r])


def test_graph_replays_match_eager(native):
    """The horizon captured as a HIP graph gives the eager result on every replay."""
    from pylcm_amd._gs_program import GsProgram

    m = ex.retirement_model(6, n_wealth=40, n_consumption=50, income_nodes=3)
    p = ex.retirement_params(6, income={"rho": 0.9, "sigma": 0.1, "mu": 0.0})
    prog = GsProgram(m, [p], np.float32)
    prog.run(use_graph=False)
    ref = prog.fetch().copy()
    for _ in range(3):
        prog.out_host[:] = 0
        prog.run(use_graph=True)
        assert np.array_equal(prog.fetch(), ref)


# the kernel's other coordinate modes and variants, each against the oracle
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_log_spaced_wealth_grid(native, dtype):
    """Log-spaced next-state grid: the coordinate goes through log."""
    from tests import models as tm

    m = tm.log_grid_model(5)
    _check(m, ex.retirement_params(5, income={"rho": 0.8, "sigma": 0.1, "mu": 0.0}), dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_irregular_wealth_grid(native, dtype):
    """Irregular next-state grid: the coordinate is a bracket search."""
    from tests import models as tm

    m = tm.irregular_grid_model(5)
    _check(m, ex.retirement_params(5), dtype)


def test_state_dependent_tables_f64(native):
    """Utility and next wealth nonlinear in wealth: tables over the full state
    grid (the general kernel variant)."""
    from tests import models as tm

    m = tm.nonaffine_model(5)
    _check(m, ex.retirement_params(5, log_utility=True), np.float64)


def test_taste_shocks_solve_f64(native):
    """EV1 taste shocks on the discrete action: logsum value, per-cell argmax."""
    from tests import models as tm

    m = tm.taste_shock_model(5)
    p = ex.retirement_params(5, income={"rho": 0.8, "sigma": 0.1, "mu": 0.0})
    p["working_life"]["taste_shocks"] = {"scale": 0.3}
    _check(m, p, np.float64, taste_regimes=("working_life",))


@pytest.mark.parametrize("dist", ["lognormal", "normal", "rouwenhorst", "tauchen", "uniform"])
def test_processes_model_solve_f64(native, dist):
    """Markov health transitions and an income process."""
    from tests import models as tm

    m = tm.processes_model(5, dist)
    _check(m, tm.processes_params(dist), np.float64)
