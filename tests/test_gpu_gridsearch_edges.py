"""GPU parity of the grid-search kernel on inputs that leave its folded loop.

The fast variant runs a folded action loop only where magnitudes rule out
overflow and NaN (|next-state coordinate| <= 2^22 over the block's states,
|EV| and its slopes <= 1e30, U + beta * cs not NaN, EV row finite); other
chunks and rows take the staged loop. Each case below first checks on the
host that its inputs reach the branch it is named for, then compares the GPU
solve with the FP64 oracle at the usual bars.
"""

import numpy as np
import pytest

from pylcm_amd import examples as ex
from pylcm_amd._gs_program import lower_problem
from tests import models as tm
from oracle import gridsearch as og
from tests.parity import compare_solution
from tests.test_gpu_gridsearch import RTOL, _check

pytestmark = pytest.mark.gpu


def _coordinate_bounds(model, params, regime, period=1):
    """max |A' w| + |B'| per action with the regime's largest |w| (host copy
    of the kernel's per-chunk bound, here over the whole grid)."""
    tabs, rg = lower_problem(model, params, np.float64)
    pt = tabs[(period, regime)]
    g = rg.w_pts[pt.next_grid]
    step = (g[-1] - g[0]) / (len(g) - 1)
    w_abs = np.abs(rg.w_pts[regime]).max()
    return np.abs(pt.A / step) * w_abs + np.abs((pt.B - g[0]) / step)


def test_huge_next_state_coordinates(native):
    """Interest rate 4999: next wealth lands up to ~6e6 grid steps beyond the
    grid, so chunks with large wealth or consumption exceed the folded loop's
    2^22 coordinate bound and run staged while the rest run folded. Float64
    only: a float32 coordinate of that size carries a rounding error of a
    quarter grid step or more in any float32 evaluation, far outside the
    float32 bar against the FP64 oracle."""
    m = ex.retirement_model(4, n_wealth=600, n_consumption=600)
    p = ex.retirement_params(4, interest_rate=4999.0)
    b = _coordinate_bounds(m, p, "working_life")
    assert b.min() < 2**22 < b.max()
    _check(m, p, np.float64)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_nan_utility_on_infeasible_actions(native, dtype):
    """U is NaN for consumption above 420 (never affordable): those chunks run
    staged, where the NaN is masked like the reference's where(F, Q, -inf)."""
    m = tm.nan_infeasible_utility_model(4)
    p = ex.retirement_params(4, log_utility=True)
    tabs, _ = lower_problem(m, p, dtype)
    assert np.isnan(tabs[(1, "working_life")].U).any()
    V, _ = _check(m, p, dtype)
    assert all(np.isfinite(V[0][r]).all() for r in V[0])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_unreachable_states_make_ev_rows_infinite(native, dtype):
    """Wealth 0.5 affords no consumption (>= 1): V = -inf there, so the EV rows
    are not finite and every chunk runs the staged zero-safe loop."""
    m = ex.retirement_model(5, n_wealth=50, n_consumption=60, wealth_range=(0.5, 400.0), income_nodes=3)
    p = ex.retirement_params(5, income={"rho": 0.9, "sigma": 0.1, "mu": 0.0})
    V, _ = _check(m, p, dtype)
    assert np.isneginf(V[1]["working_life"]).any()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_value_magnitudes_above_folded_bound(native, dtype):
    """Utility scaled by 1e31: |EV| > 1e30, so rows run staged; with scale 1e20
    the same model runs folded. V crosses zero (log 1 = 0), so the error is
    measured relative to max(|V|, scale); policies as usual."""
    p = ex.retirement_params(5, log_utility=True, income={"rho": 0.9, "sigma": 0.1, "mu": 0.0})
    for scale in (1e31, 1e20):
        m = tm.scaled_utility_model(5, scale)
        V, P = m.solve(params=p, log_level="off", dtype=dtype, return_policy=True)
        Vo, Po, Qo = og.solve(m, p, grid_dtype=dtype, return_Q=True)
        _, mism, _ = compare_solution(V, P, Vo, Po, Qo, dtype)
        worst = max(float(np.max(np.abs(np.asarray(V[t][r], np.float64) - Vo[t][r])
                                 / np.maximum(np.abs(Vo[t][r]), scale)))
                    for t in Vo for r in Vo[t])
        print(f"scale {scale:g}: max error / max(|V|, scale) {worst:.3e}, policy mismatches {mism}")
        assert worst <= RTOL[dtype]
        assert mism == 0
        if scale > 1e30:
            assert max(np.abs(V[1][r]).max() for r in ("working_life", "retirement")) > 1e30
