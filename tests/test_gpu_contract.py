"""The solver plugin contract on the device: backward induction driven
through `Solver.build_period_kernels` / `PeriodKernel.__call__`, one HIP
launch per (regime, period), gives `Model.solve`'s bits and the oracle's
values; a period kernel reads `next_regime_to_V_arr` from wherever it is
handed over (its own slot, another device buffer, a host array)."""

import numpy as np
import pytest

from oracle import gridsearch as og
from pylcm_amd import _native as nat
from pylcm_amd import backward_induction as bi
from pylcm_amd import examples as ex
from pylcm_amd.contract import KernelResult, compile_cores
from tests import models as tm
from tests.parity import rel_err

pytestmark = pytest.mark.gpu


def _same(Va, Vb):
    for t in Vb:
        for r in Vb[t]:
            assert np.array_equal(np.asarray(Va[t][r]), np.asarray(Vb[t][r]), equal_nan=True), (t, r)


@pytest.mark.parametrize("case", ["stochastic_f32", "mortality_f64", "processes_f64", "taste_f64"])
def test_contract_loop_equals_model_solve(native, case):
    if case == "stochastic_f32":
        m = ex.retirement_model(8, n_wealth=60, n_consumption=80, income_nodes=7)
        p, dt = ex.retirement_params(8, income={"rho": 0.95, "sigma": 0.1, "mu": 0.0}), np.float32
    elif case == "mortality_f64":
        m, p, dt = ex.mortality_model(5), ex.mortality_params(5), np.float64
    elif case == "processes_f64":
        m, p, dt = tm.processes_model(4, "tauchen"), tm.reference_processes_params("tauchen"), np.float64
    else:
        m = tm.taste_shock_model(5)
        p = ex.retirement_params(5, income={"rho": 0.8, "sigma": 0.1, "mu": 0.0})
        p["working_life"]["taste_shocks"] = {"scale": 0.3}
        dt = np.float64
    res = bi.solve(m, flat_params=p, log_level="debug", dtype=dt)
    _same(res.value_functions, m.solve(params=p, log_level="off", dtype=dt))
    Vo, _ = og.solve(m, p, grid_dtype=dt)
    worst = max(rel_err(res.value_functions[t][r], Vo[t][r]) for t in Vo for r in Vo[t])
    assert worst <= (1e-10 if dt == np.float64 else 1e-5)


def test_contract_loop_dcegm(native):
    m = ex.iskhakov_dcegm_model(5)
    p = ex.retirement_params(5, discount_factor=0.98, disutility_of_work=1.0, interest_rate=0.0, wage=20.0,
                             log_utility=True)
    res = bi.solve(m, flat_params=p, log_level="warning", dtype=np.float64)
    _same(res.value_functions, m.solve(params=p, log_level="off", dtype=np.float64))
    kernels = bi.build_solution_kernels(m)
    assert kernels["working_life"].period_kernels[0].cores()["main"].kind == "dcegm"


def test_period_kernel_reads_handed_over_V(native):
    """Period 0 of the working regime from V_1 handed over as a host array and
    as another device buffer: the same V as the loop's."""
    T = 6
    m = ex.retirement_model(T, n_wealth=50, n_consumption=70, income_nodes=5)
    p = ex.retirement_params(T, income={"rho": 0.9, "sigma": 0.15, "mu": 0.0})
    ref = m.solve(params=p, log_level="off", dtype=np.float64)
    kernels = bi.build_solution_kernels(m)
    compiled = compile_cores(m, kernels, p, np.float64)
    k = kernels["working_life"].period_kernels[0]
    space = bi.build_context(m, "working_life").state_action_space
    host_next = {r: np.asarray(v) for r, v in ref[1].items()}
    res = k(compiled_cores=compiled[("working_life", 0)], state_action_space=space,
            next_regime_to_V_arr=host_next, next_regime_to_continuation={}, flat_params=p, period=0, ages=m.ages)
    assert isinstance(res, KernelResult) and res.continuation is None
    assert isinstance(res.V_arr, nat.DeviceArray)
    assert np.array_equal(np.asarray(res.V_arr), ref[0]["working_life"])
    # the same V_1 from another device buffer
    ctx = compiled[("working_life", 0)]["main"].program.ctx
    bufs, dev_next = [], {}
    for r, v in host_next.items():
        a = np.ascontiguousarray(v).reshape(np.shape(v))
        b = nat.DeviceBuffer(max(a.nbytes, 8))
        nat.h2d(ctx, b.ptr, a)
        bufs.append(b)
        dev_next[r] = nat.DeviceArray(ctx, b.ptr, a.shape, a.dtype, owner=b)
    res2 = k(compiled_cores=compiled[("working_life", 0)], state_action_space=space,
             next_regime_to_V_arr=dev_next, next_regime_to_continuation={}, flat_params=p, period=0, ages=m.ages)
    assert np.array_equal(np.asarray(res2.V_arr), ref[0]["working_life"])


def test_contract_loop_nan_fail_fast(native):
    """log_level="debug" stops at the first NaN period with the partial solution."""
    from pylcm_amd.exceptions import InvalidValueFunctionError

    T = 4
    m = tm.mortality_mass_model(T)
    with pytest.raises(InvalidValueFunctionError) as info:
        bi.solve(m, flat_params=tm.mortality_mass_params(T, mass_scale=1.01), log_level="debug",
                 dtype=np.float64)
    assert set(info.value.partial_solution) == {T - 1}


def test_return_simulation_policy(native):
    """DC-EGM regimes publish their refined (grid, consumption) rows as the
    simulation policy; grid-search regimes publish none (`src/lcm/model.py:484-497`)."""
    m = ex.iskhakov_dcegm_model(4)
    p = ex.retirement_params(4, discount_factor=0.98, disutility_of_work=1.0, interest_rate=0.0, wage=20.0,
                             log_utility=True)
    V, pol = m.solve(params=p, log_level="off", dtype=np.float64, return_simulation_policy=True)
    assert set(pol) == set(V)
    assert set(pol[0]) == {"working_life", "retirement"} and not pol[3]
    g, c = pol[0]["working_life"]["grid"], pol[0]["working_life"]["policy"]
    assert g.shape == c.shape and g.shape[-2] == 2
    ok = np.isfinite(g) & np.isfinite(c)
    assert ok.any() and np.all(c[ok] > 0)
    gs = ex.retirement_model(4, n_wealth=20, n_consumption=30)
    V2, pol2 = gs.solve(params=ex.retirement_params(4), log_level="off", return_simulation_policy=True)
    assert set(pol2) == set(V2) and all(len(pol2[t]) == 0 for t in pol2)
