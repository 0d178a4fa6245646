"""The solver plugin contract on the host: each regime's solver builds one
period kernel per active period; kernels satisfy the `PeriodKernel`
protocol, expose one "main" core, bind fixed params and build lowering args
(`src/_lcm/solution/contract.py:306-426,474-489,549-619`)."""

from types import MappingProxyType

import numpy as np

from pylcm_amd import backward_induction as bi
from pylcm_amd import examples as ex
from pylcm_amd.contract import PeriodKernel, SolutionKernels
from pylcm_amd.solvers import DCEGM, GridSearch


def test_grid_search_kernels_per_active_period():
    m = ex.retirement_model(6, n_wealth=20, n_consumption=30)
    ks = bi.build_solution_kernels(m)
    assert set(ks) == set(m.regime_names)
    for r, sk in ks.items():
        assert isinstance(sk, SolutionKernels)
        active = [t for t in range(6) if m.layouts[r].regime.active(m.ages.period_to_age(t))]
        assert sorted(sk.period_kernels) == active
        for t, k in sk.period_kernels.items():
            assert isinstance(k, PeriodKernel)
            core = k.cores()["main"]
            assert core is k.core and (core.kind, core.regime_name, core.period) == ("grid_search", r, t)
    assert not GridSearch().requires_continuation


def test_dcegm_kernels_and_continuation_flag():
    m = ex.iskhakov_dcegm_model(4)
    ks = bi.build_solution_kernels(m)
    assert ks["working_life"].period_kernels[0].core.kind == "dcegm"
    assert ks["dead"].period_kernels[3].core.kind == "grid_search"
    assert m.layouts["working_life"].regime.solver.requires_continuation
    assert isinstance(m.layouts["working_life"].regime.solver, DCEGM)


def test_fixed_params_bind_into_the_core_and_lower_args():
    m = ex.retirement_model(4, n_wealth=20, n_consumption=30)
    k = bi.build_solution_kernels(m)["working_life"].period_kernels[1]
    k2 = k.with_fixed_params(fixed_flat_params={"working_life": {"wage": 3.0}})
    assert k2.core.fixed == (("wage", 3.0),) and k.core.fixed == ()
    assert k.with_fixed_params(fixed_flat_params={"retirement": {"x": 1}}) is k
    space = bi.build_context(m, "working_life").state_action_space
    args = k.build_lower_args(core_key="main", state_action_space=space, next_regime_to_V_arr=MappingProxyType({}),
                              next_regime_to_continuation=MappingProxyType({}), flat_params={}, period=1, ages=m.ages)
    assert {"wealth", "labor_supply", "consumption", "next_regime_to_V_arr", "period", "age"} <= set(args)
    assert np.allclose(args["wealth"][[0, -1]], [1.0, 400.0]) and args["age"] == 1


def test_api_keywords_and_exception_hierarchy():
    """`log_level` is required by solve and simulate, `period_to_regime_to_V_arr`
    by simulate (`src/lcm/model.py:440-449,648-660`); a regime error is a
    model initialization error (`src/lcm/exceptions.py:59-70`)."""
    import pytest

    from pylcm_amd.exceptions import ModelInitializationError, RegimeInitializationError

    m, p = ex.deterministic_workload()
    with pytest.raises(TypeError):
        m.solve(params=p)
    with pytest.raises(TypeError):
        m.simulate(params=p, initial_conditions={}, log_level="off")
    with pytest.raises(ValueError):
        m.simulate(params=p, initial_conditions={}, period_to_regime_to_V_arr=None, log_level="loud")
    assert issubclass(RegimeInitializationError, ModelInitializationError)
