"""Backward induction through the solver plugin contract.

The host loop of `src/_lcm/solution/backward_induction.py:44-302`: build each
regime's period kernels from its solver (`Solver.build_period_kernels`),
compile them (`pylcm_amd.contract.compile_cores`: lowering and one packed
device program), then for t = T-1 .. 0 call every active regime's period
kernel with the next period's value functions and continuations
(`_run_period_kernel`, `:327-388`), roll them (`_roll_continuation_inputs`,
`:391-440`), and with validation on stop at the first period whose V holds a
NaN (`:262-272`).

Every kernel call enqueues HIP launches on the program's stream; V stays on
the device (`DeviceArray`) between periods and is read to the host only for
the NaN check (log levels other than "off") and at the end.

`Model.solve` runs the same lowering as one fused launch per period, replayed
as a HIP graph; this driver launches per (regime, period) like the
reference's loop and gives the same bits (`tests/test_gpu_contract.py`).
"""

from __future__ import annotations

import logging
from dataclasses import dataclass
from types import MappingProxyType

import numpy as np

from pylcm_amd import config
from pylcm_amd.contract import KernelResult, SolutionKernels, build_context, compile_cores
from pylcm_amd.exceptions import InvalidValueFunctionError

_LOG = logging.getLogger("lcm")
_VALIDATING = ("warning", "progress", "debug")


@dataclass(frozen=True, kw_only=True)
class BackwardInductionResult:
    """The outputs of one backward-induction run (`contract.py:329-344`)."""

    value_functions: MappingProxyType
    simulation_policies: MappingProxyType


def build_solution_kernels(model) -> dict[str, SolutionKernels]:
    """{regime: SolutionKernels} from each regime's solver."""
    out = {}
    for r in model.regime_names:
        solver = model.layouts[r].regime.solver
        out[r] = solver.build_period_kernels(context=build_context(model, r))
    return out


def _run_period_kernel(*, kernel, compiled_cores, state_action_space, period, flat_params, ages,
                       next_regime_to_V_arr, next_regime_to_continuation) -> KernelResult:
    return kernel(compiled_cores=compiled_cores, state_action_space=state_action_space,
                  next_regime_to_V_arr=next_regime_to_V_arr,
                  next_regime_to_continuation=next_regime_to_continuation,
                  flat_params=flat_params, period=period, ages=ages)


def solve(model, *, flat_params, log_level: str = "off", dtype=None, fixed_flat_params=None,
          kernels: dict[str, SolutionKernels] | None = None) -> BackwardInductionResult:
    """Solve `model` by calling each regime's period kernel, period by period.

    Returns `value_functions` (period -> regime -> host V) and an empty
    `simulation_policies` (no shipped kernel publishes one).
    """
    dt = np.dtype(dtype) if dtype is not None else config.canonical_float_dtype()
    kernels = kernels if kernels is not None else build_solution_kernels(model)
    if fixed_flat_params:
        kernels = {r: SolutionKernels(
            period_kernels=MappingProxyType({t: k.with_fixed_params(fixed_flat_params=fixed_flat_params)
                                             for t, k in sk.period_kernels.items()}),
            continuation_spec=sk.continuation_spec) for r, sk in kernels.items()}
    compiled = compile_cores(model, kernels, flat_params, dt)
    ages = model.ages
    spaces = {r: build_context(model, r).state_action_space for r in model.regime_names}
    next_V: dict = {}
    next_cont: dict = {}
    solution: dict = {}
    validating = log_level in _VALIDATING
    for period in reversed(range(ages.n_periods)):
        period_solution, period_cont = {}, {}
        for r in model.regime_names:
            sk = kernels[r]
            if period not in sk.period_kernels:
                continue
            res = _run_period_kernel(kernel=sk.period_kernels[period], compiled_cores=compiled[(r, period)],
                                     state_action_space=spaces[r], period=period, flat_params=flat_params,
                                     ages=ages, next_regime_to_V_arr=next_V,
                                     next_regime_to_continuation=next_cont)
            period_solution[r] = res.V_arr
            if res.continuation is not None:
                period_cont[r] = res.continuation
        # roll: this period's outputs become the next call's inputs
        next_V = {**next_V, **period_solution}
        next_cont = {**next_cont, **period_cont}
        solution[period] = period_solution
        if validating:
            bad = [r for r, v in period_solution.items() if np.isnan(np.asarray(v)).any()]
            if bad:
                msg = f"value function of regime {bad[0]!r} at period {period} contains NaN"
                if log_level == "debug":
                    partial = {t: {r: np.asarray(v) for r, v in regs.items()}
                               for t, regs in solution.items() if t > period}
                    raise InvalidValueFunctionError(msg, partial_solution=partial)
                _LOG.warning(msg)
    V = {t: MappingProxyType({r: np.asarray(v) for r, v in solution[t].items()}) for t in sorted(solution)}
    return BackwardInductionResult(value_functions=MappingProxyType(V),
                                   simulation_policies=MappingProxyType({}))
