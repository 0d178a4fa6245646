"""The solver plugin contract, with HIP-backed grid-search and DC-EGM kernels.

Restates the reference's contract (`src/_lcm/solution/contract.py:163-426,474-489,549-619`):
a `Solver` builds a regime's `SolutionKernels` from a `SolverBuildContext`;
`SolutionKernels.period_kernels[period]` is a `PeriodKernel` whose `__call__`
returns a `KernelResult(V_arr, continuation, simulation_policy)`, called once
per (regime, period) by the backward-induction loop
(`src/_lcm/solution/backward_induction.py:155-275`, here
`pylcm_amd.backward_induction.solve`).

What "core" and "compile" mean here. The reference's core is a jitted JAX
function that `_compile_all_functions` lowers and compiles ahead of time
(`backward_induction.py:577-720`). Here a core is a `HipCore`: a key naming
one (regime, period) of the model. "Compiling" (`compile_cores`) lowers the
model for the given params into the kernels' tables, packs them into one
device program (`GsProgram` or `EgmProgram`) and binds each core to it; a
compiled core is a callable that enqueues that (regime, period)'s HIP launches
through the C ABI and returns the device-resident V (`DeviceArray`).

Differences from the reference, by necessity:

* `flat_params` is the params mapping as passed to `Model.solve` (the lowering
  resolves the model / regime / function lookup itself) rather than the
  reference's flattened per-regime pools;
* `V_arr` is a `pylcm_amd._native.DeviceArray` (`np.asarray(V_arr)` copies it
  to the host) where the reference returns a JAX array;
* a DC-EGM kernel's `continuation` is an `EgmContinuation` that names the
  device rows of its carry (`EGMCarry`, `src/_lcm/egm/carry.py:25-60`).
"""

from __future__ import annotations

from collections.abc import Callable, Mapping
from dataclasses import dataclass, field, replace
from types import MappingProxyType
from typing import Any, Protocol, runtime_checkable

import numpy as np

from pylcm_amd import _native as nat


# ---------------------------------------------------------------------------
# contract types (contract.py:163-426,474-489)
# ---------------------------------------------------------------------------


@dataclass(frozen=True, kw_only=True)
class StateActionSpace:
    """A regime's state and action grids (`StateActionSpace` of the reference):
    name -> grid values, states in V-axis order (discrete and process states,
    then the continuous state), actions discrete first."""

    states: Mapping[str, np.ndarray]
    actions: Mapping[str, np.ndarray]

    @property
    def state_names(self) -> tuple[str, ...]:
        return tuple(self.states)

    @property
    def action_names(self) -> tuple[str, ...]:
        return tuple(self.actions)


@dataclass(frozen=True, kw_only=True)
class SolverBuildContext:
    """What a solver may read to build one regime's kernels (the fields of
    `contract.py:163-303` this build's solvers use, plus the model they are
    lowered from)."""

    regime_name: str
    user_regimes: Mapping[str, Any]
    state_action_space: StateActionSpace
    grids: Mapping[str, Any]
    functions: Mapping[str, Callable]
    constraints: Mapping[str, Callable]
    transitions: Any
    regimes_to_active_periods: Mapping[str, tuple[int, ...]]
    enable_jit: bool
    has_taste_shocks: bool
    model: Any = field(repr=False)


@dataclass(frozen=True, kw_only=True)
class KernelResult:
    """One regime-period solve output (`contract.py:306-326`)."""

    V_arr: Any
    continuation: Any = None
    simulation_policy: Any = None


@runtime_checkable
class PeriodKernel(Protocol):
    """One regime's per-period solve adapter (`contract.py:347-426`)."""

    def cores(self) -> Mapping[str, Callable]: ...

    @property
    def core(self) -> Callable: ...

    def with_fixed_params(self, *, fixed_flat_params) -> "PeriodKernel": ...

    def build_lower_args(self, *, core_key: str, state_action_space, next_regime_to_V_arr,
                         next_regime_to_continuation, flat_params, period: int, ages) -> Mapping[str, object]: ...

    def __call__(self, *, compiled_cores, state_action_space, next_regime_to_V_arr,
                 next_regime_to_continuation, flat_params, period: int, ages) -> KernelResult: ...


@dataclass(frozen=True, kw_only=True)
class SolutionKernels:
    """Per-period solve adapters produced by a solver (`contract.py:474-489`)."""

    period_kernels: Mapping[int, PeriodKernel]
    continuation_spec: Any = None


# ---------------------------------------------------------------------------
# HIP cores
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class HipCore:
    """The uncompiled core of one (regime, period): `kind` is "grid_search" or
    "dcegm"; `fixed` holds fixed params bound by `with_fixed_params`."""

    kind: str
    regime_name: str
    period: int
    fixed: tuple = ()


@dataclass(frozen=True)
class EgmContinuation:
    """A DC-EGM regime's carry at one period: device rows per column
    (grid, value, marginal; `n_pad` slots each) as the next period's parents
    read them, and `counts` (kept points, valid entries) per column."""

    program: Any
    period: int
    regime_name: str

    def rows(self) -> dict:
        """The carry rows copied to the host: arrays [n_cols, ...]."""
        return self.program.columns(self.period, self.regime_name)


class _CompiledGridSearch:
    """A grid-search core bound to a packed `GsProgram` (problem 0)."""

    def __init__(self, program, regime_name: str, period: int):
        self.program = program
        self.regime_name = regime_name
        self.period = period

    def _slot(self, t, r):
        o = self.program.outs[0][(t, r)]
        return self.program.base + o.v_off, o.shape

    def __call__(self, *, next_regime_to_V_arr) -> nat.DeviceArray:
        prog, t, r = self.program, self.period, self.regime_name
        pt = prog.problems[0][(t, r)]
        if not pt.terminal:
            names = [n for n, _ in pt.targets] + [n for n, _ in pt.scalars if n is not None]
            for name in names:
                _place_next_V(prog.ctx, next_regime_to_V_arr, name, *self._slot(t + 1, name), prog.dt)
        prog.launch_regime(0, t, r)
        ptr, shape = self._slot(t, r)
        return nat.DeviceArray(prog.ctx, ptr, shape, prog.dt, owner=prog)


class _CompiledDcegm:
    """A DC-EGM core bound to a packed `EgmProgram`."""

    def __init__(self, program, regime_name: str, period: int):
        self.program = program
        self.regime_name = regime_name
        self.period = period

    def __call__(self, *, next_regime_to_V_arr, next_regime_to_continuation=None):
        prog, t, r = self.program, self.period, self.regime_name
        if (t, r) in prog.host_V:  # terminal regimes publish closed forms
            return np.asarray(prog.host_V[(t, r)]), EgmContinuation(prog, t, r)
        for name, cont in (next_regime_to_continuation or {}).items():
            if isinstance(cont, EgmContinuation) and cont.program is not prog:
                raise ValueError(f"continuation of {name!r} belongs to another DC-EGM program; "
                                 "compile the regimes of one model together")
        prog.launch_regime(t, r)
        off, shape = prog.out_V[(t, r)]
        return (nat.DeviceArray(prog.ctx, prog.base + off, shape, prog.dt, owner=prog),
                EgmContinuation(prog, t, r))


def _place_next_V(ctx, next_regime_to_V_arr, name, ptr, shape, dtype):
    """Make the slot the descriptors read hold `next_regime_to_V_arr[name]`:
    nothing when it already is that slot, a device copy from another device
    array, an upload from a host array."""
    if next_regime_to_V_arr is None or name not in next_regime_to_V_arr:
        raise KeyError(f"next_regime_to_V_arr lacks regime {name!r}")
    v = next_regime_to_V_arr[name]
    if isinstance(v, nat.DeviceArray):
        if v.ptr == ptr:
            return
        n = int(np.prod(shape, dtype=np.int64)) if shape else 1
        if v.dtype != np.dtype(dtype) or v.size != n:
            raise ValueError(f"V of {name!r}: {v.shape} {v.dtype}, expected {tuple(shape)} {np.dtype(dtype)}")
        nat.d2d(ctx, ptr, v.ptr, v.nbytes)
        return
    arr = np.ascontiguousarray(np.asarray(v, dtype=dtype).reshape(shape))
    nat.h2d(ctx, ptr, arr)


def compile_cores(model, kernels: Mapping[str, SolutionKernels], flat_params, dtype) -> dict:
    """{(regime, period): {"main": compiled core}} for every period kernel:
    the model lowered for `flat_params` and packed into one device program
    (the reference's ahead-of-time compilation, `backward_induction.py:577-720`)."""
    kinds = {k.core.kind for sk in kernels.values() for k in sk.period_kernels.values()}
    fixed = _merged_fixed(kernels)
    params = _with_fixed(flat_params, fixed)
    out = {}
    if "dcegm" in kinds:
        from pylcm_amd._egm_program import EgmProgram

        prog = EgmProgram(model, params, dtype)
        for r, sk in kernels.items():
            for t in sk.period_kernels:
                out[(r, t)] = MappingProxyType({"main": _CompiledDcegm(prog, r, t)})
        return out
    from pylcm_amd._gs_program import GsProgram

    prog = GsProgram(model, [params], dtype)
    for r, sk in kernels.items():
        for t in sk.period_kernels:
            out[(r, t)] = MappingProxyType({"main": _CompiledGridSearch(prog, r, t)})
    return out


def _merged_fixed(kernels):
    fixed = {}
    for r, sk in kernels.items():
        for k in sk.period_kernels.values():
            for name, value in k.core.fixed:
                fixed.setdefault(r, {})[name] = value
    return fixed


def _with_fixed(params, fixed):
    if not fixed:
        return params
    out = dict(params or {})
    for r, vals in fixed.items():
        reg = dict(out.get(r, {}))
        reg.update(vals)
        out[r] = reg
    return out


# ---------------------------------------------------------------------------
# period kernels
# ---------------------------------------------------------------------------


@dataclass(frozen=True, kw_only=True)
class HipGridSearchPeriodKernel:
    """The grid-search period adapter (`grid_search.py:135-212`): V on the
    full state grid from one fused Q-and-max launch; no continuation, no
    simulation policy."""

    core: HipCore
    regime_name: str

    def cores(self) -> Mapping[str, Callable]:
        return MappingProxyType({"main": self.core})

    def with_fixed_params(self, *, fixed_flat_params) -> "HipGridSearchPeriodKernel":
        mine = dict((fixed_flat_params or {}).get(self.regime_name, {}))
        if not mine:
            return self
        return replace(self, core=replace(self.core, fixed=tuple(sorted(mine.items()))))

    def build_lower_args(self, *, core_key: str = "main", state_action_space, next_regime_to_V_arr,
                         next_regime_to_continuation, flat_params, period: int, ages) -> Mapping[str, object]:
        return {**dict(state_action_space.states), **dict(state_action_space.actions),
                "next_regime_to_V_arr": next_regime_to_V_arr, "flat_params": flat_params,
                "period": int(period), "age": ages.period_to_age(period)}

    def __call__(self, *, compiled_cores, state_action_space, next_regime_to_V_arr,
                 next_regime_to_continuation, flat_params, period: int, ages) -> KernelResult:
        out = compiled_cores["main"](next_regime_to_V_arr=next_regime_to_V_arr)
        if isinstance(out, tuple):  # a terminal regime under DC-EGM parents also publishes its carry
            return KernelResult(V_arr=out[0], continuation=out[1])
        return KernelResult(V_arr=out)


@dataclass(frozen=True, kw_only=True)
class HipDCEGMPeriodKernel:
    """The DC-EGM period adapter (`dcegm.py:477-535`): Euler inversion, FUES
    and publish for every column of the regime, V and the carry rows."""

    core: HipCore
    regime_name: str

    def cores(self) -> Mapping[str, Callable]:
        return MappingProxyType({"main": self.core})

    def with_fixed_params(self, *, fixed_flat_params) -> "HipDCEGMPeriodKernel":
        mine = dict((fixed_flat_params or {}).get(self.regime_name, {}))
        if not mine:
            return self
        return replace(self, core=replace(self.core, fixed=tuple(sorted(mine.items()))))

    def build_lower_args(self, *, core_key: str = "main", state_action_space, next_regime_to_V_arr,
                         next_regime_to_continuation, flat_params, period: int, ages) -> Mapping[str, object]:
        return {**dict(state_action_space.states), "next_regime_to_continuation": next_regime_to_continuation,
                "flat_params": flat_params, "period": int(period), "age": ages.period_to_age(period)}

    def __call__(self, *, compiled_cores, state_action_space, next_regime_to_V_arr,
                 next_regime_to_continuation, flat_params, period: int, ages) -> KernelResult:
        V, cont = compiled_cores["main"](next_regime_to_V_arr=next_regime_to_V_arr,
                                         next_regime_to_continuation=next_regime_to_continuation)
        return KernelResult(V_arr=V, continuation=cont)


# ---------------------------------------------------------------------------
# building a model's kernels
# ---------------------------------------------------------------------------


def state_action_space(model, regime_name, dtype=np.float64) -> StateActionSpace:
    """The regime's grids as arrays (process states at their default nodes
    count only: their values depend on params and are resolved at compile)."""
    lay = model.layouts[regime_name]
    states = {}
    for s, g in zip(lay.sd_names, lay.sd_grids):
        states[s] = np.arange(g.n_points)
    if lay.w_grid is not None:
        states[lay.w_name] = lay.w_grid.to_numpy(dtype)
    actions = {}
    for a, g in zip(lay.da_names, lay.da_grids):
        actions[a] = np.arange(g.n_points)
    if lay.c_grid is not None:
        actions[lay.c_name] = lay.c_grid.to_numpy(dtype)
    return StateActionSpace(states=MappingProxyType(states), actions=MappingProxyType(actions))


def build_context(model, regime_name: str) -> SolverBuildContext:
    lay = model.layouts[regime_name]
    reg = lay.regime
    active = {r: tuple(t for t in range(model.n_periods) if model.layouts[r].regime.active(
        model.ages.period_to_age(t))) for r in model.regime_names}
    grids = {**dict(zip(lay.sd_names, lay.sd_grids)), **dict(zip(lay.da_names, lay.da_grids))}
    if lay.w_grid is not None:
        grids[lay.w_name] = lay.w_grid
    if lay.c_grid is not None:
        grids[lay.c_name] = lay.c_grid
    return SolverBuildContext(
        regime_name=regime_name, user_regimes=model.user_regimes,
        state_action_space=state_action_space(model, regime_name), grids=MappingProxyType(grids),
        functions=MappingProxyType({k: v for k, v in reg.functions.items() if v is not None}),
        constraints=MappingProxyType({k: v for k, v in reg.constraints.items() if v is not None}),
        transitions=reg.transition, regimes_to_active_periods=MappingProxyType(active),
        enable_jit=model.enable_jit, has_taste_shocks=reg.taste_shocks is not None, model=model)


def grid_search_kernels(context: SolverBuildContext) -> SolutionKernels:
    r = context.regime_name
    return SolutionKernels(period_kernels=MappingProxyType({
        t: HipGridSearchPeriodKernel(core=HipCore("grid_search", r, t), regime_name=r)
        for t in context.regimes_to_active_periods[r]}))


def dcegm_kernels(context: SolverBuildContext) -> SolutionKernels:
    r = context.regime_name
    return SolutionKernels(period_kernels=MappingProxyType({
        t: HipDCEGMPeriodKernel(core=HipCore("dcegm", r, t), regime_name=r)
        for t in context.regimes_to_active_periods[r]}))
