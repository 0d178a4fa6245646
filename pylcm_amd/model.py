"""`Model`: specification, `solve` and `simulate` (`src/lcm/model.py:190-955`).

`Model(...)` validates the regimes and fixes their array layouts. `solve`
lowers each period to kernel tables on the host and runs backward induction on
the GPU through the HIP library; there is no CPU solve path. The result has the
reference's shape: an immutable mapping period -> regime -> V array, with V's
axes the discrete/process states followed by the continuous state.
"""

from __future__ import annotations

import logging
from types import MappingProxyType

import numpy as np

from pylcm_amd import config
from pylcm_amd._build import build_layout
from pylcm_amd.ages import AgeGrid
from pylcm_amd.exceptions import (
    InvalidValueFunctionError,
    ModelInitializationError,
)
from pylcm_amd.regime import Regime

_LOG = logging.getLogger("lcm")
_LOG_LEVELS = ("off", "warning", "progress", "debug")


def _freeze_params(d):
    return {k: (_freeze_params(v) if isinstance(v, dict) else v) for k, v in (d or {}).items()}


class Model:
    def __init__(self, *, ages: AgeGrid, regimes, regime_id_class, description: str = "",
                 enable_jit: bool = True, fixed_params=None, derived_categoricals=None,
                 functions=None, constraints=None, states=None, state_transitions=None,
                 actions=None, koopmans_aggregator=None, certainty_equivalent=None,
                 n_subjects: int | None = None):
        if not isinstance(ages, AgeGrid):
            raise ModelInitializationError("ages must be an AgeGrid")
        if not hasattr(regime_id_class, "__categories__"):
            raise ModelInitializationError("regime_id_class must be decorated with @categorical")
        names = tuple(regime_id_class.__categories__)
        if set(names) != set(regimes):
            raise ModelInitializationError(
                f"regime_id_class fields {sorted(names)} do not match regimes {sorted(regimes)}"
            )
        for n, r in regimes.items():
            if not isinstance(r, Regime):
                raise ModelInitializationError(f"regime {n!r} is not a Regime")
        for arg, v in (("certainty_equivalent", certainty_equivalent), ("koopmans_aggregator", koopmans_aggregator)):
            if v is not None and type(v).__name__ not in ("LinearExpectation", "LinearAggregator"):
                raise ModelInitializationError(f"{arg}: only the linear form is implemented")
        # model-level functions/constraints/states/actions merge into every regime
        merged = {}
        for n, r in regimes.items():
            kw = {}
            if functions:
                kw["functions"] = {**functions, **r.functions}
            if constraints:
                kw["constraints"] = {**constraints, **r.constraints}
            if states and not r.terminal:
                kw["states"] = {**states, **r.states}
            if state_transitions and not r.terminal:
                kw["state_transitions"] = {**state_transitions, **r.state_transitions}
            if actions and not r.terminal:
                kw["actions"] = {**actions, **r.actions}
            merged[n] = r.replace(**kw) if kw else r
        self.description = description
        self.ages = ages
        self.regime_id_class = regime_id_class
        self.regime_names = names
        self.regimes = MappingProxyType(merged)
        self.user_regimes = MappingProxyType(dict(regimes))
        self.fixed_params = _freeze_params(fixed_params)
        self.enable_jit = enable_jit
        self.n_subjects = n_subjects
        self.layouts = {n: build_layout(n, i, merged[n]) for i, n in enumerate(names)}
        self.regime_names_to_ids = MappingProxyType({n: i for i, n in enumerate(names)})
        self._programs = {}

    @property
    def n_periods(self) -> int:
        return self.ages.n_periods

    def regime_code(self, name: str) -> int:
        return self.regime_names_to_ids[name]

    def active_regimes(self, period: int):
        """Regimes active in `period` (evaluated once per model, then cached)."""
        cache = self.__dict__.setdefault("_active_cache", {})
        hit = cache.get(period)
        if hit is None:
            age = self.ages.period_to_age(period)
            hit = cache[period] = tuple(r for r in self.regime_names if self.layouts[r].regime.active(age))
        return list(hit)

    def is_active(self, regime: str, period: int) -> bool:
        return regime in self.active_regimes(period)

    # ------------------------------------------------------------------ solve
    def solve(self, *, params, log_level: str, max_compilation_workers=None,
              log_path=None, log_keep_n_latest: int = 3, return_simulation_policy: bool = False,
              dtype=None, return_policy: bool = False):
        """Backward induction on the GPU (`src/lcm/model.py:440-519`).

        Returns `period -> regime -> V`. `log_level` ("off", "warning",
        "progress", "debug") is required, as in the reference. With
        `return_simulation_policy=True` also returns the solvers' published
        simulation policies, `period -> regime -> {"grid", "policy"}`: the
        refined endogenous grid and consumption rows of each DC-EGM regime
        (shape `[discrete states..., discrete actions, n_pad]`); grid-search
        regimes publish none and have no entry. With `return_policy=True`
        (an extension) also returns `period -> regime -> policy`, the flat
        argmax index over the regime's action axes, discrete actions first.
        """
        if log_level not in _LOG_LEVELS:
            raise ValueError(f"log_level must be one of {_LOG_LEVELS}")
        from pylcm_amd._solve import solve_model

        dt = np.dtype(dtype) if dtype is not None else config.canonical_float_dtype()
        V, pol, sim_pol = solve_model(self, params, dt)
        try:
            self._check_nan(V, log_level)
        except InvalidValueFunctionError as exc:
            # src/lcm/model.py:552-563: snapshot of the partial solution
            if log_path is not None and exc.partial_solution is not None:
                from pylcm_amd.persistence import save_solve_snapshot

                snap = save_solve_snapshot(model=self, params=params,
                                           period_to_regime_to_V_arr=exc.partial_solution,
                                           log_path=log_path, log_keep_n_latest=log_keep_n_latest)
                exc.add_note(f"Snapshot saved to {snap}")
            raise
        Vm = MappingProxyType({t: MappingProxyType(V[t]) for t in sorted(V)})
        # src/lcm/model.py:565-577: every solve at debug, a NaN solve at
        # warning / progress
        if log_path is not None and log_level != "off" and (
                log_level == "debug" or any(np.isnan(a).any() for t in V for a in V[t].values())):
            from pylcm_amd.persistence import save_solve_snapshot

            save_solve_snapshot(model=self, params=params, period_to_regime_to_V_arr=Vm,
                                log_path=log_path, log_keep_n_latest=log_keep_n_latest)
        out = (Vm,)
        if return_simulation_policy:
            out += (MappingProxyType({t: MappingProxyType(sim_pol.get(t, {})) for t in sorted(V)}),)
        if return_policy:
            out += (MappingProxyType({t: MappingProxyType(pol[t]) for t in sorted(pol)}),)
        return out if len(out) > 1 else Vm

    def _check_nan(self, V, log_level):
        if log_level == "off":
            return
        for t in sorted(V):
            for r, arr in V[t].items():
                if np.any(np.isnan(arr)):
                    msg = f"value function of regime {r!r} at period {t} contains NaN"
                    if log_level == "debug":
                        partial = {tt: V[tt] for tt in V if tt > t}
                        raise InvalidValueFunctionError(msg, partial_solution=partial)
                    _LOG.warning(msg)

    def solve_batch(self, params_list, *, dtype=None, tile=None, use_graph=True, return_policy=False):
        """Solve many parameter sets in one fused launch per period.

        Returns a list of `period -> regime -> V` mappings, one per params set.
        """
        from pylcm_amd._solve import solve_batch

        dt = np.dtype(dtype) if dtype is not None else config.canonical_float_dtype()
        return solve_batch(self, list(params_list), dt, tile=tile, use_graph=use_graph,
                           return_policy=return_policy)

    # --------------------------------------------------------------- simulate
    def simulate(self, *, params, initial_conditions, period_to_regime_to_V_arr,
                 log_level: str, seed: int | None = None, subject_batch_size: int = 0,
                 log_path=None, log_keep_n_latest: int = 3, max_compilation_workers=None,
                 dtype=None):
        """Forward simulation on the GPU (`src/lcm/model.py:648-848`).

        `period_to_regime_to_V_arr` (None: solve first) and `log_level` are
        required, as in the reference."""
        if log_level not in _LOG_LEVELS:
            raise ValueError(f"log_level must be one of {_LOG_LEVELS}")
        from pylcm_amd._simulate import simulate_model

        dt = np.dtype(dtype) if dtype is not None else config.canonical_float_dtype()
        snapshot = log_path is not None and log_level == "debug"
        if snapshot and period_to_regime_to_V_arr is None and not self._any_dcegm():
            period_to_regime_to_V_arr = self.solve(params=params, log_level="off", dtype=dt)
        result = simulate_model(self, params=params, initial_conditions=initial_conditions,
                                V=period_to_regime_to_V_arr, seed=seed, dtype=dt,
                                subject_batch_size=subject_batch_size, log_level=log_level)
        if snapshot:
            # src/lcm/model.py:838-847
            from pylcm_amd.persistence import save_simulate_snapshot

            save_simulate_snapshot(model=self, params=params, initial_conditions=initial_conditions,
                                   period_to_regime_to_V_arr=period_to_regime_to_V_arr or {},
                                   result=result, log_path=log_path,
                                   log_keep_n_latest=log_keep_n_latest)
        return result

    def _any_dcegm(self) -> bool:
        return any(lay.is_dcegm for lay in self.layouts.values())

    def get_params_template(self):
        """Parameter names each regime's functions read, by function."""
        from pylcm_amd._dag import param_names_of

        out = {}
        for n, lay in self.layouts.items():
            variables = set(lay.state_names) | set(lay.action_names) | {"age", "period"}
            funcs = dict(lay.regime.functions)
            extra = list(lay.regime.constraints.items())
            extra += [(f"next_{s}", f) for s, f in lay.regime.state_transitions.items()
                      if f is not None and not hasattr(f, "items")]
            if lay.regime.transition is not None and not hasattr(lay.regime.transition, "items"):
                extra.append(("next_regime", lay.regime.transition))
            out[n] = param_names_of(regime_name=n, functions=funcs, variables=variables, extra=extra)
            if not lay.terminal:
                out[n].setdefault("koopmans_aggregator", []).append("discount_factor")
        return out
