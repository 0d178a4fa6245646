"""Regime specification (`src/lcm/regime.py:63-207`).

A regime declares states, actions, functions (must include `utility`),
constraints, state transitions and the regime transition; `transition=None`
marks a terminal regime. `ConsumptionSavingsRegime` adds the liquid margin
that the DC-EGM solver binds (`src/lcm/consumption_savings_regime.py:80,332`).
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from types import MappingProxyType

from pylcm_amd.exceptions import RegimeInitializationError
from pylcm_amd.solvers import GridSearch, Solver


@dataclass(frozen=True)
class MarkovTransition:
    """Wraps a function returning a probability vector over the next values."""

    func: object

    def __call__(self, *args, **kwargs):
        return self.func(*args, **kwargs)


@dataclass(frozen=True, kw_only=True)
class ExtremeValueTasteShocks:
    """EV1 taste shocks on the discrete-action combinations.

    The scale is the runtime parameter `params[regime]["taste_shocks"]["scale"]`
    (`src/lcm/taste_shocks.py:12`); the solve replaces the hard max over discrete
    actions by `scale * logsumexp(Qc / scale)`.
    """


def _always_active(_age):
    return True


def _freeze(m):
    return MappingProxyType(dict(m or {}))


@dataclass(frozen=True, kw_only=True)
class Regime:
    transition: object = None
    active: object = _always_active
    states: object = field(default_factory=dict)
    state_transitions: object = field(default_factory=dict)
    actions: object = field(default_factory=dict)
    functions: object = field(default_factory=dict)
    constraints: object = field(default_factory=dict)
    derived_categoricals: object = field(default_factory=dict)
    solver: Solver = field(default_factory=GridSearch)
    taste_shocks: ExtremeValueTasteShocks | None = None
    koopmans_aggregator: object = None
    certainty_equivalent: object = None
    description: str = ""

    def __post_init__(self):
        for name in ("states", "state_transitions", "actions", "functions",
                     "constraints", "derived_categoricals"):
            object.__setattr__(self, name, _freeze(getattr(self, name)))
        if "utility" not in self.functions:
            raise RegimeInitializationError("a regime's functions must include 'utility'")
        if self.taste_shocks is not None and not any(
            _is_discrete(g) for g in self.actions.values()
        ):
            raise RegimeInitializationError("taste shocks need at least one discrete action")

    @property
    def terminal(self) -> bool:
        return self.transition is None

    def replace(self, **kwargs) -> "Regime":
        return dataclasses.replace(self, **kwargs)


def _is_discrete(grid) -> bool:
    from pylcm_amd.grids import DiscreteGrid

    return isinstance(grid, DiscreteGrid)


@dataclass(frozen=True, kw_only=True)
class LiquidMargin:
    """Roles of the liquid consumption-savings margin for DC-EGM.

    `state` is the Euler (cash-on-hand) state, `action` the continuous action,
    `resources` the function (or state) giving resources R with `c = R - A`,
    `post_decision_state` the function giving end-of-period savings A.
    """

    state: str
    action: str
    resources: str
    post_decision_state: str


@dataclass(frozen=True, kw_only=True)
class ConsumptionSavingsRegime(Regime):
    liquid: LiquidMargin | None = None

    def __post_init__(self):
        super().__post_init__()
        if self.liquid is None and not self.terminal:
            raise RegimeInitializationError("ConsumptionSavingsRegime needs liquid=LiquidMargin(...)")
