"""Example models, written against the public API like user code.

* `retirement_model` — the consumption-savings / retirement model family of the
  reference's `lcm_examples/iskhakov_et_al_2017.py:57-174` (working life with a
  work/retire choice, absorbing retirement, deterministic death), with CRRA or
  log utility and an optional Rouwenhorst AR(1) log-income state as in
  `lcm_examples/precautionary_savings.py:53-59`. The benchmark workloads are
  instances of it:

  - deterministic, float64, T=10, wealth 100 x consumption 200;
  - stochastic, float32, T=40, wealth 500 x consumption 500, 7 income nodes;
  - batched draws of the stochastic model with grid sizes in {200, 500, 1000}.

* `dcegm_retirement_model` — the same economics for the DC-EGM solver with EV1
  taste shocks and a log-normal income shock on Gauss-Hermite nodes.
"""

from __future__ import annotations

import numpy as np

from pylcm_amd import (
    AgeGrid,
    DCEGM,
    ConsumptionSavingsRegime,
    DiscreteGrid,
    ExtremeValueTasteShocks,
    FUESEnvelope,
    IrregSpacedGrid,
    LinSpacedGrid,
    LiquidMargin,
    LogNormalIIDProcess,
    Model,
    Regime,
    RouwenhorstAR1Process,
    categorical,
)


@categorical(ordered=True)
class LaborSupply:
    work: int
    retire: int


@categorical(ordered=False)
class RegimeId:
    working_life: int
    retirement: int
    dead: int


def is_working(labor_supply):
    return labor_supply == LaborSupply.work


def labor_income(is_working, wage):
    return np.where(is_working, wage, 0.0)


def crra_utility(consumption, crra):
    return consumption ** (1.0 - crra) / (1.0 - crra)


def utility_working(crra_utility, is_working, disutility_of_work):
    return crra_utility - np.where(is_working, disutility_of_work, 0.0)


def utility_retirement(crra_utility):
    return crra_utility


def log_utility_working(consumption, is_working, disutility_of_work):
    return np.log(consumption) - np.where(is_working, disutility_of_work, 0.0)


def log_utility_retirement(consumption):
    return np.log(consumption)


def next_wealth(wealth, consumption, labor_income, interest_rate):
    return (1 + interest_rate) * (wealth - consumption) + labor_income


def next_wealth_with_income(wealth, consumption, labor_income, income, interest_rate):
    return (1 + interest_rate) * (wealth - consumption) + np.exp(income) + labor_income


def borrowing_constraint(consumption, wealth):
    return consumption <= wealth


def next_regime_from_working(labor_supply, age, final_age_alive):
    return np.where(
        age >= final_age_alive,
        RegimeId.dead,
        np.where(labor_supply == LaborSupply.retire, RegimeId.retirement, RegimeId.working_life),
    )


def next_regime_from_retirement(age, final_age_alive):
    return np.where(age >= final_age_alive, RegimeId.dead, RegimeId.retirement)


def _dead():
    return Regime(transition=None, functions={"utility": lambda: 0.0}, active=lambda _age: True)


def retirement_model(
    n_periods: int,
    *,
    n_wealth: int = 100,
    n_consumption: int = 200,
    wealth_range=(1.0, 400.0),
    consumption_range=(1.0, 400.0),
    log_utility: bool = False,
    income_nodes: int | None = None,
) -> Model:
    """The work/retire consumption-savings model with `n_periods` periods.

    The last period is spent in the terminal `dead` regime. With
    `income_nodes`, both alive regimes carry a Rouwenhorst AR(1) log-income
    state `income` (parameters `rho`, `sigma`, `mu` under
    `params[regime]["income"]`) and next wealth adds `exp(income)`.
    """
    last_age = n_periods - 1
    W = LinSpacedGrid(start=wealth_range[0], stop=wealth_range[1], n_points=n_wealth)
    Cg = LinSpacedGrid(start=consumption_range[0], stop=consumption_range[1], n_points=n_consumption)
    states = {"wealth": W}
    nw = next_wealth
    if income_nodes:
        states["income"] = RouwenhorstAR1Process(n_points=income_nodes)
        nw = next_wealth_with_income
    if log_utility:
        fw = {"utility": log_utility_working, "labor_income": labor_income, "is_working": is_working}
        fr = {"utility": log_utility_retirement}
    else:
        fw = {"utility": utility_working, "crra_utility": crra_utility, "labor_income": labor_income,
              "is_working": is_working}
        fr = {"utility": utility_retirement, "crra_utility": crra_utility}
    working = Regime(
        actions={"labor_supply": DiscreteGrid(LaborSupply), "consumption": Cg},
        states=states,
        state_transitions={"wealth": nw},
        constraints={"borrowing_constraint": borrowing_constraint},
        transition=next_regime_from_working,
        functions=fw,
        active=lambda age, la=last_age: age < la,
    )
    retired = Regime(
        actions={"consumption": Cg},
        states=states,
        state_transitions={"wealth": nw},
        constraints={"borrowing_constraint": borrowing_constraint},
        transition=next_regime_from_retirement,
        functions=fr,
        active=lambda age, la=last_age: age < la,
    )
    return Model(
        regimes={"working_life": working, "retirement": retired, "dead": _dead()},
        ages=AgeGrid(start=0, stop=last_age, step="Y"),
        regime_id_class=RegimeId,
    )


def retirement_params(
    n_periods: int,
    *,
    discount_factor: float = 0.95,
    crra: float = 1.5,
    disutility_of_work: float = 0.5,
    interest_rate: float = 0.04,
    wage: float = 1.0,
    income=None,
    log_utility: bool = False,
) -> dict:
    """Parameters of `retirement_model`; `income` = dict(rho=, sigma=, mu=)."""
    p = {
        "discount_factor": discount_factor,
        "interest_rate": interest_rate,
        "final_age_alive": n_periods - 2,
        "working_life": {
            "utility": {"disutility_of_work": disutility_of_work},
            "labor_income": {"wage": wage},
        },
        "retirement": {"next_wealth": {"labor_income": 0.0}},
    }
    if not log_utility:
        p["crra"] = crra
    if income is not None:
        p["working_life"]["income"] = dict(income)
        p["retirement"]["income"] = dict(income)
    return p


def deterministic_workload():
    """Deterministic grid search, float64: T=10, wealth 100 x consumption 200."""
    return retirement_model(10, n_wealth=100, n_consumption=200), retirement_params(
        10, discount_factor=0.95, crra=1.5, disutility_of_work=0.5, interest_rate=0.04, wage=1.0
    )


def stochastic_workload(n_periods=40, n_grid=500):
    """Stochastic grid search, float32: T=40, 500 x 500, 7 Rouwenhorst nodes."""
    m = retirement_model(n_periods, n_wealth=n_grid, n_consumption=n_grid, income_nodes=7)
    return m, retirement_params(
        n_periods, discount_factor=0.95, crra=1.5, disutility_of_work=0.5, interest_rate=0.04,
        wage=1.0, income={"rho": 0.95, "sigma": 0.1, "mu": 0.0},
    )


def draw_batch_params(n_draws, n_periods=40, seed=0):
    """Parameter draws of the batched workload: beta~U[.90,.99], crra~U[1.1,3],
    sigma~U[.05,.2], AR rho~U[.8,.97]; grid size per draw from {200, 500, 1000}."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_draws):
        size = int(rng.choice([200, 500, 1000]))
        p = retirement_params(
            n_periods,
            discount_factor=float(rng.uniform(0.90, 0.99)),
            crra=float(rng.uniform(1.1, 3.0)),
            income={"rho": float(rng.uniform(0.8, 0.97)), "sigma": float(rng.uniform(0.05, 0.2)),
                    "mu": 0.0},
        )
        out.append((size, p))
    return out


# ---------------------------------------------------------------------------
# DC-EGM variant
# ---------------------------------------------------------------------------


def savings(wealth, consumption):
    return wealth - consumption


def next_wealth_from_savings(savings, labor_income, income, interest_rate):
    return (1 + interest_rate) * savings + labor_income * income


def dcegm_retirement_model(
    n_periods: int,
    *,
    n_savings: int = 2000,
    savings_max: float = 500.0,
    kappa: float = 3.0,
    n_wealth: int = 2000,
    wealth_range=(1e-3, 500.0),
    n_consumption: int = 4000,
    taste_shocks: bool = True,
    income_nodes: int = 5,
    n_constrained_points: int = 20,
    grid_search: bool = False,
) -> Model:
    """DC-EGM retirement model: log-normal income on GH nodes, EV1 taste shocks.

    Savings nodes are exponentially spaced on [0, savings_max]:
    `A_i = savings_max * (exp(kappa i/(n-1)) - 1) / (exp(kappa) - 1)`.
    With `grid_search=True` the same economics is solved by grid search over
    `n_consumption` consumption points (the comparison answer).
    """
    last_age = n_periods - 1
    i = np.arange(n_savings)
    A = savings_max * np.expm1(kappa * i / (n_savings - 1)) / np.expm1(kappa)
    sav_grid = IrregSpacedGrid(points=tuple(float(a) for a in A))
    W = LinSpacedGrid(start=wealth_range[0], stop=wealth_range[1], n_points=n_wealth)
    Cg = LinSpacedGrid(start=wealth_range[0], stop=wealth_range[1], n_points=n_consumption)
    solver = DCEGM(savings_grid=sav_grid, envelope=FUESEnvelope(), n_constrained_points=n_constrained_points)
    states = {"wealth": W, "income": LogNormalIIDProcess(n_points=income_nodes, gauss_hermite=True)}
    margin = LiquidMargin(state="wealth", action="consumption", resources="wealth",
                          post_decision_state="savings")
    ts = ExtremeValueTasteShocks() if taste_shocks else None
    common = dict(states=states, active=lambda age, la=last_age: age < la)
    fw = {"utility": utility_working, "crra_utility": crra_utility, "labor_income": labor_income,
          "is_working": is_working, "savings": savings}
    fr = {"utility": utility_retirement, "crra_utility": crra_utility, "savings": savings}
    if grid_search:
        working = Regime(actions={"labor_supply": DiscreteGrid(LaborSupply), "consumption": Cg},
                         state_transitions={"wealth": _nw_gs}, transition=next_regime_from_working,
                         constraints={"borrowing_constraint": borrowing_constraint},
                         functions=fw, taste_shocks=ts, **common)
        retired = Regime(actions={"consumption": Cg}, state_transitions={"wealth": _nw_gs},
                         transition=next_regime_from_retirement,
                         constraints={"borrowing_constraint": borrowing_constraint},
                         functions=fr, **common)
    else:
        working = ConsumptionSavingsRegime(
            actions={"labor_supply": DiscreteGrid(LaborSupply), "consumption": Cg},
            state_transitions={"wealth": next_wealth_from_savings}, transition=next_regime_from_working,
            functions=fw, solver=solver, liquid=margin, taste_shocks=ts, **common)
        retired = ConsumptionSavingsRegime(
            actions={"consumption": Cg}, state_transitions={"wealth": next_wealth_from_savings},
            transition=next_regime_from_retirement, functions=fr, solver=solver, liquid=margin,
            **common)
    return Model(regimes={"working_life": working, "retirement": retired, "dead": _dead()},
                 ages=AgeGrid(start=0, stop=last_age, step="Y"), regime_id_class=RegimeId)


def _nw_gs(wealth, consumption, labor_income, income, interest_rate):
    return (1 + interest_rate) * (wealth - consumption) + labor_income * income


def dcegm_params(n_periods, *, discount_factor=0.96, crra=2.0, disutility_of_work=0.5,
                 interest_rate=0.03, wage=1.0, sigma=0.15, taste_scale=0.05) -> dict:
    p = retirement_params(n_periods, discount_factor=discount_factor, crra=crra,
                          disutility_of_work=disutility_of_work, interest_rate=interest_rate,
                          wage=wage)
    inc = {"mu": 0.0, "sigma": sigma}
    p["working_life"]["income"] = dict(inc)
    p["retirement"]["income"] = dict(inc)
    p["retirement"]["next_wealth"] = {"labor_income": 0.0}
    if taste_scale is not None:
        p["working_life"]["taste_shocks"] = {"scale": taste_scale}
    return p
