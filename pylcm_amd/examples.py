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

* DC-EGM models: `iskhakov_dcegm_model` (the deterministic retirement model
  solved by DC-EGM), `dcegm_twin_model` (twin of the independent `dcegm`
  package's taste-shock model the reference cross-checks against) and
  `dcegm_retirement_model` (CRRA, taste shocks, log-normal income level on
  Gauss-Hermite nodes; the DC-EGM benchmark workload).
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
    LogSpacedGrid,
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
# mortality: stochastic regime transitions
# ---------------------------------------------------------------------------


def next_regime_from_working_mortality(labor_supply, period, survival_probs):
    """[P(working), P(retired), P(dead)]: survive with `survival_probs[period]`,
    retire if the agent chose to (`lcm_examples/mortality.py:73-94`)."""
    sp = survival_probs[period]
    retire = labor_supply == LaborSupply.retire
    return np.where(retire, np.array([0.0, sp, 1.0 - sp]), np.array([sp, 0.0, 1.0 - sp]))


def next_regime_from_retirement_mortality(period, survival_probs):
    sp = survival_probs[period]
    return np.array([0.0, sp, 1.0 - sp])


def mortality_model(n_periods: int, *, n_wealth: int = 100, n_consumption: int = 500) -> Model:
    """Working life, retirement and death with age-dependent survival
    (`lcm_examples/mortality.py:117-170`): ages 40, 50, ... in 10-year
    periods, log utility with work disutility, regime transitions given as
    probability vectors (`MarkovTransition`), so the continuation blends the
    alive regime with the dead one by the survival probability."""
    from pylcm_amd.regime import MarkovTransition

    ages = AgeGrid(start=40, stop=40 + (n_periods - 1) * 10, step="10Y")
    last_age = ages.exact_values[-1]
    W = LinSpacedGrid(start=1, stop=400, n_points=n_wealth)
    Cg = LinSpacedGrid(start=1, stop=400, n_points=n_consumption)
    working = Regime(
        actions={"labor_supply": DiscreteGrid(LaborSupply), "consumption": Cg},
        states={"wealth": W},
        state_transitions={"wealth": next_wealth},
        constraints={"borrowing_constraint": borrowing_constraint},
        transition=MarkovTransition(next_regime_from_working_mortality),
        functions={"utility": log_utility_working, "labor_income": labor_income, "is_working": is_working},
        active=lambda age, la=last_age: age < la,
    )
    retired = Regime(
        actions={"consumption": Cg},
        states={"wealth": W},
        state_transitions={"wealth": next_wealth},
        constraints={"borrowing_constraint": borrowing_constraint},
        transition=MarkovTransition(next_regime_from_retirement_mortality),
        functions={"utility": log_utility_retirement},
        active=lambda age, la=last_age: age < la,
    )
    return Model(regimes={"working_life": working, "retirement": retired, "dead": _dead()},
                 ages=ages, regime_id_class=RegimeId)


def default_survival_probs(n_periods: int) -> np.ndarray:
    """0.98 down to 0.82 linearly over the first n-2 periods, then 0 (certain
    death before the last period) (`lcm_examples/mortality.py:173-192`)."""
    if n_periods <= 2:
        return np.array([0.0])
    if n_periods == 3:
        return np.array([0.98, 0.0])
    return np.concatenate([np.linspace(0.98, 0.82, n_periods - 2), [0.0]])


def mortality_params(n_periods: int, *, discount_factor: float = 0.95, disutility_of_work: float = 0.5,
                     interest_rate: float = 0.05, wage: float = 10.0, survival_probs=None) -> dict:
    """Parameters of `mortality_model` (`lcm_examples/mortality.py:195-235`)."""
    if survival_probs is None:
        survival_probs = default_survival_probs(n_periods)
    return {
        "discount_factor": discount_factor,
        "interest_rate": interest_rate,
        "survival_probs": np.asarray(survival_probs, dtype=np.float64),
        "working_life": {
            "utility": {"disutility_of_work": disutility_of_work},
            "labor_income": {"wage": wage},
        },
        "retirement": {"next_wealth": {"labor_income": 0.0}},
    }


# ---------------------------------------------------------------------------
# DC-EGM models
# ---------------------------------------------------------------------------


def savings(wealth, consumption):
    return wealth - consumption


def next_wealth_from_savings(savings, labor_income, interest_rate):
    """`(1 + r) * (wealth - consumption) + labor_income` written in the savings."""
    return (1 + interest_rate) * savings + labor_income


def inverse_marginal_utility_log(marginal_continuation):
    """Inverse of u'(c) = 1/c (work disutility is additive)."""
    return 1.0 / marginal_continuation


def inverse_marginal_utility_crra(marginal_continuation, crra):
    """Inverse of u'(c) = c**(-crra)."""
    return marginal_continuation ** (-1.0 / crra)


def iskhakov_dcegm_model(n_periods: int, *, n_savings: int = 200, n_constrained_points: int = 64,
                         n_wealth: int = 100) -> Model:
    """DC-EGM twin of `retirement_model(log_utility=True)`: the deterministic
    Iskhakov et al. (2017) retirement model with a cubic savings grid on
    [0, 400] (`tests/test_models/deterministic/dcegm_variants.py:39-140` of the
    reference)."""
    last_age = n_periods - 1
    W = LinSpacedGrid(start=1.0, stop=400.0, n_points=n_wealth)
    Cg = LinSpacedGrid(start=1.0, stop=400.0, n_points=500)
    sav = IrregSpacedGrid(points=tuple(400.0 * (i / (n_savings - 1)) ** 3 for i in range(n_savings)))
    solver = DCEGM(savings_grid=sav, n_constrained_points=n_constrained_points)
    margin = LiquidMargin(state="wealth", action="consumption", resources="wealth",
                          post_decision_state="savings")
    working = ConsumptionSavingsRegime(
        actions={"labor_supply": DiscreteGrid(LaborSupply), "consumption": Cg},
        states={"wealth": W}, state_transitions={"wealth": next_wealth_from_savings},
        transition=next_regime_from_working,
        functions={"utility": log_utility_working, "labor_income": labor_income,
                   "is_working": is_working, "savings": savings,
                   "inverse_marginal_utility": inverse_marginal_utility_log},
        solver=solver, liquid=margin, active=lambda age, la=last_age: age < la)
    retired = ConsumptionSavingsRegime(
        actions={"consumption": Cg}, states={"wealth": W},
        state_transitions={"wealth": next_wealth_from_savings},
        transition=next_regime_from_retirement,
        functions={"utility": log_utility_retirement, "savings": savings,
                   "inverse_marginal_utility": inverse_marginal_utility_log},
        solver=solver, liquid=margin, active=lambda age, la=last_age: age < la)
    return Model(regimes={"working_life": working, "retirement": retired, "dead": _dead()},
                 ages=AgeGrid(start=0, stop=last_age, step="Y"), regime_id_class=RegimeId)


# -- twin of the independent `dcegm` package's IJRS taste-shock model ----------
# (reference: tests/test_models/dcegm_paper_twin.py; fixture
#  tests/data/dcegm_reference/ijrs_taste_shocks_reference.csv)

TWIN_MIN_AGE = 20
TWIN_N_PERIODS = 10
TWIN_LAST_ALIVE_AGE = TWIN_MIN_AGE + TWIN_N_PERIODS - 2
TWIN_WEALTH_POINTS = (
    tuple(0.001 * (0.5 / 0.001) ** (i / 120) for i in range(120))
    + tuple(0.5 + 0.05 * i for i in range(90))
    + tuple(5.0 + 0.25 * i for i in range(181))
)


@categorical(ordered=False)
class TwinRegimeId:
    working_life: int
    retirement: int
    done_from_working: int
    done_retired: int


@categorical(ordered=True)
class WorkChoice:
    work: int
    retire: int


def twin_crra(consumption, rho):
    return (consumption ** (1.0 - rho) - 1.0) / (1.0 - rho)


def twin_is_working(work_choice):
    return work_choice == WorkChoice.work


def twin_utility_working(consumption, is_working, rho, delta):
    return twin_crra(consumption, rho) - np.where(is_working, delta, 0.0)


def twin_utility_retired(consumption, rho):
    return twin_crra(consumption, rho)


def twin_utility_done_from_working(wealth, rho, delta, taste_scale):
    return twin_crra(wealth, rho) + taste_scale * np.log(1.0 + np.exp(-delta / taste_scale))


def twin_utility_done_retired(wealth, rho):
    return twin_crra(wealth, rho)


def twin_labor_income(is_working, age, constant, exp_coeff, exp_squared):
    next_age = age + 1.0
    return np.where(is_working, np.exp(constant + exp_coeff * next_age + exp_squared * next_age**2), 0.0)


def twin_next_wealth(wealth, consumption, labor_income, interest_rate, floor):
    return np.maximum((1.0 + interest_rate) * (wealth - consumption) + labor_income, floor)


def twin_next_wealth_from_savings(savings, labor_income, interest_rate, floor):
    return np.maximum((1.0 + interest_rate) * savings + labor_income, floor)


def twin_inverse_marginal_utility(marginal_continuation, rho):
    return marginal_continuation ** (-1.0 / rho)


def twin_next_regime_from_working(work_choice, age):
    last = age >= TWIN_LAST_ALIVE_AGE
    retire = work_choice == WorkChoice.retire
    return np.where(last, np.where(retire, TwinRegimeId.done_retired, TwinRegimeId.done_from_working),
                    np.where(retire, TwinRegimeId.retirement, TwinRegimeId.working_life))


def twin_next_regime_from_retirement(age):
    return np.where(age >= TWIN_LAST_ALIVE_AGE, TwinRegimeId.done_retired, TwinRegimeId.retirement)


def dcegm_twin_model(*, grid_search: bool = False, savings_points=None) -> Model:
    """The IJRS consumption-retirement model with EV1 taste shocks on the
    work/retire choice, absorbing retirement and two terminal regimes holding
    the closed-form final-period values. DC-EGM on a 500-point uniform savings
    grid on [0, 50], or grid search over 400 log-spaced consumption points."""
    W = IrregSpacedGrid(points=TWIN_WEALTH_POINTS)
    Cg = LogSpacedGrid(start=0.001, stop=50.0, n_points=400)
    active = lambda age: age <= TWIN_LAST_ALIVE_AGE  # noqa: E731
    if savings_points is None:
        sav = LinSpacedGrid(start=0.0, stop=50.0, n_points=500)
    else:
        sav = IrregSpacedGrid(points=tuple(float(x) for x in savings_points))
    solver = DCEGM(savings_grid=sav)
    margin = LiquidMargin(state="wealth", action="consumption", resources="wealth",
                          post_decision_state="savings")
    fw = {"utility": twin_utility_working, "is_working": twin_is_working,
          "labor_income": twin_labor_income}
    fr = {"utility": twin_utility_retired}
    acts_w = {"work_choice": DiscreteGrid(WorkChoice), "consumption": Cg}
    if grid_search:
        working = Regime(transition=twin_next_regime_from_working, states={"wealth": W}, actions=acts_w,
                         taste_shocks=ExtremeValueTasteShocks(), active=active,
                         state_transitions={"wealth": twin_next_wealth},
                         constraints={"borrowing_constraint": borrowing_constraint}, functions=fw)
        retired = Regime(transition=twin_next_regime_from_retirement, states={"wealth": W},
                         actions={"consumption": Cg}, active=active,
                         state_transitions={"wealth": twin_next_wealth},
                         constraints={"borrowing_constraint": borrowing_constraint}, functions=fr)
    else:
        working = ConsumptionSavingsRegime(
            transition=twin_next_regime_from_working, states={"wealth": W}, actions=acts_w,
            taste_shocks=ExtremeValueTasteShocks(), active=active,
            state_transitions={"wealth": twin_next_wealth_from_savings},
            functions={**fw, "savings": savings, "inverse_marginal_utility": twin_inverse_marginal_utility},
            solver=solver, liquid=margin)
        retired = ConsumptionSavingsRegime(
            transition=twin_next_regime_from_retirement, states={"wealth": W},
            actions={"consumption": Cg}, active=active,
            state_transitions={"wealth": twin_next_wealth_from_savings},
            functions={**fr, "savings": savings, "inverse_marginal_utility": twin_inverse_marginal_utility},
            solver=solver, liquid=margin)
    done_w = Regime(transition=None, states={"wealth": W},
                    functions={"utility": twin_utility_done_from_working})
    done_r = Regime(transition=None, states={"wealth": W}, functions={"utility": twin_utility_done_retired})
    return Model(regimes={"working_life": working, "retirement": retired,
                          "done_from_working": done_w, "done_retired": done_r},
                 ages=AgeGrid(start=TWIN_MIN_AGE, stop=TWIN_MIN_AGE + TWIN_N_PERIODS - 1, step="Y"),
                 regime_id_class=TwinRegimeId)


def dcegm_twin_params(taste_shock_scale: float = 0.2) -> dict:
    return {
        "discount_factor": 0.95, "interest_rate": 0.05, "rho": 1.95, "delta": 0.35,
        "constant": 0.75, "exp_coeff": 0.04, "exp_squared": -0.0002, "floor": 0.001,
        "done_from_working": {"utility": {"taste_scale": taste_shock_scale}},
        "working_life": {"taste_shocks": {"scale": taste_shock_scale}},
        "retirement": {"next_wealth": {"labor_income": 0.0}},
    }


# -- DC-EGM retirement model with a log-normal income shock --------------------


def next_wealth_stochastic(savings, labor_income, income, interest_rate):
    """Savings grow at 1 + r; labor income scaled by the log-normal income level."""
    return (1 + interest_rate) * savings + labor_income * income


def _nw_gs(wealth, consumption, labor_income, income, interest_rate):
    return (1 + interest_rate) * (wealth - consumption) + labor_income * income


def exp_savings_points(n_savings: int, savings_max: float, kappa: float = 3.0):
    """Exponentially spaced savings nodes on [0, savings_max]:
    `A_i = savings_max * (exp(kappa i/(n-1)) - 1) / (exp(kappa) - 1)`."""
    i = np.arange(n_savings)
    return savings_max * np.expm1(kappa * i / (n_savings - 1)) / np.expm1(kappa)


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
    """DC-EGM retirement model: CRRA utility, work/retire choice with EV1 taste
    shocks, absorbing retirement, a log-normal IID income level on Gauss-Hermite
    nodes scaling the wage, exponentially spaced savings nodes.

    With `grid_search=True` the same economics is solved by grid search over
    `n_consumption` consumption points (the comparison answer).
    """
    last_age = n_periods - 1
    A = exp_savings_points(n_savings, savings_max, kappa)
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
          "is_working": is_working}
    fr = {"utility": utility_retirement, "crra_utility": crra_utility}
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
        inv = {"savings": savings, "inverse_marginal_utility": inverse_marginal_utility_crra}
        working = ConsumptionSavingsRegime(
            actions={"labor_supply": DiscreteGrid(LaborSupply), "consumption": Cg},
            state_transitions={"wealth": next_wealth_stochastic}, transition=next_regime_from_working,
            functions={**fw, **inv}, solver=solver, liquid=margin, taste_shocks=ts, **common)
        retired = ConsumptionSavingsRegime(
            actions={"consumption": Cg}, state_transitions={"wealth": next_wealth_stochastic},
            transition=next_regime_from_retirement, functions={**fr, **inv}, solver=solver,
            liquid=margin, **common)
    return Model(regimes={"working_life": working, "retirement": retired, "dead": _dead()},
                 ages=AgeGrid(start=0, stop=last_age, step="Y"), regime_id_class=RegimeId)


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


# -- serving request mix ---------------------------------------------------------


def serving_initial_conditions(model, params, grids, n_agents: int, rng):
    """Initial conditions of a serving request's forward simulation: every
    agent in working life with wealth lognormal(log 10, 0.5) and, in the
    AR(1)-income model, the income node drawn from Binomial(nodes - 1, 1/2)
    (the Rouwenhorst chain's stationary distribution); for
    `NativeSolveServer.submit(..., simulate={"initial_conditions": ...})`."""
    ic = {"wealth": np.exp(np.log(10.0) + 0.5 * rng.standard_normal(n_agents)),
          "regime_id": np.zeros(n_agents, dtype=int)}
    key = ("working_life", "income")
    if key in grids.proc:
        nodes = np.asarray(grids.proc[key][0], dtype=np.float64)
        ic["income"] = nodes[rng.binomial(len(nodes) - 1, 0.5, n_agents)]
    return ic


def serving_request_factory(seed: int = 0, horizons=(20, 60), grids=(200, 4000), income_nodes: int = 7):
    """`make(i) -> (model, params)` drawing the serving workload's requests:
    half deterministic (the T=10 workload's economics), half with the AR(1)
    income shock, horizon U{20..60}, wealth = consumption grid U{200..4000},
    and parameters drawn as in `draw_batch_params`. Models are shared between
    requests of equal shape."""
    rng = np.random.default_rng(seed)
    models = {}

    def make(i):
        stoch = bool(rng.integers(0, 2))
        T = int(rng.integers(horizons[0], horizons[1] + 1))
        g = int(rng.integers(grids[0], grids[1] + 1))
        key = (stoch, T, g)
        if key not in models:
            models[key] = retirement_model(T, n_wealth=g, n_consumption=g,
                                           income_nodes=income_nodes if stoch else None)
        income = None
        if stoch:
            income = {"rho": float(rng.uniform(0.8, 0.97)), "sigma": float(rng.uniform(0.05, 0.2)), "mu": 0.0}
        p = retirement_params(T, discount_factor=float(rng.uniform(0.90, 0.99)),
                              crra=float(rng.uniform(1.1, 3.0)), income=income)
        return models[key], p

    return make
