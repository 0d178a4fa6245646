"""Test models that exercise more of the specification than the examples.

`processes_model` follows the reference's tests/test_models/processes.py:44-120
(wealth, a stochastic income process and a Markov health state; utility reads
health; next wealth can leave the grid so the continuation extrapolates).
"""

import numpy as np

from pylcm_amd import (
    AgeGrid,
    DiscreteGrid,
    IrregSpacedGrid,
    LinSpacedGrid,
    LogNormalIIDProcess,
    LogSpacedGrid,
    MarkovTransition,
    Model,
    NormalIIDProcess,
    Regime,
    RouwenhorstAR1Process,
    TauchenAR1Process,
    UniformIIDProcess,
    categorical,
)


@categorical(ordered=True)
class Health:
    bad: int
    good: int


@categorical(ordered=False)
class AliveDead:
    alive: int
    dead: int


def next_health(health, probs_array):
    return probs_array[health]


def pm_next_wealth(consumption, wealth):
    return wealth - consumption


def pm_next_regime(age, final_age_alive):
    return np.where(age >= final_age_alive, AliveDead.dead, AliveDead.alive)


def pm_wealth_constraint(wealth, income, consumption):
    return wealth - consumption + np.exp(income) >= 0


def pm_utility(wealth, income, health, consumption):
    return np.log(consumption) * (1.0 - (1.0 - health) * 0.3)


PROCESSES = {
    "uniform": lambda: UniformIIDProcess(n_points=5),
    "normal": lambda: NormalIIDProcess(n_points=5, gauss_hermite=True),
    "lognormal": lambda: LogNormalIIDProcess(n_points=5, gauss_hermite=True),
    "tauchen": lambda: TauchenAR1Process(n_points=5, gauss_hermite=True),
    "rouwenhorst": lambda: RouwenhorstAR1Process(n_points=5),
}

PROCESS_PARAMS = {
    "uniform": {"start": 0.0, "stop": 1.0},
    "normal": {"mu": 0.0, "sigma": 1.0},
    "lognormal": {"mu": 0.0, "sigma": 1.0},
    "tauchen": {"rho": 0.9, "sigma": 1.0, "mu": 0.0},
    "rouwenhorst": {"rho": 0.9, "sigma": 1.0, "mu": 0.0},
}


def processes_model(n_periods, distribution, wealth_grid=None):
    final_age_alive = n_periods - 2
    alive = Regime(
        active=lambda age, n=final_age_alive: age <= n,
        states={
            "wealth": wealth_grid or LinSpacedGrid(start=1, stop=5, n_points=5),
            "income": PROCESSES[distribution](),
            "health": DiscreteGrid(Health),
        },
        state_transitions={"wealth": pm_next_wealth, "health": MarkovTransition(next_health)},
        actions={"consumption": LinSpacedGrid(start=0.1, stop=2, n_points=4)},
        transition=pm_next_regime,
        constraints={"wealth_constraint": pm_wealth_constraint},
        functions={"utility": pm_utility},
    )
    dead = Regime(transition=None, functions={"utility": lambda: 0.0})
    return Model(
        regimes={"alive": alive, "dead": dead},
        regime_id_class=AliveDead,
        ages=AgeGrid(start=0, stop=n_periods - 1, step="Y"),
        fixed_params={"final_age_alive": final_age_alive},
    )


def processes_params(distribution):
    return {
        "discount_factor": 0.95,
        "alive": {
            "next_health": {"probs_array": np.array([[0.9, 0.1], [0.2, 0.8]])},
            "income": dict(PROCESS_PARAMS[distribution]),
        },
    }


def log_grid_model(n_periods):
    from pylcm_amd import examples as ex

    m = ex.retirement_model(n_periods, n_wealth=40, n_consumption=50, income_nodes=3)
    regs = {}
    for n, r in m.user_regimes.items():
        if r.terminal:
            regs[n] = r
            continue
        regs[n] = r.replace(states={"wealth": LogSpacedGrid(start=1.0, stop=400.0, n_points=40),
                                    "income": r.states["income"]})
    return Model(regimes=regs, ages=m.ages, regime_id_class=m.regime_id_class)


def irregular_grid_model(n_periods):
    from pylcm_amd import examples as ex

    m = ex.retirement_model(n_periods, n_wealth=40, n_consumption=50)
    pts = tuple(float(1 + 399 * (i / 39) ** 2) for i in range(40))
    regs = {}
    for n, r in m.user_regimes.items():
        if r.terminal:
            regs[n] = r
            continue
        regs[n] = r.replace(states={"wealth": IrregSpacedGrid(points=pts)})
    return Model(regimes=regs, ages=m.ages, regime_id_class=m.regime_id_class)


def taste_shock_model(n_periods):
    from pylcm_amd import ExtremeValueTasteShocks
    from pylcm_amd import examples as ex

    m = ex.retirement_model(n_periods, n_wealth=30, n_consumption=40, income_nodes=3)
    regs = dict(m.user_regimes)
    regs["working_life"] = regs["working_life"].replace(taste_shocks=ExtremeValueTasteShocks())
    return Model(regimes=regs, ages=m.ages, regime_id_class=m.regime_id_class)


def nonaffine_model(n_periods):
    """Utility reads wealth and next wealth is nonlinear in wealth: the host
    tabulates those over the full state grid (the general kernel variant)."""
    from pylcm_amd import examples as ex

    m = ex.retirement_model(n_periods, n_wealth=30, n_consumption=40)

    def u_w(consumption, wealth, is_working, disutility_of_work):
        return np.log(consumption) + 0.01 * np.sqrt(wealth) - np.where(is_working, disutility_of_work, 0.0)

    def u_r(consumption, wealth):
        return np.log(consumption) + 0.01 * np.sqrt(wealth)

    def nw(wealth, consumption, labor_income, interest_rate):
        return (1 + interest_rate) * (wealth - consumption) + 0.1 * np.sqrt(wealth) + labor_income

    regs = dict(m.user_regimes)
    regs["working_life"] = regs["working_life"].replace(
        functions={**regs["working_life"].functions, "utility": u_w},
        state_transitions={"wealth": nw})
    regs["retirement"] = regs["retirement"].replace(
        functions={"utility": u_r}, state_transitions={"wealth": nw})
    return Model(regimes=regs, ages=m.ages, regime_id_class=m.regime_id_class)


def nan_infeasible_utility_model(n_periods):
    """Utility is NaN for consumption above 420, which no wealth node (at most
    400) can afford: the NaN sits in staged action chunks but is never feasible."""
    from pylcm_amd import examples as ex

    m = ex.retirement_model(n_periods, n_wealth=300, n_consumption=300, consumption_range=(1.0, 500.0))

    def slack(consumption):
        return np.log(np.where(consumption < 420.0, 420.0 - consumption, np.nan))

    def u_w(consumption, is_working, disutility_of_work):
        return np.log(consumption) + 0.01 * slack(consumption) - np.where(is_working, disutility_of_work, 0.0)

    def u_r(consumption):
        return np.log(consumption) + 0.01 * slack(consumption)

    regs = dict(m.user_regimes)
    regs["working_life"] = regs["working_life"].replace(
        functions={"utility": u_w, "labor_income": ex.labor_income, "is_working": ex.is_working})
    regs["retirement"] = regs["retirement"].replace(functions={"utility": u_r})
    return Model(regimes=regs, ages=m.ages, regime_id_class=m.regime_id_class)


def scaled_utility_model(n_periods, scale):
    """Log utility times `scale` (value functions of magnitude ~scale)."""
    from pylcm_amd import examples as ex

    m = ex.retirement_model(n_periods, n_wealth=90, n_consumption=120, income_nodes=3)

    def u_w(consumption, is_working, disutility_of_work):
        return scale * (np.log(consumption) - np.where(is_working, disutility_of_work, 0.0))

    def u_r(consumption):
        return scale * np.log(consumption)

    regs = dict(m.user_regimes)
    regs["working_life"] = regs["working_life"].replace(
        functions={"utility": u_w, "labor_income": ex.labor_income, "is_working": ex.is_working})
    regs["retirement"] = regs["retirement"].replace(functions={"utility": u_r})
    return Model(regimes=regs, ages=m.ages, regime_id_class=m.regime_id_class)


# -- the reference's regression-test model ------------------------------------
# (tests/test_models/deterministic/regression.py:38-134 of the reference; its
#  solution is pinned by tests/data/regression_tests/solution.pkl)

REGRESSION_START_AGE = 18


@categorical(ordered=False)
class RegressionRegimeId:
    working_life: int
    dead: int


def regression_wage(age):
    return 1 + 0.1 * age


def regression_next_regime(age, final_age_alive):
    return np.where(age >= final_age_alive, RegressionRegimeId.dead, RegressionRegimeId.working_life)


def regression_model(n_periods, wealth_grid=None, consumption_grid=None):
    from pylcm_amd import examples as ex

    final_age_alive = REGRESSION_START_AGE + n_periods - 2
    working = Regime(
        actions={"labor_supply": DiscreteGrid(ex.LaborSupply),
                 "consumption": consumption_grid or LinSpacedGrid(start=1, stop=400, n_points=500)},
        states={"wealth": wealth_grid or LinSpacedGrid(start=1, stop=400, n_points=100)},
        state_transitions={"wealth": ex.next_wealth},
        constraints={"borrowing_constraint": ex.borrowing_constraint},
        transition=regression_next_regime,
        functions={"utility": ex.log_utility_working, "labor_income": ex.labor_income,
                   "is_working": ex.is_working, "wage": regression_wage},
        active=lambda age, n=final_age_alive: age <= n,
    )
    dead = Regime(transition=None, functions={"utility": lambda: 0.0})
    return Model(regimes={"working_life": working, "dead": dead},
                 ages=AgeGrid(start=REGRESSION_START_AGE, stop=final_age_alive + 1, step="Y"),
                 regime_id_class=RegressionRegimeId)


def regression_params(n_periods, discount_factor=0.95, disutility_of_work=1.0, interest_rate=0.05):
    return {
        "discount_factor": discount_factor,
        "working_life": {
            "utility": {"disutility_of_work": disutility_of_work},
            "next_wealth": {"interest_rate": interest_rate},
            "next_regime": {"final_age_alive": REGRESSION_START_AGE + n_periods - 2},
        },
    }


# the reference's own parameters of the processes model
# (tests/test_models/processes.py:238-260 of the reference; its solution is
#  pinned by tests/data/processes/solution_{dist}.pkl)
REFERENCE_PROCESS_PARAMS = {
    "uniform": {"start": 0.0, "stop": 1.0},
    "normal": {"mu": 0.0, "sigma": 1.0},
    "lognormal": {"mu": 0.0, "sigma": 1.0},
    "tauchen": {"rho": 0.975, "sigma": 1.0, "mu": 0.0},
    "rouwenhorst": {"rho": 0.975, "sigma": 1.0, "mu": 0.0},
}


def reference_processes_params(distribution):
    return {
        "alive": {
            "discount_factor": 1.0,
            "next_health": {"probs_array": np.full((2, 2), 0.5)},
            "income": dict(REFERENCE_PROCESS_PARAMS[distribution]),
        },
        "dead": {},
    }


# -- the reference's precautionary-savings model --------------------------------
# (src/lcm_examples/precautionary_savings.py:83-200 of the reference, Rouwenhorst
#  income; its simulation is pinned by
#  tests/data/regression_tests/{f32,f64}/precautionary_savings_simulation.pkl)


@categorical(ordered=False)
class AliveDeadPS:
    alive: int
    dead: int


def ps_next_wealth(wealth, consumption, income, interest_rate):
    return (1 + interest_rate) * (wealth - consumption) + np.exp(income)


def ps_next_regime(age, final_age_alive):
    return np.where(age >= final_age_alive, AliveDeadPS.dead, AliveDeadPS.alive)


def ps_wealth_constraint(consumption, wealth):
    return consumption <= wealth


def ps_utility(consumption):
    return np.log(consumption)


def precautionary_savings_model(n_periods, *, wealth_n_points=7, consumption_n_points=7,
                                income_n_points=5, wealth_start=1.0, wealth_stop=20.0):
    final_age_alive = 20 + (n_periods - 2) * 10
    alive = Regime(
        active=lambda age, n=final_age_alive: age <= n,
        states={"wealth": LinSpacedGrid(start=wealth_start, stop=wealth_stop, n_points=wealth_n_points),
                "income": RouwenhorstAR1Process(n_points=income_n_points)},
        state_transitions={"wealth": ps_next_wealth},
        actions={"consumption": LinSpacedGrid(start=0.1, stop=5, n_points=consumption_n_points)},
        transition=ps_next_regime,
        constraints={"wealth_constraint": ps_wealth_constraint},
        functions={"utility": ps_utility},
    )
    dead = Regime(transition=None, active=lambda age, n=final_age_alive: age > n,
                  functions={"utility": lambda: 0.0})
    return Model(regimes={"alive": alive, "dead": dead}, regime_id_class=AliveDeadPS,
                 ages=AgeGrid(start=20, stop=20 + (n_periods - 1) * 10, step="10Y"),
                 fixed_params={"final_age_alive": final_age_alive})


def precautionary_savings_params(*, sigma, mu=0.0, rho=0.0, interest_rate=0.0, discount_factor=0.95):
    return {
        "discount_factor": discount_factor,
        "alive": {"income": {"mu": mu, "sigma": sigma, "rho": rho},
                  "next_wealth": {"interest_rate": interest_rate}},
    }


# -- mortality variants for the regime-mass rule (Q_and_F.py:639-653,1395-1432) --


def scaled_next_regime_from_working(labor_supply, period, survival_probs, mass_scale, shift):
    """The mortality model's transition with every probability times
    `mass_scale`, and `shift * P(survival)` moved from P(retired) to
    P(working) (a negative P(retired) after choosing to work, in every period
    with a survival chance)."""
    from pylcm_amd import examples as ex

    p = ex.next_regime_from_working_mortality(labor_supply, period, survival_probs)
    return mass_scale * (p + survival_probs[period] * np.array([shift, -shift, 0.0]))


def scaled_next_regime_from_retirement(period, survival_probs, mass_scale):
    from pylcm_amd import examples as ex

    return mass_scale * ex.next_regime_from_retirement_mortality(period, survival_probs)


def mortality_mass_model(n_periods, *, n_wealth=40, n_consumption=60):
    from pylcm_amd import examples as ex
    from pylcm_amd.regime import MarkovTransition

    m = ex.mortality_model(n_periods, n_wealth=n_wealth, n_consumption=n_consumption)
    regs = dict(m.user_regimes)
    regs["working_life"] = regs["working_life"].replace(
        transition=MarkovTransition(scaled_next_regime_from_working))
    regs["retirement"] = regs["retirement"].replace(
        transition=MarkovTransition(scaled_next_regime_from_retirement))
    return Model(regimes=regs, ages=m.ages, regime_id_class=m.regime_id_class)


def mortality_mass_params(n_periods, mass_scale=1.0, shift=0.0):
    from pylcm_amd import examples as ex

    p = ex.mortality_params(n_periods)
    p["working_life"]["next_regime"] = {"mass_scale": mass_scale, "shift": shift}
    p["retirement"]["next_regime"] = {"mass_scale": mass_scale}
    return p
