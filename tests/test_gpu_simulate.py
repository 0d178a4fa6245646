"""GPU forward simulation against the oracle (same value functions, same
counter-based draws): integer draws and chosen actions identical, values and
continuous states within the dtype's tolerance."""

import ctypes as C

import numpy as np
import pytest

from oracle import gridsearch as og
from oracle import simulate as osim
from pylcm_amd import _native as nat
from pylcm_amd import examples as ex
from tests import models as tm

pytestmark = pytest.mark.gpu


def test_device_philox_matches_oracle(native):
    ctx = nat.Context()
    n = 4096
    buf = nat.DeviceBuffer(n * 16)
    for c1, c2, c3, k0, k1 in [(0, 0, 0, 0, 0), (7, 1, 2, 12345, 678), (0xFFFFFFFF, 3, 0, 0xDEADBEEF, 0x1)]:
        nat.call("lcm_philox", ctx.handle, C.c_void_p(buf.ptr), n, c1, c2, c3, k0, k1)
        got = np.zeros(4 * n, dtype=np.uint32)
        nat.d2h(ctx, buf.ptr, got)
        ctr = np.stack([np.arange(n), np.full(n, c1), np.full(n, c2), np.full(n, c3)], axis=1)
        exp = osim.philox4x32_10(ctr, (k0, k1)).astype(np.uint32).reshape(-1)
        assert np.array_equal(got, exp)
    # Random123 known answer at counter 0, key 0
    nat.call("lcm_philox", ctx.handle, C.c_void_p(buf.ptr), 1, 0, 0, 0, 0, 0)
    got = np.zeros(4, dtype=np.uint32)
    nat.d2h(ctx, buf.ptr, got)
    assert [int(x) for x in got] == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]


def _compare(model, params, ic, dtype, seed, rtol, V=None, oracle_model=None):
    if V is None:
        V, _ = og.solve(model, params, grid_dtype=dtype)
    res = model.simulate(params=params, initial_conditions=ic, period_to_regime_to_V_arr=V, log_level="off",
                         seed=seed, dtype=dtype)
    got = res.raw
    model = oracle_model or model
    exp = osim.simulate(model, params, V, ic, seed, grid_dtype=dtype)
    sim = exp["action"] >= 0
    assert np.array_equal(got["action"] >= 0, sim)
    assert np.array_equal(got["regime"][sim], exp["regime"][sim])
    assert np.array_equal(got["action"][sim], exp["action"][sim])
    names = list(model.regime_names)
    has = {"value": lambda lay: True, "w": lambda lay: lay.w_name is not None,
           "z": lambda lay: any(type(g).__name__ != "DiscreteGrid" for g in lay.sd_grids)}
    for k in ("value", "w", "z"):
        # compare a state only in rows whose regime carries it
        carries = np.array([has[k](model.layouts[r]) for r in names])
        rows = sim & carries[np.maximum(exp["regime"], 0)]
        a, b = got[k][rows].astype(np.float64), exp[k][rows]
        fin = np.isfinite(b)
        assert np.array_equal(np.isfinite(a), fin), k
        err = np.max(np.abs(a[fin] - b[fin]) / np.maximum(np.abs(b[fin]), 1e-12)) if fin.any() else 0
        print(f"{k}: max rel err {err:.2e}")
        assert err <= rtol, (k, err)
    return res


def test_simulate_deterministic_f64(native):
    m, p = ex.deterministic_workload()
    rng = np.random.default_rng(0)
    n = 300
    ic = {"wealth": rng.uniform(1, 400, n), "regime_id": np.zeros(n, dtype=int), "age": np.zeros(n)}
    res = _compare(m, p, ic, np.float64, seed=11, rtol=1e-10)
    df = res.to_dataframe()
    assert list(df.columns[:4]) == ["subject_id", "period", "regime_name", "value"]
    assert set(df["regime_name"].unique()) <= {"working_life", "retirement", "dead"}
    # every row of an alive regime with a feasible action (finite value)
    # satisfies the borrowing constraint and has a finite utility (additional
    # targets evaluated at the panel); rows with no feasible action hold -inf
    tg = res.to_dataframe(additional_targets=["borrowing_constraint", "utility"], use_labels=False)
    ok = (tg["regime_name"] != list(m.regime_names).index("dead")) & np.isfinite(tg["value"])
    assert ok.sum() > 0
    assert (tg.loc[ok, "borrowing_constraint"] == 1).all()
    assert np.isfinite(tg.loc[ok, "utility"].astype(float)).all()


def test_simulate_stochastic_f64(native):
    T = 7
    m = ex.retirement_model(T, n_wealth=40, n_consumption=50, income_nodes=5)
    p = ex.retirement_params(T, income={"rho": 0.9, "sigma": 0.2, "mu": 0.0})
    rng = np.random.default_rng(1)
    n = 257
    ic = {"wealth": rng.uniform(1, 200, n), "income": rng.normal(0, 0.3, n),
          "regime_id": rng.integers(0, 2, n), "age": np.zeros(n)}
    _compare(m, p, ic, np.float64, seed=5, rtol=1e-9)


@pytest.mark.parametrize("dist", ["rouwenhorst", "normal"])
def test_simulate_processes_model_f64(native, dist):
    """Markov health (discrete draws) and an income process."""
    m = tm.processes_model(5, dist)
    p = tm.processes_params(dist)
    rng = np.random.default_rng(2)
    n = 200
    ic = {"wealth": rng.uniform(1, 5, n), "income": rng.normal(0, 0.5, n),
          "health": rng.integers(0, 2, n), "regime_id": np.zeros(n, dtype=int), "age": np.zeros(n)}
    _compare(m, p, ic, np.float64, seed=9, rtol=1e-9)


def test_simulate_ev_rows_from_global_memory_f64(native):
    """7 income nodes x 3000 wealth points in float64: the EV rows (168 KB) do
    not fit in LDS, so the kernel reads them in place from global memory; the
    paths equal the oracle's."""
    from pylcm_amd import _native as nat
    from pylcm_amd._simulate import SIM_LDS_LIMIT

    T = 4
    m = ex.retirement_model(T, n_wealth=3000, n_consumption=60, income_nodes=7)
    p = ex.retirement_params(T, income={"rho": 0.9, "sigma": 0.2, "mu": 0.0})
    assert nat.load_library().lcm_sim_lds_bytes(nat.LCM_DTYPE_F64, 3000, 7, 11) > SIM_LDS_LIMIT
    rng = np.random.default_rng(4)
    n = 150
    ic = {"wealth": rng.uniform(1, 400, n), "income": rng.normal(0, 0.3, n),
          "regime_id": rng.integers(0, 2, n), "age": np.zeros(n)}
    _compare(m, p, ic, np.float64, seed=9, rtol=1e-9)


def test_simulate_ev_rows_staged_in_lds_f64(native):
    """7 income nodes x 1000 wealth points in float64: one cell's EV rows
    (56 KB) exceed a CU's L1 but fit in LDS, so the kernel stages them; the
    paths equal the oracle's. (Smaller cells are read in place.)"""
    from pylcm_amd._simulate import SIM_EV_L1_BYTES, SIM_LDS_LIMIT

    assert SIM_EV_L1_BYTES < 7 * 1000 * 8 < SIM_LDS_LIMIT
    T = 4
    m = ex.retirement_model(T, n_wealth=1000, n_consumption=60, income_nodes=7)
    p = ex.retirement_params(T, income={"rho": 0.9, "sigma": 0.2, "mu": 0.0})
    rng = np.random.default_rng(6)
    n = 150
    ic = {"wealth": rng.uniform(1, 400, n), "income": rng.normal(0, 0.3, n),
          "regime_id": rng.integers(0, 2, n), "age": np.zeros(n)}
    _compare(m, p, ic, np.float64, seed=10, rtol=1e-9)


@pytest.mark.parametrize("kind", ["log", "irregular"])
def test_simulate_nonlinear_wealth_grids_f64(native, kind):
    """Log-spaced and irregular next-state grids: the simulation's coordinate
    goes through log / a bracket search; paths equal the oracle's."""
    m = tm.log_grid_model(5) if kind == "log" else tm.irregular_grid_model(5)
    p = (ex.retirement_params(5, income={"rho": 0.8, "sigma": 0.1, "mu": 0.0}) if kind == "log"
         else ex.retirement_params(5))
    rng = np.random.default_rng(12)
    n = 200
    ic = {"wealth": rng.uniform(1, 400, n), "regime_id": rng.integers(0, 2, n), "age": np.zeros(n)}
    if kind == "log":
        ic["income"] = rng.normal(0, 0.2, n)
    _compare(m, p, ic, np.float64, seed=13, rtol=1e-9)


def test_simulate_refuses_state_dependent_utility(native):
    """The simulation evaluates each action's tables at the agent's own state
    only where utility does not read the continuous state; a model whose
    utility does is refused, not approximated."""
    from pylcm_amd.exceptions import UnsupportedModelError

    m = tm.nonaffine_model(4)
    p = ex.retirement_params(4, log_utility=True)
    ic = {"wealth": np.array([10.0, 50.0]), "regime_id": np.zeros(2, dtype=int), "age": np.zeros(2)}
    with pytest.raises(UnsupportedModelError):
        m.simulate(params=p, initial_conditions=ic, period_to_regime_to_V_arr=None, log_level="off", seed=1,
                   dtype=np.float64)


def test_simulate_large_panel_f32(native):
    """100k agents of the stochastic model: runs, first-period value equals the
    solve's V where the agent starts on the grid, income moments follow AR(1)."""
    T = 10
    m = ex.retirement_model(T, n_wealth=100, n_consumption=100, income_nodes=7)
    inc = {"rho": 0.95, "sigma": 0.1, "mu": 0.0}
    p = ex.retirement_params(T, income=inc)
    n = 100_000
    rng = np.random.default_rng(3)
    nodes, P = og.process_nodes_matrix(m.regimes["working_life"].states["income"], inc)
    pi = np.array([1, 6, 15, 20, 15, 6, 1]) / 64
    z0 = nodes[rng.choice(7, size=n, p=pi)].astype(np.float32)
    w0 = og.linspace(1, 400, 100, np.float32)[rng.integers(0, 100, n)]
    ic = {"wealth": w0, "income": z0, "regime_id": np.zeros(n, dtype=int)}
    V, _ = m.solve(params=p, log_level="off", dtype=np.float32, return_policy=True)
    res = m.simulate(params=p, initial_conditions=ic, period_to_regime_to_V_arr=V, log_level="off", seed=2026,
                     dtype=np.float32)
    raw = res.raw
    assert np.all(raw["action"][0] >= 0)
    wi = np.rint((w0 - 1) / (399 / 99)).astype(int)
    zi = np.searchsorted(nodes.astype(np.float32), z0)
    assert np.allclose(raw["value"][0], V[0]["working_life"][zi, wi], rtol=1e-6)
    z1 = raw["z"][1].astype(np.float64)
    assert abs(z1.mean() - 0.95 * z0.mean()) < 0.005


def test_simulate_taste_shocks_f64(native):
    """EV1 taste shocks on the work/retire choice: the device's Gumbel-max
    draws equal the oracle's (same Philox stream), so paths are identical."""
    T = 6
    m = tm.taste_shock_model(T)
    p = ex.retirement_params(T, income={"rho": 0.8, "sigma": 0.1, "mu": 0.0})
    p["working_life"]["taste_shocks"] = {"scale": 0.3}
    rng = np.random.default_rng(4)
    n = 300
    ic = {"wealth": rng.uniform(1, 400, n), "income": rng.normal(0, 0.2, n),
          "regime_id": np.zeros(n, dtype=int), "age": np.zeros(n)}
    res = _compare(m, p, ic, np.float64, seed=21, rtol=1e-9)
    # both discrete choices occur (the draw is not the plain argmax)
    d = res.raw["action"][0] // 40
    assert 0 < d.mean() < 1


def test_simulate_taste_shock_choice_frequencies_f32(native):
    """Many agents at one state: the share choosing each discrete action is
    the logit of the cells' continuous maxima (McFadden), computed from the
    oracle's Q at that state."""
    T = 4
    m = tm.taste_shock_model(T)
    p = ex.retirement_params(T, income={"rho": 0.8, "sigma": 0.1, "mu": 0.0})
    scale = 0.2
    p["working_life"]["taste_shocks"] = {"scale": scale}
    V, _ = og.solve(m, p, grid_dtype=np.float32)
    n = 100_000
    w0 = np.full(n, 37.0)
    z0 = np.zeros(n)
    ic = {"wealth": w0, "income": z0, "regime_id": np.zeros(n, dtype=int), "age": np.zeros(n)}
    res = m.simulate(params=p, initial_conditions=ic, period_to_regime_to_V_arr=V, log_level="off", seed=3,
                     dtype=np.float32)
    d = res.raw["action"][0] // 40
    freq = np.array([np.mean(d == 0), np.mean(d == 1)])
    one = {k: v[:1] for k, v in ic.items()}
    R = {nm: og._R(nm, m.regimes[nm], np.float32, p, m.fixed_params) for nm in m.regime_id_class.__categories__}
    Q, F = _period0_q(m, R, p, V, one)
    Qc = np.where(F, Q, -np.inf).reshape(2, -1).max(axis=1)
    logit = np.exp((Qc - Qc.max()) / scale) / np.exp((Qc - Qc.max()) / scale).sum()
    print(f"frequencies {freq}, logit {logit}")
    assert np.max(np.abs(freq - logit)) < 0.01


def _period0_q(m, R, p, V, one):
    """Q and feasibility over (d, c) for one agent in period 0 (oracle)."""
    return osim._regime_period(m, R, R["working_life"], 0, p, m.fixed_params, np.float32, V,
                               np.asarray(one["wealth"], float), np.asarray(one["income"], float),
                               np.zeros(1, dtype=np.int64), return_q=True)


def test_simulate_dcegm_model_f64(native):
    """DC-EGM regimes simulate by the argmax over their own action grids with
    the intrinsic budget mask, from the DC-EGM value functions (the
    reference's path for every shipped envelope): paths equal the oracle's
    simulation of that grid-argmax view on the same V, and every feasible row
    keeps consumption <= wealth - savings_grid[0]."""
    from pylcm_amd._simulate import grid_argmax_view

    T = 5
    m = ex.dcegm_retirement_model(T, n_savings=150, savings_max=100.0, n_wealth=70,
                                  wealth_range=(0.05, 100.0), n_consumption=90)
    p = ex.dcegm_params(T)
    V = m.solve(params=p, log_level="off", dtype=np.float64)
    rng = np.random.default_rng(6)
    n = 200
    ic = {"wealth": rng.uniform(0.5, 95, n), "income": rng.normal(0, 0.1, n),
          "regime_id": rng.integers(0, 2, n), "age": np.zeros(n)}
    res = _compare(m, p, ic, np.float64, seed=8, rtol=1e-9, V=V, oracle_model=grid_argmax_view(m))
    df = res.to_dataframe(use_labels=False)
    alive = df["regime_name"] != list(m.regime_names).index("dead")
    ok = alive & np.isfinite(df["value"])
    assert ok.sum() > 0
    assert (df.loc[ok, "consumption"] <= df.loc[ok, "wealth"] + 1e-12).all()


def test_simulate_stochastic_f32_stepwise_against_oracle(native):
    """The float32 production path against the FP64 oracle, one period at a
    time from the device's own panel (so one near-tie does not send the rest
    of a path elsewhere): at every simulated (agent, period) the chosen action
    is feasible and is the oracle's first argmax of Q at the agent's state, or
    a near-tie within 8 float32 ulps of the maximum (flagged, reported);
    the recorded value is within 1e-5 relative of the oracle's maximum."""
    T = 6
    m = ex.retirement_model(T, n_wealth=60, n_consumption=80, income_nodes=5)
    inc = {"rho": 0.9, "sigma": 0.15, "mu": 0.0}
    p = ex.retirement_params(T, income=inc)
    rng = np.random.default_rng(12)
    n = 160
    ic = {"wealth": rng.uniform(1, 400, n), "income": rng.normal(0, 0.3, n),
          "regime_id": np.zeros(n, dtype=int), "age": np.zeros(n)}
    V, _ = m.solve(params=p, log_level="off", dtype=np.float32, return_policy=True)
    res = m.simulate(params=p, initial_conditions=ic, period_to_regime_to_V_arr=V, log_level="off", seed=77,
                     dtype=np.float32)
    raw = res.raw
    V64 = {t: {r: np.asarray(v, dtype=np.float64) for r, v in rr.items()} for t, rr in V.items()}
    names = list(m.regime_id_class.__categories__)
    R = {nm: og._R(nm, m.regimes[nm], np.float32, p, m.fixed_params) for nm in names}
    checked = ties = 0
    worst = 0.0
    for t in range(T):
        for k, r in enumerate(names):
            idx = np.nonzero((raw["regime"][t] == k) & (raw["action"][t] >= 0))[0]
            if idx.size == 0:
                continue
            reg = R[r]
            w = raw["w"][t, idx].astype(np.float64)
            z = raw["z"][t, idx].astype(np.float64)
            sdd = raw["sdd"][t, idx].astype(np.int64)
            Q, F = osim._regime_period(m, R, reg, t, p, m.fixed_params, np.dtype(np.float32), V64, w, z, sdd,
                                       agents=idx, seed=77, return_q=True)
            Qa = np.where(F, Q, -np.inf).reshape(idx.size, -1)
            Fa = np.broadcast_to(F, Q.shape).reshape(idx.size, -1)
            a_o = np.argmax(Qa, axis=1)
            q_max = Qa[np.arange(idx.size), a_o]
            a_g = raw["action"][t, idx]
            assert np.all(Fa[np.arange(idx.size), a_g]), (t, r)
            q_g = Qa[np.arange(idx.size), a_g]
            ulp = np.spacing(np.abs(q_max).astype(np.float32)).astype(np.float64)
            near = (a_g != a_o) & (q_max - q_g <= 8 * ulp)
            assert np.all((a_g == a_o) | near), (t, r, int(((a_g != a_o) & ~near).sum()))
            ties += int(near.sum())
            fin = np.isfinite(q_max)
            v_g = raw["value"][t, idx].astype(np.float64)
            assert np.array_equal(np.isfinite(v_g), fin)
            if fin.any():
                worst = max(worst, float(np.max(np.abs(v_g[fin] - q_max[fin]) / np.abs(q_max[fin]))))
            checked += idx.size
    print(f"{checked} agent-periods checked, {ties} flagged near-ties, max rel value err {worst:.2e}")
    assert checked > n * (T - 2)
    assert worst <= 1e-5


def test_simulation_graph_replays_equal_eager(native):
    """The simulation captured as a HIP graph (EV launch + every period's
    kernels) gives the eager panel on every replay, with eager runs between."""
    from pylcm_amd._build import resolve_grids
    from pylcm_amd._gs_program import GsProgram
    from pylcm_amd._simulate import SimProgram, _initial_conditions

    T = 8
    m = ex.retirement_model(T, n_wealth=120, n_consumption=120, income_nodes=7)
    p = ex.retirement_params(T, income={"rho": 0.95, "sigma": 0.1, "mu": 0.0})
    prog = GsProgram(m, [p], np.float32)
    prog.run(use_graph=False)
    rg = resolve_grids(m.layouts, p, m.fixed_params, np.float32)
    nodes = np.asarray(rg.proc[("working_life", "income")][0], dtype=np.float64)
    rng = np.random.default_rng(8)
    n = 5000
    ic = _initial_conditions(m, {"wealth": rng.uniform(1, 400, n), "income": nodes[rng.integers(0, 7, n)],
                                 "regime_id": np.zeros(n, dtype=int)})
    sim = SimProgram(m, p, prog)
    sim.prepare(ic, 99)
    sim.launch()
    prog.ctx.sync()
    eager = sim.collect(ic).raw
    for _ in range(3):
        sim.launch_graph()
        prog.ctx.sync()
        got = sim.collect(ic).raw
        for k in eager:
            assert np.array_equal(got[k], eager[k], equal_nan=True), k
        sim.launch()  # eager work between the replays
        prog.ctx.sync()
