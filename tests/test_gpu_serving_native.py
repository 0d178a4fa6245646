"""GPU tests of the native solve server (the serving loop in library
threads): every request's value functions equal its plain solve bit for bit
whatever else is in flight, lane count or per-launch tile; admission respects
the device-memory budget; requests that cannot run fail alone."""

import numpy as np
import pytest

from pylcm_amd import examples as ex
from pylcm_amd._gs_program import GsProgram

pytestmark = pytest.mark.gpu


def _requests(n, seed=0, max_grid=400):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        stoch = bool(rng.integers(0, 2))
        T = int(rng.integers(4, 12))
        g = int(rng.integers(30, max_grid))
        m = ex.retirement_model(T, n_wealth=g, n_consumption=int(rng.integers(30, max_grid)),
                                income_nodes=5 if stoch else None)
        p = ex.retirement_params(T, discount_factor=float(rng.uniform(0.9, 0.99)),
                                 crra=float(rng.uniform(1.1, 3.0)),
                                 income={"rho": float(rng.uniform(0.8, 0.97)),
                                         "sigma": float(rng.uniform(0.05, 0.2)), "mu": 0.0} if stoch else None)
        out.append((m, p))
    return out


def _alone(m, p, tile, dtype=np.float32):
    prog = GsProgram(m, [p], dtype, tile=tile)
    prog.run(use_graph=False)
    prog.fetch()
    return {t: {r: v for r, (v, _) in rr.items()} for t, rr in prog.results(0).items()}


def _check(handles, reqs, tile=256, periods="all", dtype=np.float32):
    for h, (m, p) in zip(handles, reqs):
        assert h.error is None, (h.rid, h.error)
        ref = _alone(m, p, tile, dtype)
        keys = sorted(ref) if periods == "all" else [0]
        assert sorted(h.result) == keys, (h.rid, sorted(h.result))
        for t in keys:
            assert sorted(h.result[t]) == sorted(ref[t])
            for r in ref[t]:
                assert np.array_equal(h.result[t][r], ref[t][r]), (h.rid, t, r)


@pytest.mark.parametrize("lanes,order", [(1, "admission"), (1, "longest"), (2, "longest")])
def test_native_results_same_alone_and_in_the_mix(native, lanes, order):
    from pylcm_amd.serving import NativeSolveServer

    reqs = _requests(24, seed=lanes)
    server = NativeSolveServer(tile=128, fetch="all", budget_bytes=1 << 30, lower_procs=0, lower_threads=2,
                               lanes=lanes, order=order)
    try:
        handles = [server.submit(m, p) for m, p in reqs[:12]]
        handles[0].done.wait(60)
        handles += [server.submit(m, p) for m, p in reqs[12:]]
        server.drain(120)
        stats = server.stats()
    finally:
        server.close()
    assert stats["launches"] > 0 and stats["retired"] == len(reqs)
    assert stats["bytes_used"] == 0
    _check(handles, reqs, tile=128)


def test_native_worker_processes_and_first_period(native):
    """Requests lowered in worker processes; only period 0 comes back."""
    from pylcm_amd.serving import NativeSolveServer

    reqs = _requests(10, seed=5, max_grid=300)
    server = NativeSolveServer(tile=256, fetch="first", budget_bytes=1 << 30, lower_procs=2)
    try:
        handles = [server.submit(m, p) for m, p in reqs]
        server.drain(120)
    finally:
        server.close()
    _check(handles, reqs, periods="first")
    assert all(h.t_submit <= h.t_lowered <= h.t_admit <= h.t_retired <= h.t_done for h in handles)


def test_native_tile_chosen_per_launch(native):
    """tile=0 with a table that sends small launches to 64-state, middle ones
    to 128-state and large ones to 256-state tasks: several tiles run and
    every request equals its plain solve."""
    from pylcm_amd.serving import NativeSolveServer

    table = {"entries": [
        {"dtype": "f32", "units": 2, "n_actions": 600, "tile": 64},
        {"dtype": "f32", "units": 12, "n_actions": 600, "tile": 128},
        {"dtype": "f32", "units": 48, "n_actions": 600, "tile": 256},
    ]}
    reqs = _requests(20, seed=9, max_grid=600)
    server = NativeSolveServer(tile=0, fetch="all", budget_bytes=1 << 30, tile_table=table, lower_procs=0)
    try:
        handles = [server.submit(m, p) for m, p in reqs[:1]]
        handles[0].done.wait(60)
        handles += [server.submit(m, p) for m, p in reqs[1:4]]
        handles += [server.submit(m, p) for m, p in reqs[4:]]
        server.drain(120)
        tiles = server.tiles_used
    finally:
        server.close()
    assert len(tiles) >= 2, tiles
    _check(handles, reqs)


def test_native_float64(native):
    from pylcm_amd.serving import NativeSolveServer

    reqs = _requests(6, seed=11, max_grid=250)
    server = NativeSolveServer(dtype=np.float64, tile=0, fetch="all", budget_bytes=1 << 30, lower_procs=0)
    try:
        handles = [server.submit(m, p) for m, p in reqs]
        server.drain(120)
    finally:
        server.close()
    _check(handles, reqs, dtype=np.float64)


def test_native_admission_by_memory_budget(native):
    """A budget of three requests: the placers wait for memory, the peak stays
    within the budget, a request larger than the whole budget fails alone."""
    from pylcm_amd.serving import NativeSolveServer

    reqs = _requests(10, seed=3, max_grid=300)
    probe = GsProgram.of([reqs[0]], np.float32, tile=256, defer=True)
    budget = 3 * probe.total_bytes
    server = NativeSolveServer(tile=256, fetch="first", budget_bytes=budget, lower_procs=0)
    try:
        handles = [server.submit(m, p) for m, p in reqs]
        server.drain(120)
        stats = server.stats()
    finally:
        server.close()
    assert stats["bytes_peak"] <= stats["bytes_budget"]
    ok = [h for h in handles if h.error is None]
    too_big = [h for h in handles if h.error is not None]
    assert len(ok) >= 3
    assert all("device memory" in str(h.error) for h in too_big), [str(h.error) for h in too_big]
    _check(ok, [reqs[h.rid] for h in ok], periods="first")


def test_native_request_that_fails_lowering_fails_alone(native):
    """A request whose grid needs more LDS than a compute unit has is refused
    with a clear error; the others complete."""
    from pylcm_amd.serving import NativeSolveServer

    reqs = _requests(4, seed=13, max_grid=200)
    big = (ex.retirement_model(3, n_wealth=12000, n_consumption=20), ex.retirement_params(3))
    server = NativeSolveServer(dtype=np.float64, tile=256, fetch="first", budget_bytes=1 << 30, lower_procs=0)
    try:
        handles = [server.submit(m, p) for m, p in reqs[:2]]
        bad = server.submit(*big)
        handles += [server.submit(m, p) for m, p in reqs[2:]]
        server.drain(120)
    finally:
        server.close()
    assert bad.error is not None and "LDS" in str(bad.error)
    _check(handles, reqs, periods="first", dtype=np.float64)


def _plain_simulation(m, p, spec):
    """The plain path: solve alone, then SimProgram with the same initial
    conditions and seed."""
    from pylcm_amd._simulate import SimProgram, _initial_conditions

    prog = GsProgram(m, [p], np.float32)
    prog.run(use_graph=False)
    n, seed = spec["n_agents"], spec["seed"]
    ic = _initial_conditions(m, spec["initial_conditions"](m, p, prog._rg[0], n, np.random.default_rng(seed)))
    sim = SimProgram(m, p, prog)
    sim.prepare(ic, seed)
    sim.launch()
    prog.ctx.sync()
    return sim.collect_raw()


@pytest.mark.parametrize("procs", [0, 2])
def test_native_forward_simulation_per_request(native, procs):
    """Requests with a forward simulation are held on the device after their
    solve, simulated on their resident value functions and released: every
    panel equals the plain solve-then-simulate path bit for bit, requests
    without one are unaffected, and no device memory stays held."""
    from pylcm_amd.serving import NativeSolveServer

    reqs = _requests(8, seed=21, max_grid=300)
    specs = [{"n_agents": 2000 + 500 * i, "seed": 100 + i, "initial_conditions": ex.serving_initial_conditions}
             if i % 3 != 2 else None for i in range(len(reqs))]
    server = NativeSolveServer(tile=0, fetch="all", budget_bytes=1 << 30, lower_procs=procs, lanes=2)
    try:
        handles = [server.submit(m, p, simulate=sp) for (m, p), sp in zip(reqs, specs)]
        server.drain(180)
        stats = server.stats()
    finally:
        server.close()
    assert stats["bytes_used"] == 0 and stats["retired"] == len(reqs)
    _check(handles, reqs)
    for h, (m, p), sp in zip(handles, reqs, specs):
        if sp is None:
            assert h.simulation is None
            continue
        ref = _plain_simulation(m, p, sp)
        assert sorted(h.simulation) == sorted(ref)
        for k in ref:
            assert h.simulation[k].shape == (m.n_periods, sp["n_agents"])
            assert np.array_equal(h.simulation[k], ref[k], equal_nan=True), (h.rid, k)
        assert h.t_retired <= h.t_simulated
