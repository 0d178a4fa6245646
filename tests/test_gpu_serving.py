"""GPU tests of the continuous-batching solve server: every request's value
functions are bit-identical whether it is solved alone or in the mix, and
admission respects the device-memory budget."""

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


def _alone(m, p, tile):
    prog = GsProgram(m, [p], np.float32, tile=tile)
    prog.run(use_graph=False)
    prog.fetch()
    return {t: {r: v for r, (v, _) in rr.items()} for t, rr in prog.results(0).items()}


def test_results_same_alone_and_in_the_mix(native):
    from pylcm_amd.serving import SolveServer

    reqs = _requests(24)
    server = SolveServer(tile=128, fetch="all", budget_bytes=1 << 30)
    handles = [server.submit(m, p) for m, p in reqs[:12]]
    # a few steps with the first wave in flight, then more arrivals join
    for _ in range(3):
        server.step()
    handles += [server.submit(m, p) for m, p in reqs[12:]]
    server.drain()
    server.close()
    assert server.launches > 0
    for h, (m, p) in zip(handles, reqs):
        assert h.error is None, h.error
        ref = _alone(m, p, 128)
        for t in ref:
            for r in ref[t]:
                assert np.array_equal(h.result[t][r], ref[t][r]), (h.rid, t, r)


def test_admission_by_memory_budget(native):
    from pylcm_amd.serving import SolveServer

    reqs = _requests(10, seed=3, max_grid=300)
    probe = GsProgram.of([reqs[0]], np.float32, tile=128, defer=True)
    budget = 3 * probe.total_bytes
    server = SolveServer(tile=128, fetch="first", budget_bytes=budget)
    handles = [server.submit(m, p) for m, p in reqs]
    peak = 0
    while server.pending():
        server.step()
        peak = max(peak, server.slab.used)
        if not server.inflight and server.finishing:
            server._retire(block=True)
    server.drain()
    server.close()
    assert peak <= server.slab.size
    ok = [h for h in handles if h.error is None]
    too_big = [h for h in handles if h.error is not None]
    assert all(isinstance(h.error, MemoryError) for h in too_big)
    assert len(ok) >= 3
    for h in ok:
        m, p = reqs[h.rid]
        ref = _alone(m, p, 128)
        for r in ref[0]:
            assert np.array_equal(h.result[0][r], ref[0][r])


def test_worker_process_lowering_same_results(native):
    """Requests lowered in spawned worker processes give the same value
    functions, bit for bit, as each request solved alone."""
    from pylcm_amd.serving import SolveServer

    reqs = _requests(8, seed=5, max_grid=300)
    server = SolveServer(tile=128, fetch="all", budget_bytes=1 << 30, lower_procs=2)
    handles = [server.submit(m, p) for m, p in reqs]
    server.drain()
    server.close()
    for h, (m, p) in zip(handles, reqs):
        assert h.error is None, h.error
        ref = _alone(m, p, 128)
        for t in ref:
            for r in ref[t]:
                assert np.array_equal(h.result[t][r], ref[t][r]), (h.rid, t, r)


def test_tile_chosen_per_launch_same_results(native):
    """tile=0: each fused launch runs the tile the autotuner table gives its
    shape (here a table that sends small launches to 64-state, middle ones to
    128-state and large ones to 256-state tasks); every request equals its
    solve alone with any tile."""
    from pylcm_amd.serving import SolveServer

    table = {"entries": [
        {"dtype": "f32", "units": 2, "n_actions": 600, "tile": 64},
        {"dtype": "f32", "units": 12, "n_actions": 600, "tile": 128},
        {"dtype": "f32", "units": 48, "n_actions": 600, "tile": 256},
    ]}
    reqs = _requests(20, seed=9, max_grid=600)
    server = SolveServer(tile=0, fetch="all", budget_bytes=1 << 30, tile_table=table)
    handles = [server.submit(m, p) for m, p in reqs[:1]]
    for _ in range(3):
        server.step()
    handles += [server.submit(m, p) for m, p in reqs[1:4]]
    for _ in range(2):
        server.step()
    handles += [server.submit(m, p) for m, p in reqs[4:]]
    server.drain()
    server.close()
    assert len(server.tiles_used) >= 2, server.tiles_used
    for h, (m, p) in zip(handles, reqs):
        assert h.error is None, h.error
        ref = _alone(m, p, 256)
        for tile in (64, 128):
            other = _alone(m, p, tile)
            for t in ref:
                for r in ref[t]:
                    assert np.array_equal(other[t][r], ref[t][r]), (h.rid, tile, t, r)
        for t in ref:
            for r in ref[t]:
                assert np.array_equal(h.result[t][r], ref[t][r]), (h.rid, t, r)
