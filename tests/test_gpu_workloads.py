"""The batched and serving workloads at their real shapes.

* Batched solve: draws of the stochastic retirement model with wealth =
  consumption grids of 200, 500 and 1000 points (7 income nodes, T=40, float32)
  in ONE `GsProgram` (one fused launch per period), replayed through the
  captured HIP graph. Each draw against the C restatement of the oracle: V
  within 1e-5 relative of the float64 solve, and policies re-checked period by
  period from the GPU's own V_{t+1}: no mismatch beyond 8 ulp, and on sampled
  periods no more departures beyond one ulp than the reference's own float32
  arithmetic makes on the same states (`oracle.gridsearch.reference_form_q`).
* Serving: requests of the serving workload's shapes (horizons up to 60, grids
  up to 4000, half with the income shock) submitted to the server in waves;
  each result is bit-identical to the request solved alone, and a small one
  matches the oracle.
"""

import numpy as np
import pytest

from oracle import cpu_solve
from oracle import gridsearch as og
from pylcm_amd import examples as ex
from pylcm_amd._gs_program import GsProgram
from tests.parity import compare_solution, rel_err

pytestmark = pytest.mark.gpu


def _draws():
    rng = np.random.default_rng(20261204)
    out = []
    for size in (200, 500, 1000, 200, 500):
        p = ex.retirement_params(
            40, discount_factor=float(rng.uniform(0.90, 0.99)), crra=float(rng.uniform(1.1, 3.0)),
            income={"rho": float(rng.uniform(0.8, 0.97)), "sigma": float(rng.uniform(0.05, 0.2)), "mu": 0.0})
        out.append((size, p))
    return out


def test_batched_uneven_grids_under_graph_match_oracle(native):
    draws = _draws()
    models = {n: ex.retirement_model(40, n_wealth=n, n_consumption=n, income_nodes=7) for n in (200, 500, 1000)}
    prog = GsProgram.of([(models[s], p) for s, p in draws], np.float32)
    prog.run(use_graph=False)
    eager = prog.fetch().copy()
    for _ in range(3):  # graph replays give the eager bits every time
        prog.out_host[:] = 0
        prog.run(use_graph=True)
        assert np.array_equal(prog.fetch(), eager)
    print(f"{prog.n_desc} descriptors, tile {prog.tile}, steps {prog.n_steps}")
    for i, (s, p) in enumerate(draws[:3]):
        res = prog.results(i)
        Vg = {t: {r: v for r, (v, _) in rr.items()} for t, rr in res.items()}
        Pg = {t: {r: q for r, (_, q) in rr.items()} for t, rr in res.items()}
        Vc, _ = cpu_solve.solve_c(models[s], p, grid_dtype=np.float32, threads=16)
        worst = max(rel_err(Vg[t][r], Vc[t][r]) for t in Vc for r in Vc[t])
        Vg64 = {t: {r: np.asarray(v, dtype=np.float64) for r, v in rr.items()} for t, rr in Vg.items()}
        _, _, rep = cpu_solve.solve_c(models[s], p, grid_dtype=np.float32, threads=16, next_V=Vg64, cmp_policy=Pg,
                                      reference_periods=(0, 13, 26, 38))
        print(f"draw {i} (grid {s}): max rel V err {worst:.2e}, {rep}")
        assert worst <= 1e-5
        assert rep["mismatches"] == 0
        # beyond one ulp, the GPU departs from the float64 argmax no more often
        # than the reference's own float32 arithmetic does on the same states
        assert rep["compared_gt_1ulp_sampled"] <= rep["reference_f32_gt_1ulp_sampled"]
        # the batch equals the draw solved alone
        alone = GsProgram(models[s], [p], np.float32)
        alone.run(use_graph=False)
        alone.fetch()
        ra = alone.results(0)
        for t in ra:
            for r in ra[t]:
                assert np.array_equal(ra[t][r][0], res[t][r][0]), (i, t, r)
                assert np.array_equal(ra[t][r][1], res[t][r][1]), (i, t, r)


def _serving_requests():
    make = ex.serving_request_factory(seed=31)
    reqs = [make(i) for i in range(10)]
    # the workload's extremes: the longest horizon with the largest grid, with
    # and without the income shock, and a small request checked by the oracle
    reqs.append((ex.retirement_model(60, n_wealth=4000, n_consumption=4000, income_nodes=7),
                 ex.retirement_params(60, discount_factor=0.95, crra=2.0,
                                      income={"rho": 0.9, "sigma": 0.1, "mu": 0.0})))
    reqs.append((ex.retirement_model(60, n_wealth=4000, n_consumption=4000),
                 ex.retirement_params(60, discount_factor=0.97, crra=1.5)))
    reqs.append((ex.retirement_model(20, n_wealth=200, n_consumption=200, income_nodes=7),
                 ex.retirement_params(20, discount_factor=0.93, crra=2.5,
                                      income={"rho": 0.85, "sigma": 0.15, "mu": 0.0})))
    return reqs


def test_serving_real_shapes_alone_and_in_the_mix(native):
    from pylcm_amd.serving import SolveServer

    reqs = _serving_requests()
    server = SolveServer(tile=0, fetch="all", budget_bytes=8 << 30)
    handles = [server.submit(m, p) for m, p in reqs[:6]]
    for _ in range(4):
        server.step()
    handles += [server.submit(m, p) for m, p in reqs[6:]]
    server.drain()
    server.close()
    shapes = [(m.n_periods, m.layouts["working_life"].n_w) for m, _ in reqs]
    print(f"{len(reqs)} requests, (T, grid) = {shapes}, {server.launches} fused launches")
    for h, (m, p) in zip(handles, reqs):
        assert h.error is None, h.error
        alone = GsProgram(m, [p], np.float32, tile=256)
        alone.run(use_graph=False)
        alone.fetch()
        ref = {t: {r: v for r, (v, _) in rr.items()} for t, rr in alone.results(0).items()}
        for t in ref:
            for r in ref[t]:
                assert np.array_equal(h.result[t][r], ref[t][r]), (h.rid, t, r)
    m, p = reqs[-1]
    V, P = m.solve(params=p, log_level="off", dtype=np.float32, return_policy=True)
    Vo, Po, Qo = og.solve(m, p, grid_dtype=np.float32, return_Q=True)
    worst, mism, ties = compare_solution(V, P, Vo, Po, Qo, np.float32, model=m, params=p)
    print(f"small request vs oracle: max rel V err {worst:.2e}, mismatches {mism}, ties {ties}")
    assert worst <= 1e-5 and mism == 0
    assert all(np.array_equal(handles[-1].result[t][r], V[t][r]) for t in V for r in V[t])
