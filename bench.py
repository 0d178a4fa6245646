"""Benchmark: batched vs plain GPU model-solves/sec of the grid-search solve.

Workload ("batched stochastic retirement model"): 1024 parameter draws of the
stochastic work/retire consumption-savings model, T=40, float32, 7 Rouwenhorst
income nodes, wealth = consumption grid sizes drawn from {200, 500, 1000};
beta~U[.90,.99], CRRA~U[1.1,3], sigma~U[.05,.2], AR rho~U[.8,.97]. One step =
one batched backward induction of all draws (one fused launch per period,
replayed as a HIP graph) with the tables already resident in HBM. `value` is
model-solves/sec of the batched path. The same draws solved one at a time
(each its own resident program and graph) give the plain-GPU rate and p50/p99
latency. A CPU restatement (oracle/gs_ref.c, all host cores) is timed on a
bounded sample of the same draws and is also the parity check (V rel err and
argmax policies).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--draws N]
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TRAFFIC_FILE = ROOT / "profiles" / "gs_period_kernel_f32_1024draws_T40_traffic.md"


def committed_traffic(key: str):
    """HBM bytes per launch of the dominant kernel, from the committed summary
    of two counter runs (FETCH_SIZE, WRITE_SIZE) of this same workload, each
    scaled by what the counter reads for a copy kernel of known size. None when
    the summary is absent or was taken on another workload."""
    if not TRAFFIC_FILE.exists():
        return None
    fields = {}
    for line in TRAFFIC_FILE.read_text().splitlines():
        if line.startswith("- ") and ": " in line:
            k, v = line[2:].split(": ", 1)
            fields[k.strip()] = v.strip().strip("`")
    if fields.get("workload_key") != key:
        return None
    try:
        return int(fields["fetch_bytes_per_launch"]) + int(fields["write_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def _host_cores() -> int:
    return len(os.sched_getaffinity(0))


def _pool(procs):
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor

    return ProcessPoolExecutor(max_workers=procs, mp_context=multiprocessing.get_context("spawn"))


def _cpu_dcegm_solve(job):
    """One FP64 oracle DC-EGM solve (CPU baseline worker; no device code)."""
    from oracle import dcegm as od
    from pylcm_amd import examples as ex

    T, beta, crra, sigma = job
    m = ex.dcegm_retirement_model(T)
    p = ex.dcegm_params(T, discount_factor=beta, crra=crra, sigma=sigma)
    t0 = time.perf_counter()
    od.solve(m, p, grid_dtype=np.float32)
    return time.perf_counter() - t0


def _cpu_simulate(job):
    """The FP64 oracle simulation of a slice of agents (CPU baseline worker)."""
    from oracle import simulate as osim
    from pylcm_amd import examples as ex

    V, sub, seed = job
    m, p = ex.stochastic_workload()
    t0 = time.perf_counter()
    out = osim.simulate(m, p, V, sub, seed, grid_dtype=np.float32)
    return time.perf_counter() - t0, out["action"]


def dcegm_leg(args):
    """DC-EGM retirement model with FUES: T=50, 2000 exponentially spaced
    savings nodes on [0, 500], 2000 wealth nodes, taste shocks 0.05, log-normal
    income on 5 Gauss-Hermite nodes, CRRA 2, beta 0.96, r 0.03; float32 V with
    FP64 candidate rows. Compared with grid search on a 4000-point consumption
    grid, and with the FP64 oracle on a bounded sample (the same grids, T=3)."""
    from pylcm_amd import _native as nat
    from pylcm_amd import examples as ex
    from pylcm_amd._egm_program import EgmProgram
    from pylcm_amd._gs_program import GsProgram

    dt = np.float32
    T = args.dcegm_periods
    m = ex.dcegm_retirement_model(T)
    p = ex.dcegm_params(T)
    t0 = time.perf_counter()
    prog = EgmProgram(m, p, dt)
    build_s = time.perf_counter() - t0
    prog.run()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        prog.run(sync=False)
    prog.ctx.sync()
    ms_solve_eager = 1e3 * (time.perf_counter() - t0) / reps
    V_eager = prog.value_functions()
    # the same launches replayed as one HIP graph (the timed figure)
    prog.run(use_graph=True)
    t0 = time.perf_counter()
    for _ in range(reps):
        prog.run(sync=False, use_graph=True)
    prog.ctx.sync()
    ms_solve = 1e3 * (time.perf_counter() - t0) / reps
    e0, e1 = nat.Event(), nat.Event()
    e0.record(prog.ctx)
    prog.run(sync=False, use_graph=True)
    e1.record(prog.ctx)
    kernel_ms = e0.elapsed_ms(e1)
    V = prog.value_functions()
    graph_same = all(np.array_equal(V[t][r], V_eager[t][r]) for t in V for r in V[t])
    n_cols = sum(v[1] for v in prog._period.values())
    log(f"dcegm: {ms_solve:.2f} ms per solve as a graph ({ms_solve_eager:.2f} eager, {kernel_ms:.2f} ms on the stream; "
        f"graph equals eager: {graph_same}), {len(prog._period)} periods, "
        f"{n_cols} columns")
    # the grid-search answer on a 4000-point consumption grid
    g = GsProgram(ex.dcegm_retirement_model(T, grid_search=True), [p], dt)
    g.run(use_graph=False)
    g.fetch()
    res = g.results(0)
    diffs = []
    for t in V:
        for r in ("working_life", "retirement"):
            if r in V[t] and r in res.get(t, {}):
                a = np.asarray(V[t][r], np.float64)[..., 2:]
                b = np.asarray(res[t][r][0], np.float64)[..., 2:]
                diffs.append(np.abs(a - b) / np.abs(b))
    d = np.concatenate([x.ravel() for x in diffs])
    out = {
        "workload": "DC-EGM retirement model, T=50, 2000 savings x 2000 wealth nodes, 5 GH income nodes, "
                    "taste shocks 0.05, FUES on every column",
        "ms_per_solve": round(ms_solve, 3),
        "solves_per_s": round(1e3 / ms_solve, 2),
        "ms_per_solve_eager": round(ms_solve_eager, 3),
        "graph_equals_eager": bool(graph_same),
        "stream_ms_per_solve": round(kernel_ms, 3),
        "columns": n_cols,
        "host_build_s": round(build_s, 3),
        "vs_grid_search_4000": {"median_rel_diff": float(np.median(d)), "p99_rel_diff": float(np.percentile(d, 99)),
                                "note": "grid resolution; the two lowest wealth nodes excluded"},
        "columns_per_s": round(n_cols / (ms_solve * 1e-3), 1),
    }
    if args.dcegm_batch > 0:
        from pylcm_amd._egm_program import EgmBatch

        rng = np.random.default_rng(20261203)
        plist = [ex.dcegm_params(T, discount_factor=float(rng.uniform(0.93, 0.98)), crra=float(rng.uniform(1.5, 3.0)),
                                 sigma=float(rng.uniform(0.1, 0.2))) for _ in range(args.dcegm_batch)]
        plist[0] = p
        t0 = time.perf_counter()
        batch = EgmBatch(m, plist, dt, ctx=prog.ctx)
        bbuild = time.perf_counter() - t0
        batch.run()
        batch.run(use_graph=True)
        breps = 3
        t0 = time.perf_counter()
        for _ in range(breps):
            batch.run(sync=False, use_graph=True)
        prog.ctx.sync()
        ms_batch = 1e3 * (time.perf_counter() - t0) / breps
        Vb = batch.value_functions(0)
        same = all(np.array_equal(Vb[t][r], V[t][r]) for t in V for r in V[t])
        out["batch"] = {"draws": args.dcegm_batch, "ms_per_batch": round(ms_batch, 3),
                        "solves_per_s": round(args.dcegm_batch / (ms_batch * 1e-3), 2),
                        "columns_per_launch": sum(pr._period[max(pr._period)][1] for pr in batch.progs),
                        "vs_single": round((args.dcegm_batch / ms_batch) * ms_solve, 2),
                        "host_build_s": round(bbuild, 2), "device_MiB": round(batch.total_bytes / 2**20, 1),
                        "draw0_equals_single_solve": bool(same)}
        out["batch"]["columns_per_s"] = round(out["batch"]["columns_per_launch"] * (T - 1) / (ms_batch * 1e-3), 1)
        log(f"dcegm batch: {args.dcegm_batch} draws in {ms_batch:.2f} ms -> {out['batch']['solves_per_s']:.1f} solves/s "
            f"({out['batch']['vs_single']}x single); draw 0 equals the single solve: {same}")
        del batch
    # the batch's dominant kernel: its VALU issue fraction from the committed
    # counter pass of this workload (latency-bound bracket searches)
    frac = None
    prof = ROOT / "profiles" / "egm_kernels_dcegm_T50_2000x2000_sq_counters.md"
    if prof.exists():
        for line in prof.read_text().splitlines():
            if line.startswith("- valu_issue_frac_euler_batch:"):
                frac = float(line.split(":", 1)[1].strip().strip("`"))
    out["roofline"] = {"kernel": "egm_euler_kernel<256> (the batch's continuation reads + Euler inversion)",
                       "bound": "latency (dependent loads of bracket searches in the child carry rows)",
                       "columns_per_s_single": out["columns_per_s"],
                       "columns_per_s_batch": out.get("batch", {}).get("columns_per_s"),
                       "valu_issue_frac": frac, "source": str(prof.relative_to(ROOT)) if frac else None}
    if not args.no_cpu:
        from oracle import dcegm as od
        from tests.parity import rel_err

        Ts = 3
        ms_ = ex.dcegm_retirement_model(Ts)
        ps = ex.dcegm_params(Ts)
        small = EgmProgram(ms_, ps, dt)
        small.run()
        Vg = small.value_functions()
        t0 = time.perf_counter()
        Vo, Co = od.solve(ms_, ps, grid_dtype=dt)
        cpu_s = time.perf_counter() - t0
        worst_v, worst_rows, same_kept = 0.0, 0.0, True
        for t in Vo:
            for r in Vo[t]:
                worst_v = max(worst_v, rel_err(Vg[t][r], Vo[t][r]))
                if r in Co[t]:
                    cols = small.columns(t, r)
                    for c, oc in enumerate(Co[t][r]):
                        same_kept &= int(cols["counts"][c, 0]) == oc["n_kept"]
                        for k in ("grid", "policy", "value"):
                            worst_rows = max(worst_rows, rel_err(cols[k][c], oc[k]))
        out["parity_T3_full_grids"] = {"max_rel_V_err_vs_fp64": worst_v, "max_rel_carry_row_err": worst_rows,
                                       "same_kept_counts": bool(same_kept)}
        # all host cores: one oracle solve per core, each its own parameter draw
        cores = _host_cores()
        rng = np.random.default_rng(5)
        jobs = [(Ts, float(rng.uniform(0.93, 0.98)), float(rng.uniform(1.5, 3.0)), float(rng.uniform(0.1, 0.2)))
                for _ in range(cores)]
        with _pool(cores) as pool:
            t0 = time.perf_counter()
            list(pool.map(_cpu_dcegm_solve, jobs))
            wall = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(cores / (wall * (T - 1) / (Ts - 1)), 6), "unit": "model-solves/s",
                               "cores": cores, "kind": "port",
                               "sample": f"NumPy FP64 restatement (oracle/dcegm.py) of the same grids with T={Ts}, "
                                         f"{cores} parameter draws in {cores} processes ({Ts - 1} periods each, "
                                         f"{wall:.1f}s wall), scaled to {T - 1} periods; one solve alone took {cpu_s:.1f}s"}
        log(f"dcegm parity (T={Ts}): V {worst_v:.2e}, carry rows {worst_rows:.2e}, kept counts equal {same_kept}; "
            f"cpu {cpu_s:.1f}s")
    return out


def simulation_leg(args):
    """Forward simulation of the stochastic workload (f32, T=40, 500 x 500,
    7 Rouwenhorst nodes): N agents, w0 ~ exp(log 10 + 0.5 N(0,1)), the income
    node drawn from Binomial(6, 1/2), Philox4x32-10 streams per (agent,
    period, draw). The V tables come from the GPU solve and stay resident;
    the timed part is the EV rows plus the 40 period launches."""
    from pylcm_amd import _native as nat
    from pylcm_amd import examples as ex
    from pylcm_amd._gs_program import GsProgram
    from pylcm_amd._simulate import SimProgram, _initial_conditions

    dt = np.float32
    m, p = ex.stochastic_workload()
    T = m.n_periods
    prog = GsProgram(m, [p], dt)
    prog.run(use_graph=False)
    n = args.sim_agents
    rng = np.random.default_rng(20261203)
    from pylcm_amd._build import resolve_grids

    rg = resolve_grids(m.layouts, p, m.fixed_params, dt)
    nodes = np.asarray(rg.proc[("working_life", "income")][0], dtype=np.float64)
    w0 = np.exp(np.log(10.0) + 0.5 * rng.standard_normal(n))
    z0 = nodes[rng.binomial(6, 0.5, n)]
    ic_user = {"wealth": w0, "income": z0, "regime_id": np.zeros(n, dtype=int)}
    ic = _initial_conditions(m, ic_user)
    seed = 20261203
    sim = SimProgram(m, p, prog)
    t0 = time.perf_counter()
    sim.prepare(ic, seed)
    prep_s = time.perf_counter() - t0
    sim.launch()
    prog.ctx.sync()  # warm-up
    reps = max(1, args.steps)
    ev_ms, per_ms, tot_ms = [], [], []
    for _ in range(reps):
        evs = ((nat.Event(), nat.Event()), [(nat.Event(), nat.Event()) for _ in range(T)])
        sim.launch(events=evs)
        prog.ctx.sync()
        ev_ms.append(evs[0][0].elapsed_ms(evs[0][1]))
        per_ms.append([a.elapsed_ms(b) for a, b in evs[1]])
        tot_ms.append(evs[0][0].elapsed_ms(evs[1][-1][1]))
    res = sim.collect(ic)
    raw = res.raw
    eager_tot = float(np.median(tot_ms))
    # the same simulation replayed as one HIP graph (kernel nodes only)
    sim.launch_graph()
    prog.ctx.sync()
    same_graph = all(np.array_equal(sim.collect(ic).raw[k], raw[k], equal_nan=True) for k in raw)
    g_ms = []
    for _ in range(reps):
        sim._reset()
        e0, e1 = nat.Event(), nat.Event()
        e0.record(prog.ctx)
        sim.capture_graph().launch(prog.ctx)
        e1.record(prog.ctx)
        prog.ctx.sync()
        g_ms.append(e0.elapsed_ms(e1))
    tot = float(np.median(g_ms))
    agent_periods = int((raw["action"] >= 0).sum())
    # algorithmic work: n_actions x 10 flops per simulated agent-period
    names = list(m.regime_names)
    n_act = np.array([m.layouts[r].n_d * m.layouts[r].n_c for r in names], dtype=np.float64)
    simulated = raw["action"] >= 0
    flops = 10.0 * float(n_act[raw["regime"][simulated]].sum())
    per = np.median(np.array(per_ms), axis=0)
    sim_ms = float(per.sum())
    # EV rows: reads M and V_{t+1}, writes EV, per (period, regime)
    ev_bytes = 0
    for (t, r), pt in prog.problems[0].items():
        lay = m.layouts[r]
        if pt.terminal or not pt.targets:
            continue
        g_n = len(rg.w_pts[pt.next_grid])
        for _, M in pt.targets:
            ev_bytes += 4 * (M.size + M.shape[-1] * g_n)
        ev_bytes += 4 * lay.n_sd * lay.n_d * g_n
    ev = float(np.median(ev_ms))
    out = {
        "workload": f"{n} agents x {T} periods of the stochastic retirement model (f32, 500 x 500, 7 Rouwenhorst "
                    "nodes), V resident from the GPU solve, Philox4x32-10 draws",
        "agents": n, "periods": T,
        "ms_per_simulation": round(tot, 3),
        "ms_per_simulation_eager": round(eager_tot, 3),
        "graph_equals_eager": bool(same_graph),
        "agent_periods_per_s": round(agent_periods / (tot * 1e-3), 1),
        "host_prepare_s": round(prep_s, 3),
        "roofline": {"kernel": "sim_kernel<float> (+ per-period partition kernels)", "bound": "valu",
                     "achieved": round(flops / (sim_ms * 1e-3) / 1e12, 3), "peak": 157.3, "unit": "TFLOP/s",
                     "frac": round(flops / (sim_ms * 1e-3) / 1e12 / 157.3, 4), "flop_per_agent_action": 10,
                     "period_launch_ms_median": round(float(np.median(per)), 4)},
        "ev_kernel": {"ms": round(ev, 4), "bytes": ev_bytes, "achieved_GBs": round(ev_bytes / (ev * 1e-3) / 1e9, 2),
                      "peak_GBs": 8000, "bound": "hbm"},
    }
    log(f"simulation: {n} agents x {T} periods in {tot:.2f} ms as a graph ({eager_tot:.2f} eager; "
        f"{out['agent_periods_per_s']:.3g} agent-periods/s); "
        f"sim kernels {sim_ms:.2f} ms, {out['roofline']['achieved']} TFLOP/s; EV {ev:.3f} ms")
    if not args.no_cpu:
        k = min(args.sim_cpu_agents, n)
        prog.fetch()
        Vg = {t: {r: v for r, (v, _) in rr.items()} for t, rr in prog.results(0).items()}
        # all host cores: each process simulates its own slice of k agents
        # (the oracle numbers a slice's agents from 0, so slices after the
        # first draw other paths than the device's agents, with the same work)
        cores = _host_cores()
        kk = int(min(k * cores, n))
        bounds = np.linspace(0, kk, cores + 1).astype(int)
        jobs = []
        for a, b in zip(bounds[:-1], bounds[1:]):
            sub = {key: val[a:b] for key, val in ic_user.items()}
            jobs.append((Vg, sub, seed))
        with _pool(cores) as pool:
            t0 = time.perf_counter()
            parts = list(pool.map(_cpu_simulate, jobs))
            cpu_s = time.perf_counter() - t0
        # each slice's agents are numbered from 0 in its own run: compare the
        # first slice (the same agent numbers as the device's first agents)
        exp_action = parts[0][1]
        k0 = int(bounds[1])
        cpu_ap = int(sum((a >= 0).sum() for _, a in parts))
        same = simulated[:, :k0] == (exp_action >= 0)
        both = simulated[:, :k0] & (exp_action >= 0)
        act_mism = int((raw["action"][:, :k0][both] != exp_action[both]).sum())
        k = k0
        out["cpu_baseline"] = {"value": round(cpu_ap / cpu_s, 1), "unit": "agent-periods/s", "cores": cores,
                               "kind": "port",
                               "sample": f"{kk} agents in {cores} processes ({k0} each), NumPy FP64 restatement "
                                         f"oracle/simulate.py, {cpu_s:.1f}s wall"}
        out["parity_sample"] = {"agents": k, "same_simulated_mask": bool(same.all()),
                                "action_mismatches": act_mism, "agent_periods": int(both.sum())}
        log(f"simulation cpu baseline: {out['cpu_baseline']['value']} agent-periods/s; parity {out['parity_sample']}")
    return out


def _serve_window(args, rate, seconds, seed, sim_agents=0):
    from pylcm_amd import examples as ex
    from pylcm_amd.serving import SolveServer, serve_poisson

    import sys as _sys

    old_si = _sys.getswitchinterval()
    if args.serve_switch_ms > 0:
        _sys.setswitchinterval(args.serve_switch_ms * 1e-3)
    if args.serve_loop == "native":
        from pylcm_amd.serving import NativeSolveServer

        server = NativeSolveServer(dtype=np.float32, tile=args.serve_tile, fetch="first",
                                   lower_procs=args.serve_procs, lanes=args.serve_lanes,
                                   queue_depth=args.serve_depth, placers=args.serve_placers,
                                   order=args.serve_order, sim_threads=args.serve_sim_threads)
    else:
        server = SolveServer(dtype=np.float32, tile=args.serve_tile, fetch="first", lower_threads=args.serve_threads,
                             lower_procs=args.serve_procs, lanes=args.serve_lanes, queue_depth=args.serve_depth)
    base_make = ex.serving_request_factory(seed=7)
    if sim_agents > 0:  # every request also simulates sim_agents agents on its solution
        def make(i):
            m, p = base_make(i)
            return m, p, {"n_agents": sim_agents, "seed": 1000 + i,
                          "initial_conditions": ex.serving_initial_conditions}
    else:
        make = base_make
    # the client keeps every request; none of them forms a reference cycle, so
    # the cyclic collector only adds pauses while the window runs
    import gc

    gc.collect()
    gc.disable()
    try:
        t0 = time.perf_counter()
        reqs = serve_poisson(server, make, rate=rate, duration=seconds, seed=seed)
        wall = time.perf_counter() - t0
    finally:
        gc.enable()
    server.close()
    _sys.setswitchinterval(old_si)
    ok = [r for r in reqs if r.error is None]
    lat = np.array([r.latency for r in ok])
    span = max(r.t_done for r in ok) - min(r.t_submit for r in ok)
    out = {"arrival_rate": rate, "seconds": seconds, "requests": len(reqs), "completed": len(ok),
           "solves_per_s": round(len(ok) / span, 2),
           "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 2),
           "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 2),
           "fused_launches": server.launches,
           "gpu_busy_s": round(server.gpu_busy_seconds(), 3),
           "mean_tasks_per_launch": round(server.launch_tasks / max(server.launches, 1), 1),
           "launches_by_tile": {str(k): v for k, v in sorted(server.tiles_used.items())},
           "loop": args.serve_loop, "lanes": args.serve_lanes, "client_gc": "disabled during the window",
           "simulated_agents_per_request": sim_agents,
           "task_order": args.serve_order if args.serve_loop == "native" else "admission"}
    native = args.serve_loop == "native"
    host = server.host_seconds() if native else dict(server.timing)
    out["host_loop_s"] = {k: round(v, 3) for k, v in host.items()}
    # where the latency goes: lowering (worker processes), placement (device
    # memory wait + uploads), and the solve (admission to completion)
    stages = {"lower": [r.t_lowered - r.t_submit for r in ok],
              "place": [r.t_admit - r.t_lowered for r in ok],
              "solve": [(r.t_retired if native else r.t_done) - r.t_admit for r in ok]}
    if native and sim_agents > 0:  # solve retired in the library to the simulated panel in hand
        stages["simulate"] = [r.t_done - r.t_retired for r in ok]
    elif native:  # completion in the library to the result in the client's hands
        stages["collect"] = [r.t_done - r.t_retired for r in ok]
    out["stage_ms"] = {k: {"p50": round(1e3 * float(np.percentile(v, 50)), 2),
                           "p99": round(1e3 * float(np.percentile(v, 99)), 2)} for k, v in stages.items()}
    # capacities of the three resources a request passes through, in
    # requests/s: the lowering processes (their mean time per request in the
    # worker), the device (fused-launch busy time per request), and the
    # serving loop's own host time (admit + launch + retire per request; for
    # the native loop the library lanes' time assembling and issuing launches)
    wk = np.array([r.worker_s for r in ok if r.worker_s > 0])
    busy = out["gpu_busy_s"]
    loop_s = host["launch"] if native else sum(host.values())
    out["capacity_req_per_s"] = {
        "lowering": round(args.serve_procs / float(wk.mean()), 1) if wk.size else None,
        "lowering_worker_ms_mean": round(1e3 * float(wk.mean()), 2) if wk.size else None,
        "device": round(len(ok) / busy, 1) if busy > 0 else None,
        "serving_loop": round(len(ok) / loop_s, 1) if loop_s > 0 else None}
    log(f"serving at {rate:g}/s: {len(ok)}/{len(reqs)} requests in {wall:.1f}s, {out['solves_per_s']:.1f} solves/s, "
        f"p50 {out['p50_ms']:.1f} ms, p99 {out['p99_ms']:.1f} ms, {server.launches} fused launches")
    return out


def serving_leg(args):
    """Poisson arrivals into the continuous-batching server (requests lowered
    in worker processes); requests mix the deterministic and the AR(1) income
    retirement models, horizons U{20..60}, grids U{200..4000}. One window at
    the stated rate (latency), one well above capacity (saturation
    throughput). The same request sequence solved one at a time gives the
    plain baseline."""
    from pylcm_amd import examples as ex
    from pylcm_amd._gs_program import GsProgram

    dt = np.float32
    # plain: the same sequence, one request at a time (lower, place, solve,
    # fetch period 0), timed first so that nothing the windows leave behind
    # (their garbage, worker processes winding down) lands in it
    import gc

    gc.collect()
    make2 = ex.serving_request_factory(seed=7)
    n_plain = args.serve_plain_sample
    plain = []
    for i in range(n_plain):
        m, p = make2(i)
        t_a = time.perf_counter()
        pr = GsProgram(m, [p], dt)
        pr.run(use_graph=False)
        o = pr.outs[0]
        for (t, r), oo in o.items():
            if t == 0:
                n = int(np.prod(oo.shape, dtype=np.int64)) if oo.shape else 1
                v = np.empty(n, dtype=dt)
                from pylcm_amd import _native as nat
                nat.d2h(pr.ctx, pr.base + oo.v_off, v)
        plain.append(time.perf_counter() - t_a)
        del pr
    win = _serve_window(args, args.serve_rate, args.serve_seconds, seed=11,
                        sim_agents=args.serve_sim_agents if args.serve_loop == "native" else 0)
    sat = _serve_window(args, args.serve_sat_rate, args.serve_sat_seconds, seed=12) if args.serve_sat_seconds > 0 else None
    plain = np.array(plain)
    plain_sps = 1.0 / plain.mean()
    return {
        "workload": f"Poisson arrivals {args.serve_rate:g}/s for {args.serve_seconds:g}s; deterministic and AR(1)-income "
                    "retirement models, T U{20..60}, grids U{200..4000}, float32, period-0 V returned; "
                    f"requests lowered in {args.serve_procs} worker processes, {args.serve_loop} serving loop; "
                    f"each request of the {args.serve_rate:g}/s window also forward-simulates "
                    f"{win['simulated_agents_per_request']} agents on its resident solution; the saturation "
                    "window solves only",
        **win,
        "saturation": sat,
        "plain_serial": {"solves_per_s": round(plain_sps, 2),
                         "p50_ms": round(1e3 * float(np.percentile(plain, 50)), 2),
                         "p99_ms": round(1e3 * float(np.percentile(plain, 99)), 2),
                         "sample": n_plain},
        "serving_vs_plain": round(win["solves_per_s"] / plain_sps, 2),
        "saturation_vs_plain": round(sat["solves_per_s"] / plain_sps, 2) if sat else None,
    }


def tiles_leg(draws, models, lowered, ctx):
    """The autotuned tile against every tile and the untuned fallback, for one
    draw of each grid size solved alone (the plain path's launches) and for the
    whole batch: mean per-launch time of the search steps (HIP events)."""
    from pylcm_amd import autotune
    from pylcm_amd._gs_program import GsProgram

    def per_launch(probs, low, tile):
        prog = GsProgram.of(probs, np.float32, lowered=low, tile=tile, ctx=ctx)
        prog.run(use_graph=False)
        best = min(float(np.mean([m for s, m in prog.time_launches() if s > 0])) for _ in range(3))
        return prog, best

    out = {"single": {}}
    for size in sorted({s for s, _ in draws}):
        i = next(k for k, (s, _) in enumerate(draws) if s == size)
        probs, low = [(models[size], draws[i][1])], [lowered[i]]
        tuned, _ = per_launch(probs, low, None)
        units, n_actions = tuned.launch_shape()
        times = {t: per_launch(probs, low, t)[1] for t in autotune.TILES}
        fb = autotune.fallback_tile(units)
        out["single"][str(size)] = {
            "units": units, "tuned_tile": tuned.tile, "fallback_tile": fb,
            "ms_per_launch": {str(t): round(v, 4) for t, v in times.items()},
            "gain_vs_fallback": round(times[fb] / times[tuned.tile], 3)}
        log(f"tiles, one draw of grid {size}: {out['single'][str(size)]}")
    return out


def fresh_leg(args, models, T, tile):
    """Fresh draws: every timed batch holds NEW parameter draws of the batched
    workload, lowered and packed on the host by worker processes and placed
    into pinned host memory (batch k+1 while the GPU solves batch k), then
    uploaded as one image and solved with one fused launch per period."""
    from pylcm_amd._batch import FreshBatches

    n = args.fresh_batches + 1  # the first batch fills the pipeline
    batches = [[(s, p) for s, p in __import__("pylcm_amd.examples", fromlist=["x"]).draw_batch_params(
        args.draws, n_periods=T, seed=5000 + k)] for k in range(n)]
    fb = FreshBatches(models, np.float32, procs=args.fresh_procs, tile=tile, chunk=16)
    stamps = []
    try:
        stats = fb.run(batches, on_done=lambda k, ab: stamps.append(time.perf_counter()))
    finally:
        fb.close()
    rate = args.draws * (n - 1) / (stamps[-1] - stamps[0])
    med = {k: round(float(np.median([st[k] for st in stats[1:]])), 4) for k in stats[0]}
    log(f"fresh draws: {rate:.1f} solves/s over {n - 1} batches of {args.draws} "
        f"({args.fresh_procs} lowering processes); per batch median {med}")
    return {"solves_per_s": round(rate, 2), "batches": n - 1, "draws_per_batch": args.draws,
            "lowering_processes": args.fresh_procs, "per_batch_median_s": med}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--draws", type=int, default=1024)
    ap.add_argument("--periods", type=int, default=40)
    ap.add_argument("--plain-sample", type=int, default=48)
    ap.add_argument("--cpu-sample", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=int, default=None)
    ap.add_argument("--copy-calib-mib", type=int, default=256)
    ap.add_argument("--no-dcegm", action="store_true")
    ap.add_argument("--dcegm-periods", type=int, default=50)
    ap.add_argument("--dcegm-batch", type=int, default=64)
    ap.add_argument("--no-sim", action="store_true")
    ap.add_argument("--sim-agents", type=int, default=100_000)
    ap.add_argument("--sim-cpu-agents", type=int, default=1500)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-sat-rate", type=float, default=2000.0)
    ap.add_argument("--serve-sat-seconds", type=float, default=8.0)
    ap.add_argument("--serve-procs", type=int, default=12)
    ap.add_argument("--serve-plain-sample", type=int, default=40)
    ap.add_argument("--serve-tile", type=int, default=0, help="0: chosen per fused launch")
    ap.add_argument("--serve-lanes", type=int, default=2, help="streams carrying the fused launches")
    ap.add_argument("--serve-threads", type=int, default=4, help="threads placing lowered requests (python loop)")
    ap.add_argument("--serve-loop", choices=("native", "python"), default="native",
                    help="serving loop in library threads (native) or in this process (python)")
    ap.add_argument("--serve-placers", type=int, default=2, help="library threads placing requests (native loop)")
    ap.add_argument("--serve-sim-agents", type=int, default=10_000,
                    help="agents each request of the arrival-rate window simulates (native loop; 0: none)")
    ap.add_argument("--serve-sim-threads", type=int, default=4, help="simulation threads (streams) of the server")
    ap.add_argument("--serve-order", choices=("longest", "admission"), default="longest",
                    help="order of a fused launch's tasks (native loop)")
    ap.add_argument("--serve-depth", type=int, default=2, help="fused launches queued per lane")
    ap.add_argument("--serve-switch-ms", type=float, default=0.0,
                    help="interpreter thread switch interval during serving (0: unchanged)")
    ap.add_argument("--fresh-batches", type=int, default=4)
    ap.add_argument("--fresh-procs", type=int, default=14)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(independent solves would be split across GPUs with no communication).")
        sys.exit(2)

    from pylcm_amd import _native as nat
    from pylcm_amd import examples as ex
    from pylcm_amd._gs_program import GsProgram, lower_problem

    dt = np.float32
    T = args.periods
    draws = ex.draw_batch_params(args.draws, n_periods=T, seed=20261203)
    models = {n: ex.retirement_model(T, n_wealth=n, n_consumption=n, income_nodes=7)
              for n in sorted({s for s, _ in draws})}
    t0 = time.perf_counter()
    lowered = [lower_problem(models[s], p, dt) for s, p in draws]
    lower_s = time.perf_counter() - t0
    log(f"lowered {len(draws)} problems in {lower_s:.1f}s")
    problems = [(models[s], p) for s, p in draws]
    t0 = time.perf_counter()
    prog = GsProgram.of(problems, dt, lowered=lowered, tile=args.tile)
    log(f"packed {prog.total_bytes/2**20:.1f} MiB arena, {prog.n_desc} descriptors, tile {prog.tile},"
        f" in {time.perf_counter()-t0:.1f}s")

    # ---- copy kernel of known size (calibrates the HBM traffic counters) -----
    calib = args.copy_calib_mib * 2**20
    if calib:
        src, dst = nat.DeviceBuffer(calib), nat.DeviceBuffer(calib)
        for _ in range(2):
            nat.call("lcm_copy_launch", prog.ctx.handle, dst.ptr, src.ptr, C.c_size_t(calib))
        prog.ctx.sync()
        del src, dst

    # ---- batched: warmup + timed graph replays --------------------------------
    prog.capture_graph()
    for _ in range(args.warmup):
        prog.run(use_graph=True)
    prog.ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prog.run(use_graph=True, sync=False)
    prog.ctx.sync()
    batched_s = time.perf_counter() - t0
    ms_per_step = 1e3 * batched_s / args.steps
    value = args.draws * args.steps / batched_s
    log(f"batched: {ms_per_step:.2f} ms per batch of {args.draws} -> {value:.1f} solves/s")

    # ---- per-launch durations (HIP events on the program's stream) -----------
    # every launch of the batched program (one per backward step), so the
    # average matches a profiler's per-kernel average over the same launches
    per = prog.time_launches()
    ms_launch = np.array([m for _, m in per])
    flops = np.array([prog.flops_by_step[s] for s, _ in per])
    bytes_ = np.array([prog.bytes_by_step[s] for s, _ in per])
    avg_ms = float(ms_launch.mean())
    achieved_tflops = float(flops.sum() / (ms_launch.sum() * 1e-3) / 1e12)
    achieved_gbs = float(bytes_.sum() / (ms_launch.sum() * 1e-3) / 1e9)
    flops_per_launch = float(flops.mean())
    log(f"gs_period_kernel: avg {avg_ms:.3f} ms over {len(per)} launches, "
        f"{flops_per_launch/1e9:.1f} GFLOP per launch -> {achieved_tflops:.2f} TFLOP/s algorithmic, "
        f"{achieved_gbs:.1f} GB/s algorithmic HBM")

    prog.fetch()

    # ---- plain: the same draws one at a time ---------------------------------
    n_plain = min(args.plain_sample, args.draws)
    idx = list(range(n_plain))
    lat = []
    e2e = []
    for i in idx:
        s, p = draws[i]
        t_a = time.perf_counter()
        single = GsProgram.of([(models[s], p)], dt, lowered=[lowered[i]], ctx=prog.ctx)
        single.capture_graph()
        single.run(use_graph=True)  # warm (graph instantiation done above)
        t_b = time.perf_counter()
        reps = 3
        ts = []
        for _ in range(reps):
            t_c = time.perf_counter()
            single.run(use_graph=True)
            ts.append(time.perf_counter() - t_c)
        single.fetch()
        lat.append(min(ts))
        e2e.append(t_b - t_a + min(ts))
        # the batched answer must equal the one-at-a-time answer
        rb = prog.results(i)
        rs = single.results(0)
        for t in rs:
            for r in rs[t]:
                if not np.array_equal(rs[t][r][0], rb[t][r][0]):
                    raise SystemExit(f"batched != plain for draw {i} period {t} regime {r}")
        del single
    lat = np.array(lat)
    plain_sps = 1.0 / lat.mean()
    log(f"plain: {1e3*lat.mean():.3f} ms mean, p50 {1e3*np.percentile(lat,50):.3f} ms, "
        f"p99 {1e3*np.percentile(lat,99):.3f} ms -> {plain_sps:.1f} solves/s")

    # ---- CPU baseline + parity on a bounded sample ---------------------------
    cpu = None
    parity = {}
    if not args.no_cpu:
        from oracle import cpu_solve

        threads = len(os.sched_getaffinity(0))
        # a bounded sample with the batch's size mix: the first draw of each size
        order = []
        for size in sorted({s for s, _ in draws}):
            order += [i for i in range(args.draws) if draws[i][0] == size][: max(1, args.cpu_sample // 3)]
        from tests.parity import rel_err

        cpu_t = 0.0
        worst, mism, ties, tie_gap = 0.0, 0, 0, 0.0
        grading = {k: 0 for k in ("near_ties_le_1ulp", "near_ties_gt_1ulp", "gt_1ulp_same_as_reference_f32",
                                  "gt_1ulp_reference_f32_also_flips", "gt_1ulp_unexplained", "sampled_states",
                                  "reference_f32_gt_1ulp_sampled", "compared_gt_1ulp_sampled")}
        for i in order:
            s, p = draws[i]
            t0 = time.perf_counter()
            Vc, Pc = cpu_solve.solve_c(models[s], p, grid_dtype=dt, threads=threads)
            cpu_t += time.perf_counter() - t0
            rb = prog.results(i)
            Vg = {t: {r: v for r, (v, _) in rr.items()} for t, rr in rb.items()}
            Pg = {t: {r: q for r, (_, q) in rr.items()} for t, rr in rb.items()}
            for t in Vc:
                for r in Vc[t]:
                    worst = max(worst, rel_err(Vg[t][r], Vc[t][r]))
            # argmax check per period from the same inputs: the FP64 step reads
            # the GPU's V_{t+1}; a differing index is a near-tie when FP64 Q
            # at the two actions is within one float32 ulp; between one and
            # eight ulp it is graded against Q in the reference's float32
            # arithmetic (explained when that also departs from the FP64
            # argmax or picks the GPU's index), beyond eight it is a mismatch;
            # on sampled periods (every state) the departures beyond one ulp
            # are counted for the GPU and for the reference's float32 form
            Vg64 = {t: {r: np.asarray(v, dtype=np.float64) for r, v in rr.items()} for t, rr in Vg.items()}
            _, _, rep = cpu_solve.solve_c(models[s], p, grid_dtype=dt, threads=threads,
                                          next_V=Vg64, cmp_policy=Pg, reference_periods=(0, T // 2, T - 2))
            mism += rep["mismatches"]
            ties += rep["near_ties"]
            tie_gap = max(tie_gap, rep["max_tie_gap_ulps"])
            for k in grading:
                grading[k] += rep[k]
        cpu = {"value": len(order) / cpu_t, "unit": "model-solves/s", "cores": threads,
               "kind": "port",
               "sample": f"{len(order)} draws of the batch (grid sizes "
                         f"{sorted(draws[i][0] for i in order)}), C restatement oracle/gs_ref.c "
                         f"with OpenMP, FP64, {cpu_t:.1f}s"}
        parity = {"max_rel_V_err_vs_fp64": worst, "policy_mismatches": mism,
                  "flagged_near_ties": ties, "max_near_tie_gap_f32_ulps": round(tie_gap, 2),
                  **grading, "checked_draws": len(order)}
        log(f"cpu baseline: {cpu['value']:.3f} solves/s on {threads} threads; parity {parity}")

    tiles = tiles_leg(draws, models, lowered, prog.ctx)
    fresh = fresh_leg(args, models, T, prog.tile) if args.fresh_batches > 0 else None
    dcegm = None if args.no_dcegm else dcegm_leg(args)
    simulation = None if args.no_sim else simulation_leg(args)
    serving = serving_leg(args) if args.serve_seconds > 0 else None

    key = f"draws={args.draws} periods={T} tile={prog.tile} dtype=f32"
    traffic = committed_traffic(key)
    out = {
        "metric": "model-solves/sec (batched vs plain GPU), p50/p99 solve latency, V rel err <= 1e-5",
        "value": round(value, 3),
        "unit": "model-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded parameter draws; model grids built in-process)",
        "config": {
            "workload": "batched stochastic retirement model: 1024 parameter draws, T=40, "
                        "grid sizes {200,500,1000}, 7 Rouwenhorst nodes, one fused launch per "
                        "period replayed as a HIP graph",
            "draws": args.draws, "periods": T, "tile": prog.tile,
        },
        "tiles": dict(tiles, batch_tuned_tile=prog.tile, batch_shape=dict(zip(("units", "n_actions"),
                                                                             prog.launch_shape()))),
        "plain_gpu": {
            "solves_per_s": round(plain_sps, 3),
            "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 4),
            "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 4),
            "end_to_end_p50_ms": round(1e3 * float(np.percentile(e2e, 50)), 3),
            "sample": n_plain,
        },
        "batched_vs_plain": round(value / plain_sps, 2),
        "value_fresh_draws": fresh["solves_per_s"] if fresh else None,
        "fresh_draws": fresh,
        "fresh_vs_plain_end_to_end": (round(fresh["solves_per_s"] * float(np.percentile(e2e, 50)), 2)
                                      if fresh else None),
        "batched_latency_ms": round(ms_per_step, 3),
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved_tflops, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tflops / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic,
            "traffic_source": str(TRAFFIC_FILE.relative_to(ROOT)) if traffic else None,
            # the build the library picks for these launches (8 waves per SIMD
            # when the launch is large and its LDS leaves room for them)
            "kernel": f"gs_period_kernel<float, {prog.tile}, {max(prog.wdep_by_step)}, "
                      f"{str(max(prog.n_tasks) >= nat.GS_DENSE_MIN_TASKS and (160 * 1024) // max(prog.lds_by_step) > 6).lower()}>",
            "avg_launch_ms": round(avg_ms, 4),
            "launches": len(per),
            "gflop_per_launch": round(flops_per_launch / 1e9, 3),
            "flop_per_q_eval": 10,
            "hbm_algorithmic_GBs": round(achieved_gbs, 2),
            "hbm_frac": round(achieved_gbs / PEAK_HBM_GBS, 6),
        },
        "cpu_baseline": cpu,
        "parity": parity,
        "host_lowering_s": round(lower_s, 2),
        "dcegm": dcegm,
        "simulation": simulation,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
