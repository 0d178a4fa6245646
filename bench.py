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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--draws", type=int, default=1024)
    ap.add_argument("--periods", type=int, default=40)
    ap.add_argument("--plain-sample", type=int, default=48)
    ap.add_argument("--cpu-sample", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=int, default=None)
    ap.add_argument("--copy-calib-mib", type=int, default=256)
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
            # the GPU's V_{t+1}; a differing index is a flagged near-tie when
            # FP64 Q at the two actions is within 8 float32 ulp
            Vg64 = {t: {r: np.asarray(v, dtype=np.float64) for r, v in rr.items()} for t, rr in Vg.items()}
            _, _, rep = cpu_solve.solve_c(models[s], p, grid_dtype=dt, threads=threads,
                                          next_V=Vg64, cmp_policy=Pg)
            mism += rep["mismatches"]
            ties += rep["near_ties"]
            tie_gap = max(tie_gap, rep["max_tie_gap_ulps"])
        cpu = {"value": len(order) / cpu_t, "unit": "model-solves/s", "cores": threads,
               "kind": "port",
               "sample": f"{len(order)} draws of the batch (grid sizes "
                         f"{sorted(draws[i][0] for i in order)}), C restatement oracle/gs_ref.c "
                         f"with OpenMP, FP64, {cpu_t:.1f}s"}
        parity = {"max_rel_V_err_vs_fp64": worst, "policy_mismatches": mism,
                  "flagged_near_ties": ties, "max_near_tie_gap_f32_ulps": round(tie_gap, 2),
                  "checked_draws": len(order)}
        log(f"cpu baseline: {cpu['value']:.3f} solves/s on {threads} threads; parity {parity}")

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
        "plain_gpu": {
            "solves_per_s": round(plain_sps, 3),
            "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 4),
            "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 4),
            "end_to_end_p50_ms": round(1e3 * float(np.percentile(e2e, 50)), 3),
            "sample": n_plain,
        },
        "batched_vs_plain": round(value / plain_sps, 2),
        "batched_latency_ms": round(ms_per_step, 3),
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved_tflops, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tflops / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic,
            "traffic_source": str(TRAFFIC_FILE.relative_to(ROOT)) if traffic else None,
            "kernel": f"gs_period_kernel<float, {prog.tile}, {max(prog.wdep_by_step)}>",
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
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
