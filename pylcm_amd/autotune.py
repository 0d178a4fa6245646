"""State-tile autotuner of the fused grid-search launch.

The grid-search kernel is compiled for three state tiles (64, 128 and 256
states per workgroup; `gs_period_kernel<T, TILE, VAR>`). A smaller tile gives a
launch more workgroups (the 256 / TILE wave groups of a workgroup split the
action axis), which a small launch needs to fill 256 CUs; a larger tile
contracts each EV row once for more states and merges nothing, which a large
launch prefers. Which one wins depends on the launch: how many 256-state work
units it holds (problems x regimes x discrete states x wealth tiles) and how
many actions each state scans. The tile is a compile-time constant of the
kernel: the same kernel with the tile read from each task at run time measured
41 % slower on the 1024-draw batch (2.65 against 1.87 ms per launch), so a
launch runs one tile and this module picks it per launch.

`tune()` measures, on the device it runs on, the per-launch time of each tile
over a grid of launch shapes (grid size x number of problems x dtype) and
writes the fastest tile of each shape to a JSON table; the table measured on an
MI355X ships as `tile_table.json` beside this module. `choose()` returns the
tile of the measured shape nearest to a launch (same dtype, nearest in log
work units and log actions per state). The reference has no
counterpart: its grid search runs one XLA program per (regime, period) whose
tiling knob is the user's `batch_size` (src/_lcm/grids/continuous.py:37-40).

The tile does not change results: every state's maximum and first maximiser
are merged the same way whatever the split of its actions (the GPU tests check
bit equality across tiles).

Run `python -m pylcm_amd.autotune` on the GPU to re-tune.
"""

from __future__ import annotations

import json
import math
import time
from pathlib import Path

TILES = (64, 128, 256)
TABLE_PATH = Path(__file__).resolve().parent / "tile_table.json"

_table_cache: dict | None = None


def load_table(path: Path | None = None) -> dict:
    """The tuned table ({"entries": [...]}); empty when no table is present."""
    global _table_cache
    if path is None and _table_cache is not None:
        return _table_cache
    p = Path(path) if path else TABLE_PATH
    table = {"entries": []}
    if p.exists():
        table = json.loads(p.read_text())
    if path is None:
        _table_cache = table
    return table


def fallback_tile(units: int) -> int:
    """The tile used when no table entry applies: small launches split the
    action axis over more workgroups."""
    if units <= 256:
        return 64
    if units <= 2048:
        return 128
    return 256


def choose(dtype: str, units: int, n_actions: int, table: dict | None = None) -> int:
    """Tile of the measured launch shape nearest to this one (same dtype;
    distance in log2 of work units and of actions per state)."""
    table = table if table is not None else load_table()
    best, best_d = None, math.inf
    lu, la = math.log2(max(units, 1)), math.log2(max(n_actions, 1))
    for e in table.get("entries", []):
        if e["dtype"] != dtype:
            continue
        d = (math.log2(max(e["units"], 1)) - lu) ** 2 + (math.log2(max(e["n_actions"], 1)) - la) ** 2
        if d < best_d:
            best, best_d = e["tile"], d
    return int(best) if best is not None else fallback_tile(units)


# ---------------------------------------------------------------------------
# measurement (needs the device)
# ---------------------------------------------------------------------------


def _launch_units(prog) -> tuple[int, int]:
    """(work units, actions per state) of a program's largest step, at 256-state tiles."""
    s = max(range(prog.n_steps), key=lambda k: prog.flops_by_step[k])
    units = 0
    n_actions = 1
    for p, (model, tabs) in enumerate(zip(prog.models, prog.problems)):
        for (t, r) in tabs:
            if prog.step_of(p, t) != s:
                continue
            lay = model.layouts[r]
            units += lay.n_sd * -(-lay.n_w // 256)
            n_actions = max(n_actions, lay.n_d * lay.n_c)
    return units, n_actions


def measure_shape(n_grid: int, n_problems: int, dtype, *, periods: int = 8, income_nodes: int = 7,
                  reps: int = 3, seed: int = 0) -> dict:
    """Per-launch time of each tile for `n_problems` draws of the stochastic
    retirement model with wealth = consumption grids of `n_grid` points."""
    import numpy as np

    from pylcm_amd import examples as ex
    from pylcm_amd._gs_program import GsProgram, lower_problem

    dt = np.dtype(dtype)
    draws = ex.draw_batch_params(n_problems, n_periods=periods, seed=seed)
    model = ex.retirement_model(periods, n_wealth=n_grid, n_consumption=n_grid, income_nodes=income_nodes)
    lowered = [lower_problem(model, p, dt) for _, p in draws]
    times = {}
    shape = None
    for tile in TILES:
        prog = GsProgram.of([(model, p) for _, p in draws], dt, lowered=lowered, tile=tile)
        prog.run(use_graph=False)
        if shape is None:
            shape = _launch_units(prog) + (max(prog.wdep_by_step),)
        per = []
        for _ in range(reps):
            ms = [m for s, m in prog.time_launches() if s > 0]  # the terminal step runs no search
            per.append(sum(ms) / len(ms))
        times[tile] = min(per)
        del prog
    units, n_actions, variant = shape
    tile = min(times, key=times.get)
    return {"dtype": "f64" if dt == np.float64 else "f32", "variant": int(variant), "units": int(units),
            "n_actions": int(n_actions), "n_grid": int(n_grid), "n_problems": int(n_problems),
            "ms_per_launch": {str(k): round(v, 5) for k, v in times.items()}, "tile": int(tile),
            "gain_vs_fallback": round(times[fallback_tile(units)] / times[tile], 4)}


DEFAULT_SHAPES = [(n, p) for n in (200, 500, 1000, 2000, 4000) for p in (1, 4, 16, 64, 256)
                  if n * p <= 256 * 1000]


def tune(path: Path | None = None, dtypes=("f32", "f64"), shapes=None, log=print) -> dict:
    """Measure every shape and write the table."""
    import numpy as np

    from pylcm_amd import _native as nat

    shapes = shapes or DEFAULT_SHAPES
    entries = []
    t0 = time.perf_counter()
    for dts in dtypes:
        dt = np.float64 if dts == "f64" else np.float32
        for n, p in shapes:
            e = measure_shape(n, p, dt)
            entries.append(e)
            log(f"{dts} grid {n} x {p} problems: units {e['units']}, tiles {e['ms_per_launch']} ms -> {e['tile']}"
                f" ({e['gain_vs_fallback']}x the fallback)")
    _, _, total = nat.device_info()
    table = {"device_memory_bytes": int(total), "seconds": round(time.perf_counter() - t0, 1),
             "workload": "stochastic retirement model, 7 income nodes, wealth = consumption grid, 8 periods, "
                         "per-launch time of the search steps (events), best of 3",
             "entries": entries}
    out = Path(path) if path else TABLE_PATH
    out.write_text(json.dumps(table, indent=1) + "\n")
    global _table_cache
    _table_cache = None
    return table


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--dtypes", default="f32,f64")
    args = ap.parse_args()
    tune(args.out, dtypes=tuple(args.dtypes.split(",")))
