"""Continuous-batching solve server for estimation-style request streams.

Requests (model, params) arrive at any time. A worker thread lowers each into
its own device image (tables, descriptors, V outputs), waits until the image
fits in the server's share of free HBM, places it in one preallocated device
slab (its own upload stream) and copies its descriptors into a shared
descriptor pool. Every iteration of the serving loop issues ONE fused grid-search
launch holding the next backward step of every admitted request, whatever its
model, grid size or horizon, so requests that arrive while others run join at
the next iteration, and a request that finishes frees its memory for the
queue. Value functions stay in device memory; `fetch` chooses what is copied
back ("none", "first" = period 0, or "all").

The reference solves one model at a time from a host loop
(`src/_lcm/solution/backward_induction.py:155-275`); it has no server. The
result of a request does not depend on what else is in flight: the same
kernel, descriptors and tile produce it alone or in the mix (tests check
bit equality).
"""

from __future__ import annotations

import ctypes as C
import os
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from pylcm_amd import _native as nat
from pylcm_amd import autotune
from pylcm_amd._batch import spill_prefix, unlink_quiet
from pylcm_amd._gs_program import DESC_DTYPE, TASK_DTYPE, GsProgram, lower_problem, step_lds_bytes

ALIGN = 256


def _lower_in_worker(blob: bytes, dtype: str, tile: int, spill: str | None = None) -> dict:
    """Lower and pack one request in a worker process (no device code is
    loaded there): everything the server needs to place and run it. With
    `spill` (a path prefix, normally in /dev/shm) the bulky arrays (the table
    image, the descriptor records and the task rows) are written to a file of
    that prefix instead of travelling back pickled through the pool's pipe,
    whose single receiving thread in the server then only unpickles a few
    kilobytes per request."""
    import cloudpickle

    from pylcm_amd._batch import pack_problem

    t0 = time.perf_counter()
    model, params = cloudpickle.loads(blob)
    data = pack_problem(model, params, dtype, tile)
    if spill is not None:
        img, D0, cat = data.pop("image"), data.pop("D0"), data.pop("tasks_cat")
        data.pop("tasks_host")
        path = f"{spill}{os.getpid()}_{time.perf_counter_ns()}"
        with open(path, "wb") as f:
            f.write(img.tobytes())
            f.write(D0.tobytes())
            f.write(np.ascontiguousarray(cat, dtype=np.int32).tobytes())
        data["spilled"] = (path, int(img.nbytes), len(D0), int(cat.shape[0]))
    data["worker_s"] = time.perf_counter() - t0
    return data


def write_request_file(path: str, data: dict, dtype, fetch: str = "first", hold: bool = False) -> list:
    """Write one packed request (`pack_problem` output with 256-state tasks)
    as a request file of the native server (layout in `lcm_serve_submit`'s
    header: `LcmReqHeader`, per-step arrays, fetched outputs, pointer masks,
    descriptor records at base 0, task rows, table image). Returns the
    outputs it fetches, in file order: [(period, regime, shape, bytes)]."""
    dt = np.dtype(dtype)
    D0, masks = data["D0"], data["masks"]
    n = len(D0)
    ptr = np.zeros(n, dtype=np.uint64)
    for f, m in masks.items():
        w0 = DESC_DTYPE.fields[f][1] // 8
        m2 = np.asarray(m, dtype=bool).reshape(n, -1)
        for j in range(m2.shape[1]):
            ptr |= m2[:, j].astype(np.uint64) << np.uint64(w0 + j)
    outs = data["outs"][0]
    if fetch == "first":
        keys = [k for k in outs if k[0] == 0]
    elif fetch == "all":
        keys = sorted(outs)
    elif fetch == "none":
        keys = []
    else:
        raise ValueError(f"fetch must be 'first', 'all' or 'none', not {fetch!r}")
    fetched, f_off, f_bytes = [], [], []
    for t, r in keys:
        o = outs[(t, r)]
        nb = (int(np.prod(o.shape, dtype=np.int64)) if o.shape else 1) * dt.itemsize
        fetched.append((t, r, tuple(o.shape), nb))
        f_off.append(o.v_off)
        f_bytes.append(nb)
    n_steps = int(data["n_steps"])
    cat = np.ascontiguousarray(data["tasks_cat"], dtype=np.int32).reshape(-1, 3)
    img = np.ascontiguousarray(data["image"]).view(np.uint8)
    code = nat.LCM_DTYPE_F64 if dt == np.float64 else nat.LCM_DTYPE_F32
    header = np.array([nat.REQ_MAGIC, code, n_steps, n, len(cat), len(keys), img.nbytes, data["out_start"],
                       data["total_bytes"], nat.REQ_TILE, nat.REQ_HOLD if hold else 0] + [0] * 5, dtype=np.int64)
    per_step = np.concatenate([np.asarray(data[k], dtype=np.int64).reshape(n_steps)
                               for k in ("task_counts", "lds_g", "lds_k", "wdep_by_step")]
                              + [np.asarray(f_off, dtype=np.int64), np.asarray(f_bytes, dtype=np.int64)])
    with open(path, "wb") as f:
        for part in (header, per_step, ptr, D0, cat, img):
            f.write(memoryview(np.ascontiguousarray(part)).cast("B"))
    return fetched


def _lower_to_request_file(blob, dtype: str, spill: str, fetch: str) -> dict:
    """Lower one request in a worker and write it as a native request file
    (a path of the `spill` prefix); returns the path and the fetched outputs.
    `blob` is the cloudpickled (model, params[, simulate]), or the tuple
    itself. With a `simulate` spec the request is held after its solve and
    the worker also writes `path + ".sim"`: the simulation image
    (`SimProgram.build_image`), the initial conditions and the seed."""
    from pylcm_amd._batch import pack_problem

    t0 = time.perf_counter()
    if isinstance(blob, bytes):
        import cloudpickle

        blob = cloudpickle.loads(blob)
    model, params = blob[0], blob[1]
    sim = blob[2] if len(blob) > 2 else None
    data = pack_problem(model, params, dtype, nat.REQ_TILE, keep_lowered=sim is not None)
    path = f"{spill}{os.getpid()}_{threading.get_ident()}_{time.perf_counter_ns()}"
    out = {"path": path}
    if sim is not None:
        out["sim_path"] = _write_sim_image(path + ".sim", model, params, data, dtype, sim)
    out["fetched"] = write_request_file(path, data, dtype, fetch, hold=sim is not None)
    out["worker_s"] = time.perf_counter() - t0
    return out


def _write_sim_image(path: str, model, params, data: dict, dtype, sim: dict) -> str:
    """The host half of a request's forward simulation, written for the
    server's simulation thread: `sim` = {"n_agents", "seed",
    "initial_conditions"} where the initial conditions are a dict of arrays
    or `f(model, params, grids, n_agents, rng) -> dict` (`grids` the resolved
    grids, `rng` a NumPy generator seeded with `seed`)."""
    import pickle
    import types

    from pylcm_amd._simulate import SimProgram, _initial_conditions

    tabs, rg = data["lowered"]
    view = types.SimpleNamespace(_rg={0: rg}, problems={0: tabs}, dt=np.dtype(dtype), outs=data["outs"])
    img = SimProgram(model, params, view).build_image()
    n, seed = int(sim["n_agents"]), int(sim.get("seed", 0))
    ic_user = sim["initial_conditions"]
    if callable(ic_user):
        ic_user = ic_user(model, params, rg, n, np.random.default_rng(seed))
    ic = _initial_conditions(model, ic_user)
    if ic["_n"] != n:
        raise ValueError(f"initial conditions hold {ic['_n']} agents, the spec asks for {n}")
    with open(path, "wb") as f:
        pickle.dump({"img": img, "ic": ic, "seed": seed}, f, protocol=pickle.HIGHEST_PROTOCOL)
    return path


def _unspill(data: dict) -> np.ndarray:
    """Read back what `_lower_in_worker` spilled: fills data's D0 and
    tasks_host (copies) and returns the table image (a copy in host memory);
    removes the file."""
    path, n_img, n_desc, n_task = data.pop("spilled")
    size = n_img + n_desc * DESC_DTYPE.itemsize + n_task * 12
    buf = np.empty(size, dtype=np.uint8)
    try:
        with open(path, "rb", buffering=0) as f:
            got = f.readinto(memoryview(buf))
    finally:
        os.unlink(path)
    if got != size:
        raise RuntimeError(f"request spill file {path}: {got} of {size} bytes")
    data["D0"] = buf[n_img: n_img + n_desc * DESC_DTYPE.itemsize].view(DESC_DTYPE)
    cat = buf[n_img + n_desc * DESC_DTYPE.itemsize:].view(np.int32).reshape(n_task, 3)
    counts = np.asarray(data["task_counts"], dtype=np.int64)
    data["tasks_host"] = np.split(cat, np.cumsum(counts)[:-1])
    return buf[:n_img]


def _pool_worker_main(conn, spill):
    """A lowering process: receive (rid, dtype, tile, blob), send back
    (rid, data) or (rid, exception), until told to stop (None)."""
    import gc

    import cloudpickle  # noqa: F401  (imported once, before the first request)

    from pylcm_amd import examples  # noqa: F401

    # what the imports built lives as long as the worker: keep it out of the
    # cyclic collector's scans, so a collection pauses a request for less
    gc.collect()
    gc.freeze()
    conn.send(("ready", None))
    while True:
        try:
            msg = conn.recv()
        except EOFError:
            return
        if msg is None:
            return
        rid, dtype, tile, blob, fetch = msg
        try:
            if fetch is None:
                out = (rid, _lower_in_worker(blob, dtype, tile, spill))
            else:  # a request file for the native server
                out = (rid, _lower_to_request_file(blob, dtype, spill, fetch))
        except BaseException as e:  # surfaced to the client of this request
            out = (rid, e)
        conn.send(out)


class _LoweringPool:
    """Worker processes fed through one pipe each. The server's submitting
    thread writes each request to the worker with the fewest outstanding
    ones; one receiving thread waits on all pipes and hands each result to
    `on_result`. Compared with a process-pool executor there is no
    management thread with a shallow call queue between the two, so a busy
    interpreter cannot leave the workers idle."""

    def __init__(self, procs: int, spill: str, on_result):
        import multiprocessing

        ctx = multiprocessing.get_context("spawn")
        self.conns, self.procs, self.outstanding = [], [], []
        for _ in range(int(procs)):
            a, b = ctx.Pipe()
            pr = ctx.Process(target=_pool_worker_main, args=(b, spill), daemon=True)
            pr.start()
            b.close()
            self.conns.append(a)
            self.procs.append(pr)
            self.outstanding.append(set())
        for c in self.conns:  # every worker has imported what it needs
            tag, _ = c.recv()
            assert tag == "ready"
        self.on_result = on_result
        self._lock = threading.Lock()
        self._stop = False
        self._rx = threading.Thread(target=self._receive, daemon=True)
        self._rx.start()

    def submit(self, rid: int, dtype: str, tile: int, blob: bytes, fetch: str | None = None):
        """Queue a request; `fetch` set: the worker writes a native request
        file fetching those outputs instead of returning the packed arrays."""
        with self._lock:
            k = min(range(len(self.conns)), key=lambda j: len(self.outstanding[j]))
            self.outstanding[k].add(rid)
        self.conns[k].send((rid, dtype, tile, blob, fetch))

    def _receive(self):
        from multiprocessing.connection import wait

        live = list(self.conns)
        while live and not self._stop:
            for c in wait(live, timeout=0.1):
                k = self.conns.index(c)
                try:
                    rid, out = c.recv()
                except (EOFError, OSError):
                    # the worker ended: every request it still held fails
                    live.remove(c)
                    with self._lock:
                        lost, self.outstanding[k] = self.outstanding[k], set()
                    for rid in lost:
                        self.on_result(rid, RuntimeError(f"lowering process {k} ended"), k)
                    continue
                with self._lock:
                    self.outstanding[k].discard(rid)
                self.on_result(rid, out, k)

    def close(self):
        self._stop = True
        for c in self.conns:
            try:
                c.send(None)
            except (OSError, BrokenPipeError):
                pass
        for pr in self.procs:
            pr.join(timeout=5)
            if pr.is_alive():
                pr.terminate()
        self._rx.join(timeout=1)


class _PackedRequest:
    """The parts of a `GsProgram` the server uses, rebuilt from a worker's result."""

    def __init__(self, data: dict, dtype):
        self.dt = np.dtype(dtype)
        for k in ("tasks_host", "outs", "n_steps", "wdep_by_step", "total_bytes", "out_start", "n_desc"):
            setattr(self, k, data[k])
        self.lds_by_step = step_lds_bytes(self.dt, data["lds_g"], data["lds_k"])
        self.base = None


class _Slab:
    """First-fit allocator over one device buffer (offsets in bytes, every
    offset and length a multiple of `unit`)."""

    def __init__(self, nbytes: int, unit: int = ALIGN):
        self.unit = int(unit)
        self.size = int(nbytes) // self.unit * self.unit
        self.free = [(0, self.size)]  # sorted (offset, length)
        self.used = 0

    def alloc(self, nbytes: int):
        n = max(1, (int(nbytes) + self.unit - 1) // self.unit) * self.unit
        for i, (off, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (off + n, ln - n)
                self.used += n
                return off, n
        return None

    def release(self, off: int, n: int):
        self.used -= n
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for o, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((o, ln))
        self.free = merged


@dataclass
class Request:
    rid: int
    model: object
    params: dict
    t_submit: float
    prog: GsProgram | None = None
    error: BaseException | None = None
    t_lowered: float = 0.0
    worker_s: float = 0.0  # lowering time inside the worker process
    staged: list | None = None  # period-0 V copied to pinned host memory by the stream
    n_act_steps: list | None = None  # most actions per state of each step's tasks
    t_admit: float = 0.0
    t_retired: float = 0.0  # native server: its last launch retired (before collection)
    t_simulated: float = 0.0  # native server: its forward simulation done
    simulate: dict | None = None  # native server: forward-simulation spec
    sim_path: str | None = None
    simulation: dict | None = None  # [period, agent] arrays of the forward simulation
    t_done: float = 0.0
    fetched: list | None = None  # native server: [(period, regime, shape, bytes)] in file order
    slab: tuple | None = None
    desc: tuple | None = None
    tasks: list | None = None
    desc_base: int = 0
    lowered: object = None  # a worker's result (dict) or its exception
    result: dict | None = None
    done: threading.Event = field(default_factory=threading.Event)

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


class _Cohort:
    """Requests admitted together: their per-step task arrays concatenated
    once (member p runs steps 0 .. n_steps_p - 1), so a fused launch costs one
    array per cohort rather than one per request."""

    def __init__(self, members: list, n_act_of_desc):
        self.members = members
        self.n_steps = max(r.prog.n_steps for r in members)
        self.steps, self.lds, self.variant, self.n_actions = [], [], [], []
        self._ending = {}
        for r in members:
            self._ending.setdefault(r.prog.n_steps - 1, []).append(r)
        empty = np.zeros(0, dtype=TASK_DTYPE)
        for s in range(self.n_steps):
            live = [r for r in members if s < r.prog.n_steps]
            parts = [r.tasks[s] for r in live if len(r.tasks[s])]
            self.steps.append(parts[0] if len(parts) == 1 else (np.concatenate(parts) if parts else empty))
            # per-request figures computed when each request was placed
            self.n_actions.append(max([r.n_act_steps[s] for r in live] or [1]))
            self.lds.append(max([r.prog.lds_by_step[s] for r in live] or [0]))
            self.variant.append(max([r.prog.wdep_by_step[s] for r in live] or [0]))
        self.s = 0

    @property
    def remaining(self) -> int:
        return sum(1 for r in self.members if r.prog is not None and self.s < r.prog.n_steps)

    def finished_at(self, s: int) -> list:
        return self._ending.get(s, [])


class _Lane:
    """One stream of fused launches: its cohorts, its ring of task buffers and
    the events that order retirement. Lanes run concurrently on the device,
    so one lane's launch fills the CUs that another's slowest workgroups
    leave idle at the end of its launch."""

    def __init__(self, n_ring: int, max_tasks: int):
        self.ctx = nat.Context()
        self.n_ring = n_ring
        self.task_host = [nat.PinnedBuffer(max_tasks * TASK_DTYPE.itemsize) for _ in range(n_ring)]
        self.task_dev = [nat.DeviceBuffer(max_tasks * TASK_DTYPE.itemsize) for _ in range(n_ring)]
        self.ring_events = [nat.Event() for _ in range(n_ring)]
        self.ring_used = [False] * n_ring
        self.ring_i = 0
        self.inflight: list[_Cohort] = []  # cohorts with steps left to launch
        self.finishing: list = []           # (event, requests) in stream order
        # (start, end) events of the last TIMING launches, reused in a ring;
        # each half is read back in one call once the other half is being
        # recorded (its launches have long finished)
        self.tev = [(nat.Event(), nat.Event()) for _ in range(self.TIMING)]
        self.tev_i = 0
        self.tev_live = [False] * self.TIMING
        self.intervals: list[tuple[float, float]] = []  # ms from the server's reference event

    TIMING = 256

    def harvest(self, ref, lo: int, hi: int):
        """Read back the recorded (start, end) pairs of ring slots [lo, hi)."""
        idx = [j for j in range(lo, hi) if self.tev_live[j]]
        if not idx:
            return
        evs = (C.c_void_p * (2 * len(idx)))(*[e.handle.value for j in idx for e in self.tev[j]])
        out = (C.c_float * (2 * len(idx)))()
        nat.call("lcm_event_elapsed_batch", ref.handle, evs, 2 * len(idx), out)
        self.intervals += [(out[2 * i], out[2 * i + 1]) for i in range(len(idx))]
        for j in idx:
            self.tev_live[j] = False

    @property
    def load(self) -> int:
        return sum(c.remaining for c in self.inflight)


class SolveServer:
    COARSE_TILE = 256  # the tile requests are lowered with

    """Continuous batching of grid-search solves on one GPU.

    `budget_bytes` caps the device memory the server uses (default: the given
    fraction of the free HBM at start). `tile` is the state tile of the fused
    launch (one shape per launch). `tile=0` picks it per launch with the
    autotuner (`pylcm_amd.autotune.choose` over `tile_table`, default the
    table measured on this GPU model) from the launch's 256-state tasks and
    its largest action count: requests are lowered into 256-state tasks, and a
    launch that runs a finer tile splits each into its 128- or 64-state parts
    (two or four waves then share each task's actions, so a small launch has
    more workgroups). The tile does not change results: the maximum and its
    first maximiser are merged the same way whatever the split. `max_inflight`
    caps admitted requests.
    `lanes` streams carry the fused launches: each admitted cohort joins the
    lane with the fewest requests in flight, and the lanes' launches overlap
    on the device.
    `lower_procs` > 0 lowers requests in that many spawned worker processes
    (models are sent with cloudpickle); as with any spawned pool, the main
    script must guard its entry point with `if __name__ == "__main__":`.
    """

    def __init__(self, *, dtype=np.float32, tile: int = 0, budget_bytes: int | None = None,
                 free_fraction: float = 0.6, max_inflight: int = 256, max_desc: int = 1 << 16,
                 max_tasks: int = 1 << 20, fetch: str = "first", lower_threads: int = 1,
                 lower_procs: int = 0, tile_table: dict | None = None, lanes: int = 1,
                 result_pinned_bytes: int = 64 << 20, queue_depth: int = 2):
        self.dt = np.dtype(dtype)
        self.code = nat.LCM_DTYPE_F64 if self.dt == np.float64 else nat.LCM_DTYPE_F32
        self.auto_tile = int(tile) == 0
        self.tile = self.COARSE_TILE if self.auto_tile else int(tile)  # the lowering tile
        self.tile_table = tile_table  # None: the committed table
        self.tiles_used = {}  # tile -> fused launches
        self.fetch_mode = fetch
        self.copy_ctx = nat.Context()   # request images in, results out
        _, free, _ = nat.device_info()
        budget = int(free * free_fraction)
        if budget_bytes is not None:
            budget = min(budget, int(budget_bytes))
        self.slab = _Slab(budget)
        self.slab_buf = nat.DeviceBuffer(self.slab.size)
        # descriptor slots in whole records: a task addresses its descriptor by index
        self.desc_slab = _Slab(max_desc * DESC_DTYPE.itemsize, unit=DESC_DTYPE.itemsize)
        self.desc_buf = nat.DeviceBuffer(self.desc_slab.size)
        self.max_tasks = int(max_tasks)
        self.max_inflight = int(max_inflight)
        self.n_w_of_desc = np.zeros(int(max_desc), dtype=np.int32)  # for splitting tasks into finer tiles
        self.n_act_of_desc = np.ones(int(max_desc), dtype=np.int64)  # actions per state, for the tile choice
        # per lane at most `queue_depth` fused launches queued on its stream:
        # while one runs, the next one gathers every step that becomes ready,
        # so launches stay large
        self.lanes = [_Lane(int(queue_depth), self.max_tasks) for _ in range(max(1, int(lanes)))]
        # fetch="first": each request's period-0 V is copied by its lane's
        # stream, right after its last launch, into pinned host memory, so
        # retiring it is a host copy (requests that do not fit fall back to a
        # synchronous copy)
        self.res_slab = _Slab(int(result_pinned_bytes)) if fetch == "first" else None
        self.res_buf = nat.PinnedBuffer(self.res_slab.size) if self.res_slab else None
        self.ctx = self.lanes[0].ctx
        self._t0_event = nat.Event()
        self._t0_event.record(self.ctx)
        self._lowq: queue.Queue = queue.Queue()
        self._ready: list[Request] = []      # placed on the device, not yet in a launch
        self._ready_lock = threading.Lock()
        self._mem_cv = threading.Condition()  # guards the slabs; notified on retirement
        self._n_placed = 0
        self.completed: list[Request] = []
        self._next_id = 0
        self.launches = 0
        self.launch_tasks = 0
        self.timing = {"admit": 0.0, "launch": 0.0, "retire": 0.0}
        self._stop = False
        # lower_procs > 0: lowering runs in that many worker processes (the
        # model travels by cloudpickle) and the threads only place requests
        self._pool = None
        self._spill = None
        if lower_procs > 0:
            # worker results' bulky arrays travel through files of this prefix
            # (memory-backed /dev/shm when present); close() removes leftovers
            self._spill = spill_prefix()
            self._by_rid: dict[int, Request] = {}
            self._pool = _LoweringPool(int(lower_procs), self._spill, self._on_lowered)
            lower_threads = max(lower_threads, 2)
        self._workers = [threading.Thread(target=self._lower_worker, daemon=True) for _ in range(lower_threads)]
        for w in self._workers:
            w.start()

    # -- client side --------------------------------------------------------------
    def submit(self, model, params, t_submit: float | None = None) -> Request:
        req = Request(self._next_id, model, params, t_submit if t_submit is not None else time.perf_counter())
        self._next_id += 1
        if self._pool is not None:
            import cloudpickle

            with self._ready_lock:
                self._by_rid[req.rid] = req
            self._pool.submit(req.rid, self.dt.str, self.tile, cloudpickle.dumps((model, params)))
        else:
            self._lowq.put(req)
        return req

    def _on_lowered(self, rid, out, worker: int):
        """A worker's result or exception (receiving thread): queue the
        request for placement, in completion order."""
        with self._ready_lock:
            req = self._by_rid.pop(rid)
        req.lowered = out
        self._lowq.put(req)

    def close(self):
        self._stop = True
        for _ in self._workers:
            self._lowq.put(None)
        with self._mem_cv:
            self._mem_cv.notify_all()
        if self._pool is not None:
            self._pool.close()
        if self._spill is not None:
            import glob

            for path in glob.glob(self._spill + "*"):
                try:
                    os.unlink(path)
                except OSError:
                    pass

    # -- host lowering and admission (worker threads) ------------------------------
    def _lower_worker(self):
        copy_ctx = nat.Context()  # this worker's upload stream
        while True:
            req = self._lowq.get()
            if req is None:
                return
            try:
                if req.lowered is not None:
                    data, req.lowered = req.lowered, None
                    if isinstance(data, BaseException):
                        raise data
                    req.worker_s = data.get("worker_s", 0.0)
                    image = _unspill(data) if "spilled" in data else data["image"]
                    prog = _PackedRequest(data, self.dt)
                    D0, masks = data["D0"], data["masks"]
                else:
                    low = lower_problem(req.model, req.params, self.dt)
                    prog = GsProgram.of([(req.model, req.params)], self.dt, lowered=[low], tile=self.tile,
                                        ctx=self.ctx, defer=True)
                    D0, masks = prog.relocatable_records()
                    image = prog._arena.image(prog.out_start)
                req.prog = prog
                req.t_lowered = time.perf_counter()
                self._place(req, copy_ctx, image, D0, masks)
            except BaseException as e:  # surfaced to the client
                req.error = e
                req.t_done = time.perf_counter()
                req.done.set()
                with self._ready_lock:
                    self.completed.append(req)

    def _place(self, req, copy_ctx, image, D0, masks):
        """Wait for device memory, upload the tables and the descriptors, and
        hand the request to the serving loop."""
        prog = req.prog
        need_desc = prog.n_desc * DESC_DTYPE.itemsize
        with self._mem_cv:
            while True:
                if self._stop:
                    raise RuntimeError("server closed")
                sl = self.slab.alloc(prog.total_bytes) if self._n_placed < self.max_inflight else None
                ds = self.desc_slab.alloc(need_desc) if sl is not None else None
                if sl is not None and ds is not None:
                    self._n_placed += 1
                    break
                if sl is not None:
                    self.slab.release(*sl)
                if prog.total_bytes > self.slab.size:
                    raise MemoryError(
                        f"request needs {prog.total_bytes} B, more than the server's {self.slab.size} B")
                self._mem_cv.wait(timeout=0.02)
        assert ds[0] % DESC_DTYPE.itemsize == 0 and ds[1] >= need_desc
        base = self.slab_buf.ptr + sl[0]
        prog.base = base
        nat.h2d(copy_ctx, base + prog.out_start, image)
        D = GsProgram.records_at(D0, masks, base)
        nat.h2d(copy_ctx, self.desc_buf.ptr + ds[0], D)
        desc_base = ds[0] // DESC_DTYPE.itemsize
        # every step's task records from one concatenated array (a handful of
        # numpy calls per request, whatever its horizon), then split per step
        counts = np.fromiter((len(t) for t in prog.tasks_host), dtype=np.int64, count=len(prog.tasks_host))
        cat = (np.concatenate([t for t in prog.tasks_host if len(t)]) if counts.any()
               else np.zeros((0, 3), dtype=np.int32))
        rec = np.zeros(len(cat), dtype=TASK_DTYPE)
        rec["desc"] = cat[:, 0] + desc_base
        rec["sd"] = cat[:, 1]
        rec["w0"] = cat[:, 2]
        ends = np.cumsum(counts)
        req.tasks = np.split(rec, ends[:-1])
        req.desc_base = desc_base
        self.n_w_of_desc[desc_base: desc_base + prog.n_desc] = D0["n_w"]
        n_act = D0["n_d"].astype(np.int64) * D0["n_c"]
        self.n_act_of_desc[desc_base: desc_base + prog.n_desc] = n_act
        # most actions per state of each step's tasks
        act = n_act[cat[:, 0]] if len(cat) else np.zeros(0, dtype=np.int64)
        nz = counts > 0
        steps_max = np.ones(len(counts), dtype=np.int64)
        if nz.any():
            steps_max[nz] = np.maximum.reduceat(act, (ends - counts)[nz])
        req.n_act_steps = steps_max.tolist()
        req.slab, req.desc = sl, ds
        req.t_admit = time.perf_counter()
        with self._ready_lock:
            self._ready.append(req)

    # -- serving loop ---------------------------------------------------------------
    def _admit(self):
        with self._ready_lock:
            ready, self._ready = self._ready, []
        if ready:
            lane = min(self.lanes, key=lambda ln: ln.load)
            lane.inflight.append(_Cohort(ready, self.n_act_of_desc))

    @property
    def inflight(self) -> list:
        return [c for ln in self.lanes for c in ln.inflight]

    @property
    def finishing(self) -> list:
        return [f for ln in self.lanes for f in ln.finishing]

    def _inflight_requests(self) -> int:
        return sum(ln.load for ln in self.lanes)

    def _launch(self, lane: _Lane | None = None):
        """One fused launch on a lane: the next backward step of every request
        in flight there (per cohort, one precomputed task array)."""
        lane = lane or self.lanes[0]
        k = lane.ring_i
        if lane.ring_used[k]:
            lane.ring_events[k].synchronize()  # the copy from this host slot has finished
        rec = np.frombuffer(lane.task_host[k].array, dtype=TASK_DTYPE, count=self.max_tasks)
        picks, n_coarse, lds, variant, n_act = [], 0, 0, 0, 1
        for c in lane.inflight:
            arr = c.steps[c.s]
            if n_coarse + len(arr) + 8 > self.max_tasks:
                break  # the rest join the next launch
            picks.append(c)
            n_coarse += len(arr)
            lds = max(lds, c.lds[c.s])
            variant = max(variant, c.variant[c.s])
            n_act = max(n_act, c.n_actions[c.s])
        if not picks:
            return False
        if n_coarse == 0:  # nothing to run this step (no regime of these cohorts is active)
            ev = nat.Event()
            ev.record(lane.ctx)
            done = [r for c in picks for r in c.finished_at(c.s)]
            if done:
                lane.finishing.append((ev, done))
            for c in picks:
                c.s += 1
            lane.inflight = [c for c in lane.inflight if c.s < c.n_steps]
            return True
        tile = self.tile
        if self.auto_tile:
            # the autotuned tile of this launch's shape (its 256-state tasks
            # and the most actions per state), when its split tasks fit
            tile = autotune.choose("f64" if self.dt == np.float64 else "f32", n_coarse, n_act, self.tile_table)
            if (self.COARSE_TILE // tile) * n_coarse + 8 > self.max_tasks:
                tile = self.COARSE_TILE
        n = 0
        for c in picks:
            t = self._split(self.n_w_of_desc, c.steps[c.s], tile) if tile < self.COARSE_TILE else c.steps[c.s]
            rec[n: n + len(t)] = t
            n += len(t)
        self.tiles_used[tile] = self.tiles_used.get(tile, 0) + 1
        npad = (n + 7) // 8 * 8
        rec["desc"][n:npad] = -1
        lane.ring_i = (k + 1) % lane.n_ring
        done = []
        for c in picks:
            done += c.finished_at(c.s)
            c.s += 1
        copies = []
        if self.res_slab is not None:
            for r in done:
                copies += self._stage_result(r)
        j = lane.tev_i
        half = lane.TIMING // 2
        if j % half == 0:
            # entering a half of the ring: read back its pairs before reuse
            # (recorded a half ring or more of launches ago, so finished)
            lane.harvest(self._t0_event, j, j + half)
        e_start, e_end = lane.tev[j]
        lane.tev_live[j] = True
        lane.tev_i = (j + 1) % lane.TIMING
        nc = len(copies)
        dst = (C.c_void_p * max(nc, 1))(*[d for d, _, _ in copies])
        src = (C.c_void_p * max(nc, 1))(*[s_ for _, s_, _ in copies])
        nby = (C.c_size_t * max(nc, 1))(*[b for _, _, b in copies])
        # one foreign call: events, task upload, launch, result copies
        nat.call("lcm_gs_serve_launch", lane.ctx.handle, self.code, C.c_void_p(self.desc_buf.ptr),
                 C.c_void_p(lane.task_host[k].ptr), C.c_void_p(lane.task_dev[k].ptr), npad, tile,
                 C.c_size_t(lds), int(variant), e_start.handle, lane.ring_events[k].handle, dst, src, nby, nc,
                 e_end.handle)
        lane.ring_used[k] = True
        self.launches += 1
        self.launch_tasks += n
        if done:
            lane.finishing.append((e_end, done))
        lane.inflight = [c for c in lane.inflight if c.s < c.n_steps]
        return True

    @classmethod
    def _split(cls, n_w_of_desc, t, tile: int = 64):
        """The `tile`-state tasks of 256-state tasks `t`: the parts of each
        that start below its descriptor's n_w."""
        k = cls.COARSE_TILE // tile
        r = np.repeat(t, k)
        r["w0"] += np.tile(np.arange(k, dtype=r["w0"].dtype) * tile, len(t))
        return r[r["w0"] < n_w_of_desc[r["desc"]]]

    def _retire(self, block: bool = False):
        """Complete the requests whose last step has run. Each lane's events
        were recorded on its own stream, so they complete in order: polling a
        lane stops at its first one still pending."""
        freed = False
        for lane in self.lanes:
            freed |= self._retire_lane(lane, block)
        if freed:
            with self._mem_cv:
                self._mem_cv.notify_all()

    def _retire_lane(self, lane: _Lane, block: bool) -> bool:
        freed = False
        while lane.finishing:
            ev, reqs = lane.finishing[0]
            if block:
                ev.synchronize()
            elif not ev.query():
                break
            lane.finishing.pop(0)
            t_done = time.perf_counter()
            for req in reqs:
                req.t_done = t_done
                self._collect(req)
                with self._mem_cv:
                    self.slab.release(*req.slab)
                    self.desc_slab.release(*req.desc)
                    self._n_placed -= 1
                req.prog = None if self.fetch_mode != "keep" else req.prog
                req.tasks = None
                req.done.set()
            with self._ready_lock:
                self.completed.extend(reqs)
            freed = True
        return freed

    def _stage_result(self, req: Request) -> list:
        """Pinned host slots for a finished request's period-0 V; returns the
        (dst, src, bytes) copies that its last launch enqueues after the kernel."""
        prog = req.prog
        outs = prog.outs[0]
        staged, copies = [], []
        for key in [k for k in outs if k[0] == 0]:
            o = outs[key]
            n = int(np.prod(o.shape, dtype=np.int64)) if o.shape else 1
            nb = n * self.dt.itemsize
            sl = self.res_slab.alloc(nb)
            if sl is None:  # no room: this request is copied synchronously when it retires
                for _, s2, _, _ in staged:
                    self.res_slab.release(*s2)
                return []
            staged.append((key, sl, n, o.shape))
            copies.append((self.res_buf.ptr + sl[0], prog.base + o.v_off, nb))
        req.staged = staged
        return copies

    def _collect(self, req: Request):
        prog = req.prog
        if self.fetch_mode == "none":
            return
        if req.staged is not None:
            res = {}
            for (t, r), sl, n, shape in req.staged:
                v = np.frombuffer(self.res_buf.array, dtype=self.dt, count=n, offset=sl[0]).copy()
                res.setdefault(t, {})[r] = v.reshape(shape)
                self.res_slab.release(*sl)
            req.staged = None
            req.result = res
            return
        outs = prog.outs[0]
        keys = sorted(outs) if self.fetch_mode in ("all", "keep") else [k for k in outs if k[0] == 0]
        res = {}
        for (t, r) in keys:
            o = outs[(t, r)]
            n = int(np.prod(o.shape, dtype=np.int64)) if o.shape else 1
            v = np.empty(n, dtype=self.dt)
            nat.d2h(self.copy_ctx, prog.base + o.v_off, v)
            res.setdefault(t, {})[r] = v.reshape(o.shape)
        req.result = res

    def pending(self) -> int:
        """Requests submitted and not yet completed."""
        with self._ready_lock:
            return self._next_id - len(self.completed)

    def step(self) -> bool:
        """Admit, launch one fused step, retire finished requests. Returns
        whether any work was issued."""
        t0 = time.perf_counter()
        self._admit()
        t1 = time.perf_counter()
        issued = False
        for lane in self.lanes:
            issued |= self._launch(lane)
        t2 = time.perf_counter()
        self._retire()
        t3 = time.perf_counter()
        self.timing["admit"] += t1 - t0
        self.timing["launch"] += t2 - t1
        self.timing["retire"] += t3 - t2
        return issued

    def gpu_busy_seconds(self) -> float:
        """Time the device had at least one fused launch running (task upload
        included): the union of the launches' intervals on all lanes, from
        events; call after `drain`."""
        iv = []
        for lane in self.lanes:
            lane.harvest(self._t0_event, 0, lane.TIMING)
            iv += lane.intervals
        iv.sort()
        busy, end = 0.0, -1.0
        for a, b in iv:
            if a > end:
                busy += b - a
                end = b
            elif b > end:
                busy += b - end
                end = b
        return busy / 1e3

    def drain(self, timeout: float = 600.0):
        t_end = time.perf_counter() + timeout
        while self.pending() and time.perf_counter() < t_end:
            if not self.step():
                if self.finishing:
                    self._retire(block=True)
                else:
                    time.sleep(0.0005)
        self._retire(block=True)


class NativeSolveServer:
    """Continuous batching of grid-search solves with the serving loop in the
    library (`lcm_serve_*`): its threads place requests in device memory,
    admit them to lanes (one HIP stream each), issue one fused launch per lane
    holding the next backward step of every request in flight there, copy
    each finished request's outputs to pinned host memory behind its last
    launch, and retire launches as their end events fire. This process only
    lowers requests into request files (in `lower_procs` worker processes, or
    in `lower_threads` threads when `lower_procs` is 0) and collects results,
    so the launch cadence never waits on the interpreter.

    The same kernel, descriptors and per-launch tile choice as `SolveServer`
    solve each request (the tile does not change results), so a request's
    values equal its plain solve bit for bit. `fetch`: "first" (period-0 V),
    "all" (every period's V) or "none". `tile` 0 picks each launch's tile from
    `tile_table` (default: the table measured on this GPU model).
    `order="longest"` runs each launch's tasks longest first on every XCD
    (estimated work: states x actions per state), so the launch does not end
    waiting on a long task that started last; "admission" keeps request order.
    As with any spawned pool, a main script using `lower_procs` must guard its
    entry point with `if __name__ == "__main__":`.
    """

    loop_in_library = True

    def __init__(self, *, dtype=np.float32, tile: int = 0, budget_bytes: int | None = None,
                 free_fraction: float = 0.6, max_inflight: int = 256, max_desc: int = 1 << 16,
                 max_tasks: int = 1 << 20, fetch: str = "first", lower_procs: int = 4, lower_threads: int = 2,
                 tile_table: dict | None = None, lanes: int = 1, queue_depth: int = 2, placers: int = 2,
                 result_pinned_bytes: int = 64 << 20, order: str = "longest", sim_threads: int = 4):
        if order not in ("longest", "admission"):
            raise ValueError(f"order must be 'longest' or 'admission', not {order!r}")
        if fetch not in ("first", "all", "none"):
            raise ValueError(f"fetch must be 'first', 'all' or 'none', not {fetch!r}")
        self.dt = np.dtype(dtype)
        self.code = nat.LCM_DTYPE_F64 if self.dt == np.float64 else nat.LCM_DTYPE_F32
        self.fetch_mode = fetch
        self.n_lanes = max(1, int(lanes))
        _, free, _ = nat.device_info()
        budget = int(free * free_fraction)
        if budget_bytes is not None:
            budget = min(budget, int(budget_bytes))
        table = tile_table if tile_table is not None else autotune.load_table()
        dts = "f64" if self.dt == np.float64 else "f32"
        ent = [e for e in table.get("entries", []) if e["dtype"] == dts]
        self._t_units = np.array([e["units"] for e in ent], dtype=np.int64)
        self._t_acts = np.array([e["n_actions"] for e in ent], dtype=np.int64)
        self._t_tiles = np.array([e["tile"] for e in ent], dtype=np.int32)
        self._L = nat.lib()
        self._h = C.c_void_p()
        nat.call("lcm_serve_create", C.byref(self._h), self.code, C.c_size_t(budget), int(max_desc),
                 int(max_tasks), int(max_inflight), C.c_size_t(int(result_pinned_bytes)), self.n_lanes,
                 int(queue_depth), int(placers), int(tile), self._t_units.ctypes.data, self._t_acts.ctypes.data,
                 self._t_tiles.ctypes.data, len(ent), nat.SERVE_LONGEST_FIRST if order == "longest" else 0)
        self.budget = budget
        self._spill = spill_prefix()
        self._lock = threading.Lock()
        self._by_rid: dict[int, Request] = {}
        self.completed: list[Request] = []
        self._next_id = 0
        self.timing = {"collect": 0.0}
        self._closing = False
        self._pool = None
        self._lowq: queue.Queue | None = None
        self._workers = []
        if lower_procs > 0:
            self._pool = _LoweringPool(int(lower_procs), self._spill, self._on_lowered)
        else:
            self._lowq = queue.Queue()
            self._workers = [threading.Thread(target=self._lower_worker, daemon=True)
                             for _ in range(max(1, int(lower_threads)))]
            for w in self._workers:
                w.start()
        self._collector = threading.Thread(target=self._collect_main, daemon=True)
        self._collector.start()
        # forward simulations of held requests: `sim_threads` threads with a
        # stream each, so the simulations (latency-bound chains of small
        # period launches) overlap on the device
        self._simq: queue.Queue = queue.Queue()
        self._tlock = threading.Lock()
        self._simulators = [threading.Thread(target=self._simulate_main, daemon=True)
                            for _ in range(max(1, int(sim_threads)))]
        for t in self._simulators:
            t.start()

    # -- client side --------------------------------------------------------------
    def submit(self, model, params, t_submit: float | None = None, simulate: dict | None = None) -> Request:
        """Queue a solve. `simulate` = {"n_agents", "seed",
        "initial_conditions"} also forward-simulates that many agents on the
        solved value functions while they are still on the device (the
        initial conditions: a dict of arrays, or `f(model, params, grids,
        n_agents, rng) -> dict`); the request completes after both, with the
        panel in `Request.simulation`."""
        req = Request(self._next_id, model, params, t_submit if t_submit is not None else time.perf_counter())
        req.simulate = simulate
        self._next_id += 1
        with self._lock:
            self._by_rid[req.rid] = req
        if self._pool is not None:
            import cloudpickle

            blob = (model, params) if simulate is None else (model, params, simulate)
            self._pool.submit(req.rid, self.dt.str, nat.REQ_TILE, cloudpickle.dumps(blob), self.fetch_mode)
        else:
            self._lowq.put(req)
        return req

    def pending(self) -> int:
        """Requests submitted and not yet completed."""
        with self._lock:
            return self._next_id - len(self.completed)

    def step(self) -> bool:
        """Nothing to do on the caller's thread: the library runs the loop."""
        return False

    @property
    def finishing(self) -> list:
        return []

    def _retire(self, block: bool = False):
        pass

    def drain(self, timeout: float = 600.0):
        t_end = time.perf_counter() + timeout
        while self.pending() and time.perf_counter() < t_end:
            time.sleep(0.001)

    # -- lowering -----------------------------------------------------------------
    def _lower_worker(self):
        while True:
            req = self._lowq.get()
            if req is None:
                return
            try:
                blob = (req.model, req.params) if req.simulate is None else (req.model, req.params, req.simulate)
                out = _lower_to_request_file(blob, self.dt.str, self._spill, self.fetch_mode)
            except BaseException as e:  # surfaced to the client
                out = e
            self._on_lowered(req.rid, out, -1)

    def _on_lowered(self, rid, out, worker: int):
        """Hand a lowered request's file to the library (or fail the request)."""
        with self._lock:
            req = self._by_rid.get(rid)
        if req is None:
            return
        req.t_lowered = time.perf_counter()
        if not isinstance(out, BaseException):
            req.worker_s = out["worker_s"]
            req.fetched = out["fetched"]
            req.sim_path = out.get("sim_path")
            if self._L.lcm_serve_submit(self._h, rid, out["path"].encode()) == 0:
                return
            out = nat.NativeKernelError(f"lcm_serve_submit failed: {nat.host_error()}")
        with self._lock:
            self._by_rid.pop(rid, None)
        req.error = out
        req.t_done = time.perf_counter()
        req.done.set()
        with self._lock:
            self.completed.append(req)

    # -- collection ---------------------------------------------------------------
    def _collect_main(self):
        L, h = self._L, self._h
        cap = 256
        rids = (C.c_int64 * cap)()
        st = (C.c_int32 * cap)()
        times = (C.c_double * 3)()
        err = C.create_string_buffer(1024)
        while True:
            n = L.lcm_serve_wait(h, rids, st, cap, 50)
            if n <= 0:
                if self._closing:
                    return
                continue
            t0 = time.perf_counter()
            done = []
            for i in range(n):
                rid = int(rids[i])
                with self._lock:
                    req = self._by_rid.pop(rid, None)
                fetched = req.fetched if req is not None else []
                total = sum(f[3] for f in fetched)
                buf = np.empty(total, dtype=np.uint8)
                rc = L.lcm_serve_fetch(h, rid, buf.ctypes.data if total else None, total, times, err, len(err))
                if req is None:
                    continue
                req.t_admit, req.t_retired = times[1], times[2]
                if rc == 0:
                    res, o = {}, 0
                    if self.fetch_mode != "none":
                        for t, r, shape, nb in fetched:
                            res.setdefault(t, {})[r] = buf[o: o + nb].view(self.dt).reshape(shape)
                            o += nb
                    req.result = res
                    if req.sim_path is not None:  # held: the simulation thread finishes it
                        self._simq.put(req)
                        continue
                else:
                    req.error = nat.NativeKernelError(err.value.decode(errors="replace"), status=rc)
                    if req.sim_path is not None:
                        unlink_quiet(req.sim_path)
                req.t_done = time.perf_counter()
                req.done.set()
                done.append(req)
            with self._lock:
                self.completed.extend(done)
            self.timing["collect"] += time.perf_counter() - t0

    # -- forward simulation -------------------------------------------------------
    def _simulate_main(self):
        """Place each held request's simulation image on this thread's stream,
        simulate its agents on the request's resident value functions, then
        release the request's device memory. The thread's panel and table
        buffers are reused across requests (they only grow), so no free
        synchronizes the device while the lanes run."""
        import pickle
        import types

        from pylcm_amd._simulate import SimProgram

        ctx = nat.Context()
        bufs = [None, None]

        def mem(slot, nbytes):
            b = bufs[slot]
            if b is None or b.nbytes < nbytes:
                bufs[slot] = b = nat.DeviceBuffer(max(int(nbytes), 2 * (b.nbytes if b is not None else 0)))
            return b.ptr

        base, descs = C.c_uint64(), C.c_uint64()
        while True:
            req = self._simq.get()
            if req is None:
                return
            held = False
            try:
                t0 = time.perf_counter()
                with open(req.sim_path, "rb") as f:
                    payload = pickle.load(f)
                unlink_quiet(req.sim_path)
                nat.call("lcm_serve_held", self._h, req.rid, C.byref(base), C.byref(descs))
                held = True
                view = types.SimpleNamespace(ctx=ctx, base=0, desc_off=descs.value, dt=self.dt)
                sim = SimProgram.for_placement(req.model, req.params, view)
                sim.place(payload["img"], payload["ic"], payload["seed"], mem=mem)
                t1 = time.perf_counter()
                sim.launch()
                t2 = time.perf_counter()
                ctx.sync()
                t3 = time.perf_counter()
                req.simulation = sim.collect_raw()
                t4 = time.perf_counter()
                with self._tlock:
                    tm = self.timing
                    tm["sim_place"] = tm.get("sim_place", 0.0) + t1 - t0
                    tm["sim_launch"] = tm.get("sim_launch", 0.0) + t2 - t1
                    tm["sim_wait"] = tm.get("sim_wait", 0.0) + t3 - t2
                    tm["sim_collect"] = tm.get("sim_collect", 0.0) + t4 - t3
            except BaseException as e:  # surfaced to the client
                req.error = e
            finally:
                if held:
                    self._L.lcm_serve_release(self._h, req.rid)
            req.t_simulated = req.t_done = time.perf_counter()
            req.done.set()
            with self._lock:
                self.completed.append(req)

    # -- statistics ---------------------------------------------------------------
    def stats(self) -> dict:
        """The library's counters (after `close`, their final values)."""
        if self._h is None:
            return dict(self._final["stats"])
        v = (C.c_int64 * 12)()
        nat.call("lcm_serve_stats", self._h, v, 12)
        keys = ("launches", "tasks", "tile64", "tile128", "tile256", "loop_ns", "place_ns", "placed", "retired",
                "bytes_used", "bytes_budget", "bytes_peak")
        return dict(zip(keys, (int(x) for x in v)))

    @property
    def launches(self) -> int:
        return self.stats()["launches"]

    @property
    def launch_tasks(self) -> int:
        return self.stats()["tasks"]

    @property
    def tiles_used(self) -> dict:
        s = self.stats()
        return {t: s[f"tile{t}"] for t in (64, 128, 256) if s[f"tile{t}"]}

    def host_seconds(self) -> dict:
        """Seconds the library's lanes spent assembling and issuing launches,
        its placers spent placing, and this process's collector spent
        collecting."""
        s = self.stats()
        return {"launch": s["loop_ns"] / 1e9, "place": s["place_ns"] / 1e9,
                **{k: v for k, v in self.timing.items()}}

    def gpu_busy_seconds(self) -> float:
        """Time the device had at least one fused launch running (task upload
        included), the union over lanes; call after `drain`."""
        if self._h is None:
            return self._final["busy_s"]
        ms = C.c_double()
        nat.call("lcm_serve_busy_ms", self._h, C.byref(ms))
        return ms.value / 1e3

    def close(self):
        if self._h is None:
            return
        self._closing = True
        for _ in self._workers:
            self._lowq.put(None)
        if self._pool is not None:
            self._pool.close()
        self._collector.join(timeout=60)
        if self._collector.is_alive():  # never free the server under a waiting collector
            raise RuntimeError("NativeSolveServer.close: the collector thread did not stop")
        for _ in self._simulators:
            self._simq.put(None)
        for t in self._simulators:
            t.join(timeout=60)
            if t.is_alive():
                raise RuntimeError("NativeSolveServer.close: a simulation thread did not stop")
        self._final = {"stats": self.stats(), "busy_s": self.gpu_busy_seconds()}
        self._L.lcm_serve_destroy(self._h)
        self._h = None
        import glob

        for path in glob.glob(self._spill + "*"):
            try:
                os.unlink(path)
            except OSError:
                pass


def serve_poisson(server: SolveServer, make_request, rate: float, duration: float, seed: int = 0):
    """Submit requests with exponential inter-arrival times (rate per second)
    for `duration` seconds while running the serving loop; then drain.
    `make_request(i)` returns (model, params) or, for the native server,
    (model, params, simulate spec). Returns the completed requests."""
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    next_t = t0 + rng.exponential(1.0 / rate)
    submitted = []
    while True:
        now = time.perf_counter()
        while next_t <= now and next_t - t0 < duration:
            item = make_request(len(submitted))
            if len(item) > 2 and item[2] is not None:  # (model, params, simulate spec)
                submitted.append(server.submit(item[0], item[1], t_submit=next_t, simulate=item[2]))
            else:
                submitted.append(server.submit(item[0], item[1], t_submit=next_t))
            next_t += rng.exponential(1.0 / rate)
        if next_t - t0 >= duration and now - t0 >= duration:
            break
        if getattr(server, "loop_in_library", False):
            # nothing to run here: sleep until the next arrival
            time.sleep(min(max(next_t - time.perf_counter(), 0.0), 0.005))
        elif not server.step():
            if server.finishing:
                server._retire()
            time.sleep(0.0002)
    server.drain()
    return submitted
