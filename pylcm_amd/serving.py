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
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from pylcm_amd import _native as nat
from pylcm_amd import autotune
from pylcm_amd._gs_program import DESC_DTYPE, TASK_DTYPE, GsProgram, lower_problem, step_lds_bytes

ALIGN = 256


def _lower_in_worker(blob: bytes, dtype: str, tile: int) -> dict:
    """Lower and pack one request in a worker process (no device code is
    loaded there): everything the server needs to place and run it."""
    import cloudpickle

    from pylcm_amd._batch import pack_problem

    model, params = cloudpickle.loads(blob)
    return pack_problem(model, params, dtype, tile)


def _warm_worker() -> int:
    """Import what `_lower_in_worker` needs, so the first request pays no start-up."""
    import cloudpickle  # noqa: F401

    from pylcm_amd import examples  # noqa: F401

    return 0


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
    t_admit: float = 0.0
    t_done: float = 0.0
    slab: tuple | None = None
    desc: tuple | None = None
    tasks: list | None = None
    desc_base: int = 0
    future: object = None
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
        for s in range(self.n_steps):
            parts = [r.tasks[s] for r in members if s < r.prog.n_steps and len(r.tasks[s])]
            self.steps.append(np.concatenate(parts) if parts else np.zeros(0, dtype=TASK_DTYPE))
            arr = self.steps[-1]
            self.n_actions.append(int(n_act_of_desc[arr["desc"]].max()) if len(arr) else 1)
            self.lds.append(max([r.prog.lds_by_step[s] for r in members if s < r.prog.n_steps] or [0]))
            self.variant.append(max([r.prog.wdep_by_step[s] for r in members if s < r.prog.n_steps] or [0]))
        self.s = 0

    @property
    def remaining(self) -> int:
        return sum(1 for r in self.members if r.prog is not None and self.s < r.prog.n_steps)

    def finished_at(self, s: int) -> list:
        return self._ending.get(s, [])


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
    `lower_procs` > 0 lowers requests in that many spawned worker processes
    (models are sent with cloudpickle); as with any spawned pool, the main
    script must guard its entry point with `if __name__ == "__main__":`.
    """

    def __init__(self, *, dtype=np.float32, tile: int = 0, budget_bytes: int | None = None,
                 free_fraction: float = 0.6, max_inflight: int = 256, max_desc: int = 1 << 16,
                 max_tasks: int = 1 << 20, fetch: str = "first", lower_threads: int = 1,
                 lower_procs: int = 0, tile_table: dict | None = None):
        self.dt = np.dtype(dtype)
        self.code = nat.LCM_DTYPE_F64 if self.dt == np.float64 else nat.LCM_DTYPE_F32
        self.auto_tile = int(tile) == 0
        self.tile = self.COARSE_TILE if self.auto_tile else int(tile)  # the lowering tile
        self.tile_table = tile_table  # None: the committed table
        self.tiles_used = {}  # tile -> fused launches
        self.fetch_mode = fetch
        self.ctx = nat.Context()        # kernels and task uploads
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
        # at most two fused launches queued on the stream: while one runs, the
        # next one gathers every step that becomes ready, so launches stay large
        self.n_ring = 2
        self.task_host = [nat.PinnedBuffer(self.max_tasks * TASK_DTYPE.itemsize) for _ in range(self.n_ring)]
        self.task_dev = [nat.DeviceBuffer(self.max_tasks * TASK_DTYPE.itemsize) for _ in range(self.n_ring)]
        self.ring_events = [nat.Event() for _ in range(self.n_ring)]
        self.ring_used = [False] * self.n_ring
        self.ring_i = 0
        self._lowq: queue.Queue = queue.Queue()
        self._ready: list[Request] = []      # placed on the device, not yet in a launch
        self._ready_lock = threading.Lock()
        self._mem_cv = threading.Condition()  # guards the slabs; notified on retirement
        self._n_placed = 0
        self.inflight: list[_Cohort] = []   # cohorts with steps left to launch
        self.finishing: list = []            # (event, requests) in stream order
        self.completed: list[Request] = []
        self._next_id = 0
        self.launches = 0
        self.launch_tasks = 0
        self.timing = {"admit": 0.0, "launch": 0.0, "retire": 0.0}
        self._launch_events = []  # (start, end) per fused launch, for the GPU busy time
        self._stop = False
        # lower_procs > 0: lowering runs in that many worker processes (the
        # model travels by cloudpickle) and the threads only place requests
        self._pool = None
        if lower_procs > 0:
            import multiprocessing
            from concurrent.futures import ProcessPoolExecutor

            self._pool = ProcessPoolExecutor(max_workers=int(lower_procs),
                                             mp_context=multiprocessing.get_context("spawn"))
            for f in [self._pool.submit(_warm_worker) for _ in range(int(lower_procs))]:
                f.result()
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

            req.future = self._pool.submit(_lower_in_worker, cloudpickle.dumps((model, params)),
                                           self.dt.str, self.tile)
        self._lowq.put(req)
        return req

    def close(self):
        self._stop = True
        for _ in self._workers:
            self._lowq.put(None)
        with self._mem_cv:
            self._mem_cv.notify_all()
        if self._pool is not None:
            self._pool.shutdown(wait=False, cancel_futures=True)

    # -- host lowering and admission (worker threads) ------------------------------
    def _lower_worker(self):
        copy_ctx = nat.Context()  # this worker's upload stream
        while True:
            req = self._lowq.get()
            if req is None:
                return
            try:
                if req.future is not None:
                    data = req.future.result()
                    req.future = None
                    prog = _PackedRequest(data, self.dt)
                    D0, masks, image = data["D0"], data["masks"], data["image"]
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
        steps = []
        for t in prog.tasks_host:
            rec = np.zeros(len(t), dtype=TASK_DTYPE)
            if len(t):
                rec["desc"] = t[:, 0] + desc_base
                rec["sd"] = t[:, 1]
                rec["w0"] = t[:, 2]
            steps.append(rec)
        req.tasks = steps
        req.desc_base = desc_base
        self.n_w_of_desc[desc_base: desc_base + prog.n_desc] = D0["n_w"]
        self.n_act_of_desc[desc_base: desc_base + prog.n_desc] = D0["n_d"].astype(np.int64) * D0["n_c"]
        req.slab, req.desc = sl, ds
        req.t_admit = time.perf_counter()
        with self._ready_lock:
            self._ready.append(req)

    # -- serving loop ---------------------------------------------------------------
    def _admit(self):
        with self._ready_lock:
            ready, self._ready = self._ready, []
        if ready:
            self.inflight.append(_Cohort(ready, self.n_act_of_desc))

    def _inflight_requests(self) -> int:
        return sum(c.remaining for c in self.inflight)

    def _launch(self):
        """One fused launch: the next backward step of every request in flight
        (per cohort, one precomputed task array)."""
        k = self.ring_i
        if self.ring_used[k]:
            self.ring_events[k].synchronize()  # the copy from this host slot has finished
        rec = np.frombuffer(self.task_host[k].array, dtype=TASK_DTYPE, count=self.max_tasks)
        picks, n_coarse, lds, variant, n_act = [], 0, 0, 0, 1
        for c in self.inflight:
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
            ev.record(self.ctx)
            done = [r for c in picks for r in c.finished_at(c.s)]
            if done:
                self.finishing.append((ev, done))
            for c in picks:
                c.s += 1
            self.inflight = [c for c in self.inflight if c.s < c.n_steps]
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
        self.ring_i = (k + 1) % self.n_ring
        e_start = nat.Event()
        e_start.record(self.ctx)
        nat.call("lcm_memcpy_h2d_async", self.ctx.handle, C.c_void_p(self.task_dev[k].ptr),
                 C.c_void_p(self.task_host[k].ptr), C.c_size_t(npad * TASK_DTYPE.itemsize))
        nat.call("lcm_gs_launch_ex", self.ctx.handle, self.code, C.c_void_p(self.desc_buf.ptr),
                 C.c_void_p(self.task_dev[k].ptr), npad, tile, C.c_size_t(lds), int(variant))
        self.ring_events[k].record(self.ctx)
        e_end = nat.Event()
        e_end.record(self.ctx)
        self._launch_events.append((e_start, e_end))
        self.ring_used[k] = True
        self.launches += 1
        self.launch_tasks += n
        done = []
        for c in picks:
            done += c.finished_at(c.s)
            c.s += 1
        if done:
            self.finishing.append((e_end, done))
        self.inflight = [c for c in self.inflight if c.s < c.n_steps]
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
        """Complete the requests whose last step has run. Events were recorded
        on one stream, so they complete in order: polling stops at the first
        one still pending."""
        freed = False
        while self.finishing:
            ev, reqs = self.finishing[0]
            if block:
                ev.synchronize()
            elif not ev.query():
                break
            self.finishing.pop(0)
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
        if freed:
            with self._mem_cv:
                self._mem_cv.notify_all()

    def _collect(self, req: Request):
        prog = req.prog
        if self.fetch_mode == "none":
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
        issued = self._launch()
        t2 = time.perf_counter()
        self._retire()
        t3 = time.perf_counter()
        self.timing["admit"] += t1 - t0
        self.timing["launch"] += t2 - t1
        self.timing["retire"] += t3 - t2
        return issued

    def gpu_busy_seconds(self) -> float:
        """Sum of the fused launches' durations on the stream (task upload
        included), from events; call after `drain`."""
        return sum(a.elapsed_ms(b) for a, b in self._launch_events) / 1e3

    def drain(self, timeout: float = 600.0):
        t_end = time.perf_counter() + timeout
        while self.pending() and time.perf_counter() < t_end:
            if not self.step():
                if self.finishing:
                    self._retire(block=True)
                else:
                    time.sleep(0.0005)
        self._retire(block=True)


def serve_poisson(server: SolveServer, make_request, rate: float, duration: float, seed: int = 0):
    """Submit requests with exponential inter-arrival times (rate per second)
    for `duration` seconds while running the serving loop; then drain.
    Returns the completed requests."""
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    next_t = t0 + rng.exponential(1.0 / rate)
    submitted = []
    while True:
        now = time.perf_counter()
        while next_t <= now and next_t - t0 < duration:
            model, params = make_request(len(submitted))
            submitted.append(server.submit(model, params, t_submit=next_t))
            next_t += rng.exponential(1.0 / rate)
        if next_t - t0 >= duration and now - t0 >= duration:
            break
        if not server.step():
            if server.finishing:
                server._retire()
            time.sleep(0.0002)
    server.drain()
    return submitted
