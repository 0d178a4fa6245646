"""Batched solves of freshly drawn parameters, lowered in worker processes.

A draw's tables come from the same host lowering as a single solve
(`lower_problem` + `GsProgram.of(..., defer=True)`), run in a pool of worker
processes; `AssembledBatch` places the packed draws side by side in one device
buffer (each draw's descriptors relocated to its base address, the per-step
task lists concatenated with descriptor offsets) and runs backward induction
as one fused launch per backward step for all draws, as `GsProgram.of` does
for problems lowered in-process. `FreshBatches` pipelines batches: the workers
lower batch k+1 while the GPU solves batch k, so the fresh-draw rate is set by
the slower of the two.

A draw's results are bit-identical to those of the same (model, params) in
any other program: the kernel, descriptors, tile and launch variant per step
are the same (tests check this against `GsProgram`).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from pylcm_amd import _native as nat
from pylcm_amd._gs_program import ALIGN, DESC_DTYPE, TASK_DTYPE, GsProgram, lower_problem, step_lds_bytes


def _align(x: int) -> int:
    return (int(x) + ALIGN - 1) // ALIGN * ALIGN


def pack_problem(model, params, dtype, tile: int) -> dict:
    """Lower and pack one (model, params) on the host (no device code):
    everything needed to place and run it anywhere in device memory."""
    dt = np.dtype(dtype)
    low = lower_problem(model, params, dt)
    prog = GsProgram.of([(model, params)], dt, lowered=[low], tile=tile, defer=True)
    D0, masks = prog.relocatable_records()
    return {"image": prog._arena.image(prog.out_start), "D0": D0, "masks": masks,
            "tasks_host": prog.tasks_host, "outs": prog.outs, "n_steps": prog.n_steps,
            "lds_g": prog._lds_g, "lds_k": prog._lds_k, "wdep_by_step": prog.wdep_by_step,
            "total_bytes": prog.total_bytes, "out_start": prog.out_start, "n_desc": prog.n_desc,
            "flops_by_step": prog.flops_by_step}


# -- worker processes -------------------------------------------------------------

_MODELS: dict = {}


def _init_worker(models_blob: bytes):
    import cloudpickle

    _MODELS.clear()
    _MODELS.update(cloudpickle.loads(models_blob))


def _pack_chunk(items, dtype: str, tile: int) -> list:
    """Pack [(model key, params), ...] with the worker's models."""
    return [pack_problem(_MODELS[k], p, dtype, tile) for k, p in items]


class WorkerPool:
    """Spawned worker processes holding `models` (sent once, by cloudpickle);
    `submit(draws)` packs [(model key, params), ...] in chunks."""

    def __init__(self, models: dict, procs: int, dtype, tile: int, chunk: int = 16):
        import multiprocessing
        from concurrent.futures import ProcessPoolExecutor

        import cloudpickle

        self.dt = np.dtype(dtype)
        self.tile = int(tile)
        self.chunk = int(chunk)
        self.pool = ProcessPoolExecutor(max_workers=int(procs), mp_context=multiprocessing.get_context("spawn"),
                                        initializer=_init_worker, initargs=(cloudpickle.dumps(models),))
        # start every worker now, so the first batch pays no start-up
        for f in [self.pool.submit(_pack_chunk, [], self.dt.str, self.tile) for _ in range(int(procs))]:
            f.result()

    def submit(self, draws):
        """Futures of packed chunks, in the order of `draws`."""
        draws = list(draws)
        return [self.pool.submit(_pack_chunk, draws[i: i + self.chunk], self.dt.str, self.tile)
                for i in range(0, len(draws), self.chunk)]

    @staticmethod
    def gather(futures) -> list:
        out = []
        for f in futures:
            out.extend(f.result())
        return out

    def close(self):
        self.pool.shutdown(wait=True, cancel_futures=True)


# -- assembly on the device ---------------------------------------------------------


class AssembledBatch:
    """Packed problems in one device buffer, solved with one fused launch
    per backward step (problems aligned at their last period).

    Layout: every problem's outputs (V and policies, written by the kernels
    before any read, never uploaded), then every problem's tables, then the
    descriptor array and the per-step task lists; one upload covers the
    tables, descriptors and tasks."""

    def __init__(self, packed: list, dtype, tile: int, ctx=None, buf=None, *, base: int | None = None,
                 staging=None, upload: bool = True):
        """With `base` given (and no `buf`), only the layout and the records
        for that device address are built (`records()`), nothing is uploaded.
        `buf` is a `nat.DeviceBuffer` or a callable returning one of at least
        the given size. `staging` (a `nat.PinnedBuffer` of at least
        `upload_bytes`) receives the host image when given; with
        `upload=False` the image is built but `upload()` sends it."""
        self.dt = np.dtype(dtype)
        self.code = nat.LCM_DTYPE_F64 if self.dt == np.float64 else nat.LCM_DTYPE_F32
        self.tile = int(tile)
        dry = base is not None and buf is None
        self.ctx = ctx if (ctx is not None or dry) else nat.Context()
        self.packed = packed
        n = len(packed)
        self.n_steps = max(p["n_steps"] for p in packed)
        S = self.n_steps
        self.out_offs = np.zeros(n, dtype=np.int64)
        self.tab_offs = np.zeros(n, dtype=np.int64)
        self.desc_base = np.zeros(n, dtype=np.int64)
        off, nd = 0, 0
        for i, p in enumerate(packed):
            self.out_offs[i] = off
            off += _align(p["out_start"])
        self.tables_off = off
        for i, p in enumerate(packed):
            self.tab_offs[i] = off
            self.desc_base[i] = nd
            off += _align(p["image"].size)
            nd += p["n_desc"]
        self.n_desc = nd
        self.desc_off = off
        off += _align(nd * DESC_DTYPE.itemsize)
        self.tasks = []
        self.task_off = []
        self.lds_g, self.lds_k, self.variant = [1] * S, [0] * S, [0] * S
        self.flops_by_step = [0.0] * S
        for s in range(S):
            # problems in their order (sizes interleaved, see GsProgram._pack)
            parts, bases = [], []
            for i, p in enumerate(packed):
                if s < p["n_steps"] and len(p["tasks_host"][s]):
                    parts.append(p["tasks_host"][s])
                    bases.append(self.desc_base[i])
                    if p["lds_g"][s] > self.lds_g[s]:
                        self.lds_g[s] = p["lds_g"][s]
                    if p["lds_k"][s] > self.lds_k[s]:
                        self.lds_k[s] = p["lds_k"][s]
                    if p["wdep_by_step"][s] > self.variant[s]:
                        self.variant[s] = p["wdep_by_step"][s]
                    self.flops_by_step[s] += p["flops_by_step"][s]
            if parts:
                tl = np.concatenate(parts)
                n_t = len(tl)
                arr = np.zeros((n_t + 7) // 8 * 8, dtype=TASK_DTYPE)
                arr["desc"][:n_t] = tl[:, 0] + np.repeat(np.asarray(bases, dtype=np.int32), [len(t) for t in parts])
                arr["sd"][:n_t] = tl[:, 1]
                arr["w0"][:n_t] = tl[:, 2]
                arr["desc"][n_t:] = -1
            else:
                arr = np.zeros(0, dtype=TASK_DTYPE)
            self.tasks.append(arr)
            self.task_off.append(off)
            off += _align(max(arr.nbytes, 1))
        self.total_bytes = _align(off)
        self.upload_bytes = self.total_bytes - self.tables_off
        self._graph = None
        if dry:
            self.buf, self.base = None, int(base)
            return
        self.lds_by_step = step_lds_bytes(self.dt, self.lds_g, self.lds_k)
        if callable(buf):
            buf = buf(self.total_bytes)
        self.buf = buf if buf is not None else nat.DeviceBuffer(self.total_bytes)
        self.base = self.buf.ptr
        if staging is not None and staging.nbytes >= self.upload_bytes:
            self._img = self.host_image(into=staging.array)
        else:
            self._img = self.host_image()
        if upload:
            self.upload()

    def records(self) -> np.ndarray:
        """Every problem's descriptors relocated to its place, in index order:
        a pointer below the problem's `out_start` addresses its outputs, the
        others its tables."""
        counts = [p["n_desc"] for p in self.packed]
        D = np.concatenate([p["D0"] for p in self.packed])
        base = np.uint64(self.base)
        out_row = np.repeat(base + self.out_offs.astype(np.uint64), counts)
        start_row = np.repeat(np.asarray([p["out_start"] for p in self.packed], dtype=np.uint64), counts)
        tab_row = np.repeat(base + self.tab_offs.astype(np.uint64), counts) - start_row
        fields = set().union(*(p["masks"] for p in self.packed))
        for f in fields:
            # a problem without pointers in a list field has no mask for it
            m = np.concatenate([p["masks"][f] if f in p["masks"] else
                                np.zeros(p["D0"][f].shape, dtype=bool) for p in self.packed])
            o = D[f]
            if o.ndim == 1:
                sr, orow, trow = start_row, out_row, tab_row
            else:
                sr, orow, trow = start_row[:, None], out_row[:, None], tab_row[:, None]
            moved = np.where(o < sr, o + orow, o + trow)
            D[f] = np.where(m, moved, o)
        return D

    def host_image(self, into=None) -> np.ndarray:
        """Bytes [tables_off, total_bytes) of the buffer: the tables,
        descriptors and tasks (written into `into` when given)."""
        img = np.empty(self.upload_bytes, dtype=np.uint8) if into is None else into[: self.upload_bytes]
        t0 = self.tables_off
        for i, p in enumerate(self.packed):
            o = int(self.tab_offs[i]) - t0
            img[o: o + p["image"].size] = p["image"]
        D = self.records()
        o = self.desc_off - t0
        img[o: o + D.nbytes] = np.frombuffer(D.tobytes(), dtype=np.uint8)
        for s, arr in enumerate(self.tasks):
            if len(arr):
                o = self.task_off[s] - t0
                img[o: o + arr.nbytes] = np.frombuffer(arr.tobytes(), dtype=np.uint8)
        return img

    def upload(self):
        """Send the host image (waits for the copy)."""
        nat.h2d(self.ctx, self.base + self.tables_off, self._img)
        self._img = None

    def _record(self):
        L = nat.lib()
        for s, arr in enumerate(self.tasks):
            if not len(arr):
                continue
            st = L.lcm_gs_launch_ex(self.ctx.handle, self.code, C.c_void_p(self.base + self.desc_off),
                                    C.c_void_p(self.base + self.task_off[s]), len(arr), self.tile,
                                    C.c_size_t(self.lds_by_step[s]), int(self.variant[s]))
            nat._check(st, f"lcm_gs_launch (step {s})")

    def run(self, *, use_graph: bool = False, sync: bool = True):
        if use_graph:
            if self._graph is None:
                self._graph = nat.Graph.capture(self.ctx, self._record)
            self._graph.launch(self.ctx)
        else:
            self._record()
        if sync:
            self.ctx.sync()

    def results(self, i: int) -> dict:
        """{period: {regime: (V, policy)}} of problem i (copies from the device)."""
        p = self.packed[i]
        b = self.base + int(self.out_offs[i])
        out = {}
        for (t, r), o in p["outs"][0].items():
            n = int(np.prod(o.shape, dtype=np.int64)) if o.shape else 1
            pn = int(np.prod(o.pol_shape, dtype=np.int64)) if o.pol_shape else 1
            V = np.empty(n, dtype=self.dt)
            P = np.empty(pn, dtype=np.int32)
            nat.d2h(self.ctx, b + o.v_off, V)
            nat.d2h(self.ctx, b + o.p_off, P)
            out.setdefault(t, {})[r] = (V.reshape(o.shape), P.reshape(o.pol_shape))
        return out


class FreshBatches:
    """Solve a stream of parameter batches: batch k+1 is lowered by the
    worker pool and placed into host memory while the GPU solves batch k.
    `models` maps a key to a model; a batch is a list of (model key, params).

    Two slots (a device buffer and a pinned staging buffer each) alternate
    between consecutive batches; a slot grows when a batch needs more."""

    def __init__(self, models: dict, dtype, *, procs: int = 8, tile: int = 256, chunk: int = 16):
        self.dt = np.dtype(dtype)
        self.tile = int(tile)
        self.pool = WorkerPool(models, procs, dtype, tile, chunk=chunk)
        self.ctx = nat.Context()
        self._dev = [None, None]
        self._pin = [None, None]

    def _slot_buffer(self, k):
        def get(nbytes):
            b = self._dev[k % 2]
            if b is None or b.nbytes < nbytes:
                if b is not None:
                    b.free()
                b = self._dev[k % 2] = nat.DeviceBuffer(int(nbytes * 1.25))
            return b
        return get

    def _place(self, packed, k):
        pin = self._pin[k % 2]
        ab = AssembledBatch(packed, self.dt, self.tile, ctx=self.ctx, buf=self._slot_buffer(k), staging=pin,
                            upload=False)
        if pin is None or pin.nbytes < ab.upload_bytes:
            # a larger pinned buffer for this slot's next batches
            if pin is not None:
                pin.free()
            self._pin[k % 2] = nat.PinnedBuffer(int(ab.upload_bytes * 1.25))
        return ab

    def run(self, batches, on_done=None):
        """Solve every batch in order; `on_done(k, assembled)` is called after
        batch k has finished on the device (before its slot is reused).
        Returns per batch the seconds spent waiting for the workers, placing
        it in host memory (layout, records, image), uploading it, and waiting
        for the device once the next batch was placed."""
        import time

        batches = list(batches)
        stats = []
        if not batches:
            return stats

        def lower_and_place(k, pending):
            a = time.perf_counter()
            packed = self.pool.gather(pending)
            nxt = self.pool.submit(batches[k + 1]) if k + 1 < len(batches) else None
            b = time.perf_counter()
            ab = self._place(packed, k)
            return ab, nxt, {"wait_workers_s": b - a, "place_s": time.perf_counter() - b}

        ab, pending, st = lower_and_place(0, self.pool.submit(batches[0]))
        for k in range(len(batches)):
            t0 = time.perf_counter()
            ab.upload()
            st["upload_s"] = time.perf_counter() - t0
            ab.run(use_graph=False, sync=False)
            if k + 1 < len(batches):
                # lower and place batch k+1 while the device solves batch k
                nab, pending, nst = lower_and_place(k + 1, pending)
            t1 = time.perf_counter()
            self.ctx.sync()
            st["device_wait_s"] = time.perf_counter() - t1
            stats.append(st)
            if on_done is not None:
                on_done(k, ab)
            if k + 1 < len(batches):
                ab, st = nab, nst
        return stats

    def close(self):
        self.pool.close()
        for b in self._dev + self._pin:
            if b is not None:
                b.free()
        self._dev = [None, None]
        self._pin = [None, None]
