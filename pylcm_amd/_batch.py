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
import glob
import itertools
import os
import tempfile
import uuid

import numpy as np

from pylcm_amd import _native as nat
from pylcm_amd._gs_program import ALIGN, DESC_DTYPE, TASK_DTYPE, GsProgram, lower_problem, step_lds_bytes


def _align(x: int) -> int:
    return (int(x) + ALIGN - 1) // ALIGN * ALIGN


# The descriptor's pointers fill its first PTR_WORDS 8-byte words.
PTR_WORDS = DESC_DTYPE.fields["Bf"][1] // 8 + 1
WORDS = DESC_DTYPE.itemsize // 8
assert DESC_DTYPE.itemsize % 8 == 0 and all(
    DESC_DTYPE.fields[f][1] < PTR_WORDS * 8 for f in GsProgram._PTR_FIELDS + ("Fa", "Fb", "Vn", "M", "Vs", "ms"))


def pointer_classes(D0: np.ndarray, masks: dict, out_start: int) -> np.ndarray:
    """(n_desc, PTR_WORDS) uint8 per pointer word of records at base 0: 0 for
    no pointer, 1 for a pointer into the outputs (below `out_start`), 2 for a
    pointer into the tables."""
    n = len(D0)
    is_ptr = np.zeros((n, PTR_WORDS), dtype=bool)
    for f, m in masks.items():
        w0 = DESC_DTYPE.fields[f][1] // 8
        m2 = m.reshape(n, -1)
        is_ptr[:, w0: w0 + m2.shape[1]] |= m2
    words = D0.view(np.uint64).reshape(n, WORDS)[:, :PTR_WORDS]
    cls = np.where(words < np.uint64(out_start), 1, 2).astype(np.uint8)
    cls[~is_ptr] = 0
    return cls


def pack_problem(model, params, dtype, tile: int, keep_lowered: bool = False) -> dict:
    """Lower and pack one (model, params) on the host (no device code):
    everything needed to place and run it anywhere in device memory. With
    `keep_lowered` the result also holds the lowered (tables, grids)."""
    dt = np.dtype(dtype)
    low = lower_problem(model, params, dt)
    prog = GsProgram.of([(model, params)], dt, lowered=[low], tile=tile, defer=True)
    D0, masks = prog.relocatable_records()
    return {"image": prog._arena.image(prog.out_start), "D0": D0, "masks": masks,
            "ptr_cls": pointer_classes(D0, masks, prog.out_start),
            "tasks_cat": (np.concatenate(prog.tasks_host).astype(np.int32) if prog.tasks_host
                          else np.zeros((0, 3), dtype=np.int32)),
            "task_counts": np.array([len(t) for t in prog.tasks_host], dtype=np.int64),
            "tasks_host": prog.tasks_host, "outs": prog.outs, "n_steps": prog.n_steps,
            "lds_g": prog._lds_g, "lds_k": prog._lds_k, "wdep_by_step": prog.wdep_by_step,
            "total_bytes": prog.total_bytes, "out_start": prog.out_start, "n_desc": prog.n_desc,
            "flops_by_step": prog.flops_by_step, **({"lowered": low} if keep_lowered else {})}


# -- worker processes -------------------------------------------------------------

_MODELS: dict = {}


def _init_worker(models_blob: bytes):
    import cloudpickle

    _MODELS.clear()
    _MODELS.update(cloudpickle.loads(models_blob))


_SPILL_IDS = itertools.count()


SPILL_MIN_FREE = 4 << 30  # bytes /dev/shm must have free to be used for spill files


def spill_prefix() -> str:
    """A fresh path prefix for files handed between processes: memory-backed
    /dev/shm when present, writable and not small (container defaults can be
    64 MiB), else the temporary directory."""
    base = tempfile.gettempdir()
    if os.path.isdir("/dev/shm") and os.access("/dev/shm", os.W_OK):
        st = os.statvfs("/dev/shm")
        if st.f_bavail * st.f_frsize >= SPILL_MIN_FREE:
            base = "/dev/shm"
    return os.path.join(base, f"pylcm_amd_{os.getpid()}_{uuid.uuid4().hex[:8]}_")


def unlink_quiet(path: str):
    try:
        os.unlink(path)
    except OSError:
        pass


def _pack_chunk(items, dtype: str, tile: int, spill: str | None = None) -> list:
    """Pack [(model key, params), ...] with the worker's models. With `spill`
    the table images go into one file at that prefix (`spilled` = (path,
    offset, size) per problem) instead of through the result pipe: the
    images are most of a packed problem's bytes, and a file in memory is
    read by the parent without a copy."""
    packs = [pack_problem(_MODELS[k], p, dtype, tile) for k, p in items]
    if spill is None or not packs:
        return packs
    path = f"{spill}{os.getpid()}_{next(_SPILL_IDS)}"
    off = 0
    with open(path, "wb") as f:
        for p in packs:
            img = np.ascontiguousarray(p.pop("image")).view(np.uint8).reshape(-1)
            f.write(memoryview(img))
            p["spilled"] = (path, off, int(img.size))
            off += int(img.size)
    return packs


def _unspill_chunk(packs: list):
    """Map the file `_pack_chunk` spilled the chunk's images into and point
    every problem's `image` at its bytes; the file is removed at once (the
    mapping keeps its pages until the images are dropped)."""
    if not packs or "spilled" not in packs[0]:
        return
    path = packs[0]["spilled"][0]
    try:
        total = sum(p["spilled"][2] for p in packs)
        mm = np.memmap(path, dtype=np.uint8, mode="r", shape=(total,)) if total else np.zeros(0, np.uint8)
    finally:
        unlink_quiet(path)
    for p in packs:
        _, off, n = p.pop("spilled")
        p["image"] = mm[off: off + n]


class WorkerPool:
    """Spawned worker processes holding `models` (sent once, by cloudpickle);
    `submit(draws)` packs [(model key, params), ...] in chunks. The table
    images come back through memory-backed files (see `_pack_chunk`)."""

    def __init__(self, models: dict, procs: int, dtype, tile: int, chunk: int = 16):
        import multiprocessing
        from concurrent.futures import ProcessPoolExecutor

        import cloudpickle

        self.dt = np.dtype(dtype)
        self.tile = int(tile)
        self.chunk = int(chunk)
        self.spill = spill_prefix()
        self.pool = ProcessPoolExecutor(max_workers=int(procs), mp_context=multiprocessing.get_context("spawn"),
                                        initializer=_init_worker, initargs=(cloudpickle.dumps(models),))
        # start every worker now, so the first batch pays no start-up
        for f in [self.pool.submit(_pack_chunk, [], self.dt.str, self.tile) for _ in range(int(procs))]:
            f.result()

    def submit(self, draws):
        """Futures of packed chunks, in the order of `draws`."""
        draws = list(draws)
        return [self.pool.submit(_pack_chunk, draws[i: i + self.chunk], self.dt.str, self.tile, self.spill)
                for i in range(0, len(draws), self.chunk)]

    @staticmethod
    def gather(futures) -> list:
        out = []
        for f in futures:
            chunk = f.result()
            _unspill_chunk(chunk)
            out.extend(chunk)
        return out

    def close(self):
        self.pool.shutdown(wait=True, cancel_futures=True)
        # files of chunks that were packed but never gathered
        for path in glob.glob(glob.escape(self.spill) + "*"):
            unlink_quiet(path)


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
        # per-step parameters: the maximum (LDS, variant) or sum (flops) over
        # the problems that still run at that step
        def per_step(key, fill):
            m = np.full((n, S), fill, dtype=np.float64)
            for i, p in enumerate(packed):
                m[i, : p["n_steps"]] = p[key][: p["n_steps"]]
            return m

        counts = np.zeros((n, S), dtype=np.int64)
        for i, p in enumerate(packed):
            counts[i, : len(p["task_counts"])] = p["task_counts"]
        live = counts > 0
        self.lds_g = [int(v) for v in np.max(np.where(live, per_step("lds_g", 1), 1), axis=0, initial=1)]
        self.lds_k = [int(v) for v in np.max(np.where(live, per_step("lds_k", 0), 0), axis=0, initial=0)]
        self.variant = [int(v) for v in np.max(np.where(live, per_step("wdep_by_step", 0), 0), axis=0, initial=0)]
        self.flops_by_step = [float(v) for v in np.sum(np.where(live, per_step("flops_by_step", 0), 0), axis=0)]
        # every step's task list: the problems' blocks in problem order (sizes
        # interleaved, see GsProgram._pack), descriptors offset by the
        # problem's first descriptor, padded to a multiple of 8 with desc -1
        per = counts.sum(axis=0)
        padded = (per + 7) // 8 * 8
        step_start = np.concatenate([[0], np.cumsum(padded)])
        src_first = np.concatenate([[0], np.cumsum([len(p["tasks_cat"]) for p in packed])])[:-1]
        src = src_first[:, None] + np.concatenate([np.zeros((n, 1), dtype=np.int64), np.cumsum(counts, axis=1)[:, :-1]],
                                                   axis=1)
        dst = step_start[:-1][None, :] + np.concatenate([np.zeros((1, S), dtype=np.int64),
                                                          np.cumsum(counts, axis=0)[:-1]], axis=0)
        cT, sT, dT = counts.T.ravel(), src.T.ravel(), dst.T.ravel()
        total = int(step_start[-1])
        allt = np.concatenate([p["tasks_cat"] for p in packed]) if n else np.zeros((0, 3), dtype=np.int32)
        # gather: destination row d of block (s, i) reads source row d - (dst - src)
        rows = np.repeat(dT, cT) + (np.arange(int(cT.sum())) - np.repeat(np.cumsum(cT) - cT, cT))
        srows = rows - np.repeat(dT - sT, cT)
        # rows of four int32 (desc, sd, w0, pad), viewed as task records at the end
        assert TASK_DTYPE.itemsize == 16 and TASK_DTYPE.names[:3] == ("desc", "sd", "w0")
        flat4 = np.zeros((total, 4), dtype=np.int32)
        flat4[:, 0] = -1
        moved = allt[srows].astype(np.int32)
        moved[:, 0] += np.repeat(np.broadcast_to(self.desc_base[:, None], (n, S)).T.ravel(), cT).astype(np.int32)
        flat4[rows, :3] = moved
        flat = flat4.view(TASK_DTYPE).reshape(-1)
        self.tasks = [flat[step_start[s]: step_start[s + 1]] for s in range(S)]
        self.task_off = []
        for s in range(S):
            self.task_off.append(off)
            off += _align(max(self.tasks[s].nbytes, 1))
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

    def records(self, into=None) -> np.ndarray:
        """Every problem's descriptors relocated to its place, in index order
        (written into `into`, a uint8 array of n_desc records, when given): a
        pointer below the problem's `out_start` addresses its outputs, the
        others its tables."""
        raw = np.empty(self.n_desc * DESC_DTYPE.itemsize, dtype=np.uint8) if into is None else into
        self._records_of(0, len(self.packed), raw)
        return raw.view(DESC_DTYPE)

    def _records_of(self, lo: int, hi: int, raw: np.ndarray):
        """`records` of problems [lo, hi) into `raw` (all records' bytes)."""
        packed = self.packed[lo:hi]
        if not packed:
            return
        sz = DESC_DTYPE.itemsize
        d0, d1 = int(self.desc_base[lo]), int(self.desc_base[hi - 1]) + packed[-1]["n_desc"]
        out = raw[d0 * sz: d1 * sz]
        n = d1 - d0
        np.concatenate([p["D0"].view(np.uint8).reshape(-1) for p in packed], out=out)
        counts = [p["n_desc"] for p in packed]
        base = np.uint64(self.base)
        lut = np.zeros((n, 3), dtype=np.uint64)
        lut[:, 1] = np.repeat(base + self.out_offs[lo:hi].astype(np.uint64), counts)
        start = np.asarray([p["out_start"] for p in packed], dtype=np.uint64)
        lut[:, 2] = np.repeat(base + self.tab_offs[lo:hi].astype(np.uint64) - start, counts)
        cls = np.concatenate([p["ptr_cls"] for p in packed])
        W = out.view(np.uint64).reshape(n, WORDS)
        W[:, :PTR_WORDS] += np.take_along_axis(lut, cls.astype(np.intp), axis=1)

    def host_image(self, into=None) -> np.ndarray:
        """Bytes [tables_off, total_bytes) of the buffer: the tables,
        descriptors and tasks (written into `into` when given)."""
        img = np.empty(self.upload_bytes, dtype=np.uint8) if into is None else into[: self.upload_bytes]
        t0 = self.tables_off

        def copy(lo, hi):
            for i in range(lo, hi):
                p = self.packed[i]
                o = int(self.tab_offs[i]) - t0
                img[o: o + p["image"].size] = p["image"]

        n = len(self.packed)
        k = min(8, max(1, n // 16))
        bounds = np.linspace(0, n, k + 1).astype(int)
        if k > 1:
            # large copies and array operations release the GIL: the tables
            # and the records of k problem ranges go in parallel
            from concurrent.futures import ThreadPoolExecutor

            o = self.desc_off - t0
            recs = img[o: o + self.n_desc * DESC_DTYPE.itemsize]
            with ThreadPoolExecutor(k) as ex:
                jobs = [ex.submit(copy, int(a), int(b)) for a, b in zip(bounds[:-1], bounds[1:])]
                jobs += [ex.submit(self._records_of, int(a), int(b), recs) for a, b in zip(bounds[:-1], bounds[1:])]
                for j in jobs:
                    j.result()
        else:
            copy(0, n)
            o = self.desc_off - t0
            self.records(into=img[o: o + self.n_desc * DESC_DTYPE.itemsize])
        for s, arr in enumerate(self.tasks):
            if len(arr):
                o = self.task_off[s] - t0
                img[o: o + arr.nbytes] = arr.view(np.uint8).reshape(-1)
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
            # batch k + 1 is queued before batch k is gathered, so workers that
            # finish their last chunk of batch k go on with the next batch
            nxt = self.pool.submit(batches[k + 1]) if k + 1 < len(batches) else None
            a = time.perf_counter()
            packed = self.pool.gather(pending)
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
