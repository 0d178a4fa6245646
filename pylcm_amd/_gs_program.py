"""Device plan of grid-search solves: arena, descriptors, tasks, launches.

A `GsProgram` holds any number of problems — (model, params) pairs that may
differ in grid sizes and horizons — lowered to `PeriodTables`. It packs all
tables into one device arena (identical tables are stored once), writes one
`LcmGsDesc` per (problem, period, regime) and one task list per backward step,
and runs backward induction as one fused launch per step, optionally captured
once as a HIP graph over the horizon. Problems are aligned at their last
period: step s solves period `T_p - 1 - s` of problem p.

This replaces the reference's host loop
(`src/_lcm/solution/backward_induction.py:155-275`), which dispatches one
compiled core per (regime, period) and synchronises after each period.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from pylcm_amd import _native as nat
from pylcm_amd import autotune
from pylcm_amd._build import PeriodLowering, resolve_grids, table_digest
from pylcm_amd.exceptions import UnsupportedModelError

ALIGN = 256
DESC_DTYPE = np.dtype(nat.LcmGsDesc)
TASK_DTYPE = np.dtype(nat.LcmGsTask)
LDS_LIMIT = 160 * 1024  # bytes of LDS one workgroup may use on gfx950
FLOPS_PER_Q = 10  # budget 2, coordinate 2, interpolation 3, aggregator 2, max 1


def _align(x: int) -> int:
    return (x + ALIGN - 1) // ALIGN * ALIGN


class _Arena:
    """Host image of a device arena with content-deduplicated tables."""

    def __init__(self, digest: bool = True):
        self.parts: list[tuple[int, bytes]] = []
        self.size = 0
        self.index: dict[str, int] = {}
        self.by_id: dict[tuple, int] = {}
        self._keep: list = []
        # content digests find equal tables that are different objects (across
        # the problems of a batch); within one problem's lowering equal tables
        # are already the same object
        self.digest = digest

    def add(self, arr: np.ndarray, dedupe: bool = True, dtype=None) -> int:
        """Place `arr` (cast to `dtype`) in the arena; the same array object or
        the same contents are stored once."""
        if dedupe and dtype is not None:
            idk = (id(arr), np.dtype(dtype).str)
            if idk in self.by_id:
                return self.by_id[idk]
        a = np.ascontiguousarray(arr if dtype is None else arr.astype(dtype))
        key = table_digest(a) + str(a.dtype) if dedupe and self.digest else None
        if key is not None and key in self.index:
            off = self.index[key]
        else:
            off = _align(self.size)
            self.parts.append((off, a.tobytes()))
            self.size = off + a.nbytes
            if key is not None:
                self.index[key] = off
        if dedupe and dtype is not None:
            self.by_id[(id(arr), np.dtype(dtype).str)] = off
            self._keep.append(arr)
        return off

    def reserve(self, nbytes: int) -> int:
        off = _align(self.size)
        self.size = off + max(int(nbytes), 1)
        return off

    def image(self, start: int = 0) -> np.ndarray:
        """Bytes [start, size) of the arena (reserved space reads as zeros)."""
        buf = np.zeros(_align(max(self.size, 1)) - start, dtype=np.uint8)
        for off, b in self.parts:
            if off >= start:
                buf[off - start: off - start + len(b)] = np.frombuffer(b, dtype=np.uint8)
        return buf


@dataclass
class _Out:
    v_off: int
    p_off: int
    shape: tuple
    pol_shape: tuple


def _strides(tab: np.ndarray):
    n_sdx, n_wx, n_d, n_c = tab.shape
    ssd = n_wx * n_d * n_c if n_sdx > 1 else 0
    sw = n_d * n_c if n_wx > 1 else 0
    return ssd, sw


def interval_form(F, n_sd, n_d, n_c, dt):
    """Constraint rows `a*w + b >= 0` (or `> 0`) with every `a` in {-1, 0, 1}
    as one interval `lo <= w <= hi` per action, or None.

    Exact in the model dtype: with `b` rounded to it, `w + b >= 0` holds iff
    `w >= -b`, `-w + b >= 0` iff `w <= b`, and a strict bound moves to the next
    representable number.
    """
    dt = np.dtype(dt)
    n_sdx = max([1] + [max(Fa.shape[0], Fb.shape[0]) for Fa, Fb, _ in F])
    lo = np.full((n_sdx, 1, n_d, n_c), -np.inf, dtype=dt)
    hi = np.full((n_sdx, 1, n_d, n_c), np.inf, dtype=dt)
    for Fa, Fb, strict in F:
        if Fa.shape[1] != 1 or Fb.shape[1] != 1:
            return None
        a = np.broadcast_to(Fa, lo.shape)
        b = np.broadcast_to(Fb, lo.shape).astype(dt)
        if not np.all((a == 1) | (a == 0) | (a == -1)):
            return None
        with np.errstate(invalid="ignore"):
            tlo = np.where(a == 1, -b, -np.inf).astype(dt)
            thi = np.where(a == -1, b, np.inf).astype(dt)
            if strict:
                tlo = np.where(a == 1, np.nextafter(tlo, dt.type(np.inf)), tlo)
                thi = np.where(a == -1, np.nextafter(thi, dt.type(-np.inf)), thi)
            ok0 = (b > 0) if strict else (b >= 0)
        tlo = np.where((a == 0) & ~ok0, dt.type(np.inf), tlo)
        lo = np.maximum(lo, tlo)
        hi = np.minimum(hi, thi)
    if n_sdx > 1 and np.all(lo == lo[:1]) and np.all(hi == hi[:1]):
        lo, hi = lo[:1], hi[:1]
    return np.ascontiguousarray(lo), np.ascontiguousarray(hi)


def chunk_bounds(lo, hi, chunk):
    """(min lo, max hi) per (sd, d, chunk of `chunk` actions): a block whose
    states all lie outside [min lo, max hi] skips the chunk."""
    n_sdx, _, n_d, n_c = lo.shape
    n_ch = (n_c + chunk - 1) // chunk
    lo_p = np.full((n_sdx, n_d, n_ch * chunk), np.inf, dtype=lo.dtype)
    hi_p = np.full((n_sdx, n_d, n_ch * chunk), -np.inf, dtype=hi.dtype)
    lo_p[:, :, :n_c] = lo[:, 0]
    hi_p[:, :, :n_c] = hi[:, 0]
    out = np.empty((n_sdx, n_d, n_ch, 2), dtype=lo.dtype)
    out[..., 0] = lo_p.reshape(n_sdx, n_d, n_ch, chunk).min(axis=-1)
    out[..., 1] = hi_p.reshape(n_sdx, n_d, n_ch, chunk).max(axis=-1)
    return out


def folded_next_state(A, B, g_start, g_stop, g_n, dt):
    """(A / step, (B - start) / step) of a linear next grid, each rounded once
    to `dt` from the `dt`-rounded A, B, start and step (step computed in `dt`
    as the kernel computes it: (stop - start) / (n - 1)). The folded loop's
    coordinate of the next state is then one fma, Af * w + Bf."""
    dt = np.dtype(dt)
    c = dt.type
    start, stop = c(g_start), c(g_stop)
    step = c(stop - start) / c(g_n - 1) if g_n > 1 else c(1)
    A64 = np.asarray(A, dtype=dt).astype(np.float64)
    B64 = np.asarray(B, dtype=dt).astype(np.float64)
    return ((A64 / np.float64(step)).astype(dt),
            ((B64 - np.float64(start)) / np.float64(step)).astype(dt))


def lower_problem(model, params, dtype):
    """(tables by (period, regime), resolved grids) for one params set."""
    rg = resolve_grids(model.layouts, params, model.fixed_params, dtype)
    low = PeriodLowering(model, params, rg)
    tabs = {}
    same = {}  # period key -> tables shared by every period with that key
    for t in range(model.n_periods):
        for r in model.active_regimes(t):
            lay = model.layouts[r]
            if lay.is_dcegm:
                raise UnsupportedModelError("DC-EGM regimes are solved by the DC-EGM program")
            key = low.period_key(lay, t)
            if key is not None and key in same:
                tabs[(t, r)] = same[key]
                continue
            tabs[(t, r)] = low.lower(lay, t)
            if key is not None:
                same[key] = tabs[(t, r)]
    return tabs, rg


_LDS_CACHE: dict = {}


def _lds_bytes(code: int, g: int, k: int) -> int:
    key = (code, g, k)
    b = _LDS_CACHE.get(key)
    if b is None:
        b = _LDS_CACHE[key] = int(nat.load_library().lcm_gs_lds_bytes(code, g, k))
    return b


def step_lds_bytes(dtype, gmax, kmax):
    """`lcm_gs_lds_bytes` for each step's largest next grid and constraint count
    (asked of the library once per distinct pair)."""
    code = nat.LCM_DTYPE_F64 if np.dtype(dtype) == np.float64 else nat.LCM_DTYPE_F32
    out = [_lds_bytes(code, int(g), int(k)) for g, k in zip(gmax, kmax)]
    for b, g, k in zip(out, gmax, kmax):
        if b > LDS_LIMIT:
            raise UnsupportedModelError(
                f"a next-state grid of {int(g)} points ({int(k)} general constraints) needs {b} B of LDS "
                f"per workgroup in {np.dtype(dtype)}, above the {LDS_LIMIT} B of a gfx950 compute unit; "
                "use a coarser grid or float32")
    return out


class GsProgram:
    """Packed, device-resident grid-search solves.

    `GsProgram(model, [params, ...], dtype)` solves one model for many params;
    `GsProgram.of([(model, params), ...], dtype)` mixes models.
    """

    def __init__(self, model, params_list, dtype, *, tile: int | None = None, ctx=None,
                 lowered=None):
        self._init([(model, p) for p in params_list], dtype, tile=tile, ctx=ctx, lowered=lowered)

    @classmethod
    def of(cls, problems, dtype, *, tile=None, ctx=None, lowered=None, defer=False):
        """With `defer=True` only the host image is built; `materialize(ptr)`
        places it at a caller-provided device address later."""
        self = cls.__new__(cls)
        self._init(list(problems), dtype, tile=tile, ctx=ctx, lowered=lowered, defer=defer)
        return self

    def _init(self, problems, dtype, *, tile, ctx, lowered, defer=False):
        self._defer = defer
        self.dt = np.dtype(dtype)
        if self.dt not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise UnsupportedModelError(f"dtype {self.dt} is not supported")
        self.models = [m for m, _ in problems]
        self.ctx = ctx if ctx is not None else (None if defer else nat.Context())
        t0 = time.perf_counter()
        low = lowered if lowered is not None else [lower_problem(m, p, self.dt) for m, p in problems]
        self.problems = [tabs for tabs, _ in low]
        self._rg = [rg for _, rg in low]
        self.lower_seconds = time.perf_counter() - t0
        self.n_steps = max(m.n_periods for m in self.models)
        self.tile = tile or self.tuned_tile()
        self._graph = None
        t0 = time.perf_counter()
        self._pack()
        self.pack_seconds = time.perf_counter() - t0

    @property
    def model(self):
        return self.models[0]

    def launch_shape(self) -> tuple[int, int]:
        """(work units, actions per state) of a search step: 256-state tasks of
        the regimes every problem solves in its next-to-last period, and the
        largest action count."""
        units, n_actions = 0, 1
        for m, tabs in zip(self.models, self.problems):
            for (t, r) in tabs:
                if t == max(m.n_periods - 2, 0):
                    lay = m.layouts[r]
                    units += lay.n_sd * -(-lay.n_w // 256)
                    n_actions = max(n_actions, lay.n_d * lay.n_c)
        return max(units, 1), n_actions

    def tuned_tile(self) -> int:
        """The autotuned state tile of this program's launches."""
        units, n_actions = self.launch_shape()
        return autotune.choose("f64" if self.dt == np.float64 else "f32", units, n_actions)

    def step_of(self, p: int, t: int) -> int:
        return self.models[p].n_periods - 1 - t

    # -- packing ---------------------------------------------------------------
    def _pack(self):
        dt = self.dt
        isz = dt.itemsize
        S = self.n_steps
        ar = _Arena(digest=len(self.problems) > 1)
        self.outs: list[dict] = []
        # outputs are reserved first so that descriptors can point at them
        for model, tabs in zip(self.models, self.problems):
            o = {}
            for (t, r), pt in tabs.items():
                lay = model.layouts[r]
                n = lay.n_sd * lay.n_w
                taste = pt.taste_scale is not None
                pol_n = n * (lay.n_d if taste else 1)
                o[(t, r)] = _Out(ar.reserve(n * isz), ar.reserve(pol_n * 4), lay.V_shape,
                                 lay.V_shape + ((lay.n_d,) if taste else ()))
            self.outs.append(o)
        self.out_end = ar.size
        self.out_start = _align(self.out_end)  # every table is placed from here on
        zero_off = ar.add(np.zeros(1, dtype=dt))
        # descriptors: few distinct templates (tables, sizes) and per
        # descriptor only the output and next-period V offsets
        self._templates = []
        tpl_of, v_out, p_out, vn, vs = [], [], [], [], []
        tasks = [[] for _ in range(S)]  # step -> descriptor indices, in descriptor order
        wdep = [0] * S  # kernel variant per step: the most general any descriptor needs
        gmax = [1] * S
        kmax = [0] * S
        self.flops_by_step = [0.0] * S
        self.bytes_by_step = [0.0] * S
        dcache = {}
        task_grids = {}
        self.desc_of = {}     # (problem, period, regime) -> descriptor index
        self.desc_tasks = {}  # descriptor index -> its (sd, w0) task rows
        self._regime_tasks = {}
        for p, (model, tabs, rg) in enumerate(zip(self.models, self.problems, self._rg)):
            order = sorted(tabs.items(), key=lambda kv: (-kv[0][0], model.regime_code(kv[0][1])))
            for (t, r), pt in order:
                s = self.step_of(p, t)
                lay = model.layouts[r]
                ti, patch = self._describe_cached(model, p, t, r, pt, rg, ar, zero_off, dcache)
                d = self._templates[ti]
                di = len(tpl_of)
                tpl_of.append(ti)
                v_out.append(patch[0])
                p_out.append(patch[1])
                vn.append(patch[2])
                vs.append(patch[3])
                self.desc_of[(p, t, r)] = di
                if d["variant"] > wdep[s]:
                    wdep[s] = d["variant"]
                if d["n_constr"] > kmax[s]:
                    kmax[s] = d["n_constr"]
                if d["g_n"] > gmax[s]:
                    gmax[s] = d["g_n"]
                acct = d.get("_acct")
                if acct is None:
                    n_states = lay.n_sd * lay.n_w
                    n_q = n_states * lay.n_d * lay.n_c
                    ev_flops = 2.0 * lay.n_sd * lay.n_d * sum(d["nj"]) * max(d["g_n"], 0)
                    # algorithmic HBM bytes: next V rows read, V and policy written
                    n_next = sum(nj * d["g_n"] for nj in d["nj"])
                    acct = d["_acct"] = (FLOPS_PER_Q * n_q + ev_flops,
                                         isz * (n_next + n_states) + 4 * n_states)
                self.flops_by_step[s] += acct[0]
                self.bytes_by_step[s] += acct[1]
                # tasks (descriptor, discrete state, first wealth node), sd-major
                key = (lay.n_sd, lay.n_w)
                grid = task_grids.get(key)
                if grid is None:
                    sd_i, w_i = np.meshgrid(np.arange(lay.n_sd, dtype=np.int32),
                                            np.arange(0, lay.n_w, self.tile, dtype=np.int32), indexing="ij")
                    grid = task_grids[key] = np.stack([sd_i.ravel(), w_i.ravel()], axis=1)
                tasks[s].append(di)
                self.desc_tasks[di] = grid
        self.n_desc = len(tpl_of)
        self._tpl_of = np.asarray(tpl_of, dtype=np.int64)
        self._patch = (np.asarray(v_out, dtype=np.uint64), np.asarray(p_out, dtype=np.uint64),
                       np.asarray(vn, dtype=np.int64).reshape(self.n_desc, nat.MAX_TARGETS),
                       np.asarray(vs, dtype=np.int64).reshape(self.n_desc, nat.MAX_SCALAR))
        self.desc_off = ar.reserve(self.n_desc * DESC_DTYPE.itemsize)
        self.task_off, self.n_tasks = [], []
        self.tasks_host = []
        for s in range(S):
            # descriptor order: each descriptor's tasks stay contiguous (they
            # share its tables in L2) and problems of different sizes stay
            # interleaved, so the contiguous task ranges the kernel hands to
            # the eight XCDs carry similar work
            dis = tasks[s]
            if dis:
                grids = [self.desc_tasks[di] for di in dis]
                tl_a = np.empty((sum(len(g) for g in grids), 3), dtype=np.int32)
                tl_a[:, 0] = np.repeat(np.asarray(dis, dtype=np.int32), [len(g) for g in grids])
                tl_a[:, 1:] = np.concatenate(grids)
            else:
                tl_a = np.empty((0, 3), dtype=np.int32)
            n = len(tl_a)
            self.tasks_host.append(tl_a)
            if not n:
                self.task_off.append(0)
                self.n_tasks.append(0)
                continue
            npad = (n + 7) // 8 * 8
            arr = np.zeros(npad, dtype=TASK_DTYPE)
            arr["desc"][:n] = tl_a[:, 0]
            arr["sd"][:n] = tl_a[:, 1]
            arr["w0"][:n] = tl_a[:, 2]
            arr["desc"][n:] = -1
            self.task_off.append(ar.add(arr, dedupe=False))
            self.n_tasks.append(npad)
        self.wdep_by_step = wdep
        self._lds_g, self._lds_k = gmax, kmax
        self._lds_by_step = None
        self.total_bytes = _align(ar.size)
        self._arena = ar
        self.out_host = np.zeros(self.out_end, dtype=np.uint8)
        self.buf = None
        if not self._defer:
            self.buf = nat.DeviceBuffer(self.total_bytes)
            self.materialize(self.buf.ptr)

    @property
    def lds_by_step(self):
        """Dynamic LDS bytes of each step's launch (asked of the library on
        first use, so packing itself loads no native code)."""
        if self._lds_by_step is None:
            self._lds_by_step = step_lds_bytes(self.dt, self._lds_g, self._lds_k)
        return self._lds_by_step

    def materialize(self, ptr: int, upload_ctx=None):
        """Write the descriptors for device address `ptr` and upload the image
        there (a synchronous copy on `upload_ctx`, default the program's stream)."""
        self.base = int(ptr)
        self.D = self._desc_records()
        ar = self._arena
        ar.parts.append((self.desc_off, self.D.tobytes()))
        # the V and policy outputs come first and are written by the kernels
        # before any read, so only the tables behind them are uploaded
        image = ar.image(self.out_start)
        self.host_image_bytes = image.nbytes
        nat.h2d(upload_ctx or self.ctx, self.base + self.out_start, image)
        ar.parts.pop()

    def _describe_cached(self, model, p, t, r, pt, rg, ar, zero_off, cache):
        """`_describe`, reusing the work of an earlier period whose tables are
        the same objects (the lowering shares tables that do not depend on
        the age): only the output and next-period V offsets differ."""
        # periods that share their tables share the tables object itself: a
        # first lookup by its identity skips building the content key
        idk = (id(model), p, r, id(pt))
        hit = cache.get(idk)
        if hit is None:
            key = (id(model), p, r, id(pt.U), id(pt.F), id(pt.A), id(pt.B), pt.terminal, pt.beta,
                   pt.taste_scale, pt.next_grid, tuple((n, id(M)) for n, M in pt.targets),
                   tuple((n, id(m)) for n, m in pt.scalars))
            hit = cache.get(key)
            if hit is None:
                d = self._describe(model, p, t, r, pt, rg, ar, zero_off)
                self._templates.append(d)
                hit = cache[key] = (len(self._templates) - 1, pt)  # pt keeps the keyed tables alive
            cache[idk] = (hit[0], pt)  # this pt keeps its own identity key valid
        out = self.outs[p][(t, r)]
        vn = [-1] * nat.MAX_TARGETS
        vs = [-1] * nat.MAX_SCALAR
        if not pt.terminal:
            for k, (name, _) in enumerate(pt.targets):
                vn[k] = self.outs[p][(t + 1, name)].v_off
            for k, (name, _) in enumerate(pt.scalars):
                vs[k] = self.outs[p][(t + 1, name)].v_off if name is not None else zero_off
        return hit[0], (out.v_off, out.p_off, vn, vs)

    def _describe(self, model, p, t, r, pt, rg, ar, zero_off):
        dt = self.dt
        lay = model.layouts[r]
        d = {}
        out = self.outs[p][(t, r)]
        d["V_out"] = out.v_off
        d["pol_out"] = out.p_off
        d["w_pts"] = ar.add(rg.w_pts[r], dtype=dt) if lay.w_grid is not None else zero_off
        # shape checks of everything the kernel indexes (see LcmGsDesc)
        for tab in [pt.U] + [x for Fa, Fb, _ in pt.F for x in (Fa, Fb)]:
            assert tab.ndim == 4 and tab.shape[2:] == (lay.n_d, lay.n_c), tab.shape
            assert tab.shape[0] in (1, lay.n_sd) and tab.shape[1] in (1, lay.n_w), tab.shape
        d["U"] = ar.add(pt.U, dtype=dt)
        d["U_ssd"], d["U_sw"] = _strides(pt.U)
        wdep = d["U_sw"] != 0
        if len(pt.F) > nat.MAX_CONSTR:
            raise UnsupportedModelError(f"at most {nat.MAX_CONSTR} affine constraints per regime")
        fa, fb, fa_s, fb_s, strict = [], [], [], [], 0
        for k, (Fa, Fb, st) in enumerate(pt.F):
            fa.append(ar.add(Fa, dtype=dt))
            fb.append(ar.add(Fb, dtype=dt))
            fa_s.append(_strides(Fa))
            fb_s.append(_strides(Fb))
            wdep |= fa_s[-1][1] != 0 or fb_s[-1][1] != 0
            strict |= int(st) << k
        d["Fa"], d["Fb"], d["Fa_s"], d["Fb_s"] = fa, fb, fa_s, fb_s
        d["n_constr"], d["strict_mask"] = len(pt.F), strict
        d["terminal"] = int(pt.terminal)
        d["n_sd"], d["n_w"], d["n_d"], d["n_c"] = lay.n_sd, lay.n_w, lay.n_d, lay.n_c
        d["has_taste"] = int(pt.taste_scale is not None)
        d["taste_scale"] = pt.taste_scale or 0.0
        if pt.taste_scale is not None and lay.n_d > nat.MAX_TASTE_D:
            raise UnsupportedModelError(f"taste shocks over more than {nat.MAX_TASTE_D} discrete actions")
        d["beta"] = pt.beta
        d["Vn"], d["M"], d["nj"], d["Vs"], d["ms"] = [], [], [], [], []
        d["g_kind"], d["g_n"], d["g_start"], d["g_stop"], d["g_pts"] = 0, 0, 0.0, 0.0, zero_off
        d["A"] = d["B"] = d["Af"] = d["Bf"] = zero_off
        d["A_ssd"] = d["A_sw"] = d["B_ssd"] = d["B_sw"] = 0
        if not pt.terminal:
            if len(pt.targets) > nat.MAX_TARGETS or len(pt.scalars) > nat.MAX_SCALAR:
                raise UnsupportedModelError("too many target regimes for the kernel descriptor")
            for name, M in pt.targets:
                tl = model.layouts[name]
                assert M.shape == (lay.n_sd, lay.n_d, tl.n_sd), (M.shape, tl.n_sd)
                assert tl.n_w == len(rg.w_pts[pt.next_grid])
                d["Vn"].append(self.outs[p][(t + 1, name)].v_off)
                d["M"].append(ar.add(M, dtype=dt))
                d["nj"].append(M.shape[-1])
            for name, m in pt.scalars:
                assert m.shape == (lay.n_sd, lay.n_d)
                assert name is None or model.layouts[name].V_shape == ()
                d["Vs"].append(self.outs[p][(t + 1, name)].v_off if name is not None else zero_off)
                d["ms"].append(ar.add(m, dtype=dt))
            if pt.targets:
                for tab in (pt.A, pt.B):
                    assert tab.ndim == 4 and tab.shape[2:] == (lay.n_d, lay.n_c), tab.shape
                    assert tab.shape[0] in (1, lay.n_sd) and tab.shape[1] in (1, lay.n_w)
                glay = model.layouts[pt.next_grid]
                gpts = rg.w_pts[pt.next_grid]
                d["g_kind"] = int(glay.w_grid.kind)
                d["g_n"] = len(gpts)
                d["g_start"], d["g_stop"] = float(glay.w_grid.start), float(glay.w_grid.stop)
                d["g_pts"] = ar.add(gpts, dtype=dt)
                d["A"] = ar.add(pt.A, dtype=dt)
                d["B"] = ar.add(pt.B, dtype=dt)
                if d["g_kind"] == 0:
                    Af, Bf = folded_next_state(pt.A, pt.B, d["g_start"], d["g_stop"], d["g_n"], dt)
                    d["Af"] = ar.add(Af, dtype=dt)
                    d["Bf"] = ar.add(Bf, dtype=dt)
                d["A_ssd"], d["A_sw"] = _strides(pt.A)
                d["B_ssd"], d["B_sw"] = _strides(pt.B)
                wdep |= d["A_sw"] != 0 or d["B_sw"] != 0
        d["wdep"] = bool(wdep)
        iv = interval_form(pt.F, lay.n_sd, lay.n_d, lay.n_c, dt)
        d["use_interval"] = int(iv is not None)
        d["Flo"] = d["Fhi"] = d["Fchunk"] = zero_off
        d["Fi_ssd"] = d["Fc_ssd"] = 0
        if iv is not None:
            lo, hi = iv
            d["Flo"] = ar.add(lo, dtype=dt)
            d["Fhi"] = ar.add(hi, dtype=dt)
            d["Fi_ssd"] = lay.n_d * lay.n_c if lo.shape[0] > 1 else 0
            fc = chunk_bounds(lo, hi, nat.ACT_CHUNK)
            d["Fchunk"] = ar.add(fc, dtype=dt)
            d["Fc_ssd"] = fc.shape[1] * fc.shape[2] * 2 if fc.shape[0] > 1 else 0
        lin = (not d["g_n"]) or d["g_kind"] == 0
        d["variant"] = 2 if wdep else (0 if (iv is not None and lin) else 1)
        return d

    def relocatable_records(self):
        """(records for base address 0, {field: mask of its pointer entries}):
        `records_at(base)` turns them into the descriptors for any base."""
        return self._desc_records(base=0, with_masks=True)

    @staticmethod
    def records_at(D0, masks, base: int):
        D = D0.copy()
        for f, m in masks.items():
            D[f][m] += np.uint64(base)
        return D

    _PTR_FIELDS = ("V_out", "pol_out", "w_pts", "U", "A", "B", "g_pts", "Flo", "Fhi", "Fchunk", "Af", "Bf")
    _VAL_FIELDS = ("U_ssd", "U_sw", "A_ssd", "A_sw", "B_ssd", "B_sw", "n_sd", "n_w", "n_d", "n_c",
                   "n_constr", "strict_mask", "terminal", "g_kind", "g_n", "has_taste",
                   "beta", "taste_scale", "g_start", "g_stop", "Fi_ssd", "Fc_ssd", "use_interval")
    # (record field, descriptor key, slots, pointer?, element of a (ssd, sw) pair)
    _LIST_FIELDS = (("Fa", "Fa", nat.MAX_CONSTR, True, None), ("Fb", "Fb", nat.MAX_CONSTR, True, None),
                    ("Fa_ssd", "Fa_s", nat.MAX_CONSTR, False, 0), ("Fa_sw", "Fa_s", nat.MAX_CONSTR, False, 1),
                    ("Fb_ssd", "Fb_s", nat.MAX_CONSTR, False, 0), ("Fb_sw", "Fb_s", nat.MAX_CONSTR, False, 1),
                    ("Vn", "Vn", nat.MAX_TARGETS, True, None), ("M", "M", nat.MAX_TARGETS, True, None),
                    ("n_sd_next", "nj", nat.MAX_TARGETS, False, None),
                    ("Vs", "Vs", nat.MAX_SCALAR, True, None), ("ms", "ms", nat.MAX_SCALAR, True, None))

    def _desc_records(self, base=None, with_masks=False):
        """Descriptor records: the templates' records (filled column by
        column), gathered per descriptor, with each descriptor's output and
        next-period V offsets patched in; pointers are arena offsets plus
        `base` (default: the program's). With `with_masks`, also returns
        {field: mask of its pointer entries}."""
        base = self.base if base is None else base
        T, tmasks = self._template_records(base)
        D = T[self._tpl_of]
        v_out, p_out, vn, vs = self._patch
        D["V_out"] = v_out + np.uint64(base)
        D["pol_out"] = p_out + np.uint64(base)
        for field, arr in (("Vn", vn), ("Vs", vs)):
            present = arr >= 0
            if present.any():
                D[field] = np.where(present, arr.astype(np.uint64) + np.uint64(base), D[field]).astype(np.uint64)
        if not with_masks:
            return D
        return D, {f: m[self._tpl_of] for f, m in tmasks.items()}

    def _template_records(self, base):
        """Records of the descriptor templates and {field: pointer mask}."""
        descs = self._templates
        n = len(descs)
        D = np.zeros(n, dtype=DESC_DTYPE)
        masks = {}
        for f in self._PTR_FIELDS:
            D[f] = np.uint64(base) + np.fromiter((d[f] for d in descs), dtype=np.uint64, count=n)
            masks[f] = np.ones(n, dtype=bool)
        for f in self._VAL_FIELDS:
            D[f] = np.array([d[f] for d in descs], dtype=D.dtype[f])
        D["n_targets"] = np.array([len(d["Vn"]) for d in descs], dtype=np.int32)
        D["n_scalar"] = np.array([len(d["Vs"]) for d in descs], dtype=np.int32)
        for field, key, k, ptr, part in self._LIST_FIELDS:
            lens = np.fromiter((len(d[key]) for d in descs), dtype=np.int64, count=n)
            if not lens.any():
                continue
            pad = [0] * k
            if part is None:
                rows = [(list(d[key]) + pad)[:k] for d in descs]
            else:
                rows = [([v[part] for v in d[key]] + pad)[:k] for d in descs]
            vals = np.array(rows, dtype=D.dtype[field].base)
            present = np.arange(k)[None, :] < lens[:, None]
            if ptr:
                vals = np.where(present, vals + np.uint64(base), vals).astype(np.uint64)
                masks[field] = present
            D[field] = vals
        return D, masks

    # -- execution -------------------------------------------------------------
    def _launch_step(self, s: int):
        L = nat.lib()
        code = nat.LCM_DTYPE_F64 if self.dt == np.float64 else nat.LCM_DTYPE_F32
        st = L.lcm_gs_launch_ex(
            self.ctx.handle, code, C.c_void_p(self.base + self.desc_off),
            C.c_void_p(self.base + self.task_off[s]), self.n_tasks[s], self.tile,
            C.c_size_t(self.lds_by_step[s]), int(self.wdep_by_step[s]) | self._occ_flag(),
        )
        nat._check(st, f"lcm_gs_launch (step {s})")

    # the float32 kernel's register budget: "auto" lets the library pick by
    # the launch's task count; "dense" / "sparse" force one build (the same
    # bits either way)
    occupancy = "auto"

    def _occ_flag(self) -> int:
        return {"auto": 0, "dense": nat.GS_OCC_DENSE, "sparse": nat.GS_OCC_SPARSE}[self.occupancy]

    def launch_regime(self, p: int, t: int, r: str):
        """One (problem, period, regime) launch: that descriptor's tasks alone,
        with its step's tile, LDS size and kernel variant, so its arithmetic
        is that of the fused launch of the whole step."""
        di = self.desc_of[(p, t, r)]
        s = self.step_of(p, t)
        if di not in self._regime_tasks:
            grid = self.desc_tasks[di]
            n = len(grid)
            npad = (n + 7) // 8 * 8
            arr = np.zeros(npad, dtype=TASK_DTYPE)
            arr["desc"][:n] = di
            arr["sd"][:n] = grid[:, 0]
            arr["w0"][:n] = grid[:, 1]
            arr["desc"][n:] = -1
            buf = nat.DeviceBuffer(arr.nbytes)
            nat.h2d(self.ctx, buf.ptr, arr)
            self._regime_tasks[di] = (buf, npad)
        buf, npad = self._regime_tasks[di]
        code = nat.LCM_DTYPE_F64 if self.dt == np.float64 else nat.LCM_DTYPE_F32
        st = nat.lib().lcm_gs_launch_ex(
            self.ctx.handle, code, C.c_void_p(self.base + self.desc_off), C.c_void_p(buf.ptr), npad,
            self.tile, C.c_size_t(self.lds_by_step[s]), int(self.wdep_by_step[s]) | self._occ_flag())
        nat._check(st, f"lcm_gs_launch ({r!r}, period {t})")

    def _record(self):
        for s in range(self.n_steps):
            if self.n_tasks[s]:
                self._launch_step(s)

    def capture_graph(self):
        if self._graph is None:
            self._graph = nat.Graph.capture(self.ctx, self._record)
        return self._graph

    def run(self, *, use_graph: bool = True, sync: bool = True):
        if use_graph:
            self.capture_graph().launch(self.ctx)
        else:
            self._record()
        if sync:
            self.ctx.sync()

    def time_launches(self):
        """Per-step kernel durations in ms, from HIP events on the program's stream."""
        evs = [nat.Event() for _ in range(self.n_steps + 1)]
        ms = []
        steps = [s for s in range(self.n_steps) if self.n_tasks[s]]
        evs[0].record(self.ctx)
        for i, s in enumerate(steps):
            self._launch_step(s)
            evs[i + 1].record(self.ctx)
        self.ctx.sync()
        for i, s in enumerate(steps):
            ms.append((s, evs[i].elapsed_ms(evs[i + 1])))
        return ms

    def fetch(self):
        """Copy all V and policy outputs to the host."""
        nat.d2h(self.ctx, self.base, self.out_host)
        return self.out_host

    def results(self, problem: int = 0):
        """{period: {regime: (V, policy)}} for one problem (after fetch)."""
        out = {}
        h = self.out_host
        for (t, r), o in self.outs[problem].items():
            n = int(np.prod(o.shape, dtype=np.int64)) if o.shape else 1
            pn = int(np.prod(o.pol_shape, dtype=np.int64)) if o.pol_shape else 1
            V = np.frombuffer(h, dtype=self.dt, count=n, offset=o.v_off).reshape(o.shape).copy()
            P = np.frombuffer(h, dtype=np.int32, count=pn, offset=o.p_off).reshape(o.pol_shape).copy()
            out.setdefault(t, {})[r] = (V, P)
        return out
