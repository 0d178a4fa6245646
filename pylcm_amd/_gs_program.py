"""Device plan of grid-search solves: arena, descriptors, tasks, launches.

A `GsProgram` holds any number of problems (one params set each, possibly of
different models' grid sizes) for one model structure. It lowers every
(problem, period, regime) to `PeriodTables`, packs all tables into one device
arena (identical tables are stored once), writes one `LcmGsDesc` per
(problem, period, regime) and one task list per period, and then runs backward
induction as one fused launch per period, optionally captured once as a HIP
graph over the horizon. This replaces the reference's host loop
(`src/_lcm/solution/backward_induction.py:155-275`), which dispatches one
compiled core per (regime, period) and synchronises after each period.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from pylcm_amd import _native as nat
from pylcm_amd._build import PeriodLowering, PeriodTables, resolve_grids, table_digest
from pylcm_amd.exceptions import UnsupportedModelError

ALIGN = 256
DESC_DTYPE = np.dtype(nat.LcmGsDesc)
TASK_DTYPE = np.dtype(nat.LcmGsTask)


def _align(x: int) -> int:
    return (x + ALIGN - 1) // ALIGN * ALIGN


class _Arena:
    """Host image of a device arena with content-deduplicated tables."""

    def __init__(self):
        self.parts: list[tuple[int, bytes]] = []
        self.size = 0
        self.index: dict[str, int] = {}

    def add(self, arr: np.ndarray, dedupe: bool = True) -> int:
        a = np.ascontiguousarray(arr)
        key = table_digest(a) + str(a.dtype) if dedupe else None
        if key is not None and key in self.index:
            return self.index[key]
        off = _align(self.size)
        self.parts.append((off, a.tobytes()))
        self.size = off + a.nbytes
        if key is not None:
            self.index[key] = off
        return off

    def reserve(self, nbytes: int) -> int:
        off = _align(self.size)
        self.size = off + max(int(nbytes), 1)
        return off

    def image(self, upto: int | None = None) -> np.ndarray:
        n = self.size if upto is None else upto
        buf = np.zeros(_align(max(n, 1)), dtype=np.uint8)
        for off, b in self.parts:
            if off < n:
                buf[off: off + len(b)] = np.frombuffer(b, dtype=np.uint8)
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


def default_tile(n_states_total: int) -> int:
    """States per workgroup: small workloads split actions over more waves."""
    if n_states_total <= 64 * 1024:
        return 64
    if n_states_total <= 512 * 1024:
        return 128
    return 256


class GsProgram:
    def __init__(self, model, params_list, dtype, *, tile: int | None = None, ctx=None,
                 tables=None):
        self.model = model
        self.dt = np.dtype(dtype)
        if self.dt not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise UnsupportedModelError(f"dtype {self.dt} is not supported")
        self.T = model.n_periods
        self.ctx = ctx if ctx is not None else nat.Context()
        t0 = time.perf_counter()
        lowered = tables if tables is not None else [
            lower_problem(model, p, self.dt) for p in params_list
        ]
        self.problems = [tabs for tabs, _ in lowered]
        self._rg = [rg for _, rg in lowered]
        self.lower_seconds = time.perf_counter() - t0
        n_states = sum(
            model.layouts[r].n_sd * model.layouts[r].n_w
            for tabs in self.problems for (t, r) in tabs if t == 0
        ) or 1
        self.tile = tile or default_tile(n_states)
        self._graph = None
        self._pack()

    # -- packing ---------------------------------------------------------------
    def _pack(self):
        model, dt, T = self.model, self.dt, self.T
        isz = dt.itemsize
        ar = _Arena()
        zero_off = ar.add(np.zeros(1, dtype=dt))
        self.outs: list[dict] = []
        # outputs are reserved first so that descriptors can point at them
        for tabs in self.problems:
            o = {}
            for (t, r), pt in tabs.items():
                lay = model.layouts[r]
                n = lay.n_sd * lay.n_w
                pol_n = n * (lay.n_d if pt.taste_scale is not None else 1)
                o[(t, r)] = _Out(ar.reserve(n * isz), ar.reserve(pol_n * 4), lay.V_shape,
                                 lay.V_shape + ((lay.n_d,) if pt.taste_scale is not None else ()))
            self.outs.append(o)
        self.out_end = ar.size
        descs = []
        desc_period = []
        tasks_by_t = [[] for _ in range(T)]
        wdep_by_t = [False] * T
        gmax_by_t = [1] * T
        for p, (tabs, rg) in enumerate(zip(self.problems, self._rg)):
            for (t, r), pt in sorted(tabs.items(), key=lambda kv: (kv[0][0], model.regime_code(kv[0][1]))):
                lay = model.layouts[r]
                d = {}
                out = self.outs[p][(t, r)]
                d["V_out"] = out.v_off
                d["pol_out"] = out.p_off
                d["w_pts"] = ar.add(rg.w_pts[r].astype(dt)) if lay.w_grid is not None else zero_off
                d["U"] = ar.add(pt.U.astype(dt))
                d["U_ssd"], d["U_sw"] = _strides(pt.U)
                wdep = d["U_sw"] != 0
                if len(pt.F) > nat.MAX_CONSTR:
                    raise UnsupportedModelError(f"at most {nat.MAX_CONSTR} affine constraints per regime")
                fa, fb, fa_s, fb_s, strict = [], [], [], [], 0
                for k, (Fa, Fb, st) in enumerate(pt.F):
                    fa.append(ar.add(Fa.astype(dt)))
                    fb.append(ar.add(Fb.astype(dt)))
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
                d["A"] = d["B"] = zero_off
                d["A_ssd"] = d["A_sw"] = d["B_ssd"] = d["B_sw"] = 0
                # shape checks of everything the kernel indexes (see LcmGsDesc)
                for tab in [pt.U] + [x for Fa, Fb, _ in pt.F for x in (Fa, Fb)]:
                    assert tab.ndim == 4 and tab.shape[2:] == (lay.n_d, lay.n_c), tab.shape
                    assert tab.shape[0] in (1, lay.n_sd) and tab.shape[1] in (1, lay.n_w), tab.shape
                if not pt.terminal:
                    if len(pt.targets) > nat.MAX_TARGETS or len(pt.scalars) > nat.MAX_SCALAR:
                        raise UnsupportedModelError("too many target regimes for the kernel descriptor")
                    for name, M in pt.targets:
                        tl = model.layouts[name]
                        assert M.shape == (lay.n_sd, lay.n_d, tl.n_sd), (M.shape, tl.n_sd)
                        assert tl.n_w == len(rg.w_pts[pt.next_grid])
                    for name, m in pt.scalars:
                        assert m.shape == (lay.n_sd, lay.n_d)
                        assert name is None or model.layouts[name].V_shape == ()
                    if pt.targets:
                        for tab in (pt.A, pt.B):
                            assert tab.ndim == 4 and tab.shape[2:] == (lay.n_d, lay.n_c), tab.shape
                            assert tab.shape[0] in (1, lay.n_sd) and tab.shape[1] in (1, lay.n_w)
                    for name, M in pt.targets:
                        d["Vn"].append(self.outs[p][(t + 1, name)].v_off)
                        d["M"].append(ar.add(M.astype(dt)))
                        d["nj"].append(M.shape[-1])
                    for name, m in pt.scalars:
                        d["Vs"].append(self.outs[p][(t + 1, name)].v_off if name is not None else zero_off)
                        d["ms"].append(ar.add(m.astype(dt)))
                    if pt.targets:
                        glay = model.layouts[pt.next_grid]
                        gpts = rg.w_pts[pt.next_grid]
                        d["g_kind"] = int(glay.w_grid.kind)
                        d["g_n"] = len(gpts)
                        d["g_start"], d["g_stop"] = float(glay.w_grid.start), float(glay.w_grid.stop)
                        d["g_pts"] = ar.add(gpts.astype(dt))
                        d["A"] = ar.add(pt.A.astype(dt))
                        d["B"] = ar.add(pt.B.astype(dt))
                        d["A_ssd"], d["A_sw"] = _strides(pt.A)
                        d["B_ssd"], d["B_sw"] = _strides(pt.B)
                        wdep |= d["A_sw"] != 0 or d["B_sw"] != 0
                        gmax_by_t[t] = max(gmax_by_t[t], d["g_n"])
                di = len(descs)
                descs.append(d)
                desc_period.append(t)
                wdep_by_t[t] |= wdep
                for sd in range(lay.n_sd):
                    for w0 in range(0, lay.n_w, self.tile):
                        tasks_by_t[t].append((di, sd, w0))
        self.n_desc = len(descs)
        self.desc_off = ar.reserve(self.n_desc * DESC_DTYPE.itemsize)
        self.task_off, self.n_tasks = [], []
        task_arrays = []
        for t in range(T):
            tl = tasks_by_t[t]
            if not tl:
                self.task_off.append(0)
                self.n_tasks.append(0)
                task_arrays.append(None)
                continue
            n = len(tl)
            npad = (n + 7) // 8 * 8
            arr = np.zeros(npad, dtype=TASK_DTYPE)
            tl_a = np.asarray(tl, dtype=np.int32)
            arr["desc"][:n] = tl_a[:, 0]
            arr["sd"][:n] = tl_a[:, 1]
            arr["w0"][:n] = tl_a[:, 2]
            arr["desc"][n:] = -1
            self.task_off.append(ar.add(arr, dedupe=False))
            self.n_tasks.append(npad)
            task_arrays.append(arr)
        self.wdep_by_t = wdep_by_t
        self.lds_by_t = [int(2 * g * self.dt.itemsize + 256 * (self.dt.itemsize + 8) + 64) for g in gmax_by_t]
        self.total_bytes = _align(ar.size)
        self.buf = nat.DeviceBuffer(self.total_bytes)
        base = self.buf.ptr
        self.base = base
        D = np.zeros(self.n_desc, dtype=DESC_DTYPE)

        def ptr(off):
            return base + off

        for i, d in enumerate(descs):
            rec = D[i]
            rec["V_out"] = ptr(d["V_out"])
            rec["pol_out"] = ptr(d["pol_out"])
            rec["w_pts"] = ptr(d["w_pts"])
            rec["U"] = ptr(d["U"])
            rec["A"] = ptr(d["A"])
            rec["B"] = ptr(d["B"])
            for k in range(nat.MAX_CONSTR):
                if k < d["n_constr"]:
                    rec["Fa"][k] = ptr(d["Fa"][k])
                    rec["Fb"][k] = ptr(d["Fb"][k])
                    rec["Fa_ssd"][k], rec["Fa_sw"][k] = d["Fa_s"][k]
                    rec["Fb_ssd"][k], rec["Fb_sw"][k] = d["Fb_s"][k]
            for k, (v, m, nj) in enumerate(zip(d["Vn"], d["M"], d["nj"])):
                rec["Vn"][k] = ptr(v)
                rec["M"][k] = ptr(m)
                rec["n_sd_next"][k] = nj
            for k, (v, m) in enumerate(zip(d["Vs"], d["ms"])):
                rec["Vs"][k] = ptr(v)
                rec["ms"][k] = ptr(m)
            rec["g_pts"] = ptr(d["g_pts"])
            for f in ("U_ssd", "U_sw", "A_ssd", "A_sw", "B_ssd", "B_sw", "n_sd", "n_w", "n_d", "n_c",
                      "n_constr", "strict_mask", "terminal", "g_kind", "g_n", "has_taste",
                      "beta", "taste_scale", "g_start", "g_stop"):
                rec[f] = d[f]
            rec["n_targets"] = len(d["Vn"])
            rec["n_scalar"] = len(d["Vs"])
        ar.parts.append((self.desc_off, D.tobytes()))
        image = ar.image()
        self.host_image_bytes = image.nbytes
        # upload everything (outputs are overwritten by the kernels)
        nat.h2d(self.ctx, base, image)
        self.out_host = np.zeros(self.out_end, dtype=np.uint8)

    # -- execution -------------------------------------------------------------
    def _record(self):
        L = nat.lib()
        code = nat.LCM_DTYPE_F64 if self.dt == np.float64 else nat.LCM_DTYPE_F32
        for t in reversed(range(self.T)):
            if not self.n_tasks[t]:
                continue
            st = L.lcm_gs_launch_ex(
                self.ctx.handle, code, C.c_void_p(self.base + self.desc_off),
                C.c_void_p(self.base + self.task_off[t]), self.n_tasks[t], self.tile,
                C.c_size_t(self.lds_by_t[t]), int(self.wdep_by_t[t]),
            )
            nat._check(st, f"lcm_gs_launch (period {t})")

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


def lower_problem(model, params, dtype):
    """(tables by (period, regime), resolved grids) for one params set."""
    rg = resolve_grids(model.layouts, params, model.fixed_params, dtype)
    low = PeriodLowering(model, params, rg)
    tabs = {}
    for t in range(model.n_periods):
        age = model.ages.period_to_age(t)
        for r in model.regime_names:
            lay = model.layouts[r]
            if not lay.regime.active(age):
                continue
            if lay.is_dcegm:
                raise UnsupportedModelError("DC-EGM regimes are solved by the DC-EGM program")
            tabs[(t, r)] = low.lower(lay, t)
    return tabs, rg
