"""Device plan of a DC-EGM solve: lowering, arena, descriptors, launches.

A regime solved by `DCEGM` (a `ConsumptionSavingsRegime` whose liquid margin
has resources equal to its continuous state) is lowered per period into one
`LcmEgmCol` per (discrete state, discrete action) column:

* the utility and the inverse marginal utility as expression tapes with the
  column's constants (`_tape.py`);
* per target regime reachable next period: its probability, the next
  resources R'(A_j) = law(savings=A_j) and dR'/dA at every savings node (host
  tapes), the child's node weights, and the child's carry rows;
* output slots for the refined carry rows, the counts and the column's V row.

Terminal targets with a continuous state carry (grid, utility, d utility /
d state) rows computed on the host (`src/_lcm/egm/terminal.py:85-180`);
stateless terminals enter as a scalar value (`continuation.py:488-497`).
Each period is one `lcm_egm_launch` over all its DC-EGM columns plus one
`lcm_egm_aggregate` (logsum or max over the discrete actions,
`src/_lcm/egm/step.py:731-738`). The reference's equivalent host loop is
`src/_lcm/solution/backward_induction.py:155-275` calling
`_DCEGMPeriodKernel.__call__` (`src/_lcm/solution/dcegm.py:510-535`).

The candidate rows are FP64 on the device whatever the model dtype; V is
published in the model dtype.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from pylcm_amd import _native as nat
from pylcm_amd import _tape
from pylcm_amd._build import PeriodLowering, resolve_grids
from pylcm_amd._dag import Evaluator
from pylcm_amd._gs_program import _Arena
from pylcm_amd.exceptions import InvalidParamsError, UnsupportedModelError

ROW = np.dtype(nat.LcmEgmRow)
COL = np.dtype(nat.LcmEgmCol)
AGG = np.dtype(nat.LcmEgmAgg)
CONSTRAINED_OFFSET_FRACTION = 1e-4


def _abi_check():
    L = nat.load_library()
    sizes = [C.c_size_t() for _ in range(3)]
    nat._check(L.lcm_abi_sizes_egm(*[C.byref(s) for s in sizes]), "lcm_abi_sizes_egm")
    got = tuple(s.value for s in sizes)
    want = (C.sizeof(nat.LcmEgmRow), C.sizeof(nat.LcmEgmCol), C.sizeof(nat.LcmEgmAgg))
    if got != want:
        raise nat.NativeKernelError(f"EGM struct layout mismatch: library {got}, binding {want}")


def _columns(lay, rg):
    """Per-column variables of a regime: arrays over (sd combo, da combo)."""
    sd_idx = (np.indices(lay.sd_sizes).reshape(len(lay.sd_sizes), -1).T
              if lay.sd_sizes else np.zeros((1, 0), dtype=np.int64))
    da_idx = (np.indices(lay.da_sizes).reshape(len(lay.da_sizes), -1).T
              if lay.da_sizes else np.zeros((1, 0), dtype=np.int64))
    n_sd, n_d = sd_idx.shape[0], da_idx.shape[0]
    sd_rep = np.repeat(sd_idx, n_d, axis=0)
    da_rep = np.tile(da_idx, (n_sd, 1))
    vars_ = {}
    vals = rg.sd_values[lay.name]
    from pylcm_amd.grids import DiscreteGrid

    for k, (s, g) in enumerate(zip(lay.sd_names, lay.sd_grids)):
        v = vals[k][sd_rep[:, k]]
        vars_[s] = v.astype(np.int64) if isinstance(g, DiscreteGrid) else v.astype(np.float64)
    for k, a in enumerate(lay.da_names):
        vars_[a] = da_rep[:, k].astype(np.int64)
    return vars_, n_sd, n_d


class EgmProgram:
    """Backward induction of one (model, params) with DC-EGM regimes."""

    def __init__(self, model, params, dtype, ctx=None, device=True, nodes_per_block: int = 0):
        if device:
            _abi_check()
        self.nodes_per_block = int(nodes_per_block)  # 0: chosen per launch by the library
        self._graph = None  # the period loop captured as a HIP graph (`run(use_graph=True)`)
        self.model = model
        self.params = params
        self.dt = np.dtype(dtype)
        if self.dt not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise UnsupportedModelError("dtype must be float32 or float64")
        self.ctx = (ctx or nat.Context()) if device else None
        self.layouts = model.layouts
        self.rg = resolve_grids(self.layouts, params, model.fixed_params, self.dt)
        self.low = PeriodLowering(model, params, self.rg)
        self.ar = _Arena()
        self.T = model.n_periods
        self.names = list(model.regime_names)
        self.host_V = {}        # (t, r) -> host V (terminal regimes)
        self.carry = {}         # (t, r) -> dict describing rows for parents
        self.out_V = {}         # (t, r) -> (offset, shape) of the device V in dtype
        self.col_out = {}       # (t, r) -> offsets of the carry rows of its columns
        self._col_specs = {}    # t -> list of column spec dicts (filled into descriptors later)
        self._agg_specs = {}
        self.max_cand = 0
        self.tapes = {}         # (t, r) -> (utility, inverse, resources law) tapes
        self._lower_all()
        if device:
            self._pack()

    # ------------------------------------------------------------------ lowering
    def _evaluator(self, lay, variables, t):
        v = dict(variables)
        v["age"] = self.model.ages.period_to_age(t)
        v["period"] = t
        funcs = {k: f for k, f in lay.regime.functions.items() if f is not None}
        return Evaluator(regime_name=lay.name, functions=funcs, variables=v, params=self.params,
                         fixed_params=self.model.fixed_params, param_dtype=self.dt)

    def _round(self, x):
        return float(np.asarray(x, dtype=self.dt))

    def _lower_all(self):
        model = self.model
        for t in reversed(range(self.T)):
            age = model.ages.period_to_age(t)
            for r in self.names:
                lay = self.layouts[r]
                if not lay.regime.active(age):
                    continue
                if lay.terminal:
                    self._lower_terminal(lay, t)
                elif lay.is_dcegm:
                    self._lower_dcegm(lay, t)
                else:
                    raise UnsupportedModelError(
                        f"regime {r!r}: grid-search regimes in a model with DC-EGM regimes are not "
                        "implemented on the device")

    def _lower_terminal(self, lay, t):
        if lay.V_shape == ():
            ev = self._evaluator(lay, {}, t)
            v = float(np.asarray(ev.call("utility"), dtype=np.float64))
            self.host_V[(t, lay.name)] = np.asarray(v, dtype=self.dt)
            self.carry[(t, lay.name)] = {"kind": "scalar", "value": float(np.asarray(v, dtype=self.dt))}
            return
        if lay.da_names or lay.c_name:
            raise UnsupportedModelError(f"terminal regime {lay.name!r} with actions under DC-EGM parents")
        vars_, n_sd, _ = _columns(lay, self.rg)
        w = self.rg.w_pts[lay.name]
        ev = self._evaluator(lay, vars_, t)

        def call(sym):
            ev.variables[lay.w_name] = sym
            ev._memo.clear()
            return ev.call("utility")

        tp = _tape.trace(n_sd, call)
        val, der = tp.eval(np.broadcast_to(w, (n_sd, len(w))), derivative=True)
        val = np.ascontiguousarray(val, dtype=np.float64)
        der = np.where(np.isneginf(val), 0.0, der)
        self.host_V[(t, lay.name)] = val.reshape(lay.V_shape).astype(self.dt)
        g_off = self.ar.add(np.ascontiguousarray(w, dtype=np.float64))
        rows = []
        nvalid = int(np.sum(~np.isnan(w)))
        nv_off = self.ar.add(np.asarray([nvalid, 0], dtype=np.int32))
        for sd in range(n_sd):
            v_off = self.ar.add(val[sd])
            m_off = self.ar.add(np.ascontiguousarray(der[sd], dtype=np.float64))
            rows.append((g_off, v_off, m_off, nv_off, len(w)))
        self.carry[(t, lay.name)] = {"kind": "rows", "rows": rows, "N": 1, "n_sd": n_sd, "scale": 0.0}

    def _lower_dcegm(self, lay, t):
        model = self.model
        reg = lay.regime
        liq = reg.liquid
        solver = reg.solver
        r = lay.name
        if liq is None or liq.resources != lay.w_name or liq.state != lay.w_name:
            raise UnsupportedModelError(
                f"regime {r!r}: DC-EGM on the device takes resources equal to the continuous state")
        if liq.action != lay.c_name:
            raise UnsupportedModelError(f"regime {r!r}: the liquid action must be the continuous action")
        if type(solver.envelope).__name__ != "FUESEnvelope":
            raise UnsupportedModelError(f"regime {r!r}: only the FUES envelope is implemented")
        if "inverse_marginal_utility" not in reg.functions:
            raise UnsupportedModelError(
                f"regime {r!r}: supply functions['inverse_marginal_utility']; the numeric inverse is "
                "not implemented on the device")
        if t == self.T - 1:
            raise InvalidParamsError(f"non-terminal regime {r!r} is active in the last period")
        vars_, n_sd, n_d = _columns(lay, self.rg)
        n_cols = n_sd * n_d
        A = solver.savings_grid.to_numpy(self.dt).astype(np.float64)
        W = self.rg.w_pts[r]
        n_s, n_w = len(A), len(W)
        if n_s < 2 or n_w < 1:
            raise UnsupportedModelError(f"regime {r!r}: need at least 2 savings nodes and 1 state node")
        if not np.all(np.diff(A) > 0):
            raise UnsupportedModelError(f"regime {r!r}: the savings grid must be strictly increasing")
        n_con = int(solver.n_constrained_points)
        n_cand = n_s + n_con
        if n_cand > nat.EGM_MAX_CAND:
            raise UnsupportedModelError(
                f"regime {r!r}: {n_cand} candidates per column exceed the device limit {nat.EGM_MAX_CAND}")
        n_pad = int(math.ceil(solver.refined_grid_factor * n_cand))
        ev = self._evaluator(lay, vars_, t)
        beta = float(ev.param("koopmans_aggregator", "discount_factor"))
        scale = 0.0
        if reg.taste_shocks is not None:
            rp = (self.params or {}).get(r, {})
            ts = rp.get("taste_shocks", {}) if isinstance(rp, dict) else {}
            if "scale" not in ts:
                raise InvalidParamsError(f"regime {r!r}: missing params[...]['taste_shocks']['scale']")
            scale = self._round(ts["scale"])
            if not scale > 0:
                raise InvalidParamsError("taste shock scale must be positive")

        def traced(fname, var_name, func=None):
            e = self._evaluator(lay, vars_, t)

            def call(sym):
                e.variables[var_name] = sym
                return e.call(fname, func)

            try:
                return _tape.trace(n_cols, call)
            except UnsupportedModelError as err:
                raise UnsupportedModelError(f"regime {r!r}, {fname}: {err}") from err

        u_tape = traced("utility", lay.c_name)
        i_tape = traced("inverse_marginal_utility", "marginal_continuation")
        law = reg.state_transitions.get(lay.w_name)
        if law is None or hasattr(law, "items"):
            raise UnsupportedModelError(f"regime {r!r}: one law for {lay.w_name!r} is required")
        r_tape = traced(f"next_{lay.w_name}", liq.post_decision_state, law)
        Rn, G = r_tape.eval(np.broadcast_to(A, (n_cols, n_s)), derivative=True)
        Rn = np.ascontiguousarray(Rn, dtype=np.float64)
        G = np.ascontiguousarray(G, dtype=np.float64)

        ar = self.ar
        A_off = ar.add(A)
        W_off = ar.add(np.ascontiguousarray(W, dtype=np.float64))
        up_off = ar.add(u_tape.program())
        ip_off = ar.add(i_tape.program())
        uc = u_tape.const_table()
        ic = i_tape.const_table()
        uc_off = ar.add(uc, dedupe=False)
        ic_off = ar.add(ic, dedupe=False)
        self.tapes[(t, r)] = (u_tape, i_tape, r_tape)
        R_off = ar.add(Rn, dedupe=False)
        G_off = ar.add(G, dedupe=False)

        # targets reachable next period
        age_next = model.ages.period_to_age(t + 1)
        active_next = [x for x in self.names if self.layouts[x].regime.active(age_next)]
        tr = reg.transition
        tr_funcs = list(tr.values()) if hasattr(tr, "items") else [tr]
        p = self.low._cached("P", lay, tr_funcs, t, lambda: self.low.regime_probs(lay, t))
        for x in self.names:
            if x not in active_next and np.any(p[:, :, model.regime_code(x)] != 0):
                raise InvalidParamsError(f"regime {r!r} at period {t} moves to {x!r}, inactive next period")
        if len(active_next) > nat.EGM_MAX_TARGETS:
            raise UnsupportedModelError(f"regime {r!r}: more than {nat.EGM_MAX_TARGETS} target regimes")
        targets = []
        for x in active_next:
            info = self.carry[(t + 1, x)]
            pr = np.asarray(p[:, :, model.regime_code(x)], dtype=np.float64).reshape(n_cols)
            if info["kind"] == "scalar":
                targets.append({"kind": 0, "prob": pr, "scalar": info["value"]})
                continue
            tl = self.layouts[x]
            if tl.w_name != lay.w_name:
                raise UnsupportedModelError(f"target {x!r} does not carry {lay.w_name!r}")
            Wt = self.low.node_weights(lay, tl, t).reshape(n_cols, -1)
            Wt = np.ascontiguousarray(Wt, dtype=np.float64)
            if Wt.shape[1] != info["n_sd"]:
                raise UnsupportedModelError(f"target {x!r}: node count mismatch")
            if info["N"] > nat.EGM_MAX_ROWS:
                raise UnsupportedModelError(f"target {x!r}: more than {nat.EGM_MAX_ROWS} discrete-action rows")
            targets.append({"kind": 1, "prob": pr, "W_off": ar.add(Wt, dedupe=False), "K": Wt.shape[1],
                            "N": info["N"], "scale": info["scale"], "child": (t + 1, x)})
        n_rows = sum(tg["K"] * tg["N"] for tg in targets if tg["kind"] == 1)
        if n_rows > nat.EGM_MAX_ROWS_PER_COL:
            raise UnsupportedModelError(
                f"regime {r!r}: its columns read {n_rows} child rows; the device limit is "
                f"{nat.EGM_MAX_ROWS_PER_COL}")

        # outputs
        slot = 8 * n_pad
        grid_off = ar.reserve(n_cols * slot)
        pol_off = ar.reserve(n_cols * slot)
        val_off = ar.reserve(n_cols * slot)
        mar_off = ar.reserve(n_cols * slot)
        cnt_off = ar.reserve(n_cols * 8)
        vcol_off = ar.reserve(n_cols * n_w * 8)
        vout_off = ar.reserve(n_sd * n_w * self.dt.itemsize)
        self.out_V[(t, r)] = (vout_off, lay.V_shape)
        self.col_out[(t, r)] = dict(grid=grid_off, policy=pol_off, value=val_off, marg=mar_off,
                                    counts=cnt_off, vcol=vcol_off, n_cols=n_cols, n_pad=n_pad, n_w=n_w,
                                    n_sd=n_sd, n_d=n_d)
        rows = [(grid_off + c * slot, val_off + c * slot, mar_off + c * slot, cnt_off + c * 8 + 4, n_pad)
                for c in range(n_cols)]
        self.carry[(t, r)] = {"kind": "rows", "rows": rows, "N": n_d, "n_sd": n_sd, "scale": scale}
        ratio = (1.0 / CONSTRAINED_OFFSET_FRACTION) ** (1.0 / max(n_con - 1, 1))
        n_scan = solver.envelope.n_points_to_scan
        spec = dict(A=A_off, W=W_off, up=up_off, ip=ip_off, uc=uc_off, uc_n=uc.shape[1], ic=ic_off,
                    ic_n=ic.shape[1], R=R_off, G=G_off, n_s=n_s, n_w=n_w, n_con=n_con, n_pad=n_pad,
                    n_cols=n_cols, beta=beta, ratio=ratio, jump=float(solver.envelope.jump_thresh),
                    n_scan=0 if n_scan is None else int(n_scan), w_max=float(np.max(W)),
                    targets=targets, out=self.col_out[(t, r)], n_rows=n_rows, regime=r)
        self._col_specs.setdefault(t, []).append(spec)
        self._agg_specs.setdefault(t, []).append((t, r, scale))
        self.max_cand = max(self.max_cand, n_cand)

    # ------------------------------------------------------------------ packing
    def _reserve(self):
        """Reserve the child row tables, the candidate scratch and the
        per-period descriptor arrays in the arena (no device work)."""
        ar = self.ar
        # child row tables: one LcmEgmRow array per (t, regime) that has rows
        self._rowtab = {}
        for key, info in self.carry.items():
            if info["kind"] == "rows":
                self._rowtab[key] = ar.reserve(len(info["rows"]) * ROW.itemsize)
        self._period = {}
        # candidate rows g, p, v, s of every column, then EV and the all-dead flag
        self._cand_per_col = ((4 * self.max_cand + 2) * 8 + 255) // 256 * 256
        max_cols = max((sum(s["n_cols"] for s in specs) for specs in self._col_specs.values()), default=0)
        self.cand_off = ar.reserve(max(max_cols, 1) * max(self._cand_per_col, 8))
        for t, specs in self._col_specs.items():
            n_cols = sum(s["n_cols"] for s in specs)
            aggs = sum(self.col_out[(tt, r)]["n_sd"] for tt, r, _ in self._agg_specs[t])
            self._period[t] = (ar.reserve(n_cols * COL.itemsize), n_cols, ar.reserve(aggs * AGG.itemsize), aggs)
        self.total_bytes = ar.size

    def _image(self, base):
        """The arena image for device address `base` (row tables and
        descriptors written), with {t: (columns, aggregates)} ctypes arrays
        and {t: launch sizes}."""
        ar = self.ar
        img = ar.image()
        if img.size < ar.size:
            img = np.concatenate([img, np.zeros(ar.size - img.size, np.uint8)])

        def put(off, struct_arr):
            raw = np.frombuffer(bytes(struct_arr), dtype=np.uint8)
            img[off: off + raw.size] = raw

        for key, off in self._rowtab.items():
            rows = self.carry[key]["rows"]
            arr = (nat.LcmEgmRow * len(rows))()
            for i, (g, v, m, nv, ln) in enumerate(rows):
                arr[i] = nat.LcmEgmRow(base + g, base + v, base + m, base + nv, ln, 0)
            put(off, arr)
        descs, sizes = {}, {}
        cand_per_col = self._cand_per_col
        for t, specs in self._col_specs.items():
            col_off, n_cols, agg_off, n_aggs = self._period[t]
            arr = (nat.LcmEgmCol * n_cols)()
            ci = 0
            max_cand = 0
            for s in specs:
                o = s["out"]
                for c in range(s["n_cols"]):
                    d = arr[ci]
                    d.A, d.W = base + s["A"], base + s["W"]
                    d.u_prog, d.i_prog = base + s["up"], base + s["ip"]
                    d.u_c = base + s["uc"] + 8 * c * s["uc_n"]
                    d.i_c = base + s["ic"] + 8 * c * s["ic_n"]
                    d.n_targets = len(s["targets"])
                    for k, tg in enumerate(s["targets"]):
                        d.tkind[k] = tg["kind"]
                        d.tprob[k] = float(tg["prob"][c])
                        if tg["kind"] == 0:
                            d.tscalar[k] = tg["scalar"]
                            continue
                        d.tR[k] = base + s["R"] + 8 * c * s["n_s"]
                        d.tG[k] = base + s["G"] + 8 * c * s["n_s"]
                        d.tW[k] = base + tg["W_off"] + 8 * c * tg["K"]
                        d.tRows[k] = base + self._rowtab[tg["child"]]
                        d.tK[k], d.tN[k], d.tscale[k] = tg["K"], tg["N"], tg["scale"]
                    slot = 8 * s["n_pad"]
                    d.out_grid = base + o["grid"] + c * slot
                    d.out_policy = base + o["policy"] + c * slot
                    d.out_value = base + o["value"] + c * slot
                    d.out_marg = base + o["marg"] + c * slot
                    d.out_counts = base + o["counts"] + 8 * c
                    d.out_V = base + o["vcol"] + 8 * c * s["n_w"]
                    d.cand = base + self.cand_off + ci * cand_per_col
                    d.beta, d.ratio, d.jump_thresh, d.w_max = s["beta"], s["ratio"], s["jump"], s["w_max"]
                    d.n_s, d.n_w, d.n_con, d.n_pad = s["n_s"], s["n_w"], s["n_con"], s["n_pad"]
                    d.n_scan = s["n_scan"]
                    max_cand = max(max_cand, s["n_s"] + s["n_con"])
                    ci += 1
            put(col_off, arr)
            aggs = (nat.LcmEgmAgg * n_aggs)()
            ai = 0
            max_w = 0
            for tt, r, scale in self._agg_specs[t]:
                o = self.col_out[(tt, r)]
                vout_off, _ = self.out_V[(tt, r)]
                for sd in range(o["n_sd"]):
                    aggs[ai] = nat.LcmEgmAgg(base + o["vcol"] + 8 * sd * o["n_d"] * o["n_w"],
                                             base + vout_off + self.dt.itemsize * sd * o["n_w"],
                                             scale, o["n_d"], o["n_w"])
                    ai += 1
                max_w = max(max_w, o["n_w"])
            put(agg_off, aggs)
            descs[t] = (arr, aggs)
            sizes[t] = (int(nat.lib().lcm_egm_lds_bytes(max_cand)), max_w,
                        max(sp["n_rows"] for sp in specs), max(sp["n_s"] for sp in specs))
        return img, descs, sizes

    def _pack(self):
        self._reserve()
        self.buf = nat.DeviceBuffer(max(self.ar.size, 256))
        self._graph = None  # a captured graph addresses the previous buffer
        self.base = self.buf.ptr
        img, _, self.lds_by_t = self._image(self.base)
        nat.h2d(self.ctx, self.base, img[: self.ar.size])
        self.ctx.sync()

    # ------------------------------------------------------------------ running
    def run(self, sync=True, use_graph=False):
        """Every period's launches on the program's stream, newest period
        first. `use_graph`: replay them as one HIP graph (captured on first
        use; kernel nodes only), which saves the host one foreign call per
        launch; the arithmetic is the same launches'."""
        if use_graph:
            if self._graph is None:
                self._graph = nat.Graph.capture(self.ctx, self._enqueue)
            self._graph.launch(self.ctx)
        else:
            self._enqueue()
        if sync:
            self.ctx.sync()

    def _enqueue(self):
        L = nat.lib()
        code = nat.LCM_DTYPE_F64 if self.dt == np.float64 else nat.LCM_DTYPE_F32
        base = self.base
        for t in sorted(self._period, reverse=True):
            col_off, n_cols, agg_off, n_aggs = self._period[t]
            lds, max_w, max_rows, max_s = self.lds_by_t[t]
            nat._check(L.lcm_egm_launch_ex(self.ctx.handle, C.c_void_p(base + col_off), n_cols, C.c_size_t(lds),
                                           max_rows, max_s, self.nodes_per_block), f"lcm_egm_launch (period {t})")
            nat._check(L.lcm_egm_aggregate(self.ctx.handle, code, C.c_void_p(base + agg_off), n_aggs, max_w),
                       f"lcm_egm_aggregate (period {t})")

    def launch_regime(self, t: int, r: str):
        """The columns and aggregates of one (period, regime) alone: the
        regime's contiguous range of the period's descriptor arrays, launched
        with the period's sizes (so its arithmetic is that of the whole
        period's launch)."""
        L = nat.lib()
        code = nat.LCM_DTYPE_F64 if self.dt == np.float64 else nat.LCM_DTYPE_F32
        col_off, _, agg_off, _ = self._period[t]
        c0 = 0
        spec = None
        for sp in self._col_specs[t]:
            if sp["regime"] == r:
                spec = sp
                break
            c0 += sp["n_cols"]
        if spec is None:
            raise KeyError(f"regime {r!r} has no DC-EGM columns at period {t}")
        a0 = 0
        for tt, rr, _ in self._agg_specs[t]:
            if rr == r:
                break
            a0 += self.col_out[(tt, rr)]["n_sd"]
        lds, max_w, max_rows, max_s = self.lds_by_t[t]
        nat._check(L.lcm_egm_launch_ex(self.ctx.handle, C.c_void_p(self.base + col_off + c0 * COL.itemsize),
                                       spec["n_cols"], C.c_size_t(lds), max_rows, max_s, self.nodes_per_block),
                   f"lcm_egm_launch ({r!r}, period {t})")
        nat._check(L.lcm_egm_aggregate(self.ctx.handle, code, C.c_void_p(self.base + agg_off + a0 * AGG.itemsize),
                                       self.col_out[(t, r)]["n_sd"], max_w),
                   f"lcm_egm_aggregate ({r!r}, period {t})")

    def _fetch(self, off, nbytes, dtype):
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        nat.d2h(self.ctx, self.base + off, out)
        return out

    def value_functions(self):
        V = {}
        for t in range(self.T):
            age = self.model.ages.period_to_age(t)
            for r in self.names:
                if not self.layouts[r].regime.active(age):
                    continue
                if (t, r) in self.host_V:
                    V.setdefault(t, {})[r] = np.asarray(self.host_V[(t, r)])
                else:
                    off, shape = self.out_V[(t, r)]
                    n = int(np.prod(shape))
                    V.setdefault(t, {})[r] = self._fetch(off, n * self.dt.itemsize, self.dt).reshape(shape)
        return V

    def columns(self, t, r):
        """FP64 carry rows and V rows of every column of (t, r):
        dict of arrays [n_cols, ...] plus counts [n_cols, 2]."""
        o = self.col_out[(t, r)]
        nc, npad, nw = o["n_cols"], o["n_pad"], o["n_w"]
        out = {k: self._fetch(o[k], nc * npad * 8, np.float64).reshape(nc, npad)
               for k in ("grid", "policy", "value", "marg")}
        out["counts"] = self._fetch(o["counts"], nc * 8, np.int32).reshape(nc, 2)
        out["V"] = self._fetch(o["vcol"], nc * nw * 8, np.float64).reshape(nc, nw)
        return out

    def policies(self):
        """Refined consumption policy rows: period -> regime -> (grid, policy)
        arrays of shape [n_sd..., n_d, n_pad]."""
        P = {}
        for (t, r), o in self.col_out.items():
            cols = self.columns(t, r)
            lay = self.layouts[r]
            shape = lay.sd_sizes + (o["n_d"], o["n_pad"])
            P.setdefault(t, {})[r] = (cols["grid"].reshape(shape), cols["policy"].reshape(shape))
        return P


def solve_dcegm_model(model, params, dtype):
    """One DC-EGM solve on the device: (V, policies)."""
    prog = EgmProgram(model, params, dtype)
    prog.run()
    return prog.value_functions(), prog.policies()


class EgmBatch:
    """Many parameter sets of one DC-EGM model: per period ONE column launch
    and ONE aggregate launch hold the columns of every set. Each set keeps
    its own arena (tables, carry rows, scratch) inside one device buffer, so
    its results are those of its own `EgmProgram`."""

    def __init__(self, model, params_list, dtype, ctx=None, nodes_per_block: int = 0):
        _abi_check()
        self.nodes_per_block = int(nodes_per_block)  # 0: chosen per launch by the library
        self._graph = None  # the period loop captured as a HIP graph (`run(use_graph=True)`)
        self.dt = np.dtype(dtype)
        self.ctx = ctx or nat.Context()
        self.progs = [EgmProgram(model, p, dtype, device=False) for p in params_list]
        offs, total = [], 0
        for pr in self.progs:
            pr._reserve()
            offs.append(total)
            total += (pr.ar.size + 255) // 256 * 256
        periods = sorted(self.progs[0]._period) if self.progs else []
        # the fused per-period descriptor arrays follow the programs' arenas
        self._fused = {}
        for t in periods:
            n_cols = sum(pr._period[t][1] for pr in self.progs)
            n_aggs = sum(pr._period[t][3] for pr in self.progs)
            col_off = total
            total += (n_cols * COL.itemsize + 255) // 256 * 256
            agg_off = total
            total += (n_aggs * AGG.itemsize + 255) // 256 * 256
            self._fused[t] = (col_off, n_cols, agg_off, n_aggs)
        self.total_bytes = total
        self.buf = nat.DeviceBuffer(max(total, 256))
        self._graph = None  # a captured graph addresses the previous buffer
        base = self.buf.ptr
        cols = {t: [] for t in periods}
        aggs = {t: [] for t in periods}
        self.lds_by_t = {}
        for pr, off in zip(self.progs, offs):
            pr.ctx = self.ctx
            pr.buf = None
            pr.base = base + off
            img, descs, sizes = pr._image(pr.base)
            nat.h2d(self.ctx, pr.base, img[: pr.ar.size])
            for t in periods:
                cols[t].append(bytes(descs[t][0]))
                aggs[t].append(bytes(descs[t][1]))
                lds, mw, mr, ms = sizes[t]
                o = self.lds_by_t.get(t, (0, 0, 0, 0))
                self.lds_by_t[t] = (max(o[0], lds), max(o[1], mw), max(o[2], mr), max(o[3], ms))
        for t in periods:
            col_off, _, agg_off, _ = self._fused[t]
            c = np.frombuffer(b"".join(cols[t]), dtype=np.uint8)
            a = np.frombuffer(b"".join(aggs[t]), dtype=np.uint8)
            if c.size:
                nat.h2d(self.ctx, base + col_off, c)
            if a.size:
                nat.h2d(self.ctx, base + agg_off, a)
        self.ctx.sync()

    def run(self, sync=True, use_graph=False):
        """The batch's fused launches, newest period first; `use_graph` as
        `EgmProgram.run`."""
        if use_graph:
            if self._graph is None:
                self._graph = nat.Graph.capture(self.ctx, self._enqueue)
            self._graph.launch(self.ctx)
        else:
            self._enqueue()
        if sync:
            self.ctx.sync()

    def _enqueue(self):
        L = nat.lib()
        code = nat.LCM_DTYPE_F64 if self.dt == np.float64 else nat.LCM_DTYPE_F32
        base = self.buf.ptr
        for t in sorted(self._fused, reverse=True):
            col_off, n_cols, agg_off, n_aggs = self._fused[t]
            lds, max_w, max_rows, max_s = self.lds_by_t[t]
            nat._check(L.lcm_egm_launch_ex(self.ctx.handle, C.c_void_p(base + col_off), n_cols, C.c_size_t(lds),
                                           max_rows, max_s, self.nodes_per_block), f"lcm_egm_launch (period {t})")
            nat._check(L.lcm_egm_aggregate(self.ctx.handle, code, C.c_void_p(base + agg_off), n_aggs, max_w),
                       f"lcm_egm_aggregate (period {t})")

    def value_functions(self, i):
        return self.progs[i].value_functions()

    def policies(self, i):
        return self.progs[i].policies()
