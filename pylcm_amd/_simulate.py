"""Forward simulation on the GPU (`src/lcm/model.py:648-848`,
`src/_lcm/simulation/simulate.py:74-620`).

Per (regime, period) the host lowers the regime's functions with the agent's
continuous state `w` and its process value `z` as symbols (`_affine.Affine`
with features), giving tables over (discrete state, discrete action,
continuous action). The device then, per period:

1. contracts next-period value rows into EV rows (`lcm_ev_launch`) from the
   solve's own descriptors and value arrays, which stay in HBM;
2. per agent, takes the first argmax of Q over the action grid at the agent's
   off-grid state, mixing the EV rows of the two process nodes around `z` with
   the interpolated transition-row weights, and draws the next state and regime
   with Philox4x32-10 (`lcm_sim_period_ex`).

Random draws use a counter-based generator keyed by `seed` with counter
(agent, period, stream), so results do not depend on how agents are ordered or
batched; they are not the reference's JAX threefry streams.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import pandas as pd

from pylcm_amd import _native as nat
from pylcm_amd._affine import Affine, Constraint, FeatureSet, NonAffine
from pylcm_amd._build import PeriodLowering, _functions_of, interpolated_row
from pylcm_amd._gs_program import GsProgram, _align, interval_form, lower_problem
from pylcm_amd.exceptions import (
    InvalidInitialConditionsError,
    UnsupportedModelError,
)
from pylcm_amd.grids import DiscreteGrid
from pylcm_amd.processes import StochasticProcess, draw_params
from pylcm_amd.regime import MarkovTransition

MAX_FEAT = nat.SIM_MAX_FEAT


# ---------------------------------------------------------------------------
# lowering
# ---------------------------------------------------------------------------


@dataclass
class SimTables:
    regime: str
    period: int
    terminal: bool
    n_sdd: int
    n_d: int
    n_c: int
    U0: np.ndarray
    Uf: list
    A: np.ndarray | None
    B0: np.ndarray | None
    Bf: list
    lo: np.ndarray
    hi: np.ndarray
    feats: list          # [(op, coef[n_sdd, MAX_FEAT], const[n_sdd])]
    preg: np.ndarray | None = None       # [n_sdd, n_d, n_regimes]
    Wd: dict = field(default_factory=dict)  # target regime -> [n_sdd, n_d, n_sdd_r]
    interp: bool = False
    LoB: np.ndarray | None = None
    HiB: np.ndarray | None = None
    LoF: dict = field(default_factory=dict)
    HiF: dict = field(default_factory=dict)
    lo_strict: bool = False
    hi_strict: bool = False


class SimLayout:
    """Agent-side axes of a regime: discrete non-process states `sdd`, at most
    one process state, the continuous state, discrete and continuous actions."""

    def __init__(self, lay):
        self.lay = lay
        self.sdd_idx = [i for i, g in enumerate(lay.sd_grids) if isinstance(g, DiscreteGrid)]
        self.proc_idx = [i for i, g in enumerate(lay.sd_grids) if isinstance(g, StochasticProcess)]
        if len(self.proc_idx) > 1:
            raise UnsupportedModelError(
                f"regime {lay.name!r}: simulation supports at most one process state"
            )
        self.sdd_names = [lay.sd_names[i] for i in self.sdd_idx]
        self.sdd_sizes = tuple(lay.sd_grids[i].n_points for i in self.sdd_idx)
        self.n_sdd = int(np.prod(self.sdd_sizes, dtype=np.int64)) if self.sdd_sizes else 1
        self.proc_name = lay.sd_names[self.proc_idx[0]] if self.proc_idx else None
        # flat full-sd index = sum idx_k * stride_k (row-major over lay.sd_sizes)
        sizes = lay.sd_sizes
        strides = [int(np.prod(sizes[k + 1:], dtype=np.int64)) for k in range(len(sizes))]
        self.p_stride = strides[self.proc_idx[0]] if self.proc_idx else 0
        base = np.zeros(self.n_sdd, dtype=np.int32)
        if self.sdd_sizes:
            for flat, idx in enumerate(np.ndindex(*self.sdd_sizes)):
                base[flat] = sum(i * strides[k] for i, k in zip(idx, self.sdd_idx))
        self.sd_base = base


class SimLowering(PeriodLowering):
    def _sim_axes(self, sl: SimLayout, fs: FeatureSet):
        lay = sl.lay
        n_sdax = len(sl.sdd_idx)
        ndim = n_sdax + 1 + len(lay.da_names) + 1
        v = {}

        def on_axis(vals, ax):
            shape = [1] * ndim
            shape[ax] = len(vals)
            return np.asarray(vals).reshape(shape)

        for j, i in enumerate(sl.sdd_idx):
            v[lay.sd_names[i]] = on_axis(np.arange(lay.sd_grids[i].n_points, dtype=np.int64), j)
        for j, a in enumerate(lay.da_names):
            v[a] = on_axis(np.arange(lay.da_grids[j].n_points, dtype=np.int64), n_sdax + 1 + j)
        if lay.c_name:
            v[lay.c_name] = on_axis(self.rg.c_pts[lay.name], ndim - 1)
        if lay.w_name:
            v[lay.w_name] = Affine(1.0, 0.0, None, fs)
        if sl.proc_name:
            fid = fs.identity(sl.proc_name)
            v[sl.proc_name] = Affine(0.0, 0.0, {fid: 1.0}, fs)
        shape = sl.sdd_sizes + (1,) + lay.da_sizes + (lay.n_c,)
        return v, shape

    def _to3(self, sl, arr, shape):
        a = np.broadcast_to(np.asarray(arr, dtype=np.float64), shape)
        return np.ascontiguousarray(a.reshape(sl.n_sdd, sl.lay.n_d, sl.lay.n_c))

    def lower_sim(self, sl: SimLayout, period: int) -> SimTables:
        """The simulation tables of (regime, period). Periods with the same
        `period_key` (every function the tables read is age-free, and the
        regime transition's ids and the active next regimes agree) share one
        lowering, as the solve's tables do."""
        key = self.period_key(sl.lay, period)
        if key is None:
            return self._lower_sim(sl, period)
        memo_key = ("sim", sl.lay.name, key)
        if memo_key not in self._memo:
            self._memo[memo_key] = self._lower_sim(sl, period)
        return self._memo[memo_key]

    def _lower_sim(self, sl: SimLayout, period: int) -> SimTables:
        lay = sl.lay
        reg = lay.regime
        n_sdax = len(sl.sdd_idx)
        action_axes = tuple(range(n_sdax + 1, n_sdax + 2 + len(lay.da_names)))
        fs = FeatureSet(action_axes=action_axes)
        variables, shape = self._sim_axes(sl, fs)
        ev = self._evaluator(lay, variables, period)
        try:
            u = ev.call("utility")
        except NonAffine as e:
            raise UnsupportedModelError(f"regime {lay.name!r}: utility is not simulatable here ({e})") from e
        if isinstance(u, Affine) and u.has_w:
            raise UnsupportedModelError(f"regime {lay.name!r}: utility reads the continuous state")
        U0 = self._to3(sl, u.b if isinstance(u, Affine) else u, shape)
        ufeats = u.feats if isinstance(u, Affine) else {}
        rows = []
        feat_rows = []  # (a, b, strict, {fid: coef}) of rows reading per-agent features
        for cname, cf in reg.constraints.items():
            if cf is None:
                continue
            try:
                out = ev.call(cname, cf)
            except NonAffine as e:
                raise UnsupportedModelError(f"constraint {cname!r} is not simulatable here ({e})") from e
            if isinstance(out, Constraint):
                for (a, b, strict), fts in zip(out.terms, out.feats):
                    if fts and any(np.any(v != 0) for v in fts.values()):
                        feat_rows.append((self._to3(sl, a, shape), self._to3(sl, b, shape), strict,
                                          {k: self._to3(sl, v, shape) for k, v in fts.items()}))
                        continue
                    rows.append((self._to3(sl, a, shape)[:, None], self._to3(sl, b, shape)[:, None], strict))
            else:
                m = self._to3(sl, np.asarray(out, dtype=bool), shape)
                rows.append((np.zeros_like(m)[:, None], np.where(m > 0, 0.0, -np.inf)[:, None], False))
        dt = self.rg.dtype
        iv = interval_form(rows, sl.n_sdd, lay.n_d, lay.n_c, dt)
        if iv is None:
            raise UnsupportedModelError(f"regime {lay.name!r}: constraints have no interval form")
        lo = np.broadcast_to(iv[0][:, 0], (sl.n_sdd, lay.n_d, lay.n_c))
        hi = np.broadcast_to(iv[1][:, 0], (sl.n_sdd, lay.n_d, lay.n_c))
        st = SimTables(regime=lay.name, period=period, terminal=lay.terminal, n_sdd=sl.n_sdd,
                       n_d=lay.n_d, n_c=lay.n_c, U0=U0, Uf=[], A=None, B0=None, Bf=[],
                       lo=np.ascontiguousarray(lo), hi=np.ascontiguousarray(hi), feats=[])
        # a row a*w + b + sum_k f_k phi_k >= 0 with a = +1 bounds w from below
        # (w >= -b - sum f phi), with a = -1 from above (w <= b + sum f phi)
        for a, b, strict, fts in feat_rows:
            if np.all(a == 1):
                if st.LoB is not None:
                    raise UnsupportedModelError("two lower bounds read per-agent values")
                st.LoB, st.LoF, st.lo_strict = -b, {k: -v for k, v in fts.items()}, strict
            elif np.all(a == -1):
                if st.HiB is not None:
                    raise UnsupportedModelError("two upper bounds read per-agent values")
                st.HiB, st.HiF, st.hi_strict = b, dict(fts), strict
            else:
                raise UnsupportedModelError("a constraint reading per-agent values needs coefficient +-1 on w")
        bfeats = {}
        if not lay.terminal and lay.w_name is not None:
            law = reg.state_transitions.get(lay.w_name)
            try:
                x = ev.call(f"next_{lay.w_name}", law)
            except NonAffine as e:
                raise UnsupportedModelError(f"next {lay.w_name!r} is not simulatable here ({e})") from e
            if not isinstance(x, Affine):
                x = Affine(0.0, x)
            st.A = self._to3(sl, x.a, shape)
            st.B0 = self._to3(sl, x.b, shape)
            bfeats = x.feats
        n_feat = len(fs.recipes)
        if n_feat > MAX_FEAT:
            raise UnsupportedModelError(f"regime {lay.name!r}: more than {MAX_FEAT} per-agent features")
        for k, (op, coefs, const) in enumerate(fs.recipes):
            coef = np.zeros((sl.n_sdd, MAX_FEAT))
            cst = np.zeros(sl.n_sdd)
            if op != 0:
                sdd_shape = sl.sdd_sizes + (1,) * (len(shape) - n_sdax)
                for fid, cv in coefs.items():
                    coef[:, fid] = np.broadcast_to(cv, sdd_shape).reshape(sl.n_sdd) if np.ndim(cv) else cv
                cst[:] = np.broadcast_to(const, sdd_shape).reshape(sl.n_sdd) if np.ndim(const) else const
            st.feats.append((op, coef, cst))
            st.Uf.append(self._to3(sl, ufeats[k], shape) if k in ufeats else None)
            st.Bf.append(self._to3(sl, bfeats[k], shape) if k in bfeats else None)
        if lay.terminal:
            return st
        model = self.model
        n_reg = len(model.regime_names)
        # regime probabilities and next discrete states, per (sdd, d)
        preg = np.zeros((sl.n_sdd, lay.n_d, n_reg))
        tr = reg.transition
        for i, j, v in self._sim_points(sl, fs):
            if hasattr(tr, "items"):
                for tname, f in tr.items():
                    preg[i, j, model.regime_code(tname)] = float(np.asarray(self._const(lay, period, "next_regime", f, v)))
            elif isinstance(tr, MarkovTransition):
                preg[i, j] = np.asarray(self._const(lay, period, "next_regime", tr, v), dtype=np.float64)
            else:
                preg[i, j, int(np.asarray(self._const(lay, period, "next_regime", tr, v)))] = 1.0
        st.preg = preg
        for r in model.regime_names:
            if not np.any(preg[:, :, model.regime_code(r)] != 0):
                continue
            tl = model.layouts[r]
            tsl = SimLayout(tl)
            per_axis = []
            for kk in tsl.sdd_idx:
                s, g = tl.sd_names[kk], tl.sd_grids[kk]
                W = np.zeros((sl.n_sdd, lay.n_d, g.n_points))
                law = reg.state_transitions.get(s)
                if law is None:
                    raise UnsupportedModelError(f"no law for state {s!r} entering {r!r}")
                if hasattr(law, "items"):
                    law = law[r]
                for i, j, v in self._sim_points(sl, fs):
                    out = np.asarray(self._const(lay, period, f"next_{s}", law, v), dtype=np.float64)
                    if isinstance(law, MarkovTransition):
                        W[i, j] = out
                    else:
                        W[i, j, int(out)] = 1.0
                per_axis.append(W)
            joint = np.ones((sl.n_sdd, lay.n_d, 1))
            for W in per_axis:
                joint = (joint[..., :, None] * W[..., None, :]).reshape(sl.n_sdd, lay.n_d, -1)
            st.Wd[r] = joint
        return st

    def _sim_points(self, sl, fs):
        lay = sl.lay
        for i, sdi in enumerate(np.ndindex(*sl.sdd_sizes) if sl.sdd_sizes else [()]):
            for j, dai in enumerate(np.ndindex(*lay.da_sizes) if lay.da_sizes else [()]):
                v = {sl.sdd_names[k]: int(sdi[k]) for k in range(len(sdi))}
                v.update({lay.da_names[k]: int(dai[k]) for k in range(len(dai))})
                if lay.w_name:
                    v[lay.w_name] = Affine(1.0, 0.0, None, fs)
                if sl.proc_name:
                    v[sl.proc_name] = Affine(0.0, 0.0, {0: 1.0}, fs)
                yield i, j, v

    def _const(self, lay, period, fname, func, variables):
        out = self._pointwise(lay, period, fname, func, variables)
        if isinstance(out, Affine):
            if out.is_const:
                return out.b
            raise UnsupportedModelError(f"{fname} depends on the agent's continuous values")
        return out


# ---------------------------------------------------------------------------
# result
# ---------------------------------------------------------------------------


# LDS one simulation workgroup may use (gfx950: 160 KiB per workgroup)
SIM_LDS_LIMIT = 160 * 1024
# EV rows of one cell up to this size are read from global memory (L1-resident)
SIM_EV_L1_BYTES = 32 * 1024


class SimulationResult:
    """Simulated panel (`src/lcm/result.py:125-185`)."""

    def __init__(self, frame: pd.DataFrame, model, terminal_names, params=None, dtype=None):
        self._frame = frame
        self._model = model
        self._terminal = terminal_names
        self._params = params
        self._dtype = dtype

    # -- additional targets (src/_lcm/simulation/additional_targets.py:54-265) --
    def _targets_of(self, r):
        """User functions and constraints of regime r that can be evaluated
        from a row of the panel (the inverse marginal utility of a DC-EGM
        regime reads the continuation, so it is not one)."""
        lay = self._model.layouts[r]
        reg = lay.regime
        names = {k for k, f in reg.functions.items() if f is not None} | set(reg.constraints)
        if lay.is_dcegm:
            names.discard("inverse_marginal_utility")
        return names

    @property
    def available_targets(self) -> list[str]:
        """Names accepted by `to_dataframe(additional_targets=...)`."""
        out = set()
        for r in self._model.regime_names:
            out |= self._targets_of(r)
        return sorted(out)

    def _compute_targets(self, df, targets):
        from pylcm_amd._dag import Evaluator
        from pylcm_amd.exceptions import InvalidAdditionalTargetsError

        if targets == "all":
            targets = self.available_targets
        bad = set(targets) - set(self.available_targets)
        if bad:
            raise InvalidAdditionalTargetsError(
                f"Targets {bad} not found in any regime. Available targets: {self.available_targets}")
        cols = {t: np.full(len(df), np.nan, dtype=object) for t in targets}
        model = self._model
        for r in model.regime_names:
            rows = np.nonzero((df["regime_name"] == r).to_numpy())[0]
            mine = [t for t in targets if t in self._targets_of(r)]
            if not rows.size or not mine:
                continue
            lay = model.layouts[r]
            reg = lay.regime
            variables = {"age": df["age"].to_numpy()[rows], "period": df["period"].to_numpy()[rows]}
            for name in list(lay.state_names) + list(lay.action_names):
                v = df[name].to_numpy()[rows]
                discrete = name in lay.sd_names and isinstance(lay.sd_grids[lay.sd_names.index(name)], DiscreteGrid)
                discrete |= name in lay.da_names
                variables[name] = v.astype(np.int64) if discrete else v.astype(np.float64)
            funcs = {k: f for k, f in reg.functions.items() if f is not None}
            ev = Evaluator(regime_name=r, functions=funcs, variables=variables, params=self._params,
                           fixed_params=model.fixed_params, param_dtype=self._dtype)
            for t in mine:
                val = ev.call(t, reg.constraints[t]) if t in reg.constraints else ev.call(t)
                cols[t][rows] = np.broadcast_to(np.asarray(val), rows.shape)
        out = df.copy()
        for t, v in cols.items():
            try:
                out[t] = v.astype(np.float64)
            except (TypeError, ValueError):
                out[t] = v
        return out

    def to_dataframe(self, additional_targets=None, *, use_labels: bool = True,
                     terminal_rows: str = "first") -> pd.DataFrame:
        df = self._frame
        if additional_targets:
            df = self._compute_targets(df, additional_targets)
        if terminal_rows == "first" and self._terminal:
            is_term = df["regime_name"].isin(list(self._terminal))
            first = df[is_term].groupby("subject_id")["period"].transform("min")
            keep = ~is_term
            keep.loc[is_term] = df.loc[is_term, "period"].values == first.values
            df = df[keep]
        df = df.reset_index(drop=True)
        if not use_labels:
            out = df.copy()
            names = list(self._model.regime_names)
            out["regime_name"] = out["regime_name"].map({n: i for i, n in enumerate(names)})
            return out
        out = df.copy()
        out["regime_name"] = pd.Categorical(out["regime_name"], categories=list(self._model.regime_names))
        for col, cats in _discrete_labels(self._model).items():
            if col in out.columns:
                codes = out[col]
                out[col] = pd.Categorical.from_codes(
                    np.where(codes.isna(), -1, codes.fillna(-1)).astype(int), categories=list(cats))
        return out


def _discrete_labels(model):
    out = {}
    for lay in model.layouts.values():
        for n, g in list(zip(lay.sd_names, lay.sd_grids)) + list(zip(lay.da_names, lay.da_grids)):
            if isinstance(g, DiscreteGrid):
                out[n] = g.categories
    return out


# ---------------------------------------------------------------------------
# driver
# ---------------------------------------------------------------------------


DCEGM_BUDGET_CONSTRAINT_NAME = "dcegm_budget_constraint"


def _budget_constraint(action: str, resources: str, limit: float):
    """`action <= resources - limit` as a constraint reading those two names
    (src/_lcm/egm/budget.py:36-95)."""
    import inspect

    def dcegm_budget_constraint(**kw):
        return kw[action] <= kw[resources] - limit

    P = inspect.Parameter
    dcegm_budget_constraint.__signature__ = inspect.Signature(
        [P(action, P.POSITIONAL_OR_KEYWORD), P(resources, P.POSITIONAL_OR_KEYWORD)])
    return dcegm_budget_constraint


def grid_argmax_view(model):
    """The model with every DC-EGM regime simulated by the argmax over its own
    action grids, as the reference's simulation does for every shipped
    envelope (src/_lcm/regime_building/processing.py:1662-1700): the regime
    keeps its functions (minus the inverse marginal utility, which only the
    Euler inversion reads) and gains the budget mask the EGM solve enforces
    intrinsically, `action <= resources - savings_grid[0]`
    (src/_lcm/egm/budget.py:36-95)."""
    from pylcm_amd.model import Model
    from pylcm_amd.regime import Regime
    from pylcm_amd.solvers import DCEGM, GridSearch

    regimes = {}
    for name, reg in model.user_regimes.items():
        if not isinstance(reg.solver, DCEGM) or reg.terminal:
            regimes[name] = reg
            continue
        mg = reg.liquid
        limit = float(reg.solver.savings_grid.to_numpy(np.float64)[0])
        regimes[name] = Regime(
            transition=reg.transition, active=reg.active, states=reg.states,
            state_transitions=reg.state_transitions, actions=reg.actions,
            functions={k: f for k, f in reg.functions.items() if k != "inverse_marginal_utility"},
            constraints={**reg.constraints,
                         DCEGM_BUDGET_CONSTRAINT_NAME: _budget_constraint(mg.action, mg.resources, limit)},
            derived_categoricals=reg.derived_categoricals, solver=GridSearch(),
            taste_shocks=reg.taste_shocks, koopmans_aggregator=reg.koopmans_aggregator,
            certainty_equivalent=reg.certainty_equivalent)
    return Model(regimes=regimes, ages=model.ages, regime_id_class=model.regime_id_class,
                 fixed_params=model.fixed_params)


def simulate_model(model, *, params, initial_conditions, V, seed, dtype, subject_batch_size=0,
                   log_level="off", prog: GsProgram | None = None):
    dt = np.dtype(dtype)
    if any(lay.is_dcegm for lay in model.layouts.values()):
        # DC-EGM regimes: value functions from the DC-EGM solve, choices by
        # the grid argmax of the view
        if V is None:
            from pylcm_amd._egm_program import solve_dcegm_model

            V, _ = solve_dcegm_model(model, params, dt)
        model = grid_argmax_view(model)
    ic = _initial_conditions(model, initial_conditions)
    if prog is None:
        prog = GsProgram(model, [params], dt)
    if V is None:
        prog.run(use_graph=False)
    else:
        _load_V(prog, V)
    sim = SimProgram(model, params, prog)
    return sim.run(ic, seed=0 if seed is None else int(seed))


def _load_V(prog: GsProgram, V):
    for (t, r), o in prog.outs[0].items():
        if t not in V or r not in V[t]:
            raise InvalidInitialConditionsError(f"period_to_regime_to_V_arr lacks period {t}, regime {r!r}")
        arr = np.ascontiguousarray(np.asarray(V[t][r], dtype=prog.dt).reshape(o.shape))
        nat.h2d(prog.ctx, prog.base + o.v_off, arr)


def _initial_conditions(model, ic):
    if isinstance(ic, pd.DataFrame):
        ic = {c: ic[c].to_numpy() for c in ic.columns}
        if "regime_name" in ic:
            ic["regime_id"] = np.array([model.regime_code(n) for n in ic.pop("regime_name")])
    ic = {k: np.asarray(v) for k, v in ic.items()}
    if "regime_id" not in ic:
        raise InvalidInitialConditionsError("initial_conditions need a 'regime_id' entry")
    n = len(ic["regime_id"])
    for k, v in ic.items():
        if len(v) != n:
            raise InvalidInitialConditionsError(f"initial condition {k!r} has length {len(v)}, expected {n}")
    ages = ic.get("age", np.full(n, model.ages.period_to_age(0)))
    start = np.array([model.ages.age_to_period(a) for a in np.asarray(ages).tolist()], dtype=np.int32)
    ic["_start"] = start
    ic["_n"] = n
    return ic


class SimProgram:
    """Device simulation of one solved program.

    With `static_scratch=True` every period goes through
    `lcm_sim_period_noalloc`, whose partition scratch is the library's static
    device array (for callers that cannot allocate device memory: at most
    `nat.SIM_NOALLOC_MAX_ORDER` regimes x agents); otherwise through
    `lcm_sim_period_ex` with scratch in the program's own buffer.
    """

    def __init__(self, model, params, prog: GsProgram, *, static_scratch: bool = False):
        self.static_scratch = static_scratch
        self.model = model
        self.params = params
        self.prog = prog
        self.dt = prog.dt
        self.rg = prog._rg[0]
        self.low = SimLowering(model, params, self.rg)
        self.sl = {r: SimLayout(model.layouts[r]) for r in model.regime_names}
        if len(model.regime_names) > nat.SIM_MAX_REGIMES:
            raise UnsupportedModelError(f"simulation supports at most {nat.SIM_MAX_REGIMES} regimes")

    @classmethod
    def for_placement(cls, model, params, prog):
        """A program that only places an image `build_image` made elsewhere
        (in a lowering process): `prog` provides the stream (`ctx`), `dt`
        and the solve's descriptors at `base + desc_off`."""
        self = cls.__new__(cls)
        self.static_scratch = False
        self.model = model
        self.params = params
        self.prog = prog
        self.dt = prog.dt
        self.sl = {r: SimLayout(model.layouts[r]) for r in model.regime_names}
        if len(model.regime_names) > nat.SIM_MAX_REGIMES:
            raise UnsupportedModelError(f"simulation supports at most {nat.SIM_MAX_REGIMES} regimes")
        return self

    def run(self, ic, seed):
        self.prepare(ic, seed)
        self.launch()
        self.prog.ctx.sync()
        return self.collect(ic)

    def prepare(self, ic, seed, mem=None):
        """Allocate the agent panels, lower every period's tables and upload
        them with the descriptors; `launch` then enqueues the simulation.
        `build_image` (host only) followed by `place` (device)."""
        self.place(self.build_image(), ic, seed, mem=mem)

    # pointer bases of a simulation image: its own table arena, the arena of
    # EV rows / CS / finite flags the device writes, the solve's descriptors
    BASE_TABLES, BASE_EV, BASE_SOLVE = 0, 1, 2

    def build_image(self) -> dict:
        """The host half of `prepare`: every period's simulation tables and
        the EV and simulation descriptors, with each pointer written relative
        to one of three bases and listed by byte offset in `patch[kind]`, so
        the image can be built where the solve was lowered (another process)
        and placed anywhere. The agent panels are not part of it."""
        model, dt, prog = self.model, self.dt, self.prog
        T = model.n_periods
        isz = dt.itemsize
        names = list(model.regime_names)
        R = len(names)
        from pylcm_amd._gs_program import _Arena, DESC_DTYPE

        ar = _Arena()  # tables and descriptors (uploaded)
        rs = _Arena()  # EV rows, CS and per-row finite flags (written on the device)
        zero = ar.add(np.zeros(4, dtype=dt))
        ev_meta = {}
        for (t, r), o in prog.outs[0].items():
            lay = model.layouts[r]
            pt = prog.problems[0][(t, r)]
            if lay.terminal:
                continue
            # g_n = 0: only stateless targets, the kernel fills CS alone
            g_n = len(self.rg.w_pts[pt.next_grid]) if pt.targets else 0
            ev_meta[(t, r)] = (rs.reserve(max(lay.n_sd * lay.n_d * g_n, 1) * isz),
                               rs.reserve(lay.n_sd * lay.n_d * isz), g_n,
                               rs.reserve(lay.n_sd * lay.n_d * 4))  # per-row finite flags
        sim_desc = {}
        lds_max = 0
        keep_tables = []  # the arena dedupes by object identity while it is built
        L = nat.load_library()
        code = nat.LCM_DTYPE_F64 if dt == np.float64 else nat.LCM_DTYPE_F32
        for t in range(T):
            for r in model.active_regimes(t):
                sl = self.sl[r]
                lay = sl.lay
                st = self.low.lower_sim(sl, t)
                keep_tables.append(st)
                sim_desc[(t, r)] = self._desc(st, sl, t, r, ar, zero, ev_meta)
                g_n = ev_meta.get((t, r), (0, 0, 0, 0))[2]
                ntab = st.n_sdd * (7 + 4 * len(st.feats))
                need = int(L.lcm_sim_lds_bytes(code, g_n, lay.n_sd if g_n else 0, ntab))
                # EV rows are read in place from global memory when one cell's
                # rows fit in a CU's L1 (no staging pass) or do not fit in LDS
                # next to the tables; otherwise they are staged in LDS
                cell_bytes = lay.n_sd * g_n * dt.itemsize
                ev_global = need > SIM_LDS_LIMIT or cell_bytes <= SIM_EV_L1_BYTES
                sim_desc[(t, r)]["ev_global"] = int(ev_global)
                if ev_global:
                    need = int(L.lcm_sim_lds_bytes(code, g_n, 0, ntab))
                lds_max = max(lds_max, need)
        if lds_max > SIM_LDS_LIMIT:
            raise UnsupportedModelError("simulation action tables and next-state grid do not fit in LDS "
                                        f"({lds_max} B > {SIM_LDS_LIMIT} B)")
        # EV descriptors point at the solve program's own descriptors
        gs_index = self._gs_desc_index()
        ev_dt = np.dtype(nat.LcmEvDesc)
        sd_dt = np.dtype(nat.LcmSimDesc)
        ev_descs = np.zeros(len(ev_meta), dtype=ev_dt)
        ev_keys = sorted(ev_meta)
        ev_desc_off = _align(ar.size)
        sim_desc_off = ev_desc_off + _align(len(ev_keys) * ev_dt.itemsize)
        patch = {self.BASE_TABLES: [], self.BASE_EV: [], self.BASE_SOLVE: []}

        for i, key in enumerate(ev_keys):
            t, r = key
            lay = model.layouts[r]
            e_off, c_off, g_n, f_off = ev_meta[key]
            ev_descs[i]["gs"] = gs_index[key] * DESC_DTYPE.itemsize
            ev_descs[i]["EV"], ev_descs[i]["CS"], ev_descs[i]["fin"] = e_off, c_off, f_off
            ev_descs[i]["n_sd"], ev_descs[i]["n_d"], ev_descs[i]["g_n"] = lay.n_sd, lay.n_d, g_n
            base_i = ev_desc_off + i * ev_dt.itemsize
            patch[self.BASE_SOLVE].append(base_i + ev_dt.fields["gs"][1])
            for f in ("EV", "CS", "fin"):
                patch[self.BASE_EV].append(base_i + ev_dt.fields[f][1])
        SD = np.zeros(T * R, dtype=sd_dt)
        for (t, r), d in sim_desc.items():
            i = t * R + model.regime_code(r)
            rec = SD[i]
            base_i = sim_desc_off + i * sd_dt.itemsize
            for k, v in d.items():
                if k in ("Uf", "Bf", "fcoef", "fconst", "Wd", "LoF", "HiF"):
                    for j, p in enumerate(v):
                        rec[k][j] = p if p is not None else 0
                        if p is not None:
                            patch[self.BASE_TABLES].append(base_i + sd_dt.fields[k][1] + 8 * j)
                elif k in ("n_sdd_next", "feat_op"):
                    for j, p in enumerate(v):
                        rec[k][j] = p
                elif k in ("U0", "A", "B0", "lo", "hi", "sd_base", "nodes", "g_pts", "preg", "LoB", "HiB"):
                    rec[k] = v if v is not None else 0
                    if v is not None:
                        patch[self.BASE_TABLES].append(base_i + sd_dt.fields[k][1])
                elif k in ("EV", "CS", "ev_fin"):
                    rec[k] = v if v is not None else 0
                    if v is not None:
                        patch[self.BASE_EV].append(base_i + sd_dt.fields[k][1])
                else:
                    rec[k] = v
        ar.parts.append((ev_desc_off, ev_descs.tobytes()))
        ar.parts.append((sim_desc_off, SD.tobytes()))
        ar.size = max(ar.size, sim_desc_off + SD.nbytes)
        del keep_tables
        return {"T": T, "R": R, "image": ar.image(), "ev_bytes": _align(max(rs.size, 1)),
                "patch": {k: np.asarray(v, dtype=np.int64) for k, v in patch.items()},
                "ev_desc_off": ev_desc_off, "sim_desc_off": sim_desc_off, "ev_keys": ev_keys,
                "lds_max": lds_max, "sim_desc_size": sd_dt.itemsize}

    def place(self, img: dict, ic, seed, mem=None):
        """The device half of `prepare`: allocate the agent panels (n = the
        initial conditions' agents) and the image's two arenas, write the
        initial state, relocate the image's pointers and upload it. `mem`,
        when given, is `mem(slot, nbytes) -> device pointer` (slot 0: panels,
        1: tables and EV rows), so a caller can reuse its own memory;
        otherwise the program allocates its buffers."""
        model, dt, prog = self.model, self.dt, self.prog
        T, R = img["T"], img["R"]
        n = ic["_n"]
        isz = dt.itemsize
        names = list(model.regime_names)
        # -- device buffers ------------------------------------------------------
        parts = []

        def alloc(nbytes):
            parts.append(int(nbytes))
            return sum(_align(p) for p in parts[:-1])

        off = {}
        for side in ("cur", "nxt"):
            off[side] = dict(w=alloc(n * isz), z=alloc(n * isz), sdd=alloc(n * 4),
                             regime=alloc(n * 4), start=alloc(n * 4))
        off["out"] = dict(regime=alloc(T * n * 4), action=alloc(T * n * 4), value=alloc(T * n * isz),
                          w=alloc(T * n * isz), z=alloc(T * n * isz), sdd=alloc(T * n * 4))
        off["counts"] = alloc(2 * R * 4)  # two halves, alternating by period
        off["order"] = alloc(R * n * 4)
        panel_bytes = sum(_align(p) for p in parts)
        arena_bytes = _align(img["image"].nbytes) + img["ev_bytes"]
        if mem is None:
            self.buf = nat.DeviceBuffer(panel_bytes)
            self.buf2 = nat.DeviceBuffer(arena_bytes)
            base, b2 = self.buf.ptr, self.buf2.ptr
        else:
            base, b2 = mem(0, panel_bytes), mem(1, arena_bytes)
        ctx = prog.ctx
        # -- initial agent state -------------------------------------------------
        w0 = np.zeros(n)
        z0 = np.zeros(n)
        sdd0 = np.zeros(n, dtype=np.int32)
        reg0 = np.asarray(ic["regime_id"], dtype=np.int32)
        for r in names:
            sl = self.sl[r]
            lay = sl.lay
            mask = reg0 == model.regime_code(r)
            if not mask.any():
                continue
            if lay.w_name:
                if lay.w_name not in ic:
                    raise InvalidInitialConditionsError(f"missing initial state {lay.w_name!r}")
                w0[mask] = ic[lay.w_name][mask]
            if sl.proc_name:
                if sl.proc_name not in ic:
                    raise InvalidInitialConditionsError(f"missing initial state {sl.proc_name!r}")
                z0[mask] = ic[sl.proc_name][mask]
            if sl.sdd_names:
                codes = [np.asarray(ic[s])[mask].astype(np.int64) for s in sl.sdd_names]
                sdd0[mask] = np.ravel_multi_index(codes, sl.sdd_sizes)
        nat.h2d(ctx, base + off["cur"]["w"], w0.astype(dt))
        nat.h2d(ctx, base + off["cur"]["z"], z0.astype(dt))
        nat.h2d(ctx, base + off["cur"]["sdd"], sdd0)
        nat.h2d(ctx, base + off["cur"]["regime"], reg0)
        nat.h2d(ctx, base + off["cur"]["start"], ic["_start"])
        nat.h2d(ctx, base + off["nxt"]["start"], ic["_start"])
        # -- the image, relocated ------------------------------------------------
        ev_base = b2 + _align(img["image"].nbytes)
        image = np.array(img["image"], dtype=np.uint8, copy=True)
        words = image.view(np.uint64)
        for kind, add in ((self.BASE_TABLES, b2), (self.BASE_EV, ev_base), (self.BASE_SOLVE, prog.base + prog.desc_off)):
            words[img["patch"][kind] // 8] += np.uint64(add)
        nat.h2d(ctx, b2, image)
        self._st = dict(T=T, n=n, R=R, seed=seed, base=base, off=off, b2=b2, ev_desc_off=img["ev_desc_off"],
                        sim_desc_off=img["sim_desc_off"], ev_keys=img["ev_keys"], lds_max=img["lds_max"],
                        sim_desc_size=img["sim_desc_size"])
        # agent panels are rewritten by `launch`; keep the initial state to restore
        self._init_state = [(off[side][k], arr) for side, k, arr in (
            ("cur", "w", w0.astype(dt)), ("cur", "z", z0.astype(dt)), ("cur", "sdd", sdd0),
            ("cur", "regime", reg0), ("cur", "start", ic["_start"]), ("nxt", "start", ic["_start"]))]
        self._launched = False
        self._graph = None

    def launch(self, events=None):
        """Enqueue the EV rows and the T period launches on the solve
        program's stream (no host sync). With `events` = (ev_pair, [pair per
        period]) of `nat.Event`, records them around the EV launch and each
        period."""
        self._reset()
        self._enqueue(events)

    def _reset(self):
        """A repeated run starts from the same initial panel (synchronous
        uploads, never part of a captured graph)."""
        if self._launched:
            base = self._st["base"]
            for o, arr in self._init_state:
                nat.h2d(self.prog.ctx, base + o, arr)
        self._launched = True

    def capture_graph(self):
        """The EV launch and the T period launches captured once as a HIP graph
        (kernel nodes only: the scratch counters are cleared by a kernel)."""
        if self._graph is None:
            self._graph = nat.Graph.capture(self.prog.ctx, self._enqueue)
        return self._graph

    def launch_graph(self):
        """`launch` as one graph launch (no host sync)."""
        self._reset()
        self.capture_graph().launch(self.prog.ctx)

    def _enqueue(self, events=None):
        st, model, dt, prog = self._st, self.model, self.dt, self.prog
        ctx = prog.ctx
        T, n, R, seed, base, off, b2 = (st[k] for k in ("T", "n", "R", "seed", "base", "off", "b2"))
        L = nat.lib()
        code = nat.LCM_DTYPE_F64 if dt == np.float64 else nat.LCM_DTYPE_F32
        # EV rows for every (period, regime) with a stateful continuation
        if events:
            events[0][0].record(ctx)
        if st["ev_keys"]:
            max_rows = max(model.layouts[r].n_sd * model.layouts[r].n_d for _, r in st["ev_keys"])
            nat._check(L.lcm_ev_launch(ctx.handle, code, C.c_void_p(b2 + st["ev_desc_off"]), len(st["ev_keys"]),
                                       max_rows), "lcm_ev_launch")
        if events:
            events[0][1].record(ctx)

        def agents(side):
            o = off[side]
            return nat.LcmAgents(base + o["w"], base + o["z"], base + o["sdd"], base + o["regime"], base + o["start"])

        A = [agents("cur"), agents("nxt")]
        O = nat.LcmSimOut(*(base + off["out"][k] for k in ("regime", "action", "value", "w", "z", "sdd")))
        for t in range(T):
            cur, nxt = A[t % 2], A[(t + 1) % 2]
            if events:
                events[1][t][0].record(ctx)
            descs = C.c_void_p(b2 + st["sim_desc_off"] + t * R * st["sim_desc_size"])
            if self.static_scratch:
                rc = L.lcm_sim_period_noalloc(ctx.handle, code, descs, R, C.byref(cur), C.byref(nxt), C.byref(O),
                                              n, t, C.c_uint64(seed), C.c_size_t(st["lds_max"]))
            else:
                rc = L.lcm_sim_period_ex(ctx.handle, code, descs, R, C.byref(cur), C.byref(nxt), C.byref(O), n, t,
                                         C.c_uint64(seed), C.c_size_t(st["lds_max"]),
                                         C.c_void_p(base + off["counts"]), C.c_void_p(base + off["order"]))
            nat._check(rc, f"lcm_sim_period (period {t})")
            if events:
                events[1][t][1].record(ctx)

    def collect_raw(self) -> dict:
        """The panel's [period, agent] arrays as the device wrote them (after
        the stream has finished): regime, action, sdd (int32), value, w, z."""
        st, dt, ctx = self._st, self.dt, self.prog.ctx
        T, n, base, off = st["T"], st["n"], st["base"], st["off"]

        def get(name, dtype_):
            arr = np.zeros(T * n, dtype=dtype_)
            nat.d2h(ctx, base + off["out"][name], arr)
            return arr.reshape(T, n)

        out = {k: get(k, np.int32) for k in ("regime", "action", "sdd")}
        out.update({k: get(k, dt) for k in ("value", "w", "z")})
        return out

    def collect(self, ic):
        """Download the panel (after the stream has finished) and build the result."""
        model = self.model
        names = list(model.regime_names)
        out = self.collect_raw()
        self.raw = out
        res = SimulationResult(self._frame(out, ic), model,
                               frozenset(r for r in names if model.layouts[r].terminal),
                               params=self.params, dtype=self.dt)
        res.raw = out  # [period, agent] arrays as the device wrote them
        return res

    def _gs_desc_index(self):
        """Index of each (period, regime) descriptor in the solve program."""
        model = self.model
        tabs = self.prog.problems[0]
        order = sorted(tabs.items(), key=lambda kv: (-kv[0][0], model.regime_code(kv[0][1])))
        return {k: i for i, (k, _) in enumerate(order)}

    def _desc(self, st: SimTables, sl: SimLayout, t, r, ar, zero, ev_meta):
        dt = self.dt
        lay = sl.lay
        model = self.model
        d = {"U0": ar.add(st.U0, dtype=dt), "lo": ar.add(st.lo, dtype=dt), "hi": ar.add(st.hi, dtype=dt)}
        d["A"] = ar.add(st.A, dtype=dt) if st.A is not None else zero
        d["B0"] = ar.add(st.B0, dtype=dt) if st.B0 is not None else zero
        d["Uf"] = [ar.add(u, dtype=dt) if u is not None else None for u in st.Uf]
        d["Bf"] = [ar.add(b, dtype=dt) if b is not None else None for b in st.Bf]
        nf = len(st.feats)
        d["LoB"] = ar.add(st.LoB, dtype=dt) if st.LoB is not None else None
        d["HiB"] = ar.add(st.HiB, dtype=dt) if st.HiB is not None else None
        d["LoF"] = [ar.add(st.LoF[k], dtype=dt) if k in st.LoF else None for k in range(nf)]
        d["HiF"] = [ar.add(st.HiF[k], dtype=dt) if k in st.HiF else None for k in range(nf)]
        d["lo_strict"], d["hi_strict"] = int(st.lo_strict), int(st.hi_strict)
        d["fcoef"] = [ar.add(c, dtype=dt) for _, c, _ in st.feats]
        d["fconst"] = [ar.add(c, dtype=dt) for _, _, c in st.feats]
        d["feat_op"] = [op for op, _, _ in st.feats]
        d["n_feat"] = len(st.feats)
        d["n_sdd"], d["n_d"], d["n_c"] = st.n_sdd, st.n_d, st.n_c
        d["terminal"] = int(st.terminal)
        d["regime_code"] = model.regime_code(r)
        d["n_regimes"] = len(model.regime_names)
        d["n_rows"] = lay.n_sd
        d["EV"] = d["CS"] = d["ev_fin"] = None
        d["sd_base"] = ar.add(sl.sd_base)
        d["nodes"] = zero
        d["g_pts"] = zero
        d["preg"] = None
        d["interp"] = 0
        d["n_nodes"] = 0
        d["p_stride"] = sl.p_stride
        d["proc_kind"] = 0
        d["pa"] = d["pb"] = d["pc"] = 0.0
        d["beta"] = 0.0
        d["g_kind"], d["g_n"], d["g_start"], d["g_stop"] = 0, 0, 0.0, 0.0
        if sl.proc_name:
            nodes, _ = self.rg.proc[(r, sl.proc_name)]
            d["nodes"] = ar.add(nodes, dtype=dt)
            d["n_nodes"] = len(nodes)
            proc = lay.sd_grids[sl.proc_idx[0]]
            from pylcm_amd._build import process_runtime_params

            kind, a, b, c = draw_params(proc, process_runtime_params(self.params, model.fixed_params, r,
                                                                     sl.proc_name))
            d["proc_kind"], d["pa"], d["pb"], d["pc"] = kind, a, b, c
        d["taste_scale"] = 0.0
        if not st.terminal:
            pt = self.prog.problems[0][(t, r)]
            d["beta"] = pt.beta
            d["taste_scale"] = pt.taste_scale or 0.0
            d["preg"] = ar.add(st.preg, dtype=dt)
            d["Wd"] = [None] * nat.SIM_MAX_REGIMES
            d["n_sdd_next"] = [0] * nat.SIM_MAX_REGIMES
            for tr_name, W in st.Wd.items():
                code = model.regime_code(tr_name)
                d["Wd"][code] = ar.add(W, dtype=dt)
                d["n_sdd_next"][code] = W.shape[-1]
            e_off, c_off, g_n, f_off = ev_meta[(t, r)]
            d["EV"], d["CS"], d["ev_fin"] = e_off, c_off, f_off
            if g_n > 0:
                d["interp"] = 1
                glay = model.layouts[pt.next_grid]
                d["g_kind"], d["g_n"] = int(glay.w_grid.kind), g_n
                d["g_start"], d["g_stop"] = float(glay.w_grid.start), float(glay.w_grid.stop)
                d["g_pts"] = ar.add(self.rg.w_pts[pt.next_grid], dtype=dt)
        return d

    def _frame(self, out, ic):
        model = self.model
        T, n = out["regime"].shape
        names = list(model.regime_names)
        simulated = out["action"] >= 0
        tt, ii = np.nonzero(simulated)
        df = {"subject_id": ii.astype(np.int64), "period": tt.astype(np.int64)}
        reg = out["regime"][tt, ii]
        df["regime_name"] = np.array(names, dtype=object)[reg]
        df["value"] = out["value"][tt, ii].astype(np.float64)
        cols = {}
        for r in names:
            sl = self.sl[r]
            lay = sl.lay
            m = reg == model.regime_code(r)
            if not m.any():
                continue
            if lay.w_name:
                cols.setdefault(lay.w_name, np.full(len(tt), np.nan))[m] = out["w"][tt, ii][m]
            if sl.proc_name:
                cols.setdefault(sl.proc_name, np.full(len(tt), np.nan))[m] = out["z"][tt, ii][m]
            if sl.sdd_names:
                codes = np.unravel_index(out["sdd"][tt, ii][m], sl.sdd_sizes)
                for s, cdx in zip(sl.sdd_names, codes):
                    cols.setdefault(s, np.full(len(tt), np.nan))[m] = cdx
            a = out["action"][tt, ii][m]
            d_idx, c_idx = np.divmod(a, lay.n_c)
            if lay.c_name:
                cols.setdefault(lay.c_name, np.full(len(tt), np.nan))[m] = self.rg.c_pts[r][c_idx]
            if lay.da_names:
                codes = np.unravel_index(d_idx, lay.da_sizes)
                for an, cdx in zip(lay.da_names, codes):
                    cols.setdefault(an, np.full(len(tt), np.nan))[m] = cdx
        state_names = []
        action_names = []
        for r in names:
            lay = model.layouts[r]
            for s in lay.state_names:
                if s not in state_names:
                    state_names.append(s)
            for a in lay.action_names:
                if a not in action_names:
                    action_names.append(a)
        for c in state_names + action_names:
            df[c] = cols.get(c, np.full(len(tt), np.nan))
        ages = np.asarray([float(model.ages.period_to_age(p)) for p in range(T)])
        df["age"] = ages[tt]
        frame = pd.DataFrame(df)
        return frame.sort_values(["subject_id", "period"]).reset_index(drop=True)
