"""FP64 oracle of the grid-search solve (TEST INFRASTRUCTURE ONLY).

Restates, in NumPy over the full state x action product, the reference's
per-period computation:

* grid points: `jnp.linspace` (`src/_lcm/grids/coordinates.py:15-27`),
  log grids (`coordinates.py:47-70`), irregular points;
* coordinates: `get_linspace_coordinate` (`coordinates.py:30-44`),
  `get_logspace_coordinate` (`:73-114`), `get_irreg_coordinate` (`:117-167`);
* `map_coordinates` with linear extrapolation and zero-safe corners
  (`src/_lcm/regime_building/ndimage.py:28-109`, `src/_lcm/zero_safe.py:161-251`);
* processes: Rouwenhorst by the Kopecky-Suen binomial sum
  (`src/_lcm/processes/ar1.py:157-189`), Tauchen (`ar1.py:59-99`), Gauss-Hermite
  normal/lognormal (`src/_lcm/processes/base.py:16-33`, `iid.py:71-210`),
  uniform (`iid.py:34-69`); interpolated transition rows
  (`src/_lcm/regime_building/processing.py:2780-2840`);
* `Q_and_F` (`src/_lcm/regime_building/Q_and_F.py:58-170`) with
  `compute_CE` (`:466-681`): per target regime the node expectation
  `sum w v / sum w` (`:1145-1209`), `CE = sum_r p_r E_r / mass` with NaN when the
  mass is off by more than 1e-3 or a probability is negative (`:1398-1432`);
* `Q = U + beta * CE` (`src/lcm/koopmans_aggregation.py:84-98`);
* `V = max(Q, where=F, initial=-inf)` (`max_Q_over_a.py:152-160`), the first
  argmax over the flattened action axes (`argmax.py:6-69`), and with taste
  shocks the logsum over discrete actions (`max_Q_over_a.py:120-138`,
  `src/_lcm/logsum.py:32-64`).

Unlike the product, the oracle evaluates the expectation node by node, exactly
in the reference's order (interpolate each node, then weight), on concrete
arrays; it never contracts V first. All arithmetic is float64; grid points and
float parameters are first rounded to the model dtype so the oracle solves the
same discretised model as a float32 solve, in exact-as-possible arithmetic.

Parity pinning (see tests/test_oracle_*.py): closed forms for the last decision
period and the cake-eating model, brute force over tiny grids with a
hand-written loop, scipy's `map_coordinates(order=1)` inside the grid, the
Kopecky-Suen matrix against the product's independent recursion, the
reference's analytical-solution CSV fixtures (Iskhakov et al. 2017), and the
reference's own argmax tie rule.
"""

from __future__ import annotations

import inspect
from math import comb, erf, sqrt

import numpy as np

# ----------------------------------------------------------------------------
# grids and coordinates
# ----------------------------------------------------------------------------


def linspace(start, stop, n, dtype=np.float64):
    """`jnp.linspace(start, stop, n)` in `dtype` (jax: start*(1-s) + stop*s)."""
    dt = np.dtype(dtype)
    if n == 1:
        return np.array([start], dtype=dt)
    div = n - 1
    s = (np.arange(div, dtype=dt) / dt.type(div)).astype(dt)
    out = dt.type(start) * (dt.type(1) - s) + dt.type(stop) * s
    return np.append(out, dt.type(stop)).astype(dt)


def grid_points(grid, dtype=np.float64, runtime=None):
    name = type(grid).__name__
    if name == "LinSpacedGrid":
        return linspace(grid.start, grid.stop, grid.n_points, dtype).astype(np.float64)
    if name == "LogSpacedGrid":
        dt = np.dtype(dtype)
        pts = np.exp(linspace(np.log(dt.type(grid.start)), np.log(dt.type(grid.stop)),
                              grid.n_points, dt)).astype(dt)
        pts[0], pts[-1] = dt.type(grid.start), dt.type(grid.stop)
        return pts.astype(np.float64)
    if name == "IrregSpacedGrid":
        return np.asarray(grid.points, dtype=dtype).astype(np.float64)
    if name == "DiscreteGrid":
        return np.arange(len(grid.category_class.__categories__), dtype=np.float64)
    nodes, _ = process_nodes_matrix(grid, runtime)
    return nodes.astype(dtype).astype(np.float64)


def lin_coordinate(value, start, stop, n):
    """`get_linspace_coordinate` (coordinates.py:30-44)."""
    n_steps = n - 1
    if n_steps <= 0:
        return np.zeros_like(value)
    step = (stop - start) / n_steps
    return (value - start) / step


def log_coordinate(value, start, stop, n):
    """`get_logspace_coordinate` (coordinates.py:73-114)."""
    sl, el = np.log(start), np.log(stop)
    cl = lin_coordinate(np.log(value), sl, el, n)
    rlo = np.floor(cl)
    step = (el - sl) / (n - 1)
    lower = np.exp(sl + step * rlo)
    upper = np.exp(sl + step * (rlo + 1))
    return rlo + (value - lower) / (upper - lower)


def irreg_coordinate(value, points):
    """`get_irreg_coordinate` (coordinates.py:117-167)."""
    n = len(points)
    if n == 1:
        return np.zeros_like(value)
    hi = np.clip(np.searchsorted(points, value, side="right"), 1, n - 1)
    lo = hi - 1
    return lo + (value - points[lo]) / (points[hi] - points[lo])


def coordinate(grid, pts, value):
    name = type(grid).__name__
    if name == "LinSpacedGrid":
        return lin_coordinate(value, np.float64(pts[0]), np.float64(pts[-1]), len(pts))
    if name == "LogSpacedGrid":
        return log_coordinate(value, np.float64(pts[0]), np.float64(pts[-1]), len(pts))
    return irreg_coordinate(value, pts)


def zero_safe(weight, value):
    """`w * v` with `v` replaced by 0 where `w == 0` and `v` is not finite."""
    safe_v = np.where((weight == 0) & ~np.isfinite(value), 0.0, value)
    return weight * safe_v


def interp_1d(values, coord):
    """`map_coordinates` on one axis (ndimage.py:28-109): clip lower index to
    [0, n-2], weights (1-t, t), zero-safe corners, lower term first."""
    n = values.shape[-1]
    if n == 1:
        return np.broadcast_to(values[..., 0], np.shape(coord)).astype(np.float64)
    with np.errstate(invalid="ignore"):
        lo = np.clip(np.floor(coord), 0, n - 2)
    lo_i = np.where(np.isfinite(lo), lo, 0).astype(np.int64)
    t = coord - lo
    v0 = np.take(values, lo_i)
    v1 = np.take(values, lo_i + 1)
    return zero_safe(1 - t, v0) + zero_safe(t, v1)


# ----------------------------------------------------------------------------
# processes
# ----------------------------------------------------------------------------


def _ncdf(x):
    return 0.5 * (1.0 + np.vectorize(erf)(np.asarray(x, dtype=np.float64) / sqrt(2.0)))


def rouwenhorst_kopecky_suen(n, rho):
    """P[i,j] = sum_k C(i,k) C(n-1-i,j-k) q^(n-1-i-j+2k) (1-q)^(i+j-2k) (ar1.py:164-189)."""
    q = (rho + 1) / 2
    P = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            s = 0.0
            for k in range(max(0, i + j - n + 1), min(i, j) + 1):
                s += comb(i, k) * comb(n - 1 - i, j - k) * q ** (n - 1 - i - j + 2 * k) * (1 - q) ** (i + j - 2 * k)
            P[i, j] = s
    return P


def _proc_params(proc, runtime):
    runtime = runtime or {}
    fields = ("rho", "sigma", "mu", "n_std", "start", "stop")
    out = {}
    for f in fields:
        if hasattr(proc, f):
            v = getattr(proc, f)
            out[f] = runtime.get(f) if v is None else v
    return out


def process_nodes_matrix(proc, runtime=None):
    name = type(proc).__name__
    p = _proc_params(proc, runtime)
    n = proc.n_points
    if name == "RouwenhorstAR1Process":
        nu = np.sqrt((n - 1) / (1 - p["rho"] ** 2)) * p["sigma"]
        m = p["mu"] / (1 - p["rho"])
        return linspace(m - nu, m + nu, n), rouwenhorst_kopecky_suen(n, p["rho"])
    if name == "TauchenAR1Process":
        std_y = np.sqrt(p["sigma"] ** 2 / (1 - p["rho"] ** 2))
        if proc.gauss_hermite:
            x, _ = np.polynomial.hermite.hermgauss(n)
            c = std_y * np.sqrt(2) * x
        else:
            c = linspace(-p["n_std"] * std_y, p["n_std"] * std_y, n)
        mid = (c[:-1] + c[1:]) / 2
        cdf = _ncdf((mid[None, :] - p["rho"] * c[:, None]) / p["sigma"])
        P = np.concatenate([cdf[:, :1], np.diff(cdf, axis=1), 1 - cdf[:, -1:]], axis=1)
        return c + p["mu"] / (1 - p["rho"]), P
    if name in ("NormalIIDProcess", "LogNormalIIDProcess"):
        if proc.gauss_hermite:
            x, w = np.polynomial.hermite.hermgauss(n)
            nodes = p["mu"] + p["sigma"] * np.sqrt(2) * x
            w = w / np.sqrt(np.pi)
        else:
            lo = p["mu"] - p["n_std"] * p["sigma"]
            hi = p["mu"] + p["n_std"] * p["sigma"]
            nodes = linspace(lo, hi, n)
            step = (hi - lo) / (n - 1)
            w = np.zeros(n)
            w[1:] = np.diff(_ncdf((nodes + 0.5 * step - p["mu"]) / p["sigma"]))
            w[0] = _ncdf((lo + 0.5 * step - p["mu"]) / p["sigma"])
            w[-1] = 1 - _ncdf((hi - 0.5 * step - p["mu"]) / p["sigma"])
        if name == "LogNormalIIDProcess":
            nodes = np.exp(nodes)
        return nodes, np.tile(w, (n, 1))
    if name == "UniformIIDProcess":
        return linspace(p["start"], p["stop"], n), np.full((n, n), 1.0 / n)
    raise NotImplementedError(name)


def transition_row(P, nodes, value):
    """Interpolated transition row at `value` (processing.py:2780-2840)."""
    n = len(nodes)
    coord = irreg_coordinate(np.float64(value), nodes)
    rows = np.stack([interp_1d(P[:, j], np.float64(coord)) for j in range(n)])
    return rows


# ----------------------------------------------------------------------------
# function evaluation (dags-like name resolution)
# ----------------------------------------------------------------------------


def _unwrap(f):
    return getattr(f, "func", f) if type(f).__name__ == "MarkovTransition" else f


def _lookup(src, regime, fname, arg):
    if not src:
        return None
    r = src.get(regime)
    if isinstance(r, dict):
        f = r.get(fname)
        if isinstance(f, dict) and arg in f:
            return (f[arg],)
        if arg in r and not isinstance(r[arg], dict):
            return (r[arg],)
    if arg in src and not isinstance(src[arg], dict):
        return (src[arg],)
    return None


class Env:
    def __init__(self, regime_name, functions, variables, params, fixed, pdtype):
        self.rn, self.fn, self.vars = regime_name, functions, variables
        self.params, self.fixed, self.pdtype = params or {}, fixed or {}, np.dtype(pdtype)
        self.memo = {}

    def param(self, fname, arg):
        for src in (self.params, self.fixed):
            hit = _lookup(src, self.rn, fname, arg)
            if hit is not None:
                v = hit[0]
                if isinstance(v, float):
                    return float(np.asarray(v, dtype=self.pdtype))
                if isinstance(v, np.ndarray) and v.dtype.kind == "f":
                    return v.astype(self.pdtype).astype(np.float64)
                return v
        return None

    def call(self, fname, func=None):
        if func is None and fname in self.memo:
            return self.memo[fname]
        f = _unwrap(func if func is not None else self.fn[fname])
        kw = {}
        for p in inspect.signature(f).parameters.values():
            a = p.name
            if a in self.vars:
                kw[a] = self.vars[a]
            elif a in self.fn and a != fname:
                kw[a] = self.call(a)
            else:
                v = self.param(fname, a)
                if v is None:
                    if p.default is inspect.Parameter.empty:
                        raise KeyError(f"cannot resolve {a!r} of {fname!r}")
                    v = p.default
                kw[a] = v
        out = f(**kw)
        if func is None:
            self.memo[fname] = out
        return out


# ----------------------------------------------------------------------------
# regime structure
# ----------------------------------------------------------------------------


class _R:
    """Axes of one regime: discrete/process states, continuous state,
    discrete actions, continuous action (the reference's V/argmax layout)."""

    def __init__(self, name, regime, gdt, params, fixed):
        self.name, self.regime = name, regime
        self.sd, self.w, self.da, self.c = [], None, [], None
        for s, g in regime.states.items():
            k = type(g).__name__
            if k in ("LinSpacedGrid", "LogSpacedGrid", "IrregSpacedGrid"):
                self.w = (s, g)
            else:
                self.sd.append((s, g))
        for a, g in regime.actions.items():
            if type(g).__name__ == "DiscreteGrid":
                self.da.append((a, g))
            else:
                self.c = (a, g)
        self.sd_vals, self.proc = [], {}
        for s, g in self.sd:
            if type(g).__name__ == "DiscreteGrid":
                self.sd_vals.append(grid_points(g))
            else:
                rt = _proc_runtime(params, fixed, name, s)
                nodes, P = process_nodes_matrix(g, rt)
                nodes = nodes.astype(gdt).astype(np.float64)
                self.proc[s] = (nodes, P)
                self.sd_vals.append(nodes)
        self.w_pts = grid_points(self.w[1], gdt) if self.w else None
        self.c_pts = grid_points(self.c[1], gdt) if self.c else None
        self.da_vals = [grid_points(g) for _, g in self.da]

    @property
    def V_shape(self):
        return tuple(len(v) for v in self.sd_vals) + ((len(self.w_pts),) if self.w else ())

    @property
    def terminal(self):
        return self.regime.transition is None

    def full_vars(self):
        """Variables on the broadcast grid [sd..., w, da..., c]."""
        nsd, nda = len(self.sd), len(self.da)
        ndim = nsd + 1 + nda + 1
        out = {}

        def ax(v, i):
            sh = [1] * ndim
            sh[i] = len(v)
            return np.asarray(v).reshape(sh)

        for i, ((s, g), v) in enumerate(zip(self.sd, self.sd_vals)):
            out[s] = ax(v.astype(np.int64) if type(g).__name__ == "DiscreteGrid" else v, i)
        if self.w:
            out[self.w[0]] = ax(self.w_pts, nsd)
        for j, ((a, _), v) in enumerate(zip(self.da, self.da_vals)):
            out[a] = ax(v.astype(np.int64), nsd + 1 + j)
        if self.c:
            out[self.c[0]] = ax(self.c_pts, ndim - 1)
        return out, ndim

    def full_shape(self):
        return (tuple(len(v) for v in self.sd_vals) + ((len(self.w_pts),) if self.w else (1,))
                + tuple(len(v) for v in self.da_vals) + ((len(self.c_pts),) if self.c else (1,)))


def _proc_runtime(params, fixed, regime, state):
    out = {}
    for src in (fixed or {}, params or {}):
        r = src.get(regime)
        if isinstance(r, dict) and isinstance(r.get(state), dict):
            out.update(r[state])
    return out


# ----------------------------------------------------------------------------
# solve
# ----------------------------------------------------------------------------


def _active(regime, age):
    return bool(regime.active(age))


def solve(model, params, *, grid_dtype=np.float64, return_Q=False):
    """Backward induction. Returns (V, policy): period -> regime -> array.

    `policy` is the first argmax over the flattened (discrete..., continuous)
    action axes, or with taste shocks the continuous argmax per discrete-action
    combination (flat index), matching the product's output convention.
    """
    gdt = np.dtype(grid_dtype)
    fixed = model.fixed_params
    names = list(model.regime_id_class.__categories__)
    R = {n: _R(n, model.regimes[n], gdt, params, fixed) for n in names}
    T = model.ages.n_periods
    V, POL, QS = {}, {}, {}
    for t in reversed(range(T)):
        age = model.ages.period_to_age(t)
        V[t], POL[t] = {}, {}
        for n in names:
            reg = R[n]
            if not _active(reg.regime, age):
                continue
            Q, F, taste = _q_and_f(model, R, reg, t, params, fixed, gdt, V.get(t + 1))
            if return_Q:
                QS[(t, n)] = (Q, F)
            Vt, pol = _max_over_actions(Q, F, reg, taste)
            V[t][n] = Vt.reshape(reg.V_shape)
            POL[t][n] = pol.reshape(reg.V_shape + ((pol.shape[-1],) if taste is not None else ()))
    if return_Q:
        return V, POL, QS
    return V, POL


def _env(model, reg, t, params, fixed, gdt, variables):
    v = dict(variables)
    v["age"] = model.ages.period_to_age(t)
    v["period"] = t
    funcs = {k: f for k, f in reg.regime.functions.items() if f is not None}
    return Env(reg.name, funcs, v, params, fixed, gdt)


def _pointwise_points(reg):
    sd_sizes = [len(v) for v in reg.sd_vals]
    da_sizes = [len(v) for v in reg.da_vals]
    for sdi in np.ndindex(*sd_sizes) if sd_sizes else [()]:
        for dai in np.ndindex(*da_sizes) if da_sizes else [()]:
            v = {}
            for k, (s, g) in enumerate(reg.sd):
                val = reg.sd_vals[k][sdi[k]]
                v[s] = int(val) if type(g).__name__ == "DiscreteGrid" else float(val)
            for k, (a, _) in enumerate(reg.da):
                v[a] = int(dai[k])
            yield sdi, dai, v


def tabulate(model, R, reg, t, params, fixed, gdt, Vnext):
    """Evaluate the user's functions of one (regime, period) on the full
    state x action grid: utility, feasibility, next continuous state, and per
    target regime its probability and node weights (Q_and_F.py:466-681)."""
    variables, ndim = reg.full_vars()
    shape = reg.full_shape()
    ev = _env(model, reg, t, params, fixed, gdt, variables)
    tab = {"shape": shape, "terminal": reg.terminal, "targets": [], "scalars": []}
    tab["U"] = np.broadcast_to(np.asarray(ev.call("utility"), dtype=np.float64), shape)
    F = np.ones(shape, dtype=bool)
    for cname, cf in reg.regime.constraints.items():
        F = F & np.broadcast_to(np.asarray(ev.call(cname, cf), dtype=bool), shape)
    tab["F"] = F
    tab["taste"] = None
    if reg.regime.taste_shocks is not None:
        tab["taste"] = float(params[reg.name]["taste_shocks"]["scale"])
    if reg.terminal:
        return tab
    tab["beta"] = ev.param("koopmans_aggregator", "discount_factor")
    names = list(model.regime_id_class.__categories__)
    age_next = model.ages.period_to_age(t + 1)
    active_next = [n for n in names if _active(R[n].regime, age_next)]
    sd_sizes = tuple(len(v) for v in reg.sd_vals)
    da_sizes = tuple(len(v) for v in reg.da_vals)
    # regime transition probabilities per (sd, da) point
    P = np.zeros(sd_sizes + da_sizes + (len(names),))
    tr = reg.regime.transition
    for sdi, dai, v in _pointwise_points(reg):
        e = _env(model, reg, t, params, fixed, gdt, v)
        if hasattr(tr, "items"):
            for tn, f in tr.items():
                P[sdi + dai + (names.index(tn),)] = float(np.asarray(e.call("next_regime", f)))
        elif type(tr).__name__ == "MarkovTransition":
            P[sdi + dai] = np.asarray(e.call("next_regime", tr), dtype=np.float64)
        else:
            P[sdi + dai + (int(np.asarray(e.call("next_regime", tr))),)] = 1.0
    for n in active_next:
        tgt = R[n]
        p = P[..., names.index(n)]
        if tgt.V_shape == ():
            tab["scalars"].append((n, p, np.float64(Vnext[n])))
            continue
        law = reg.regime.state_transitions[tgt.w[0]]
        wn = np.broadcast_to(np.asarray(ev.call(f"next_{tgt.w[0]}", law), dtype=np.float64), shape)
        tab["wn"] = wn
        tab["grid"] = (tgt.w[1], tgt.w_pts)
        # joint weights over the target's discrete/process axes, per (sd, da)
        axes_w = []
        for (s, g) in tgt.sd:
            n_pts = len(tgt.sd_vals[[x for x, _ in tgt.sd].index(s)])
            Wax = np.zeros(sd_sizes + da_sizes + (n_pts,))
            for sdi, dai, v in _pointwise_points(reg):
                if type(g).__name__ == "DiscreteGrid":
                    law_s = reg.regime.state_transitions[s]
                    if hasattr(law_s, "items"):
                        law_s = law_s[n]
                    e = _env(model, reg, t, params, fixed, gdt, v)
                    out = np.asarray(e.call(f"next_{s}", law_s), dtype=np.float64)
                    if type(law_s).__name__ == "MarkovTransition":
                        Wax[sdi + dai] = out
                    else:
                        Wax[sdi + dai + (int(out),)] = 1.0
                else:
                    nodes_t, P_t = tgt.proc[s]
                    Wax[sdi + dai] = transition_row(P_t, nodes_t, v[s])
            axes_w.append(Wax)
        node_index = list(np.ndindex(*[a.shape[-1] for a in axes_w])) if axes_w else [()]
        W = np.ones(sd_sizes + da_sizes + (len(node_index),))
        for flat, idx in enumerate(node_index):
            for ax_i, k in enumerate(idx):
                W[..., flat] = W[..., flat] * axes_w[ax_i][..., k]
        Vn = np.asarray(Vnext[n], dtype=np.float64).reshape(-1, tgt.V_shape[-1])
        tab["targets"].append((n, p, W, Vn))
    return tab


def _q_and_f(model, R, reg, t, params, fixed, gdt, Vnext):
    tab = tabulate(model, R, reg, t, params, fixed, gdt, Vnext)
    U, F, shape = tab["U"], tab["F"], tab["shape"]
    if tab["terminal"]:
        return U, F, tab["taste"]
    sd_sizes = tuple(len(v) for v in reg.sd_vals)
    da_sizes = tuple(len(v) for v in reg.da_vals)

    def on_grid(arr_sd_da):
        """(sd..., da...) -> broadcast over [sd..., w, da..., c]."""
        a = np.asarray(arr_sd_da).reshape(sd_sizes + (1,) + da_sizes + (1,))
        return np.broadcast_to(a, shape)

    CE = np.zeros(shape)
    mass = np.zeros(shape)
    neg = np.zeros(shape, dtype=bool)
    for n, p, v in tab["scalars"]:
        pg = on_grid(p)
        mass = mass + pg
        neg = neg | (np.signbit(pg) & (pg != 0))
        CE = CE + zero_safe(pg, v)
    for n, p, W, Vn in tab["targets"]:
        pg = on_grid(p)
        mass = mass + pg
        neg = neg | (np.signbit(pg) & (pg != 0))
        grid, pts = tab["grid"]
        coord = coordinate(grid, pts, tab["wn"])
        num = np.zeros(shape)
        wsum = np.zeros(shape)
        for k in range(W.shape[-1]):
            wk = on_grid(W[..., k])
            num = num + zero_safe(wk, interp_1d(Vn[k], coord))
            wsum = wsum + wk
        E = num / np.where(wsum > 0, wsum, 1.0)
        CE = CE + zero_safe(pg, E)
    ok = (np.abs(mass - 1.0) <= 1e-3) & ~neg
    CE = CE / np.where(ok, mass, np.nan)
    Q = U + tab["beta"] * CE
    return Q, F, tab["taste"]


def _max_over_actions(Q, F, reg, taste):
    nsd = len(reg.sd)
    nda = len(reg.da)
    n_state_axes = nsd + 1
    st_shape = Q.shape[:n_state_axes]
    n_states = int(np.prod(st_shape))
    n_d = int(np.prod(Q.shape[n_state_axes:n_state_axes + nda])) if nda else 1
    n_c = Q.shape[-1]
    Qf = Q.reshape(n_states, n_d, n_c)
    Ff = F.reshape(n_states, n_d, n_c)
    if taste is None:
        Qa = Qf.reshape(n_states, -1)
        Fa = Ff.reshape(n_states, -1)
        m = np.max(np.where(Fa, Qa, -np.inf), axis=1)
        # NaN at a feasible entry propagates (jnp.max)
        nan = np.any(Fa & np.isnan(Qa), axis=1)
        m = np.where(nan, np.nan, m)
        mask = (Qa == m[:, None]) & Fa
        pol = np.argmax(mask, axis=1).astype(np.int32)
        return m, pol
    Qm = np.where(Ff, Qf, -np.inf)
    qc = np.max(Qm, axis=2)
    nanc = np.any(Ff & np.isnan(Qf), axis=2)
    qc = np.where(nanc, np.nan, qc)
    mask = (Qf == qc[:, :, None]) & Ff
    cidx = np.argmax(mask, axis=2)
    pol = (np.arange(n_d)[None, :] * n_c + cidx).astype(np.int32)
    # a discrete action with no feasible continuous choice (or a NaN max) has policy 0
    pol = np.where(np.isnan(qc) | ~np.any(Ff, axis=2), 0, pol).astype(np.int32)
    vmax = np.max(qc, axis=1)
    fm = np.where(np.isneginf(vmax), 0.0, vmax)
    with np.errstate(divide="ignore", invalid="ignore"):
        sh = (qc - fm[:, None]) / taste
        s = np.sum(np.exp(sh), axis=1)
        V = fm + taste * np.log(s)
    V = np.where(np.any(np.isnan(qc), axis=1), np.nan, V)
    return V, pol


# ----------------------------------------------------------------------------
# the reference's arithmetic in a lower precision (for grading argmax ties)
# ----------------------------------------------------------------------------


def reference_form_q(tab, reg, states, compute_dtype=np.float32):
    """Q and F of one (regime, period) at flat states `states` (index
    `sd * n_w + w`), every arithmetic step rounded to `compute_dtype`, in the
    reference's order of operations (`Q_and_F.py:466-681,1145-1209`): per
    target and node the coordinate of the next state, the interpolation
    `(1 - t) * V[lo] + t * V[lo + 1]` with zero-safe corners
    (`ndimage.py:73-109`), the node-weighted sum divided by the weight sum,
    the probability-weighted blend over targets divided by the mass, and
    `Q = U + beta * CE`.

    `tab` is `tabulate(...)` of the same (regime, period); the user's
    functions enter as their float64 values rounded to `compute_dtype`. This
    is how the reference's float32 solve evaluates Q, up to its compiler's
    choice of fused multiply-adds; argmax ties whose float64 gap exceeds one
    float32 ulp are graded against it."""
    cdt = np.dtype(compute_dtype)
    c = cdt.type
    shape = tab["shape"]
    n_sdax = len(reg.sd)
    nsd = int(np.prod(shape[:n_sdax])) if n_sdax else 1
    nw = shape[n_sdax]
    nd = int(np.prod(shape[n_sdax + 1:-1])) if len(shape) > n_sdax + 2 else 1
    nc = shape[-1]
    states = np.asarray(states, dtype=np.int64)
    sd_i, w_i = states // nw, states % nw

    def at_states(a):
        return np.broadcast_to(a, shape).reshape(nsd, nw, nd, nc)[sd_i, w_i]

    def per_sd_da(p):
        return np.asarray(p, dtype=np.float64).reshape(nsd, nd)[sd_i][:, :, None].astype(cdt)

    def zs(w, v):
        return w * np.where((w == c(0)) & ~np.isfinite(v), c(0), v)

    U = at_states(tab["U"]).astype(cdt)
    F = at_states(tab["F"])
    if tab["terminal"]:
        return U, F
    CE = np.zeros(U.shape, dtype=cdt)
    mass = np.zeros(U.shape, dtype=cdt)
    neg = np.zeros(U.shape, dtype=bool)
    for _, p, v in tab["scalars"]:
        pg = per_sd_da(p)
        mass = mass + pg
        neg |= np.signbit(pg) & (pg != c(0))
        CE = CE + zs(pg, c(v))
    if tab["targets"]:
        grid, pts = tab["grid"]
        pts = pts.astype(cdt)
        x = at_states(tab["wn"]).astype(cdt)
        n = len(pts)
        kind = type(grid).__name__
        if n == 1:
            coord = np.zeros_like(x)
        elif kind == "LinSpacedGrid":
            step = (pts[-1] - pts[0]) / c(n - 1)
            coord = (x - pts[0]) / step
        elif kind == "LogSpacedGrid":
            sl, el = np.log(pts[0]), np.log(pts[-1])
            step = (el - sl) / c(n - 1)
            cl = (np.log(x) - sl) / step
            rlo = np.floor(cl)
            lower, upper = np.exp(sl + step * rlo), np.exp(sl + step * (rlo + c(1)))
            coord = rlo + (x - lower) / (upper - lower)
        else:
            hi = np.clip(np.searchsorted(pts, x, side="right"), 1, n - 1)
            lo_ = hi - 1
            coord = lo_.astype(cdt) + (x - pts[lo_]) / (pts[hi] - pts[lo_])
        with np.errstate(invalid="ignore"):
            lo = np.clip(np.floor(coord), c(0), c(max(n - 2, 0)))
        lo_i = np.where(np.isfinite(lo), lo, 0).astype(np.int64)
        t = coord - lo
    for _, p, W, Vn in tab["targets"]:
        pg = per_sd_da(p)
        mass = mass + pg
        neg |= np.signbit(pg) & (pg != c(0))
        Wk = np.asarray(W, dtype=np.float64).reshape(nsd, nd, -1)[sd_i].astype(cdt)
        num = np.zeros(U.shape, dtype=cdt)
        wsum = np.zeros(U.shape, dtype=cdt)
        for k in range(Wk.shape[-1]):
            row = np.asarray(Vn[k]).astype(cdt)
            if len(row) == 1:
                v = np.broadcast_to(row[0], U.shape)
            else:
                v = zs(c(1) - t, row[lo_i]) + zs(t, row[lo_i + 1])
            wk = Wk[:, :, k][:, :, None]
            num = num + zs(wk, v)
            wsum = wsum + wk
        E = num / np.where(wsum > c(0), wsum, c(1))
        CE = CE + zs(pg, E)
    ok = (np.abs(mass.astype(np.float64) - 1.0) <= 1e-3) & ~neg
    CE = CE / np.where(ok, mass, c(np.nan))
    return U + c(tab["beta"]) * CE, F


def first_argmax(Q, F):
    """First maximiser over the flattened action axes of each state row
    (`argmax.py:6-69`); -1 where no action is feasible."""
    k = Q.shape[0]
    Qa = np.where(F, Q, -np.inf).reshape(k, -1)
    m = np.max(Qa, axis=1)
    pol = np.argmax(Qa == m[:, None], axis=1)
    return np.where(np.isfinite(m), pol, -1)
