"""FP64 oracle of the DC-EGM solve with the FUES upper envelope (TEST
INFRASTRUCTURE ONLY: only tests/, smoke() and bench.py's CPU leg import it).

Restates, in NumPy, the reference's per-period DC-EGM step for a regime whose
liquid margin is (Euler state = resources, consumption, savings = resources -
consumption):

* the continuation at each savings node A_j: per reachable target regime, per
  child stochastic node, per child discrete-action row, a Hermite read of the
  child's value row and a linear read of its marginal row at R'(A_j), the
  marginal times dR'/dA (`src/_lcm/egm/continuation.py:1353-1534`); the child's
  rows aggregated by logsum/softmax or the hard max with a one-hot first argmax
  (`continuation.py:1317-1351`, `src/_lcm/logsum.py:32-64`); the node weights
  applied zero-safely without renormalisation (`continuation.py:959-1140`); the
  targets blended with `where(p > 0, p x, p 0)` and stateless targets adding
  `p V` to the value only (`continuation.py:473-498`);
* the Euler inversion `c = (u')^{-1}(max(beta E[mu'], eps))`, NaN for a negative
  target (`src/_lcm/egm/euler.py:18-68`), `R = A + c`, `v = u(c) + beta E[V']`
  (`src/_lcm/egm/step_core.py:397-409`);
* the closed-form constrained candidates (`step_core.py:443-514`), the candidate
  assembly with exogenous source savings and the `-inf` -> NaN masking
  (`step_core.py:218-266`);
* FUES: the NaN-stable lexsort, the coincident-abscissa reduction and the serial
  scan with exhaustive same-segment searches, crossing insertion and endpoint
  snapping (`src/_lcm/egm/upper_envelope/fues.py:151-424,549-1189`);
* the publish: marginal utility `u'(policy)`, the Hermite read of the refined row
  at the publish resources, the constrained floor and the overflow poison
  (`step_core.py:517-616`), the NaN-padded interpolation rules
  (`src/_lcm/egm/interp.py:141-196,460-530,889-985,1237-1309`);
* the regime's V: logsum over its discrete-action combos with taste shocks, else
  the max (`src/_lcm/egm/step.py:731-738`);
* terminal targets: a wealth-state terminal carries (wealth grid, utility,
  d utility / d wealth) and a stateless terminal is a scalar target
  (`src/_lcm/egm/terminal.py:47-180`).

Derivatives of the user's functions (u'(c), dR'/dA, a terminal utility's slope)
are taken by the complex step `Im f(x + i h) / h` with h = 1e-30, which is exact
to rounding for the analytic NumPy expressions these models use. Arithmetic is
FP64 throughout; grid points and float parameters are first rounded to the model
dtype (the product solves the same discretised model).

Parity pinning (tests/test_oracle_dcegm.py): the reference's FUES unit-test rows
(crossing segments with an analytic envelope, node-aligned crossings, a
three-branch chain, dominated coincident pairs, savings provenance, overflow),
the Iskhakov et al. (2017) analytical value functions, the independent `dcegm`
package's taste-shock fixture of the reference, closed forms for the last
decision period, and a fine consumption-grid search of the same model.
"""

from __future__ import annotations

import math

import numpy as np

from oracle import gridsearch as og

CONSTRAINED_OFFSET_FRACTION = 1e-4
_H = 1e-30


def _np_errstate():
    return np.errstate(all="ignore")


# ----------------------------------------------------------------------------
# FUES (fues.py)
# ----------------------------------------------------------------------------


def _slope(xa, ya, xb, yb):
    with _np_errstate():
        dx = xb - xa
        return np.where(dx == 0.0, 0.0, (yb - ya) / np.where(dx == 0.0, 1.0, dx))


def _has_policy_jump(ga, pa, gb, pb, thresh, sa=None, sb=None):
    with _np_errstate():
        ia = ga - pa
        ib = gb - pb
        s = _slope(ga, ia, gb, ib)
        coincident = (sa != sb) if sa is not None else (ia != ib)
        return np.where(ga == gb, coincident, np.abs(s) > thresh)


def _intersect(xa, ya, sa, xb, yb, sb):
    with _np_errstate():
        den = sa - sb
        x = np.where(den == 0.0, np.nan, (yb - ya + sa * xa - sb * xb) / np.where(den == 0.0, 1.0, den))
        safe_x = np.where(den == 0.0, xa, x)
        y = ya + sa * (safe_x - xa)
        return float(x), float(y)


def _find_same_segment(g, p, v, anchor_g, anchor_p, idx, direction, n_scan, thresh):
    n = g.shape[0]
    win = idx + direction * (1 + np.arange(n_scan))
    inb = (win >= 0) & (win < n)
    cl = np.clip(win, 0, n - 1)
    wg, wp, wv = g[cl], p[cl], v[cl]
    same = inb & ~np.isnan(wg) & ~np.isnan(wv) & ~_has_policy_jump(anchor_g, anchor_p, wg, wp, thresh)
    found = bool(np.any(same))
    pos = int(np.argmax(same))
    return found, float(wg[pos]), float(wp[pos]), float(wv[pos])


def _reduce_coincident(g, p, v, s, use_savings):
    n = g.shape[0]
    finite = ~np.isnan(g)
    new = np.ones(n, dtype=bool)
    new[1:] = (g[1:] != g[:-1]) & finite[1:]
    gid = np.cumsum(new) - 1
    vneg = np.where(finite, v, -np.inf)
    gmax = np.full(n, -np.inf)
    for k in range(n):  # segment max with NaN propagation
        a = gmax[gid[k]]
        b = vneg[k]
        gmax[gid[k]] = np.nan if (np.isnan(a) or np.isnan(b)) else max(a, b)
    gm = gmax[gid]
    is_max = finite & (v == gm)
    src = s if use_savings else p
    collapse = np.zeros(n, dtype=bool)
    for i in range(n):
        if not is_max[i]:
            continue
        for j in range(i):
            if gid[j] == gid[i] and is_max[j] and src[j] == src[i]:
                collapse[i] = True
                break
    keep = finite & is_max & ~collapse
    nan = np.nan
    return (np.where(keep, g, nan), np.where(keep, p, nan), np.where(keep, v, nan),
            np.where(keep, s, nan))


def refine_envelope(grid, policy, value, n_refined, jump_thresh=2.0, n_points_to_scan=None,
                    savings=None):
    """FUES refinement (fues.py:151-338). Returns (grid, policy, value, n_kept)
    as NaN-padded float64 arrays of length n_refined."""
    grid = np.asarray(grid, dtype=np.float64)
    policy = np.asarray(policy, dtype=np.float64)
    value = np.asarray(value, dtype=np.float64)
    gk = np.where(np.isnan(grid), np.inf, grid)
    vk = np.where(np.isnan(value), -np.inf, value)
    pk = np.where(np.isnan(policy), -np.inf, policy)
    order = np.lexsort((-pk, -vk, gk))
    g, p, v = grid[order], policy[order], value[order]
    n = g.shape[0]
    n_scan = max(n - 1, 1) if n_points_to_scan is None else int(n_points_to_scan)
    use_sav = savings is not None
    s = np.asarray(savings, dtype=np.float64)[order] if use_sav else np.zeros(n)
    g, p, v, s = _reduce_coincident(g, p, v, s, use_sav)
    eps = np.finfo(np.float64).eps
    th = jump_thresh
    # carry: k then j as (grid, policy, value), seg_k = seg_j = 0, savings_j
    Gk, Pk, Vk = float(g[0]), float(p[0]), float(v[0])
    Gj, Pj, Vj = Gk, Pk, Vk
    Sj = float(s[0])
    out_g, out_p, out_v = [], [], []
    with _np_errstate():
        for idx in range(1, n):
            Gi, Pi, Vi, Si = float(g[idx]), float(p[idx]), float(v[idx]), float(s[idx])
            valid = not (math.isnan(Gi) or math.isnan(Vi))
            switches = bool(_has_policy_jump(Gj, Pj, Gi, Pi, th, Sj if use_sav else None,
                                             Si if use_sav else None))
            node_cross = valid and switches and (Gi == Gj) and (Vi == Vj)
            secant = float(_slope(Gj, Vj, Gi, Vi))
            grad_before = float(_slope(Gk, Vk, Gj, Vj))
            jf, jg, jp, jv = _find_same_segment(g, p, v, Gj, Pj, idx, 1, n_scan, th)
            sec_i_jseg = float(_slope(Gi, Vi, jg, jv))
            below_j = switches and jf and ((Vi < jv) if Gi == jg else (secant < sec_i_jseg))
            if use_sav:
                sav_dec = Si < Sj
            else:
                scale = max(max(abs(Gi), abs(Pi)), max(abs(Gj), abs(Pj)))
                sav_dec = (Gi - Pi) < (Gj - Pj) - 16.0 * eps * scale
            dropped = ((not valid) or ((Vi < Vj) and not switches)
                       or (sav_dec and secant < grad_before) or below_j) and not node_cross
            ff, fg, fp_, fv = _find_same_segment(g, p, v, Gi, Pi, idx, 1, n_scan, th)
            bf, bg, bp, bv = _find_same_segment(g, p, v, Gi, Pi, idx, -1, n_scan, th)
            partner = ff or bf
            pg, pp, pv = (fg, fp_, fv) if ff else (bg, bp, bv)
            i_slope = float(_slope(Gi, Vi, pg, pv))
            i_pslope = float(_slope(Gi, Pi, pg, pp))
            j_dom = (not dropped) and switches and partner and (secant > i_slope) and not node_cross
            pslope_kj = float(_slope(Gk, Pk, Gj, Pj))
            kg6, kv6 = _intersect(Gj, Vj, grad_before, Gi, Vi, i_slope)
            kpl6 = Pj + pslope_kj * (kg6 - Gj)
            kpr6 = Pi + i_pslope * (kg6 - Gi)
            kink6 = j_dom and (kg6 > Gk) and (kg6 < Gj) and (kg6 < Gi)
            j_slope = float(_slope(Gj, Vj, jg, jv)) if jf else grad_before
            j_pslope = float(_slope(Gj, Pj, jg, jp)) if jf else pslope_kj
            kg5, kv5 = _intersect(Gj, Vj, j_slope, Gi, Vi, i_slope)
            kpl5 = Pj + j_pslope * (kg5 - Gj)
            kpr5 = Pi + i_pslope * (kg5 - Gi)
            crossing = (not dropped) and (not j_dom) and (not node_cross) and switches and partner
            tol = 16.0 * eps * max(max(abs(Gj), abs(Gi)), 1.0)
            on_j = crossing and abs(kg5 - Gj) <= tol
            on_i = crossing and (not on_j) and abs(kg5 - Gi) <= tol
            kink5 = crossing and (not on_j) and (not on_i) and (kg5 > Gj) and (kg5 < Gi)
            plain = (not dropped) and (not j_dom) and not (kink5 or on_j or on_i)
            # emission block (fues.py:549-637)
            emits_j = plain or kink5 or on_j or on_i
            rows = []
            if kink6:
                rows = [(kg6, kpl6, kv6), (kg6, kpr6, kv6)]
            elif kink5:
                rows = [(Gj, Pj, Vj), (kg5, kpl5, kv5), (kg5, kpr5, kv5)]
            elif on_j:
                rows = [(Gj, Pj, Vj), (Gj, kpr5, kv5)]
            elif on_i:
                rows = [(Gj, Pj, Vj), (Gi, kpl5, kv5)]
            elif emits_j:
                rows = [(Gj, Pj, Vj)]
            for r in rows:
                out_g.append(r[0])
                out_p.append(r[1])
                out_v.append(r[2])
            if dropped:
                continue
            keeps_k = j_dom and not kink6
            if not keeps_k:
                if kink5:
                    Gk, Pk, Vk = kg5, kpr5, kv5
                elif kink6:
                    Gk, Pk, Vk = kg6, kpr6, kv6
                elif on_j:
                    Gk, Pk, Vk = Gj, kpr5, kv5
                else:
                    Gk, Pk, Vk = Gj, Pj, Vj
            Gj, Pj, Vj, Sj = Gi, Pi, Vi, Si
    if not (math.isnan(Gj) or math.isnan(Vj)):
        out_g.append(Gj)
        out_p.append(Pj)
        out_v.append(Vj)
    n_kept = len(out_g)
    rg = np.full(n_refined, np.nan)
    rp = np.full(n_refined, np.nan)
    rv = np.full(n_refined, np.nan)
    m = min(n_kept, n_refined)
    rg[:m], rp[:m], rv[:m] = out_g[:m], out_p[:m], out_v[:m]
    return rg, rp, rv, n_kept


# ----------------------------------------------------------------------------
# NaN-padded row reads (interp.py)
# ----------------------------------------------------------------------------


def _between(xq, xl, xu, fl, fu, sl=None, su=None):
    with _np_errstate():
        w = xu - xl
        safe = np.where(w == 0.0, 1.0, w)
        raw = (xq - xl) / safe
        rel = np.where(w == 0.0, 1.0, np.clip(raw, 0.0, 1.0))
        wl = 1.0 - rel
        lin = np.where(wl > 0.0, wl * fl, 0.0) + np.where(rel > 0.0, rel * fu, 0.0)
        fin = np.isfinite(fl) & np.isfinite(fu)
        ext = np.where(fin & (w != 0.0) & (raw < 0.0), (raw - rel) * (fu - fl), 0.0)
        lin = lin + ext
        if sl is None:
            return lin
        df = fu - fl
        sdf = np.where(np.isfinite(df), df, 0.0)
        sec = sdf / safe

        def limit(m):
            same = m * sec > 0.0
            lim = np.sign(sec) * np.minimum(np.abs(m), 3.0 * np.abs(sec))
            return np.where(same, lim, 0.0)

        a = safe * limit(sl) - sdf
        b = sdf - safe * limit(su)
        ok = (w > 0.0) & np.isfinite(fl) & np.isfinite(fu) & np.isfinite(sl) & np.isfinite(su)
        corr = rel * (1.0 - rel) * ((1.0 - rel) * a + rel * b)
        return lin + np.where(ok, corr, 0.0)


def read_row(xq, xp, fp, slopes=None):
    """Read a NaN-padded row at queries xq (interp.py:141-196,490-530)."""
    xq = np.asarray(xq, dtype=np.float64)
    xp = np.asarray(xp, dtype=np.float64)
    fp = np.asarray(fp, dtype=np.float64)
    search = np.where(np.isnan(xp), np.inf, xp)
    valid = int(np.sum(~np.isnan(xp)))
    up = np.clip(np.searchsorted(search, xq, side="right"), 1, max(valid - 1, 1))
    lo = up - 1
    deg = valid < 2
    xl = np.where(deg, 0.0, xp[lo])
    xu = np.where(deg, 1.0, xp[up])
    fl = np.where(deg, 0.0, fp[lo])
    fu = np.where(deg, 0.0, fp[up])
    if slopes is None:
        res = _between(xq, xl, xu, fl, fu)
    else:
        sp = np.asarray(slopes, dtype=np.float64)
        res = _between(xq, xl, xu, fl, fu, np.where(deg, 0.0, sp[lo]), np.where(deg, 0.0, sp[up]))
    if valid == 1:
        res = np.full_like(res, fp[0])
    if valid == 0:
        res = np.full_like(res, np.nan)
    return np.where(np.isnan(xq), np.nan, res)


# ----------------------------------------------------------------------------
# per-column EGM step (step_core.py)
# ----------------------------------------------------------------------------


def egm_column(*, A, publish, u, du, inv, beta, EV, EMU, n_con, n_pad, jump_thresh=2.0,
               n_points_to_scan=None):
    """One (discrete state, discrete action) column given E[V'], E[mu'] at the
    savings nodes A. u, du, inv act on float64 arrays. Returns a dict with the
    published V row and the carry rows (grid, policy, value, marginal)."""
    with _np_errstate():
        disc = beta * EMU
        eps = np.finfo(np.float64).eps
        target = np.where(disc < 0.0, np.nan, np.maximum(disc, eps))
        act = np.asarray(inv(target), dtype=np.float64) * np.ones_like(A)
        endog = A + act
        val = u(act) + beta * EV
        limit = float(A[0])
        span = min(float(endog[0]), float(np.max(publish))) - limit
        empty = not (span > 0.0 and math.isfinite(span))
        ratio = (1.0 / CONSTRAINED_OFFSET_FRACTION) ** (1.0 / max(n_con - 1, 1))
        safe = 1.0 if empty else span
        cact = safe * CONSTRAINED_OFFSET_FRACTION * ratio ** np.arange(n_con, dtype=np.float64)
        dev0 = beta * EV[0]
        cval = u(cact) + dev0
        if empty:
            cact = np.full(n_con, np.nan)
            cval = np.full(n_con, np.nan)
        cg = np.concatenate([limit + cact, endog])
        cp = np.concatenate([cact, act])
        cv = np.concatenate([cval, val])
        cs = np.concatenate([np.full(n_con, limit), A])
        dead = np.isneginf(cv)
        rg, rp, rv, n_kept = refine_envelope(
            np.where(dead, np.nan, cg), np.where(dead, np.nan, cp), np.where(dead, np.nan, cv),
            n_pad, jump_thresh, n_points_to_scan, savings=np.where(dead, np.nan, cs))
        over = n_kept > n_pad
        mu = du(np.where(np.isnan(rp), 1.0, rp))
        mu = np.where(np.isnan(rp), np.nan, mu)
        mu = np.where(np.isneginf(rv), 0.0, mu)
        vint = read_row(publish, rg, rv, mu)
        cf = publish - limit
        vcon = np.where(cf > 0.0, u(np.maximum(cf, np.finfo(np.float64).tiny)) + dev0, -np.inf)
        V = np.where((cf > 0.0) & (publish <= rg[0]), vcon, np.maximum(vint, vcon))
        if over:
            V = np.full_like(V, np.nan)
            rv = np.full_like(rv, np.nan)
            mu = np.full_like(mu, np.nan)
        if np.all(dead):
            V = np.full_like(V, -np.inf)
            rv = np.full_like(rv, -np.inf)
            mu = np.zeros_like(mu)
    return {"V": V, "grid": rg, "policy": rp, "value": rv, "marginal": mu, "n_kept": n_kept,
            "EV": EV, "EMU": EMU, "action": act}


def aggregate_rows(values, marginals, scale):
    """Child choice aggregation over the leading axis (continuation.py:1317-1351,
    logsum.py:32-64): smoothed value and sum_d P_d mu_d."""
    values = np.asarray(values, dtype=np.float64)
    marginals = np.asarray(marginals, dtype=np.float64)
    with _np_errstate():
        if scale is not None and scale > 0.0:
            m = np.max(values, axis=0)
            fm = np.where(np.isneginf(m), 0.0, m)
            sh = (values - fm) / scale
            sm = fm + scale * np.log(np.sum(np.exp(sh), axis=0))
            allm = np.all(np.isneginf(values), axis=0)
            e = np.exp(sh - np.max(sh, axis=0))
            probs = np.where(allm, 0.0, e / np.sum(e, axis=0))
            return sm, np.sum(probs * marginals, axis=0)
        hm = np.max(values, axis=0)
        arg = np.argmax(values, axis=0)  # first index of the max (NaN -> its index)
        onehot = np.arange(values.shape[0])[:, None] == arg[None, :]
        return hm, np.sum(onehot * marginals, axis=0)


# ----------------------------------------------------------------------------
# model level
# ----------------------------------------------------------------------------


def _is_dcegm(regime):
    return type(getattr(regime, "solver", None)).__name__ == "DCEGM"


def _cstep(f, x):
    """Complex-step derivative of a NumPy function at float64 points x."""
    with _np_errstate():
        out = f(np.asarray(x, dtype=np.float64) + 1j * _H)
        return np.imag(np.asarray(out, dtype=np.complex128)) / _H


def _env_call(model, reg, t, params, fixed, gdt, variables, fname, func=None):
    ev = og._env(model, reg, t, params, fixed, gdt, variables)
    return ev.call(fname, func)


def _child_reads(model, R, reg, t, params, fixed, gdt, v, carries, Vnext):
    """Per target regime: (prob, kind, payload) for one combo `v`."""
    names = list(model.regime_id_class.__categories__)
    age_next = model.ages.period_to_age(t + 1)
    tr = reg.regime.transition
    e = og._env(model, reg, t, params, fixed, gdt, v)
    P = np.zeros(len(names))
    if hasattr(tr, "items"):
        for tn, f in tr.items():
            P[names.index(tn)] = float(np.asarray(e.call("next_regime", f)))
    elif type(tr).__name__ == "MarkovTransition":
        P[:] = np.asarray(e.call("next_regime", tr), dtype=np.float64)
    else:
        P[int(np.asarray(e.call("next_regime", tr)))] = 1.0
    out = []
    for n in names:
        tgt = R[n]
        if not og._active(tgt.regime, age_next):
            continue
        p = float(P[names.index(n)])
        if tgt.V_shape == ():
            out.append((p, "scalar", float(np.asarray(Vnext[n]))))
            continue
        # node weights over the target's discrete/process states
        axes_w = []
        for (s, g) in tgt.sd:
            k = [x for x, _ in tgt.sd].index(s)
            n_pts = len(tgt.sd_vals[k])
            if type(g).__name__ == "DiscreteGrid":
                law = reg.regime.state_transitions[s]
                if hasattr(law, "items"):
                    law = law[n]
                o = np.asarray(e.call(f"next_{s}", law), dtype=np.float64)
                if type(law).__name__ == "MarkovTransition":
                    axes_w.append(o)
                else:
                    w = np.zeros(n_pts)
                    w[int(o)] = 1.0
                    axes_w.append(w)
            else:
                nodes_t, P_t = tgt.proc[s]
                axes_w.append(og.transition_row(P_t, nodes_t, v[s]))
        idx = list(np.ndindex(*[len(a) for a in axes_w])) if axes_w else [()]
        W = np.array([np.prod([axes_w[i][k] for i, k in enumerate(ix)]) if ix else 1.0 for ix in idx])
        out.append((p, "rows", (n, W, idx)))
    return out


def solve(model, params, *, grid_dtype=np.float64):
    """Backward induction with DC-EGM regimes. Returns (V, COLS): period ->
    regime -> V array, and period -> regime -> list of per-combo column dicts
    (V row, refined carry rows) in (discrete state, discrete action) order."""
    gdt = np.dtype(grid_dtype)
    fixed = model.fixed_params
    names = list(model.regime_id_class.__categories__)
    R = {n: og._R(n, model.regimes[n], gdt, params, fixed) for n in names}
    T = model.ages.n_periods
    V, COLS = {}, {}
    carries = {}  # regime -> {child sd combo index: [rows per action combo]} of period t+1
    for t in reversed(range(T)):
        age = model.ages.period_to_age(t)
        V[t], COLS[t] = {}, {}
        new_carries = {}
        for n in names:
            reg = R[n]
            if not og._active(reg.regime, age):
                continue
            if reg.terminal:
                V[t][n], new_carries[n] = _terminal(model, reg, t, params, fixed, gdt)
                continue
            if not _is_dcegm(reg.regime):
                raise NotImplementedError("the DC-EGM oracle solves DC-EGM and terminal regimes only")
            V[t][n], COLS[t][n], new_carries[n] = _dcegm_regime(
                model, R, reg, t, params, fixed, gdt, carries, V.get(t + 1))
        carries = new_carries
    return V, COLS


def _terminal(model, reg, t, params, fixed, gdt):
    if not reg.w:
        val = _env_call(model, reg, t, params, fixed, gdt, {}, "utility")
        Vt = np.asarray(val, dtype=np.float64)
        return Vt, None
    wname = reg.w[0]
    rows = {}
    Vt = np.zeros(reg.V_shape)
    for sdi, dai, v in og._pointwise_points(reg):
        def util(x, v=v):
            vv = dict(v)
            vv[wname] = x
            return _env_call(model, reg, t, params, fixed, gdt, vv, "utility")
        val = np.broadcast_to(np.asarray(util(reg.w_pts), dtype=np.float64), reg.w_pts.shape)
        mu = np.broadcast_to(_cstep(util, reg.w_pts), reg.w_pts.shape)
        mu = np.where(np.isneginf(val), 0.0, mu)
        Vt[sdi] = val
        rows[sdi] = [(reg.w_pts.copy(), val.copy(), mu.copy())]
    return Vt, {"rows": rows, "scale": 0.0}


def _dcegm_regime(model, R, reg, t, params, fixed, gdt, carries, Vnext):
    rg = reg.regime
    solver = rg.solver
    liq = rg.liquid
    if liq.resources != liq.state:
        raise NotImplementedError("the DC-EGM oracle takes resources equal to the Euler state")
    A = og.grid_points(solver.savings_grid, gdt)
    n_con = int(solver.n_constrained_points)
    n_pad = int(math.ceil(solver.refined_grid_factor * (len(A) + n_con)))
    env_scale = None
    if rg.taste_shocks is not None:
        env_scale = float(np.asarray(params[reg.name]["taste_shocks"]["scale"], dtype=gdt))
    wname, cname = liq.state, liq.action
    law = rg.state_transitions[wname]
    cols = []
    da_sizes = tuple(len(x) for x in reg.da_vals)
    n_da = int(np.prod(da_sizes)) if da_sizes else 1
    Vstack = np.zeros(reg.V_shape[:-1] + (n_da, len(reg.w_pts)))
    rows_out = {}
    for sdi, dai, v in og._pointwise_points(reg):
        ev = og._env(model, reg, t, params, fixed, gdt, v)
        beta = ev.param("koopmans_aggregator", "discount_factor")

        def u(c, v=v):
            vv = dict(v)
            vv[cname] = c
            return np.asarray(_env_call(model, reg, t, params, fixed, gdt, vv, "utility"))

        def inv(x, v=v):
            vv = dict(v)
            vv["marginal_continuation"] = x
            return np.asarray(_env_call(model, reg, t, params, fixed, gdt, vv, "inverse_marginal_utility"))

        def du(c, u=u):
            return _cstep(u, c)

        def wnext(a, v=v):
            vv = dict(v)
            vv[liq.post_decision_state] = a
            return np.asarray(_env_call(model, reg, t, params, fixed, gdt, vv, f"next_{wname}", law))

        Rn = np.broadcast_to(np.asarray(wnext(A), dtype=np.float64), A.shape)
        gR = np.broadcast_to(_cstep(wnext, A), A.shape)
        EV = np.zeros_like(A)
        EMU = np.zeros_like(A)
        for p, kind, pay in _child_reads(model, R, reg, t, params, fixed, gdt, v, carries, Vnext):
            if kind == "scalar":
                EV = EV + (p * pay if p > 0.0 else p * 0.0)
                continue
            child, W, idx = pay
            cc = carries[child]
            vsum = np.zeros_like(A)
            msum = np.zeros_like(A)
            for w, ix in zip(W, idx):
                rows = cc["rows"][ix]
                vals = []
                mars = []
                for (xg, xv, xm) in rows:
                    val = read_row(Rn, xg, xv, xm)
                    mar = read_row(Rn, xg, xm) * gR
                    mar = np.where(np.isneginf(val), 0.0, mar)
                    vals.append(val)
                    mars.append(mar)
                sv, sm = aggregate_rows(np.array(vals), np.array(mars), cc["scale"])
                vsum = vsum + og.zero_safe(w, sv)
                msum = msum + og.zero_safe(w, sm)
            EV = EV + (p * vsum if p > 0.0 else p * 0.0)
            EMU = EMU + (p * msum if p > 0.0 else p * 0.0)
        col = egm_column(A=A, publish=reg.w_pts, u=u, du=du, inv=inv, beta=beta, EV=EV, EMU=EMU,
                         n_con=n_con, n_pad=n_pad, jump_thresh=solver.envelope.jump_thresh,
                         n_points_to_scan=solver.envelope.n_points_to_scan)
        cols.append(col)
        flat_d = int(np.ravel_multi_index(dai, da_sizes)) if da_sizes else 0
        Vstack[sdi + (flat_d,)] = col["V"]
        rows_out.setdefault(sdi, []).append((col["grid"], col["value"], col["marginal"]))
    if n_da > 1 and env_scale is not None:
        Vt, _ = aggregate_rows(np.moveaxis(Vstack, -2, 0), np.zeros_like(np.moveaxis(Vstack, -2, 0)),
                               env_scale)
    elif n_da > 1:
        Vt = np.max(Vstack, axis=-2)
    else:
        Vt = Vstack[..., 0, :]
    return Vt.reshape(reg.V_shape), cols, {"rows": rows_out, "scale": env_scale or 0.0}
