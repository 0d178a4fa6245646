"""FP64 oracle of the forward simulation (TEST INFRASTRUCTURE ONLY).

Restates `_simulate_regime_in_period` (src/_lcm/simulation/simulate.py:465-620)
on concrete arrays: per agent, Q over the full action grid at the agent's own
state — utility, feasibility and next continuous state from the user's
functions evaluated at the agent's values; the continuation per target regime
as the node expectation with the process weights interpolated at the agent's
off-grid process value (src/_lcm/regime_building/processing.py:2780-2840) —
then the first argmax (src/_lcm/regime_building/argmax.py:6-69) and the value.

Draws follow this project's counter-based scheme (the reference's JAX threefry
streams cannot be reproduced): Philox4x32-10 (Salmon et al., SC'11), restated
here independently of the device code and pinned by the Random123 known-answer
vectors, with counter (agent, period, stream, 0) and key (seed low, seed high);
stream 0 draws the regime, 1 the process innovation (Box-Muller), 2 the
discrete states, 3 the taste shocks (counter word 3 = the discrete cell).
Parity of a whole path is therefore exact in the integer draws and within
float tolerance in the continuous states.
"""

from __future__ import annotations

import numpy as np

from oracle import gridsearch as og

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    """ctr: (n, 4) uint64 holding 32-bit words; key: (k0, k1). Returns (n, 4)."""
    c = np.array(ctr, dtype=np.uint64).reshape(-1, 4) & MASK
    k0, k1 = np.uint64(key[0] & MASK), np.uint64(key[1] & MASK)
    for r in range(10):
        if r:
            k0 = (k0 + np.uint64(W0)) & np.uint64(MASK)
            k1 = (k1 + np.uint64(W1)) & np.uint64(MASK)
        p0 = np.uint64(M0) * c[:, 0]
        p1 = np.uint64(M1) * c[:, 2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(MASK)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(MASK)
        c = np.stack([hi1 ^ c[:, 1] ^ k0, lo1, hi0 ^ c[:, 3] ^ k1, lo0], axis=1)
    return c


def uniform53(a, b):
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    return ((a >> np.uint64(5)).astype(np.float64) * 67108864.0 + (b >> np.uint64(6)).astype(np.float64)) / 9007199254740992.0


def draws(agents, period, stream, seed, word3=0):
    ctr = np.stack([np.asarray(agents, dtype=np.uint64), np.full(len(agents), period, np.uint64),
                    np.full(len(agents), stream, np.uint64), np.full(len(agents), word3, np.uint64)], axis=1)
    return philox4x32_10(ctr, (seed & MASK, (seed >> 32) & MASK))


EULER_GAMMA = 0.5772156649015329


def taste_noise(agents, period, d, seed, scale):
    """Mean-zero EV1 noise of discrete cell d: scale * (Gumbel - Euler's gamma)
    (src/_lcm/regime_building/max_Q_over_a.py:326-346), the Gumbel drawn as
    -log(-log(u)) from stream 3 with counter word 3 = d, u in (0, 1)."""
    blk = draws(agents, period, 3, seed, word3=d)
    u = uniform53(blk[:, 0], blk[:, 1]) + 2.0 ** -54
    return scale * (-np.log(-np.log(u)) - EULER_GAMMA)


def normal_bm(block):
    u1 = 1.0 - uniform53(block[:, 0], block[:, 1])
    u2 = uniform53(block[:, 2], block[:, 3])
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def draw_categorical(p, u):
    """Row-wise: first i with p_i > 0 and u * sum(p) < cumsum(p)_i; else last positive."""
    p = np.asarray(p, dtype=np.float64)
    total = p.sum(axis=1, keepdims=True)
    acc = np.cumsum(p, axis=1)
    hit = (p > 0) & (u[:, None] * total < acc)
    last = np.where(p > 0, np.arange(p.shape[1]), -1).max(axis=1)
    first = np.where(hit.any(axis=1), np.argmax(hit, axis=1), np.maximum(last, 0))
    return first


def _axes(reg):
    sdd = [(s, g) for s, g in reg.sd if type(g).__name__ == "DiscreteGrid"]
    proc = [(s, g) for s, g in reg.sd if type(g).__name__ != "DiscreteGrid"]
    return sdd, (proc[0] if proc else None)


def simulate(model, params, V, initial, seed, *, grid_dtype=np.float64):
    """Returns dict of arrays [T, n]: regime, action, value, w, z, sdd (sdd = flat
    code over the regime's discrete non-process states); action -1 where an
    agent is not simulated."""
    gdt = np.dtype(grid_dtype)
    fixed = model.fixed_params
    names = list(model.regime_id_class.__categories__)
    R = {n: og._R(n, model.regimes[n], gdt, params, fixed) for n in names}
    T = model.ages.n_periods
    n = len(initial["regime_id"])
    start = np.array([model.ages.age_to_period(a) for a in initial.get("age", np.full(n, model.ages.period_to_age(0)))])
    regime = np.asarray(initial["regime_id"], dtype=np.int64).copy()
    w = np.zeros(n)
    z = np.zeros(n)
    sdd = np.zeros(n, dtype=np.int64)
    for r in names:
        reg = R[r]
        m = regime == names.index(r)
        if not m.any():
            continue
        s_d, proc = _axes(reg)
        if reg.w:
            w[m] = np.asarray(initial[reg.w[0]], dtype=np.float64)[m].astype(gdt)
        if proc:
            z[m] = np.asarray(initial[proc[0]], dtype=np.float64)[m].astype(gdt)
        if s_d:
            sizes = [len(reg.sd_vals[[x for x, _ in reg.sd].index(s)]) for s, _ in s_d]
            sdd[m] = np.ravel_multi_index([np.asarray(initial[s])[m].astype(int) for s, _ in s_d], sizes)
    out = {k: np.zeros((T, n)) for k in ("value", "w", "z")}
    out.update({k: np.full((T, n), -1, dtype=np.int64) for k in ("regime", "action", "sdd")})
    for t in range(T):
        age = model.ages.period_to_age(t)
        new_regime, new_w, new_z, new_sdd = regime.copy(), w.copy(), z.copy(), sdd.copy()
        for r in names:
            reg = R[r]
            if not og._active(reg.regime, age):
                continue
            idx = np.nonzero((regime == names.index(r)) & (start <= t))[0]
            if idx.size == 0:
                continue
            res = _regime_period(model, R, reg, t, params, fixed, gdt, V, w[idx], z[idx], sdd[idx],
                                 agents=idx, seed=seed)
            a, val, nxt = res
            out["regime"][t, idx] = names.index(r)
            out["action"][t, idx] = a
            out["value"][t, idx] = val
            out["w"][t, idx] = w[idx]
            out["z"][t, idx] = z[idx]
            out["sdd"][t, idx] = sdd[idx]
            if reg.terminal:
                continue
            preg, wd, wn = nxt
            ru = draws(idx, t, 0, seed)
            nr = draw_categorical(preg, uniform53(ru[:, 0], ru[:, 1]))
            new_regime[idx] = nr
            ds = np.zeros(idx.size, dtype=np.int64)
            du = draws(idx, t, 2, seed)
            u2 = uniform53(du[:, 0], du[:, 1])
            for tr, W in wd.items():
                m = nr == names.index(tr)
                if m.any():
                    ds[m] = draw_categorical(W[m], u2[m])
            new_sdd[idx] = ds
            new_w[idx] = wn
            s_d, proc = _axes(reg)
            if proc:
                pu = draws(idx, t, 1, seed)
                nrm = normal_bm(pu)
                u = uniform53(pu[:, 2], pu[:, 3])
                pname, pg = proc
                rt = og._proc_runtime(params, fixed, r, pname)
                p = og._proc_params(pg, rt)
                kind = type(pg).__name__
                if kind in ("RouwenhorstAR1Process", "TauchenAR1Process"):
                    new_z[idx] = p["mu"] + p["rho"] * z[idx] + p["sigma"] * nrm
                elif kind == "NormalIIDProcess":
                    new_z[idx] = p["mu"] + p["sigma"] * nrm
                elif kind == "LogNormalIIDProcess":
                    new_z[idx] = np.exp(p["mu"] + p["sigma"] * nrm)
                else:
                    new_z[idx] = p["start"] + (p["stop"] - p["start"]) * u
        regime, w, z, sdd = new_regime, new_w, new_z, new_sdd
    return out


def _regime_period(model, R, reg, t, params, fixed, gdt, V, w, z, sdd, agents=None, seed=0, return_q=False):
    """Argmax over actions for agents (arrays w, z, sdd) of one regime
    (with `return_q`: Q and F over the agents' action grids instead)."""
    names = list(model.regime_id_class.__categories__)
    n = len(w)
    s_d, proc = _axes(reg)
    sd_sizes = [len(reg.sd_vals[[x for x, _ in reg.sd].index(s)]) for s, _ in s_d]
    codes = np.unravel_index(sdd, sd_sizes) if s_d else ()
    da_sizes = [len(v) for v in reg.da_vals]
    nda = len(reg.da)
    ndim = 1 + nda + 1
    v = {}

    def ax(vals, i):
        sh = [1] * ndim
        sh[i] = len(vals)
        return np.asarray(vals).reshape(sh)

    v_agent = {}
    for (s, _), cd in zip(s_d, codes):
        v_agent[s] = np.asarray(cd).reshape([n] + [1] * (ndim - 1))
    if reg.w:
        v_agent[reg.w[0]] = np.asarray(w).reshape([n] + [1] * (ndim - 1))
    if proc:
        v_agent[proc[0]] = np.asarray(z).reshape([n] + [1] * (ndim - 1))
    v.update(v_agent)
    for j, ((a, _), vals) in enumerate(zip(reg.da, reg.da_vals)):
        v[a] = ax(vals.astype(np.int64), 1 + j)
    if reg.c:
        v[reg.c[0]] = ax(reg.c_pts, ndim - 1)
    shape = (n,) + tuple(da_sizes) + ((len(reg.c_pts),) if reg.c else (1,))
    ev = og._env(model, reg, t, params, fixed, gdt, v)
    U = np.broadcast_to(np.asarray(ev.call("utility"), dtype=np.float64), shape)
    F = np.ones(shape, dtype=bool)
    for cn, cf in reg.regime.constraints.items():
        F = F & np.broadcast_to(np.asarray(ev.call(cn, cf), dtype=bool), shape)
    if reg.terminal:
        Q = U
        nxt = None
    else:
        beta = ev.param("koopmans_aggregator", "discount_factor")
        age_next = model.ages.period_to_age(t + 1)
        active_next = [r for r in names if og._active(R[r].regime, age_next)]
        n_d = int(np.prod(da_sizes)) if da_sizes else 1
        preg = np.zeros((n, n_d, len(names)))
        wd = {}
        # per-agent, per discrete action: regime probabilities and next discrete weights
        for i in range(n):
            for dflat, dai in enumerate(np.ndindex(*da_sizes) if da_sizes else [()]):
                pv = {s: int(np.asarray(v_agent[s]).ravel()[i]) for s, _ in s_d}
                if proc:
                    pv[proc[0]] = float(z[i])
                if reg.w:
                    pv[reg.w[0]] = float(w[i])
                pv.update({a: int(k) for (a, _), k in zip(reg.da, dai)})
                e = og._env(model, reg, t, params, fixed, gdt, pv)
                tr = reg.regime.transition
                if hasattr(tr, "items"):
                    for tn, f in tr.items():
                        preg[i, dflat, names.index(tn)] = float(np.asarray(e.call("next_regime", f)))
                elif type(tr).__name__ == "MarkovTransition":
                    preg[i, dflat] = np.asarray(e.call("next_regime", tr), dtype=np.float64)
                else:
                    preg[i, dflat, int(np.asarray(e.call("next_regime", tr)))] = 1.0
        CE = np.zeros(shape)
        mass = np.zeros(shape)

        def per_d(arr):  # (n, n_d) -> broadcast over shape
            return np.broadcast_to(arr.reshape((n,) + tuple(da_sizes) + (1,)), shape)

        wn_all = None
        for tr_name in active_next:
            tgt = R[tr_name]
            p = preg[:, :, names.index(tr_name)]
            pg = per_d(p)
            mass = mass + pg
            if tgt.V_shape == ():
                CE = CE + og.zero_safe(pg, np.float64(V[t + 1][tr_name]))
                continue
            law = reg.regime.state_transitions[tgt.w[0]]
            wn = np.broadcast_to(np.asarray(ev.call(f"next_{tgt.w[0]}", law), dtype=np.float64), shape)
            wn_all = wn
            coord = og.coordinate(tgt.w[1], tgt.w_pts, wn)
            t_sdd, t_proc = _axes(tgt)
            # next discrete (non-process) weights per agent and d
            Wdisc = np.ones((n, n_d, 1))
            for s, g in t_sdd:
                npts = len(g.category_class.__categories__)
                Wax = np.zeros((n, n_d, npts))
                law_s = reg.regime.state_transitions[s]
                if hasattr(law_s, "items"):
                    law_s = law_s[tr_name]
                for i in range(n):
                    for dflat, dai in enumerate(np.ndindex(*da_sizes) if da_sizes else [()]):
                        pv = {s2: int(np.asarray(v_agent[s2]).ravel()[i]) for s2, _ in s_d}
                        if proc:
                            pv[proc[0]] = float(z[i])
                        pv.update({a: int(k) for (a, _), k in zip(reg.da, dai)})
                        e = og._env(model, reg, t, params, fixed, gdt, pv)
                        o = np.asarray(e.call(f"next_{s}", law_s), dtype=np.float64)
                        if type(law_s).__name__ == "MarkovTransition":
                            Wax[i, dflat] = o
                        else:
                            Wax[i, dflat, int(o)] = 1.0
                Wdisc = (Wdisc[..., :, None] * Wax[..., None, :]).reshape(n, n_d, -1)
            wd[tr_name] = Wdisc
            # process weights: interpolated transition row at the agent's value
            if t_proc:
                nodes_t, P_t = tgt.proc[t_proc[0]]
                Wp = np.stack([og.transition_row(P_t, nodes_t, zi) for zi in z])  # (n, K)
            else:
                Wp = np.ones((n, 1))
            Vn = np.asarray(V[t + 1][tr_name], dtype=np.float64)
            # V axes: target sd axes in declaration order, then w
            tsd_names = [x for x, _ in tgt.sd]
            num = np.zeros(shape)
            wsum = np.zeros(shape)
            Vflat = Vn.reshape(-1, Vn.shape[-1])
            sizes_t = [len(g.category_class.__categories__) for _, g in t_sdd]
            for flat, idxs in enumerate(np.ndindex(*Vn.shape[:-1])):
                wk = np.ones((n, n_d))
                disc_idx = []
                for axn, k in zip(tsd_names, idxs):
                    if t_proc and axn == t_proc[0]:
                        wk = wk * Wp[:, k][:, None]
                    else:
                        disc_idx.append(k)
                if t_sdd:
                    wk = wk * Wdisc[:, :, int(np.ravel_multi_index(disc_idx, sizes_t))]
                wkg = per_d(wk)
                num = num + og.zero_safe(wkg, og.interp_1d(Vflat[flat], coord))
                wsum = wsum + wkg
            E = num / np.where(wsum > 0, wsum, 1.0)
            CE = CE + og.zero_safe(pg, E)
        ok = np.abs(mass - 1.0) <= 1e-3
        CE = CE / np.where(ok, mass, np.nan)
        Q = U + beta * CE
        nxt = [preg, wd, wn_all]
    if return_q:
        return Q, F
    Qa = np.where(F, Q, -np.inf).reshape(n, -1)
    Fa = F.reshape(n, -1)
    mx = Qa.max(axis=1)
    nanm = np.any(Fa & np.isnan(Q.reshape(n, -1)), axis=1)
    a = np.argmax((Q.reshape(n, -1) == mx[:, None]) & Fa, axis=1)
    val = np.where(nanm, np.nan, mx)
    if reg.regime.taste_shocks is not None and not reg.terminal:
        # Gumbel-max over the discrete cells of the continuous maxima
        # (max_Q_over_a.py:282-304): Qc per cell, its first maximiser (the
        # first NaN if any: numpy's argmax), noise per cell, first argmax
        scale = float(params[reg.name]["taste_shocks"]["scale"])
        n_dd = int(np.prod(da_sizes)) if da_sizes else 1
        Qm = Qa.reshape(n, n_dd, -1)
        Qc = Qm.max(axis=2)
        ci = np.argmax(Qm, axis=2)
        noisy = Qc + np.stack([taste_noise(agents, t, dd, seed, scale) for dd in range(n_dd)], axis=1)
        dsel = np.argmax(noisy, axis=1)
        a = dsel * Qm.shape[2] + ci[np.arange(n), dsel]
        val = Qc[np.arange(n), dsel]
    if reg.terminal:
        return a, val, None
    preg, wd, wn_all = nxt
    d_idx = a // (Q.reshape(n, -1).shape[1] // int(np.prod(da_sizes) if da_sizes else 1)) if da_sizes else np.zeros(n, int)
    n_d = int(np.prod(da_sizes)) if da_sizes else 1
    nc = Q.reshape(n, -1).shape[1] // n_d
    d_idx, c_idx = np.divmod(a, nc)
    p_sel = preg[np.arange(n), d_idx]
    wd_sel = {k: W[np.arange(n), d_idx] for k, W in wd.items()}
    if wn_all is not None:
        wn_sel = wn_all.reshape(n, -1)[np.arange(n), a]
    else:
        wn_sel = np.zeros(n)
    return a, val, (p_sel, wd_sel, wn_sel)
