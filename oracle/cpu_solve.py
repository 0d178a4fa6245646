"""Grid-search solve with the C restatement (TEST INFRASTRUCTURE ONLY).

The user's functions are tabulated over the full state x action grid by
`oracle.gridsearch.tabulate`; the per-period Bellman step runs in
`oracle/gs_ref.c` (OpenMP over states) in the reference's order of operations.
This is the CPU baseline `bench.py` times, and a cross-check of the NumPy
oracle.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from oracle import build_c
from oracle import gridsearch as og

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not build_c.OUT.exists():
            build_c.build()
        _lib = C.CDLL(str(build_c.OUT))
        _lib.gs_step_ref.restype = C.c_int
    return _lib


def _dp(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


def solve_c(model, params, *, grid_dtype=np.float64, threads=0, next_V=None, cmp_policy=None,
            tie_ulps=8):
    """FP64 backward induction with the C Bellman step. Returns (V, POL).

    With `next_V` (period -> regime -> V) each period reads that V for period
    t+1 instead of its own, so a solve computed elsewhere is re-checked one
    period at a time from the same inputs. With `cmp_policy` the return is
    (V, POL, report): `report` counts the policy entries that differ from
    `cmp_policy` as near-ties (this step's FP64 Q at the two actions within
    `tie_ulps` units in the last place of `grid_dtype`) or genuine mismatches.
    """
    lib = _load()
    report = {"mismatches": 0, "near_ties": 0, "max_tie_gap_ulps": 0.0}
    eps = float(np.finfo(np.dtype(grid_dtype)).eps)
    gdt = np.dtype(grid_dtype)
    fixed = model.fixed_params
    names = list(model.regime_id_class.__categories__)
    R = {n: og._R(n, model.regimes[n], gdt, params, fixed) for n in names}
    T = model.ages.n_periods
    V, POL = {}, {}
    keep = []
    for t in reversed(range(T)):
        age = model.ages.period_to_age(t)
        V[t], POL[t] = {}, {}
        for n in names:
            reg = R[n]
            if not og._active(reg.regime, age):
                continue
            Vnext = next_V.get(t + 1) if next_V is not None else V.get(t + 1)
            tab = og.tabulate(model, R, reg, t, params, fixed, gdt, Vnext)
            shape = tab["shape"]
            nsd = int(np.prod([len(v) for v in reg.sd_vals])) if reg.sd_vals else 1
            nw = len(reg.w_pts) if reg.w else 1
            nd = int(np.prod([len(v) for v in reg.da_vals])) if reg.da_vals else 1
            nc = len(reg.c_pts) if reg.c else 1
            U = np.ascontiguousarray(tab["U"], dtype=np.float64)
            F = np.ascontiguousarray(tab["F"], dtype=np.uint8)
            wn = np.ascontiguousarray(tab.get("wn", np.zeros(shape)), dtype=np.float64)
            tg = tab["targets"]
            sc = tab["scalars"]
            Vt = (C.POINTER(C.c_double) * max(len(tg), 1))()
            Wt = (C.POINTER(C.c_double) * max(len(tg), 1))()
            pt = (C.POINTER(C.c_double) * max(len(tg), 1))()
            Kt = (C.c_int * max(len(tg), 1))()
            for i, (name, p, W, Vn) in enumerate(tg):
                arrs = [np.ascontiguousarray(Vn, dtype=np.float64),
                        np.ascontiguousarray(W.reshape(nsd * nd, -1), dtype=np.float64),
                        np.ascontiguousarray(p.reshape(-1), dtype=np.float64)]
                keep += arrs
                Vt[i], Wt[i], pt[i] = (_dp(a) for a in arrs)
                Kt[i] = W.shape[-1]
            qs = (C.POINTER(C.c_double) * max(len(sc), 1))()
            vs = np.zeros(max(len(sc), 1))
            for i, (name, p, v) in enumerate(sc):
                a = np.ascontiguousarray(p.reshape(-1), dtype=np.float64)
                keep.append(a)
                qs[i] = _dp(a)
                vs[i] = v
            gkind, G, gpts = 0, 1, np.zeros(1)
            if tg:
                grid, gpts = tab["grid"]
                gkind = {"LinSpacedGrid": 0, "LogSpacedGrid": 1}.get(type(grid).__name__, 2)
                G = len(gpts)
            gpts = np.ascontiguousarray(gpts, dtype=np.float64)
            taste = tab["taste"] or 0.0
            Vout = np.zeros(nsd * nw)
            pol = np.zeros(nsd * nw * (nd if taste > 0 else 1), dtype=np.int32)
            pcmp = qcmp = None
            if cmp_policy is not None:
                pcmp = np.ascontiguousarray(np.asarray(cmp_policy[t][n]).reshape(-1), dtype=np.int32)
                assert pcmp.size == pol.size
                qcmp = np.zeros(2 * pol.size)
            lib.gs_step_ref(nsd, nw, nd, nc, _dp(U), F.ctypes.data_as(C.POINTER(C.c_uint8)), _dp(wn),
                            int(tab["terminal"]), len(tg), Vt, Kt, Wt, pt, len(sc), _dp(vs), qs,
                            gkind, G, _dp(gpts), C.c_double(tab.get("beta", 0.0) or 0.0),
                            C.c_double(taste), _dp(Vout) if False else Vout.ctypes.data_as(C.POINTER(C.c_double)),
                            pol.ctypes.data_as(C.POINTER(C.c_int32)),
                            None if pcmp is None else pcmp.ctypes.data_as(C.POINTER(C.c_int32)),
                            None if qcmp is None else qcmp.ctypes.data_as(C.POINTER(C.c_double)),
                            int(threads))
            if pcmp is not None:
                bad = np.nonzero(pcmp != pol)[0]
                q_at, q_best = qcmp[0::2][bad], qcmp[1::2][bad]
                with np.errstate(invalid="ignore"):
                    gap = np.abs(q_best - q_at) / (eps * np.maximum(np.maximum(np.abs(q_best), np.abs(q_at)), 1.0))
                tie = np.isfinite(gap) & (gap <= tie_ulps)
                report["mismatches"] += int(np.sum(~tie))
                report["near_ties"] += int(np.sum(tie))
                if np.any(tie):
                    report["max_tie_gap_ulps"] = max(report["max_tie_gap_ulps"], float(np.max(gap[tie])))
            V[t][n] = Vout.reshape(reg.V_shape)
            POL[t][n] = pol.reshape(reg.V_shape + ((nd,) if taste > 0 else ()))
    if cmp_policy is not None:
        return V, POL, report
    return V, POL
