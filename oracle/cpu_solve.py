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
            tie_ulps=8, reference_periods=()):
    """FP64 backward induction with the C Bellman step. Returns (V, POL).

    With `next_V` (period -> regime -> V) each period reads that V for period
    t+1 instead of its own, so a solve computed elsewhere is re-checked one
    period at a time from the same inputs. With `cmp_policy` the return is
    (V, POL, report): `report` counts the policy entries that differ from
    `cmp_policy` as near-ties (this step's FP64 Q at the two actions within
    `tie_ulps` units in the last place of `grid_dtype`) or genuine mismatches,
    ties split at one ulp and those above graded against the argmax of Q in
    the reference's own arithmetic at `grid_dtype` (`reference_form_q`). For
    the periods in `reference_periods` it also counts, over every state, how
    often that reference-arithmetic argmax and `cmp_policy` each depart from
    the FP64 argmax by more than one ulp.
    """
    lib = _load()
    report = {"mismatches": 0, "near_ties": 0, "max_tie_gap_ulps": 0.0, "near_ties_le_1ulp": 0,
              "near_ties_gt_1ulp": 0, "gt_1ulp_same_as_reference_f32": 0, "gt_1ulp_reference_f32_also_flips": 0,
              "gt_1ulp_unexplained": 0, "sampled_states": 0, "reference_f32_gt_1ulp_sampled": 0,
              "compared_gt_1ulp_sampled": 0, "reference_f32_max_gap_ulps_sampled": 0.0}
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
                gt1 = tie & (gap > 1.0)
                report["near_ties_le_1ulp"] += int(np.sum(tie & ~gt1))
                report["near_ties_gt_1ulp"] += int(np.sum(gt1))
                if np.any(gt1) and taste == 0:
                    # ties beyond one ulp: the argmax of Q in the reference's
                    # own arithmetic at the compute precision, from the same V
                    states = bad[gt1]
                    Qr, Fr = og.reference_form_q(tab, reg, states, compute_dtype=gdt)
                    a_ref = og.first_argmax(Qr, Fr)
                    same = a_ref == pcmp[states]
                    flips = a_ref != pol[states]
                    report["gt_1ulp_same_as_reference_f32"] += int(np.sum(same))
                    report["gt_1ulp_reference_f32_also_flips"] += int(np.sum(flips))
                    report["gt_1ulp_unexplained"] += int(np.sum(~(same | flips)))
                elif np.any(gt1):
                    report["gt_1ulp_unexplained"] += int(np.sum(gt1))
                if t in reference_periods and taste == 0 and not tab["terminal"]:
                    n_st = nsd * nw
                    states = np.arange(n_st)
                    Q64, F64 = og.reference_form_q(tab, reg, states, compute_dtype=np.float64)
                    Qr, _ = og.reference_form_q(tab, reg, states, compute_dtype=gdt)
                    a_ref = og.first_argmax(Qr, F64)
                    Qf = np.where(F64, Q64, -np.inf).reshape(n_st, -1)
                    pol64 = og.first_argmax(Q64, F64)

                    def n_gt1(a):
                        diff = np.nonzero((a != pol64) & (a >= 0))[0]
                        qa, qb = Qf[diff, a[diff]], Qf[diff, pol64[diff]]
                        with np.errstate(invalid="ignore"):
                            g = np.abs(qb - qa) / (eps * np.maximum(np.maximum(np.abs(qa), np.abs(qb)), 1.0))
                        return int(np.sum(g > 1.0)), (float(np.max(g)) if g.size else 0.0)

                    nr, gr = n_gt1(a_ref)
                    nc_, _ = n_gt1(pcmp.astype(np.int64))
                    report["sampled_states"] += n_st
                    report["reference_f32_gt_1ulp_sampled"] += nr
                    report["compared_gt_1ulp_sampled"] += nc_
                    report["reference_f32_max_gap_ulps_sampled"] = max(report["reference_f32_max_gap_ulps_sampled"], gr)
            V[t][n] = Vout.reshape(reg.V_shape)
            POL[t][n] = pol.reshape(reg.V_shape + ((nd,) if taste > 0 else ()))
    if cmp_policy is not None:
        return V, POL, report
    return V, POL
