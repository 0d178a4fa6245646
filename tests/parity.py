"""Parity helpers: compare a solve against the oracle (test infrastructure).

Policies must match exactly. A differing index is a flagged near-tie when the
oracle's Q at the two indices differ by at most one unit in the last place of
the compute dtype (relative to |Q|), the north star's rule. A difference of
more than one ulp (and at most `tie_ulps`) counts as explained only when Q
evaluated in the reference's own arithmetic at the compute precision, from the
GPU's V_{t+1} (`oracle.gridsearch.reference_form_q`), also departs from the
float64 argmax or picks the GPU's index; otherwise it is a mismatch.
"""

from __future__ import annotations

import numpy as np


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.abs(a - b) / np.maximum(np.abs(b), np.finfo(np.float64).tiny)
    r = np.where(same, 0.0, r)
    return float(np.max(r)) if r.size else 0.0


def tie_gaps(P_gpu, P_or, Q_or, dtype, taste=False):
    """(indices where the policies differ, their Q gap in ulps of `dtype`)."""
    pg = np.asarray(P_gpu).reshape(-1)
    po = np.asarray(P_or).reshape(-1)
    bad = np.nonzero(pg != po)[0]
    if bad.size == 0:
        return bad, np.zeros(0)
    Q, F = Q_or
    n_states = int(np.prod(np.asarray(P_or).shape[: (np.asarray(P_or).ndim - (1 if taste else 0))]))
    Qs = np.where(F, Q, -np.inf).reshape(n_states, -1)
    eps = np.finfo(dtype).eps
    gaps = np.empty(bad.size)
    for k, i in enumerate(bad):
        s = i // (np.asarray(P_or).shape[-1]) if taste else i
        qa, qb = Qs[s, pg[i]], Qs[s, po[i]]
        ok = np.isfinite(qa) and np.isfinite(qb)
        gaps[k] = abs(qa - qb) / (eps * max(abs(qa), abs(qb), 1.0)) if ok else np.inf
    return bad, gaps


def policy_report(P_gpu, P_or, Q_or, dtype, tie_ulps=8, taste=False, reference_argmax=None):
    """(n_mismatch, n_tie_le_1ulp, n_tie_gt_1ulp) for one (period, regime).

    `reference_argmax(states) -> actions` grades the ties above one ulp (see
    the module docstring); without it they count as ties."""
    bad, gaps = tie_gaps(P_gpu, P_or, Q_or, dtype, taste)
    if bad.size == 0:
        return 0, 0, 0
    le1 = gaps <= 1.0
    gt1 = (gaps > 1.0) & (gaps <= tie_ulps)
    mism = int(np.sum(~(le1 | gt1)))
    if np.any(gt1) and reference_argmax is not None and not taste:
        states = bad[gt1]
        a_ref = reference_argmax(states)
        pg = np.asarray(P_gpu).reshape(-1)[states]
        po = np.asarray(P_or).reshape(-1)[states]
        explained = (a_ref == pg) | (a_ref != po)
        mism += int(np.sum(~explained))
    return mism, int(np.sum(le1)), int(np.sum(gt1))


def compare_solution(V_gpu, P_gpu, V_or, P_or, Q_or, dtype, taste_regimes=(), *, model=None, params=None,
                     detail=False):
    """Max relative V error, policy mismatches and flagged near-ties, over all
    periods and regimes of the oracle's solution. With `model` and `params`,
    ties above one ulp are graded against the reference's arithmetic at the
    compute precision. `detail=True` also returns the ties split at one ulp."""
    worst = 0.0
    mism = 0
    le1 = gt1 = 0
    for t in V_or:
        for r in V_or[t]:
            worst = max(worst, rel_err(V_gpu[t][r], V_or[t][r]))
            ref = None
            if model is not None and t + 1 in V_gpu:
                ref = _reference_argmax_fn(model, params, t, r, V_gpu[t + 1], dtype)
            m, a, b = policy_report(P_gpu[t][r], P_or[t][r], Q_or[(t, r)], dtype,
                                    taste=r in taste_regimes, reference_argmax=ref)
            mism += m
            le1 += a
            gt1 += b
    if detail:
        return worst, mism, le1 + gt1, {"near_ties_le_1ulp": le1, "near_ties_gt_1ulp": gt1}
    return worst, mism, le1 + gt1


def _reference_argmax_fn(model, params, t, r, V_next, dtype):
    """states -> the first argmax of Q in the reference's arithmetic at
    `dtype`, from the GPU's V_{t+1} (tabulated lazily on first use)."""
    from oracle import gridsearch as og

    cache = {}

    def fn(states):
        if "tab" not in cache:
            names = list(model.regime_id_class.__categories__)
            R = {n: og._R(n, model.regimes[n], np.dtype(dtype), params, model.fixed_params) for n in names}
            Vn = {k: np.asarray(v, dtype=np.float64) for k, v in V_next.items()}
            cache["reg"] = R[r]
            cache["tab"] = og.tabulate(model, R, R[r], t, params, model.fixed_params, np.dtype(dtype), Vn)
        Qr, Fr = og.reference_form_q(cache["tab"], cache["reg"], states, compute_dtype=dtype)
        return og.first_argmax(Qr, Fr)

    return fn
