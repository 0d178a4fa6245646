"""Parity helpers: compare a solve against the oracle (test infrastructure).

Policies must match exactly; a mismatch is tolerated only as a flagged near-tie:
the oracle's Q at the two indices differ by at most `tie_ulps` units in the last
place of the compute dtype (relative to |Q|).
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


def policy_report(P_gpu, P_or, Q_or, dtype, tie_ulps=8, taste=False):
    """(n_mismatch, n_near_tie) for one (period, regime)."""
    pg = np.asarray(P_gpu).reshape(-1)
    po = np.asarray(P_or).reshape(-1)
    bad = np.nonzero(pg != po)[0]
    if bad.size == 0:
        return 0, 0
    Q, F = Q_or
    n_states = int(np.prod(np.asarray(P_or).shape[: (np.asarray(P_or).ndim - (1 if taste else 0))]))
    Qs = np.where(F, Q, -np.inf).reshape(n_states, -1)
    eps = np.finfo(dtype).eps
    ties = 0
    for i in bad:
        s = i // (np.asarray(P_or).shape[-1]) if taste else i
        qa, qb = Qs[s, pg[i]], Qs[s, po[i]]
        if np.isfinite(qa) and np.isfinite(qb) and abs(qa - qb) <= tie_ulps * eps * max(abs(qa), abs(qb), 1.0):
            ties += 1
    return int(bad.size - ties), int(ties)


def compare_solution(V_gpu, P_gpu, V_or, P_or, Q_or, dtype, taste_regimes=()):
    """Max relative V error, policy mismatches and flagged near-ties, over all
    periods and regimes of the oracle's solution."""
    worst = 0.0
    mism = 0
    ties = 0
    for t in V_or:
        for r in V_or[t]:
            worst = max(worst, rel_err(V_gpu[t][r], V_or[t][r]))
            m, k = policy_report(P_gpu[t][r], P_or[t][r], Q_or[(t, r)], dtype,
                                 taste=r in taste_regimes)
            mism += m
            ties += k
    return worst, mism, ties
