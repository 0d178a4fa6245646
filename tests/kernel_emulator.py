"""NumPy emulation of the fused grid-search kernel's arithmetic on the lowered
tables (test infrastructure).

It executes `PeriodTables` exactly as `gs_period_kernel` specifies them
(contract V_{t+1} with M into EV rows, interpolate at `A*w + B`, add utility,
mask with the affine constraints, first argmax), so the CPU test suite checks
the host lowering against the oracle without a GPU. The GPU tests then check
the kernel against the same oracle.
"""

from __future__ import annotations

import numpy as np

from oracle.gridsearch import coordinate as o_coordinate
from oracle.gridsearch import zero_safe


def _full(tab, n_sd, n_w):
    return np.broadcast_to(tab, (n_sd, n_w) + tab.shape[2:])


def emulate(model, tabs, rg, dtype=np.float64):
    V, POL = {}, {}
    T = model.n_periods
    for t in reversed(range(T)):
        V[t], POL[t] = {}, {}
        for (tt, r), pt in tabs.items():
            if tt != t:
                continue
            lay = model.layouts[r]
            n_sd, n_w, n_d, n_c = lay.n_sd, lay.n_w, lay.n_d, lay.n_c
            w = rg.w_pts[r] if lay.w_grid is not None else np.zeros(1)
            U = _full(pt.U, n_sd, n_w)
            F = np.ones((n_sd, n_w, n_d, n_c), dtype=bool)
            for Fa, Fb, strict in pt.F:
                f = _full(Fa, n_sd, n_w) * w[None, :, None, None] + _full(Fb, n_sd, n_w)
                F &= (f > 0) if strict else (f >= 0)
            if pt.terminal:
                Q = np.array(U, dtype=np.float64)
            else:
                cs = np.zeros((n_sd, n_d))
                for name, m in pt.scalars:
                    v = 0.0 if name is None else float(V[t + 1][name])
                    cs = cs + zero_safe(m, v)
                Q = np.array(U, dtype=np.float64)
                if pt.targets:
                    gl = model.layouts[pt.next_grid]
                    gpts = rg.w_pts[pt.next_grid]
                    EV = np.zeros((n_sd, n_d, len(gpts)))
                    for name, M in pt.targets:
                        Vn = V[t + 1][name].reshape(M.shape[-1], -1)
                        for j in range(M.shape[-1]):
                            EV += zero_safe(M[:, :, j][..., None], Vn[j][None, None, :])
                    x = _full(pt.A, n_sd, n_w) * w[None, :, None, None] + _full(pt.B, n_sd, n_w)
                    coord = o_coordinate(gl.w_grid, gpts, x)
                    n = len(gpts)
                    lo = np.clip(np.floor(np.nan_to_num(coord, nan=0.0)), 0, n - 2).astype(int)
                    tt_ = coord - lo
                    sd_i = np.arange(n_sd)[:, None, None, None]
                    d_i = np.arange(n_d)[None, None, :, None]
                    v0 = EV[sd_i, d_i, lo]
                    v1 = EV[sd_i, d_i, lo + 1]
                    val = zero_safe(1 - tt_, v0) + zero_safe(tt_, v1)
                    Q = Q + pt.beta * (cs[:, None, :, None] + val)
                else:
                    Q = Q + pt.beta * cs[:, None, :, None]
            Qa = Q.reshape(n_sd * n_w, n_d * n_c)
            Fa_ = F.reshape(n_sd * n_w, n_d * n_c)
            if pt.taste_scale is None:
                m = np.max(np.where(Fa_, Qa, -np.inf), axis=1)
                m = np.where(np.any(Fa_ & np.isnan(Qa), axis=1), np.nan, m)
                pol = np.argmax((Qa == m[:, None]) & Fa_, axis=1)
                V[t][r] = m.reshape(lay.V_shape)
                POL[t][r] = pol.astype(np.int32).reshape(lay.V_shape)
            else:
                Q3 = Qa.reshape(-1, n_d, n_c)
                F3 = Fa_.reshape(-1, n_d, n_c)
                qc = np.max(np.where(F3, Q3, -np.inf), axis=2)
                ci = np.argmax((Q3 == qc[:, :, None]) & F3, axis=2)
                pol = np.where(np.any(F3, axis=2), np.arange(n_d)[None] * n_c + ci, 0)
                mx = np.max(qc, axis=1)
                fm = np.where(np.isneginf(mx), 0.0, mx)
                with np.errstate(divide="ignore"):
                    Vv = fm + pt.taste_scale * np.log(np.sum(np.exp((qc - fm[:, None]) / pt.taste_scale), axis=1))
                V[t][r] = Vv.reshape(lay.V_shape)
                POL[t][r] = pol.astype(np.int32).reshape(lay.V_shape + (n_d,))
    return V, POL
