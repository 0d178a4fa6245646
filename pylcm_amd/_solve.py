"""Host drivers of the GPU solve (plain and batched)."""

from __future__ import annotations

import numpy as np

from pylcm_amd._gs_program import GsProgram, lower_problem


def _split(results):
    V, P = {}, {}
    for t, regs in results.items():
        V[t] = {r: v for r, (v, _) in regs.items()}
        P[t] = {r: p for r, (_, p) in regs.items()}
    return V, P


def _has_dcegm(model):
    return any(lay.is_dcegm for lay in model.layouts.values())


def solve_model(model, params, dtype):
    """One solve: (period->regime->V, period->regime->policy,
    period->regime->published simulation policy) on the host. The policy is
    the flat argmax index for grid-search regimes; DC-EGM regimes publish
    their refined (grid, consumption) rows, grid-search regimes none."""
    if _has_dcegm(model):
        from pylcm_amd._egm_program import solve_dcegm_model

        V, P = solve_dcegm_model(model, params, dtype)
        sim = {t: {r: {"grid": g, "policy": c} for r, (g, c) in regs.items()} for t, regs in P.items()}
        return V, P, sim
    prog = GsProgram(model, [params], dtype)
    prog.run(use_graph=False)
    prog.fetch()
    V, P = _split(prog.results(0))
    return V, P, {}


def solve_batch(model, params_list, dtype, *, tile=None, use_graph=True, return_policy=False):
    if _has_dcegm(model):
        from pylcm_amd._egm_program import EgmBatch

        batch = EgmBatch(model, params_list, dtype)
        batch.run()
        out = []
        for i in range(len(params_list)):
            V = batch.value_functions(i)
            out.append((V, batch.policies(i)) if return_policy else V)
        return out
    prog = GsProgram(model, params_list, dtype, tile=tile)
    prog.run(use_graph=use_graph)
    prog.fetch()
    out = []
    for p in range(len(params_list)):
        V, P = _split(prog.results(p))
        out.append((V, P) if return_policy else V)
    return out
