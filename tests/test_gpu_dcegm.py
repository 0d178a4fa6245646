"""GPU parity of the DC-EGM / FUES kernel against the FP64 oracle.

Bars: float64 V and the refined carry rows (endogenous grid, consumption
policy, value) within 1e-10 relative of the oracle with the same kept-point
counts; float32 V within 1e-5 relative (the device keeps the candidate rows
in FP64 and rounds only the published V).
"""

import numpy as np
import pytest

from oracle import dcegm as od
from pylcm_amd import examples as ex
from tests.parity import rel_err

pytestmark = pytest.mark.gpu


def _compare(model, params, dtype, rtol):
    from pylcm_amd._egm_program import EgmProgram

    prog = EgmProgram(model, params, dtype)
    prog.run()
    V = prog.value_functions()
    Vo, Co = od.solve(model, params, grid_dtype=dtype)
    worst_v, worst_rows, n_cols = 0.0, 0.0, 0
    for t in Vo:
        for r in Vo[t]:
            worst_v = max(worst_v, rel_err(V[t][r], Vo[t][r]))
            if r not in Co[t]:
                continue
            cols = prog.columns(t, r)
            for c, oc in enumerate(Co[t][r]):
                n_cols += 1
                assert cols["counts"][c, 0] == oc["n_kept"], (t, r, c)
                for k in ("grid", "policy", "value"):
                    worst_rows = max(worst_rows, rel_err(cols[k][c], oc[k]))
                worst_rows = max(worst_rows, rel_err(cols["V"][c], oc["V"]))
    print(f"{n_cols} columns: max rel V err {worst_v:.3e}, carry rows {worst_rows:.3e}")
    assert worst_v <= rtol
    if dtype == np.float64:
        assert worst_rows <= rtol
    return prog, V


def test_device_tape_matches_host(native):
    import ctypes as C

    from pylcm_amd import _native as nat
    from pylcm_amd import _tape

    k = np.array([0.5, 1.95, 2.5])
    ctx = nat.Context()
    x = np.linspace(0.2, 3.0, 37)
    for f in (lambda z: (z ** (1.0 - k) - 1.0) / (1.0 - k) - np.where(k > 1, 0.35, 0.0),
              lambda z: np.log(z) + np.exp(-k * z) * np.sqrt(z),
              lambda z: np.maximum(1.05 * z + k, 0.9) / (1.0 + z) - np.log1p(z) + np.expm1(-z)):
        tp = _tape.trace(3, f)
        prog = tp.program()
        consts = tp.const_table()
        for c in range(3):
            hv, hd = tp.eval(np.broadcast_to(x, (3, x.size)), derivative=True)
            bufs = [nat.DeviceBuffer(max(a.nbytes, 8)) for a in (prog, consts[c], x, x, x)]
            nat.h2d(ctx, bufs[0].ptr, prog)
            nat.h2d(ctx, bufs[1].ptr, np.ascontiguousarray(consts[c]))
            nat.h2d(ctx, bufs[2].ptr, x)
            nat.call("lcm_tape_eval", ctx.handle, C.c_void_p(bufs[0].ptr), C.c_void_p(bufs[1].ptr),
                     C.c_void_p(bufs[2].ptr), C.c_void_p(bufs[3].ptr), C.c_void_p(bufs[4].ptr), x.size)
            ctx.sync()
            dv, dd = np.empty_like(x), np.empty_like(x)
            nat.d2h(ctx, bufs[3].ptr, dv)
            nat.d2h(ctx, bufs[4].ptr, dd)
            np.testing.assert_allclose(dv, hv[c], rtol=1e-14)
            np.testing.assert_allclose(dd, hd[c], rtol=1e-13)


def test_iskhakov_dcegm_f64(native):
    m = ex.iskhakov_dcegm_model(6)
    p = ex.retirement_params(6, discount_factor=0.98, disutility_of_work=1.0, interest_rate=0.0, wage=20.0,
                             log_utility=True)
    _, V = _compare(m, p, np.float64, 1e-10)
    from pathlib import Path

    g = Path(__file__).parent / "golden" / "analytical_solution"
    an = np.genfromtxt(g / "iskhakov_2017_five_periods__values_worker.csv", delimiter=",")
    np.testing.assert_allclose(np.stack([V[t]["working_life"] for t in range(5)]), an, atol=1e-3)


def test_dcegm_twin_taste_shocks_f64(native):
    _compare(ex.dcegm_twin_model(), ex.dcegm_twin_params(0.2), np.float64, 1e-10)


def test_dcegm_stochastic_income_f64(native):
    m = ex.dcegm_retirement_model(6, n_savings=300, savings_max=100.0, n_wealth=120, wealth_range=(0.05, 100.0))
    _compare(m, ex.dcegm_params(6), np.float64, 1e-10)


def test_dcegm_stochastic_income_f32(native):
    m = ex.dcegm_retirement_model(6, n_savings=300, savings_max=100.0, n_wealth=120, wealth_range=(0.05, 100.0))
    _compare(m, ex.dcegm_params(6), np.float32, 1e-5)


def test_dcegm_model_solve_api(native):
    """Model.solve routes DC-EGM regimes to the device EGM path."""
    m = ex.iskhakov_dcegm_model(4)
    p = ex.retirement_params(4, discount_factor=0.98, disutility_of_work=0.1, interest_rate=0.0, wage=20.0,
                             log_utility=True)
    V = m.solve(params=p, log_level="off", dtype=np.float64)
    Vo, _ = od.solve(m, p)
    for t in Vo:
        for r in Vo[t]:
            assert rel_err(V[t][r], Vo[t][r]) <= 1e-10


@pytest.mark.parametrize("nodes", [0, 64, 128, 256])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dcegm_batch_matches_single_solves(native, dtype, nodes):
    """Parameter sets solved in one fused launch per period give, bit for bit,
    the V, carry rows and kept counts of each set solved alone; the first set
    also matches the FP64 oracle. `nodes` fixes the savings nodes per
    continuation workgroup of the batch (0: chosen per launch), so each
    kernel variant is compared with the single solve's."""
    from pylcm_amd._egm_program import EgmBatch, EgmProgram

    m = ex.dcegm_retirement_model(5, n_savings=200, savings_max=100.0, n_wealth=90, wealth_range=(0.05, 100.0))
    plist = [ex.dcegm_params(5, discount_factor=b, crra=c, sigma=s)
             for b, c, s in [(0.96, 2.0, 0.15), (0.93, 1.5, 0.1), (0.98, 3.0, 0.2)]]
    batch = EgmBatch(m, plist, dtype, nodes_per_block=nodes)
    batch.run()
    for i, p in enumerate(plist):
        single = EgmProgram(m, p, dtype)
        single.run()
        Vb, Vs = batch.value_functions(i), single.value_functions()
        for t in Vs:
            for r in Vs[t]:
                assert np.array_equal(Vb[t][r], Vs[t][r]), (i, t, r)
        for (t, r) in single.col_out:
            cb, cs = batch.progs[i].columns(t, r), single.columns(t, r)
            for k in cs:
                assert np.array_equal(cb[k], cs[k], equal_nan=True), (i, t, r, k)
    Vo, _ = od.solve(m, plist[0], grid_dtype=dtype)
    worst = max(rel_err(batch.value_functions(0)[t][r], Vo[t][r]) for t in Vo for r in Vo[t])
    assert worst <= (1e-10 if dtype == np.float64 else 1e-5)
    # the model API routes batched DC-EGM solves here
    Vapi = m.solve_batch(plist[:2], dtype=dtype)
    assert all(np.array_equal(Vapi[1][t][r], batch.value_functions(1)[t][r]) for t in Vapi[1] for r in Vapi[1][t])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_dcegm_graph_replays_equal_eager(native, dtype):
    """The DC-EGM solve replayed as one HIP graph gives the eager solve's V
    and carry rows bit for bit: a program that only ever runs the graph, and
    replays interleaved with the same program's eager runs; likewise for a
    batch."""
    from pylcm_amd._egm_program import EgmBatch, EgmProgram

    m = ex.dcegm_retirement_model(5, n_savings=200, savings_max=100.0, n_wealth=90, wealth_range=(0.05, 100.0))
    plist = [ex.dcegm_params(5, discount_factor=b, crra=c, sigma=s)
             for b, c, s in [(0.96, 2.0, 0.15), (0.93, 1.5, 0.1)]]
    eager = EgmProgram(m, plist[0], dtype)
    eager.run()
    V0 = eager.value_functions()
    C0 = {k: eager.columns(*k) for k in eager.col_out}
    graph = EgmProgram(m, plist[0], dtype)
    for k in range(3):
        graph.run(use_graph=True)
        V, Cols = graph.value_functions(), {kk: graph.columns(*kk) for kk in graph.col_out}
        assert all(np.array_equal(V[t][r], V0[t][r]) for t in V0 for r in V0[t]), k
        assert all(np.array_equal(Cols[kk][f], C0[kk][f], equal_nan=True) for kk in C0 for f in C0[kk]), k
        graph.run()  # the program's own eager work between replays
    be = EgmBatch(m, plist, dtype)
    be.run()
    bg = EgmBatch(m, plist, dtype)
    for k in range(2):
        bg.run(use_graph=True)
        for i in range(len(plist)):
            Vg, Ve = bg.value_functions(i), be.value_functions(i)
            assert all(np.array_equal(Vg[t][r], Ve[t][r]) for t in Ve for r in Ve[t]), (k, i)
        bg.run()
