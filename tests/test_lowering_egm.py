"""DC-EGM lowering on the host (CPU): expression tapes against the user's
functions, derivatives against the complex step, and the refusals."""

import numpy as np
import pytest

from pylcm_amd import _tape
from pylcm_amd import examples as ex
from pylcm_amd._egm_program import EgmProgram, _columns
from pylcm_amd.exceptions import UnsupportedModelError


def test_tape_ops_and_derivatives():
    n = 3
    k = np.array([0.5, 1.5, 2.5])
    mask = np.array([True, False, True])
    funcs = [
        lambda x: (x ** (1.0 - k) - 1.0) / (1.0 - k),
        lambda x: np.log(x) - np.where(mask, 0.3, 0.0),
        lambda x: np.exp(-k * x) + np.sqrt(x) * k,
        lambda x: 1.0 / x + np.log1p(x) - np.expm1(-x),
        lambda x: np.where(mask, x * x, 2.0 * x) - (k - x) / k,
        lambda x: np.maximum(1.03 * x + k, 0.9),
        lambda x: 2.0 ** x - x ** 2.0,
    ]
    x = np.linspace(0.2, 3.0, 11)
    X = np.broadcast_to(x, (n, 11))
    for f in funcs:
        tp = _tape.trace(n, f)
        v, d = tp.eval(X, derivative=True)
        for i in range(n):
            def fi(z, i=i):
                out = f(np.broadcast_to(z, (n,) + z.shape).T)  # columns on the last axis
                return out.T[i]
            np.testing.assert_allclose(v[i], fi(x), rtol=1e-15, atol=1e-300)
            np.testing.assert_allclose(d[i], fi(x + 1e-30j).imag / 1e-30, rtol=1e-13, atol=1e-15)


def test_tape_abs_derivative():
    tp = _tape.trace(1, lambda x: abs(x - 1.0))
    v, d = tp.eval(np.array([[0.5, 2.0]]), derivative=True)
    np.testing.assert_array_equal(v, [[0.5, 1.0]])
    np.testing.assert_array_equal(d, [[-1.0, 1.0]])


def test_tape_refuses_branches_on_the_variable():
    with pytest.raises(UnsupportedModelError):
        _tape.trace(2, lambda x: np.where(x > 1.0, x, 0.0))
    with pytest.raises(UnsupportedModelError):
        _tape.trace(2, lambda x: x if x > 1 else 2 * x)


CASES = [
    ("iskhakov", lambda: ex.iskhakov_dcegm_model(5),
     lambda: ex.retirement_params(5, discount_factor=0.98, disutility_of_work=1.0, interest_rate=0.0,
                                  wage=20.0, log_utility=True)),
    ("twin", ex.dcegm_twin_model, lambda: ex.dcegm_twin_params(0.2)),
    ("stochastic", lambda: ex.dcegm_retirement_model(4, n_savings=50, n_wealth=30), lambda: ex.dcegm_params(4)),
]


@pytest.mark.parametrize("name,mk,pk", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_tapes_reproduce_user_functions(name, mk, pk, dtype):
    m, p = mk(), pk()
    prog = EgmProgram(m, p, dtype, device=False)
    assert prog.tapes
    for (t, r), (ut, it, rt) in prog.tapes.items():
        lay = m.layouts[r]
        vars_, n_sd, n_d = _columns(lay, prog.rg)
        n_cols = n_sd * n_d
        c = np.linspace(0.05, 30.0, 17)
        A = lay.regime.solver.savings_grid.to_numpy(dtype).astype(np.float64)

        def direct(fname, var, val, func=None):
            e = prog._evaluator(lay, {k: v[:, None] for k, v in vars_.items()}, t)
            e.variables[var] = val
            return np.broadcast_to(np.asarray(e.call(fname, func), dtype=np.complex128 if np.iscomplexobj(val)
                                              else np.float64), (n_cols, val.shape[-1]))

        v, d = ut.eval(np.broadcast_to(c, (n_cols, c.size)), derivative=True)
        np.testing.assert_allclose(v, direct("utility", lay.c_name, c[None, :]), rtol=1e-14)
        np.testing.assert_allclose(d, direct("utility", lay.c_name, c[None, :] + 1e-30j).imag / 1e-30, rtol=1e-12)
        x = np.geomspace(1e-3, 10.0, 9)
        np.testing.assert_allclose(it.eval(np.broadcast_to(x, (n_cols, x.size))),
                                   direct("inverse_marginal_utility", "marginal_continuation", x[None, :]),
                                   rtol=1e-14)
        law = lay.regime.state_transitions[lay.w_name]
        Rv, Rd = rt.eval(np.broadcast_to(A, (n_cols, A.size)), derivative=True)
        np.testing.assert_allclose(Rv, direct(f"next_{lay.w_name}", "savings", A[None, :], law), rtol=1e-15)
        np.testing.assert_allclose(
            Rd, direct(f"next_{lay.w_name}", "savings", A[None, :] + 1e-30j, law).imag / 1e-30, rtol=1e-13)


def test_refusals():
    m = ex.dcegm_retirement_model(4, n_savings=50, n_wealth=30)
    p = ex.dcegm_params(4)
    reg = m.regimes["working_life"]
    funcs = {k: v for k, v in reg.functions.items() if k != "inverse_marginal_utility"}
    bad = m.__class__(regimes={**m.regimes, "working_life": reg.replace(functions=funcs)},
                      ages=m.ages, regime_id_class=m.regime_id_class)
    with pytest.raises(UnsupportedModelError, match="inverse_marginal_utility"):
        EgmProgram(bad, p, np.float64, device=False)
    big = ex.dcegm_retirement_model(4, n_savings=4100, n_wealth=30)
    with pytest.raises(UnsupportedModelError, match="candidates"):
        EgmProgram(big, p, np.float64, device=False)
