"""`lcm_sim_period_noalloc` (partition scratch in the library's static device
array) against `lcm_sim_period_ex` (scratch in a caller buffer): identical
panels on two calls in a row, the second with fewer agents, and a clear error
above the static capacity."""

import numpy as np
import pytest

from pylcm_amd import _native as nat
from pylcm_amd import examples as ex
from pylcm_amd._gs_program import GsProgram
from pylcm_amd._simulate import SimProgram, _initial_conditions
from pylcm_amd.exceptions import NativeKernelError

pytestmark = pytest.mark.gpu


def _panel(model, params, prog, ic, static):
    sim = SimProgram(model, params, prog, static_scratch=static)
    res = sim.run(_initial_conditions(model, ic), seed=17)
    return res.raw


def _ic(n, seed):
    rng = np.random.default_rng(seed)
    return {"wealth": np.exp(np.log(10.0) + 0.5 * rng.standard_normal(n)),
            "income": rng.choice(np.linspace(-0.3, 0.3, 7), size=n),
            "regime_id": rng.choice([0, 1], size=n).astype(np.int32)}


def test_noalloc_matches_caller_scratch(native):
    T = 8
    m = ex.retirement_model(T, n_wealth=60, n_consumption=80, income_nodes=7)
    p = ex.retirement_params(T, income={"rho": 0.95, "sigma": 0.1, "mu": 0.0})
    prog = GsProgram(m, [p], np.float32)
    prog.run(use_graph=False)
    for n, seed in ((20_000, 1), (3_001, 2)):  # the second call is smaller
        ic = _ic(n, seed)
        a = _panel(m, p, prog, ic, static=False)
        b = _panel(m, p, prog, ic, static=True)
        for k in a:
            assert np.array_equal(a[k], b[k], equal_nan=True), (n, k)
        assert (a["action"] >= 0).any()


def test_noalloc_rejects_sizes_above_the_static_scratch(native):
    T = 3
    m = ex.retirement_model(T, n_wealth=20, n_consumption=20)
    p = ex.retirement_params(T)
    prog = GsProgram(m, [p], np.float64)
    prog.run(use_graph=False)
    n = nat.SIM_NOALLOC_MAX_ORDER // len(m.regime_names) + 1
    ic = {"wealth": np.full(n, 50.0), "regime_id": np.zeros(n, dtype=np.int32)}
    with pytest.raises(NativeKernelError, match="static scratch"):
        _panel(m, p, prog, ic, static=True)
