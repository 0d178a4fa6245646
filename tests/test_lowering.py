"""Host lowering (CPU): lowered tables, executed with the kernel's arithmetic,
reproduce the oracle; the affine tracer and the interval form are exact; the
C ABI library loads and exports every declared symbol."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import gridsearch as og
from pylcm_amd import LinSpacedGrid, Model, Regime, UnsupportedModelError
from pylcm_amd import examples as ex
from pylcm_amd._affine import Affine, Constraint, NonAffine
from pylcm_amd._gs_program import interval_form, lower_problem
from tests import models as tm
from tests.kernel_emulator import emulate
from tests.parity import compare_solution

ROOT = Path(__file__).resolve().parents[1]


def _emulated_matches_oracle(model, params, dtype, taste=()):
    tabs, rg = lower_problem(model, params, dtype)
    V, P = emulate(model, tabs, rg)
    Vo, Po, Qo = og.solve(model, params, grid_dtype=dtype, return_Q=True)
    worst, mism, ties = compare_solution(V, P, Vo, Po, Qo, np.float64, taste)
    assert worst <= 1e-12, worst
    assert mism == 0


def test_deterministic():
    m, p = ex.deterministic_workload()
    _emulated_matches_oracle(m, p, np.float64)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_stochastic(dt):
    m = ex.retirement_model(6, n_wealth=25, n_consumption=30, income_nodes=5)
    p = ex.retirement_params(6, income={"rho": 0.9, "sigma": 0.15, "mu": 0.05})
    _emulated_matches_oracle(m, p, dt)


@pytest.mark.parametrize("dist", sorted(tm.PROCESSES))
def test_processes_model(dist):
    m = tm.processes_model(4, dist)
    _emulated_matches_oracle(m, tm.processes_params(dist), np.float64)


def test_log_and_irregular_grids():
    m = tm.log_grid_model(5)
    _emulated_matches_oracle(m, ex.retirement_params(5, income={"rho": 0.8, "sigma": 0.1, "mu": 0.0}),
                             np.float64)
    m = tm.irregular_grid_model(5)
    _emulated_matches_oracle(m, ex.retirement_params(5), np.float64)


def test_taste_shocks():
    m = tm.taste_shock_model(5)
    p = ex.retirement_params(5, income={"rho": 0.8, "sigma": 0.1, "mu": 0.0})
    p["working_life"]["taste_shocks"] = {"scale": 0.3}
    _emulated_matches_oracle(m, p, np.float64, taste=("working_life",))


def test_nonaffine_functions_are_tabulated():
    m = tm.nonaffine_model(5)
    tabs, _ = lower_problem(m, ex.retirement_params(5, log_utility=True), np.float64)
    pt = tabs[(0, "working_life")]
    assert pt.U.shape[1] == 30 and pt.B.shape[1] == 30  # full state grid
    _emulated_matches_oracle(m, ex.retirement_params(5, log_utility=True), np.float64)


def test_affine_tracer():
    w = Affine(1.0, 0.0)
    e = 1.04 * (w - np.array([1.0, 2.0])) + 3.0
    assert np.allclose(e.a, 1.04) and np.allclose(e.b, [1.04 * -1 + 3, 1.04 * -2 + 3])
    c = np.array([1.0, 5.0]) <= w
    assert isinstance(c, Constraint) and np.allclose(c.terms[0][0], 1) and np.allclose(c.terms[0][1], [-1, -5])
    with pytest.raises(NonAffine):
        np.log(w)
    with pytest.raises(NonAffine):
        _ = w * w
    assert np.allclose(np.where(np.array([True, False]), w, 2 * w).a, [1, 2])


def test_interval_form_is_exact():
    rng = np.random.default_rng(1)
    for dt in (np.float32, np.float64):
        c = rng.uniform(0, 10, size=(1, 1, 2, 50))
        rows = [(np.ones_like(c), -c, False), (-np.ones_like(c), c + 3.0, True)]
        lo, hi = interval_form(rows, 1, 2, 50, dt)
        w = np.concatenate([rng.uniform(-1, 15, 500), c.ravel(), (c + 3).ravel()]).astype(dt)
        for x in w:
            f = np.ones(c.shape, bool)
            for a, b, strict in rows:
                v = (a.astype(dt) * x + b.astype(dt)).astype(dt)
                f &= (v > 0) if strict else (v >= 0)
            assert np.array_equal(f, (lo <= x) & (x <= hi))
    # a coefficient other than -1, 0, 1 has no exact interval form
    assert interval_form([(np.full((1, 1, 1, 3), 0.5), np.zeros((1, 1, 1, 3)), False)], 1, 1, 3, np.float32) is None


def test_unsupported_model_is_refused():
    W = LinSpacedGrid(start=1, stop=2, n_points=3)
    r = Regime(states={"a": W, "b": W}, actions={"c": W}, transition=lambda: 0,
               functions={"utility": lambda c: c}, state_transitions={"a": lambda a: a, "b": lambda b: b})
    from pylcm_amd import categorical

    @categorical
    class Rid:
        only: int

    with pytest.raises(UnsupportedModelError):
        Model(regimes={"only": r}, ages=tm.AgeGrid(start=0, stop=2, step="Y"), regime_id_class=Rid)


def test_library_exports_every_declared_symbol():
    from pylcm_amd import _native

    header = (ROOT / "include" / "pylcm_amd.h").read_text()
    names = set(re.findall(r"^(?:int|int64_t|double|const char\*|size_t)\s+(lcm_\w+)\(", header, re.M))
    assert names, "no declarations found"
    lib = _native.load_library()
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(_native.declared_symbols()) >= names
    a, b = ctypes.c_size_t(), ctypes.c_size_t()
    lib.lcm_abi_sizes(ctypes.byref(a), ctypes.byref(b))
    assert a.value == ctypes.sizeof(_native.LcmGsDesc) and b.value == ctypes.sizeof(_native.LcmGsTask)


def test_struct_layouts_match_the_library():
    """Every ctypes mirror has the size the library was compiled with."""
    from pylcm_amd import _native

    lib = _native.load_library()
    s = [ctypes.c_size_t() for _ in range(4)]
    assert lib.lcm_abi_sizes_sim(*(ctypes.byref(x) for x in s)) == 0
    assert [x.value for x in s] == [ctypes.sizeof(_native.LcmEvDesc), ctypes.sizeof(_native.LcmSimDesc),
                                    ctypes.sizeof(_native.LcmAgents), ctypes.sizeof(_native.LcmSimOut)]
    e = [ctypes.c_size_t() for _ in range(3)]
    assert lib.lcm_abi_sizes_egm(*(ctypes.byref(x) for x in e)) == 0
    assert [x.value for x in e] == [ctypes.sizeof(_native.LcmEgmRow), ctypes.sizeof(_native.LcmEgmCol),
                                    ctypes.sizeof(_native.LcmEgmAgg)]


def test_missing_library_fails_loudly(tmp_path):
    from pylcm_amd import _native
    from pylcm_amd.exceptions import NativeKernelError

    with pytest.raises(NativeKernelError, match="missing"):
        _native.load_library(tmp_path / "liblcm_amd.so")


def test_solve_without_a_device_raises_instead_of_falling_back():
    """No CPU execution path: where the device cannot be initialised (this
    suite's hosts have no GPU) a solve raises NativeKernelError."""
    from pylcm_amd import _native
    from pylcm_amd import examples as ex
    from pylcm_amd.exceptions import NativeKernelError

    if _native.load_library().lcm_init(0) == 0:
        pytest.skip("a GPU is present")
    m, p = ex.deterministic_workload()
    with pytest.raises(NativeKernelError):
        m.solve(params=p, log_level="off")


def test_descriptor_pointers_address_the_right_outputs():
    """Every packed descriptor writes its own (problem, period, regime) output
    and reads the next period's outputs of its targets; table pointers lie
    inside the arena. Checked on the host before any launch."""
    from pylcm_amd import examples as ex
    from pylcm_amd import _native as nat
    from pylcm_amd._gs_program import GsProgram

    problems = []
    for T, n, inc in ((6, 30, 5), (4, 45, None), (6, 30, 5)):
        m = ex.retirement_model(T, n_wealth=n, n_consumption=n + 7, income_nodes=inc)
        inc_p = {"rho": 0.9, "sigma": 0.1, "mu": 0.0} if inc else None
        problems.append((m, ex.retirement_params(T, crra=1.5 + 0.1 * T, income=inc_p)))
    prog = GsProgram.of(problems, np.float32, defer=True)
    base = 1 << 40
    prog.base = base
    D = prog._desc_records()
    assert len(D) == prog.n_desc
    for (p, t, r), di in prog.desc_of.items():
        o = prog.outs[p][(t, r)]
        assert D["V_out"][di] == base + o.v_off and D["pol_out"][di] == base + o.p_off
        pt = prog.problems[p][(t, r)]
        assert D["n_targets"][di] == len(pt.targets)
        for k, (name, _) in enumerate(pt.targets):
            assert D["Vn"][di][k] == base + prog.outs[p][(t + 1, name)].v_off
        for k, (name, _) in enumerate(pt.scalars):
            if name is not None:
                assert D["Vs"][di][k] == base + prog.outs[p][(t + 1, name)].v_off
        for f in ("U", "A", "B", "Af", "Bf", "Flo", "Fhi", "Fchunk", "g_pts", "w_pts"):
            assert base <= int(D[f][di]) < base + prog.total_bytes, (f, p, t, r)
    # relocated records differ from these by the base only
    D0, masks = prog.relocatable_records()
    assert np.array_equal(GsProgram.records_at(D0, masks, base), D)
