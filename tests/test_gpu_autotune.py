"""GPU tests of the tile autotuner and of the kernel builds chosen per
launch: every tile and both register budgets give bit-identical V and
policies, the tuned tile included, and the tuner measures a launch shape on
the device."""

import numpy as np
import pytest

from pylcm_amd import autotune
from pylcm_amd import examples as ex
from pylcm_amd._gs_program import GsProgram, lower_problem

pytestmark = pytest.mark.gpu


def test_every_tile_gives_the_same_bits(native):
    draws = ex.draw_batch_params(6, n_periods=6, seed=5)
    models = {n: ex.retirement_model(6, n_wealth=n, n_consumption=n, income_nodes=7) for n in (200, 500, 1000)}
    lowered = [lower_problem(models[s], p, np.float32) for s, p in draws]
    probs = [(models[s], p) for s, p in draws]
    outs = {}
    for tile in (None, 64, 128, 256):
        prog = GsProgram.of(probs, np.float32, lowered=lowered, tile=tile)
        prog.run(use_graph=False)
        outs[tile if tile else f"tuned {prog.tile}"] = prog.fetch().copy()
    ref = outs[256]
    for k, v in outs.items():
        assert np.array_equal(v, ref), k


def test_tuner_measures_a_shape(native):
    e = autotune.measure_shape(200, 2, np.float32, periods=4, reps=1)
    assert e["tile"] in autotune.TILES
    assert set(e["ms_per_launch"]) == {"64", "128", "256"}
    assert all(v > 0 for v in e["ms_per_launch"].values())
    print(e)


@pytest.mark.parametrize("tile", [64, 128, 256])
def test_both_register_budgets_give_the_same_bits(native, tile):
    """The float32 kernel's two builds (8 waves per SIMD, and unconstrained)
    give bit-identical V and policies on every kernel variant: the fast one
    (linear grids, interval constraints), the staged one (an irregular grid)
    and the general one (tables over the state grid)."""
    from tests import models as tm

    cases = [
        (ex.retirement_model(5, n_wealth=300, n_consumption=260, income_nodes=5),
         ex.retirement_params(5, income={"rho": 0.9, "sigma": 0.1, "mu": 0.0})),
        (tm.irregular_grid_model(5), ex.retirement_params(5)),
        (tm.nonaffine_model(5), ex.retirement_params(5, log_utility=True)),
    ]
    variants = set()
    for m, p in cases:
        prog = GsProgram(m, [p], np.float32, tile=tile)
        variants |= set(prog.wdep_by_step)
        outs = {}
        for occ in ("sparse", "dense", "auto"):
            prog.occupancy = occ
            prog.out_host[:] = 0
            prog.run(use_graph=False)
            outs[occ] = prog.fetch().copy()
        assert np.array_equal(outs["dense"], outs["sparse"]), (m, tile)
        assert np.array_equal(outs["auto"], outs["sparse"]), (m, tile)
    assert variants == {0, 1, 2}


def test_conflicting_register_budget_flags_are_refused(native):
    import ctypes as C

    from pylcm_amd import _native as nat
    from pylcm_amd.exceptions import NativeKernelError

    m = ex.retirement_model(3, n_wealth=50, n_consumption=40)
    prog = GsProgram(m, [ex.retirement_params(3)], np.float32)
    prog.run(use_graph=False)
    s = next(k for k in range(prog.n_steps) if prog.n_tasks[k])
    st = nat.lib().lcm_gs_launch_ex(
        prog.ctx.handle, nat.LCM_DTYPE_F32, C.c_void_p(prog.base + prog.desc_off),
        C.c_void_p(prog.base + prog.task_off[s]), prog.n_tasks[s], prog.tile,
        C.c_size_t(prog.lds_by_step[s]), int(prog.wdep_by_step[s]) | nat.GS_OCC_DENSE | nat.GS_OCC_SPARSE)
    with pytest.raises(NativeKernelError, match="register budgets"):
        nat._check(st, "lcm_gs_launch_ex")
