"""GPU tests of the tile autotuner: every tile gives bit-identical V and
policies on a batch of uneven grids, the tuned tile included, and the tuner
measures a launch shape on the device."""

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
