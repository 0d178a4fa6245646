"""Fresh-draw batches: draws lowered and packed in worker processes and
assembled into one device program give, bit for bit, the results of each
draw solved in its own program, batch after batch (the third batch needs a
larger slot than the first, so its buffers grow)."""

import numpy as np
import pytest

from pylcm_amd import examples as ex
from pylcm_amd._batch import FreshBatches
from pylcm_amd._gs_program import GsProgram

pytestmark = pytest.mark.gpu


def test_fresh_batches_equal_single_programs(native):
    T = 8
    models = {n: ex.retirement_model(T, n_wealth=n, n_consumption=n, income_nodes=7) for n in (40, 90, 160)}
    batches = []
    for k in range(4):
        draws = ex.draw_batch_params(6, n_periods=T, seed=100 + k)
        big = 160 if k == 2 else 90
        batches.append([(40 if i % 2 else big, p) for i, (_, p) in enumerate(draws)])
    fb = FreshBatches(models, np.float32, procs=2, tile=256, chunk=2)
    seen = []

    def check(k, ab):
        for i, (key, p) in enumerate(batches[k]):
            got = ab.results(i)
            single = GsProgram(models[key], [p], np.float32, tile=256)
            single.run(use_graph=False)
            single.fetch()
            ref = single.results(0)
            for t in ref:
                for r in ref[t]:
                    assert np.array_equal(got[t][r][0], ref[t][r][0]), (k, i, t, r)
                    assert np.array_equal(got[t][r][1], ref[t][r][1]), (k, i, t, r)
        seen.append(k)

    try:
        fb.run(batches, on_done=check)
    finally:
        fb.close()
    assert seen == [0, 1, 2, 3]
