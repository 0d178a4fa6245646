"""Host-side pieces of the solve server (CPU): the device-slab allocator."""

from pylcm_amd.serving import _Slab


def test_slab_first_fit_and_coalescing():
    s = _Slab(10 * 256)
    a = s.alloc(300)      # 512 bytes
    b = s.alloc(256)
    c = s.alloc(1000)     # 1024
    assert a == (0, 512) and b == (512, 256) and c == (768, 1024)
    assert s.alloc(2000) is None
    s.release(*b)
    assert s.alloc(256) == (512, 256)
    s.release(512, 256)
    s.release(*a)
    s.release(*c)
    assert s.free == [(0, 2560)] and s.used == 0
    assert s.alloc(2560) == (0, 2560)


def test_descriptor_slots_are_whole_records():
    """Descriptor slots must start at multiples of the record size: a task
    names its descriptor by index into the pool."""
    from pylcm_amd._gs_program import DESC_DTYPE

    rec = DESC_DTYPE.itemsize
    s = _Slab(100 * rec, unit=rec)
    offs = [s.alloc(k * rec)[0] for k in (3, 7, 1, 5)]
    assert all(o % rec == 0 for o in offs)
    assert offs == [0, 3 * rec, 10 * rec, 11 * rec]


def test_fused_tasks_address_each_requests_own_descriptors():
    """Byte-level rehearsal of the descriptor pool: records copied at the slot
    offsets the server allocates, read back by the index every task carries,
    point into the request's own device image and cover the task's state."""
    import numpy as np

    from pylcm_amd import examples as ex
    from pylcm_amd._gs_program import DESC_DTYPE, GsProgram

    rng = np.random.default_rng(5)
    reqs = []
    for _ in range(12):
        T, g = int(rng.integers(4, 9)), int(rng.integers(30, 300))
        stoch = bool(rng.integers(0, 2))
        reqs.append((ex.retirement_model(T, n_wealth=g, n_consumption=g, income_nodes=3 if stoch else None),
                     ex.retirement_params(T, income={"rho": 0.9, "sigma": 0.1, "mu": 0.0} if stoch else None)))
    rec = DESC_DTYPE.itemsize
    slab, dslab = _Slab(1 << 30), _Slab(4096 * rec, unit=rec)
    pool = np.zeros(4096 * rec, dtype=np.uint8)
    fake = 1 << 40
    placed = []
    for m, p in reqs:
        prog = GsProgram.of([(m, p)], np.float32, tile=128, defer=True)
        sl, ds = slab.alloc(prog.total_bytes), dslab.alloc(prog.n_desc * rec)
        prog.base = fake + sl[0]
        raw = np.frombuffer(prog._desc_records(prog._descs).tobytes(), dtype=np.uint8)
        pool[ds[0]: ds[0] + raw.size] = raw
        placed.append((prog, sl, ds[0] // rec))
    table = np.frombuffer(pool.tobytes(), dtype=DESC_DTYPE)
    for prog, sl, db in placed:
        lo, hi = fake + sl[0], fake + sl[0] + sl[1]
        for tasks in prog.tasks_host:
            if len(tasks) == 0:
                continue
            r = table[tasks[:, 0] + db]
            assert np.all((r["V_out"] >= lo) & (r["V_out"] < hi))
            assert np.all((r["U"] >= lo) & (r["U"] < hi))
            assert np.all(tasks[:, 1] < r["n_sd"]) and np.all(tasks[:, 2] < r["n_w"])


def test_worker_process_lowering_matches_in_process():
    """Lowering in a spawned worker (the model travels by cloudpickle) gives
    the same device image, relocatable descriptor records and task lists as
    lowering in the server's own process."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor

    import cloudpickle
    import numpy as np

    from pylcm_amd import examples as ex
    from pylcm_amd._gs_program import GsProgram, lower_problem
    from pylcm_amd.serving import _lower_in_worker

    make = ex.serving_request_factory(seed=3, horizons=(4, 6), grids=(40, 300))
    reqs = [make(i) for i in range(3)]
    with ProcessPoolExecutor(max_workers=2, mp_context=multiprocessing.get_context("spawn")) as pool:
        outs = list(pool.map(_lower_in_worker, [cloudpickle.dumps(r) for r in reqs],
                             [np.dtype(np.float32).str] * 3, [128] * 3))
    for (m, p), got in zip(reqs, outs):
        prog = GsProgram.of([(m, p)], np.float32, lowered=[lower_problem(m, p, np.float32)], tile=128, defer=True)
        D0, masks = prog.relocatable_records()
        assert got["image"].tobytes() == prog._arena.image(prog.out_start).tobytes()
        assert got["D0"].tobytes() == D0.tobytes()
        assert sorted(got["masks"]) == sorted(masks)
        assert all(np.array_equal(got["masks"][f], masks[f]) for f in masks)
        assert len(got["tasks_host"]) == len(prog.tasks_host)
        assert all(np.array_equal(a, b) for a, b in zip(got["tasks_host"], prog.tasks_host))
        assert (got["total_bytes"], got["out_start"], got["n_desc"], got["n_steps"]) == (
            prog.total_bytes, prog.out_start, prog.n_desc, prog.n_steps)
        assert got["outs"][0].keys() == prog.outs[0].keys()


def test_split_tasks_are_the_64_state_tiling():
    """With the tile chosen per launch the server splits a request's
    256-state tasks into their 64-state quarters for small launches: the same
    task set as lowering with tile 64."""
    from types import SimpleNamespace

    import numpy as np

    from pylcm_amd import examples as ex
    from pylcm_amd._gs_program import TASK_DTYPE, GsProgram
    from pylcm_amd.serving import SolveServer

    for T, g, nodes in ((5, 700, 3), (4, 257, None), (6, 64, 5), (3, 1000, None)):
        m = ex.retirement_model(T, n_wealth=g, n_consumption=50, income_nodes=nodes)
        p = ex.retirement_params(T, income={"rho": 0.9, "sigma": 0.1, "mu": 0.0} if nodes else None)
        fine = GsProgram.of([(m, p)], np.float32, tile=64, defer=True)
        coarse = GsProgram.of([(m, p)], np.float32, tile=256, defer=True)
        D0, _ = coarse.relocatable_records()
        base = 17  # the request's first descriptor slot in the server's table
        req = SimpleNamespace(desc_base=base, n_w=np.asarray(D0["n_w"], dtype=np.int64))
        assert fine.n_steps == coarse.n_steps
        for tf, tc in zip(fine.tasks_host, coarse.tasks_host):
            want = {(int(d) + base, int(s_), int(w)) for d, s_, w in np.asarray(tf)} if len(tf) else set()
            got = set()
            if len(tc):
                rec = np.zeros(len(tc), dtype=TASK_DTYPE)
                rec["desc"], rec["sd"], rec["w0"] = tc[:, 0] + base, tc[:, 1], tc[:, 2]
                out = SolveServer._split(req, rec)
                got = {(int(a["desc"]), int(a["sd"]), int(a["w0"])) for a in out}
            assert got == want


def test_grid_above_lds_capacity_is_a_clear_error():
    """A next-state grid whose EV row does not fit the workgroup's LDS is
    refused with UnsupportedModelError before anything is launched."""
    import numpy as np
    import pytest

    from pylcm_amd import examples as ex
    from pylcm_amd._gs_program import GsProgram
    from pylcm_amd.exceptions import UnsupportedModelError

    m = ex.retirement_model(3, n_wealth=12000, n_consumption=20)
    p = ex.retirement_params(3)
    prog = GsProgram.of([(m, p)], np.float64, tile=256, defer=True)
    with pytest.raises(UnsupportedModelError, match="LDS"):
        prog.lds_by_step
    ok = GsProgram.of([(ex.retirement_model(3, n_wealth=2000, n_consumption=20), p)], np.float64,
                      tile=256, defer=True)
    assert max(ok.lds_by_step) <= 160 * 1024
