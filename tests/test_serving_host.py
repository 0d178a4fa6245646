"""Host-side pieces of the solve server (CPU): the device-slab allocator."""

import pytest

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
        raw = np.frombuffer(prog._desc_records().tobytes(), dtype=np.uint8)
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


@pytest.mark.parametrize("tile", [64, 128])
def test_split_tasks_are_the_finer_tiling(tile):
    """With the tile chosen per launch the server splits a request's
    256-state tasks into their 64- or 128-state parts for launches that run a
    finer tile: the same task set as lowering with that tile."""

    import numpy as np

    from pylcm_amd import examples as ex
    from pylcm_amd._gs_program import TASK_DTYPE, GsProgram
    from pylcm_amd.serving import SolveServer

    for T, g, nodes in ((5, 700, 3), (4, 257, None), (6, 64, 5), (3, 1000, None)):
        m = ex.retirement_model(T, n_wealth=g, n_consumption=50, income_nodes=nodes)
        p = ex.retirement_params(T, income={"rho": 0.9, "sigma": 0.1, "mu": 0.0} if nodes else None)
        fine = GsProgram.of([(m, p)], np.float32, tile=tile, defer=True)
        coarse = GsProgram.of([(m, p)], np.float32, tile=256, defer=True)
        D0, _ = coarse.relocatable_records()
        base = 17  # the request's first descriptor slot in the server's table
        n_w_of_desc = np.zeros(base + len(D0), dtype=np.int32)
        n_w_of_desc[base:] = D0["n_w"]
        assert fine.n_steps == coarse.n_steps
        for tf, tc in zip(fine.tasks_host, coarse.tasks_host):
            want = {(int(d) + base, int(s_), int(w)) for d, s_, w in np.asarray(tf)} if len(tf) else set()
            got = set()
            if len(tc):
                rec = np.zeros(len(tc), dtype=TASK_DTYPE)
                rec["desc"], rec["sd"], rec["w0"] = tc[:, 0] + base, tc[:, 1], tc[:, 2]
                out = SolveServer._split(n_w_of_desc, rec, tile)
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


def test_assembled_batch_layout():
    """Draws packed separately and assembled into one program: every task
    addresses a descriptor of its own draw, whose pointers lie in that draw's
    region; step s of every draw is its period T - 1 - s."""
    import numpy as np

    from pylcm_amd import examples as ex
    from pylcm_amd._batch import AssembledBatch, pack_problem
    from pylcm_amd._gs_program import GsProgram

    draws = []
    for T, n, inc in ((5, 40, 3), (3, 60, None), (5, 35, 3)):
        m = ex.retirement_model(T, n_wealth=n, n_consumption=n, income_nodes=inc)
        draws.append((m, ex.retirement_params(T, income={"rho": 0.9, "sigma": 0.1, "mu": 0.0} if inc else None)))
    packed = [pack_problem(m, p, np.float32, 128) for m, p in draws]
    ab = AssembledBatch(packed, np.float32, 128, base=1 << 40)
    D = ab.records()
    assert len(D) == ab.n_desc == sum(p["n_desc"] for p in packed)
    owner = np.repeat(np.arange(len(packed)), [p["n_desc"] for p in packed])
    for s, arr in enumerate(ab.tasks):
        real = arr[arr["desc"] >= 0]
        assert len(arr) % 8 == 0
        # exactly the problems' blocks in problem order, then desc -1 padding
        blocks = [np.asarray(p["tasks_host"][s]) + [int(ab.desc_base[i]), 0, 0]
                  for i, p in enumerate(packed) if s < p["n_steps"] and len(p["tasks_host"][s])]
        expect_rows = np.concatenate(blocks) if blocks else np.zeros((0, 3), dtype=np.int64)
        got_rows = np.stack([arr["desc"], arr["sd"], arr["w0"]], axis=1)
        assert np.array_equal(got_rows[: len(expect_rows)], expect_rows), s
        assert np.all(arr["desc"][len(expect_rows):] == -1) and len(arr) - len(expect_rows) < 8
        for i, p in enumerate(packed):
            mine = real[owner[real["desc"]] == i]
            expect = len(p["tasks_host"][s]) if s < p["n_steps"] else 0
            assert len(mine) == expect, (s, i)
            if not expect:
                continue
            olo = ab.base + int(ab.out_offs[i])
            tlo = ab.base + int(ab.tab_offs[i])
            r = D[mine["desc"]]
            assert np.all((r["V_out"] >= olo) & (r["V_out"] < olo + p["out_start"]))
            assert np.all((r["U"] >= tlo) & (r["U"] < tlo + p["image"].size))
            assert np.all(mine["sd"] < r["n_sd"]) and np.all(mine["w0"] < r["n_w"])
    # a draw's records are its single-program records with outputs and tables moved
    prog = GsProgram.of([draws[1]], np.float32, tile=128, defer=True)
    prog.base = 0
    single = prog._desc_records()
    start = int(ab.desc_base[1])
    mine = D[start: start + prog.n_desc]
    out0 = ab.base + int(ab.out_offs[1])
    tab0 = ab.base + int(ab.tab_offs[1]) - prog.out_start
    assert np.array_equal(mine["V_out"], single["V_out"] + np.uint64(out0))
    assert np.array_equal(mine["U"], single["U"] + np.uint64(tab0))
    assert np.array_equal(mine["beta"], single["beta"]) and np.array_equal(mine["n_w"], single["n_w"])


def test_assembled_batch_mixes_problems_with_different_pointer_fields():
    """A deterministic problem (no income process) next to stochastic ones:
    records relocate every pointer whatever fields each problem uses."""
    import numpy as np

    from pylcm_amd import examples as ex
    from pylcm_amd._batch import AssembledBatch, pack_problem
    from pylcm_amd._gs_program import GsProgram

    draws = [(ex.retirement_model(4, n_wealth=30, n_consumption=30), ex.retirement_params(4)),
             (tm_model := ex.retirement_model(3, n_wealth=20, n_consumption=25, income_nodes=3),
              ex.retirement_params(3, income={"rho": 0.8, "sigma": 0.1, "mu": 0.0}))]
    packed = [pack_problem(m, p, np.float64, 64) for m, p in draws]
    ab = AssembledBatch(packed, np.float64, 64, base=1 << 40)
    D = ab.records()
    for i, (m, p) in enumerate(draws):
        prog = GsProgram.of([(m, p)], np.float64, tile=64, defer=True)
        prog.base = 0
        single = prog._desc_records()
        mine = D[int(ab.desc_base[i]): int(ab.desc_base[i]) + prog.n_desc]
        for f in ("V_out", "pol_out", "U", "A", "Fchunk"):
            delta = (mine[f].astype(np.int64) - single[f].astype(np.int64))
            lo = int(ab.out_offs[i]) if f in ("V_out", "pol_out") else int(ab.tab_offs[i]) - prog.out_start
            assert np.all(delta == ab.base + lo), f
        for k in range(4):
            got, ref = mine["Vn"][:, k].astype(np.int64), single["Vn"][:, k].astype(np.int64)
            has = ref != 0
            assert np.all(got[has] - ref[has] == ab.base + int(ab.out_offs[i]))
    assert tm_model.n_periods == 3


def test_worker_result_spill_round_trip(tmp_path):
    """A worker's bulky arrays written to a spill file and read back equal the
    pickled ones (table image, descriptor records, per-step task rows), and
    the file is removed."""
    import cloudpickle
    import numpy as np

    from pylcm_amd import examples as ex
    from pylcm_amd.serving import _lower_in_worker, _unspill

    m = ex.retirement_model(7, n_wealth=300, n_consumption=300, income_nodes=5)
    p = ex.retirement_params(7, income={"rho": 0.9, "sigma": 0.1, "mu": 0.0})
    blob = cloudpickle.dumps((m, p))
    plain = _lower_in_worker(blob, "<f4", 256)
    spilled = _lower_in_worker(blob, "<f4", 256, str(tmp_path / "req_"))
    assert "image" not in spilled and len(list(tmp_path.iterdir())) == 1
    image = _unspill(spilled)
    assert list(tmp_path.iterdir()) == []
    assert np.array_equal(image, plain["image"].view(np.uint8).ravel())
    assert spilled["D0"].tobytes() == plain["D0"].tobytes()
    assert len(spilled["tasks_host"]) == len(plain["tasks_host"])
    for a, b in zip(spilled["tasks_host"], plain["tasks_host"]):
        assert np.array_equal(a, b)


def test_fresh_batch_workers_hand_images_over_in_files():
    """The fresh-batch worker pool packs chunks in spawned processes and hands
    the table images back through memory-backed files: every packed problem
    equals packing it in this process, and no file is left behind, also for
    chunks that were packed but never gathered."""
    import glob

    import numpy as np

    from pylcm_amd import examples as ex
    from pylcm_amd._batch import WorkerPool, pack_problem

    T = 6
    models = {n: ex.retirement_model(T, n_wealth=n, n_consumption=n, income_nodes=3) for n in (20, 40)}
    draws = [(20 if i % 2 else 40, p) for i, (_, p) in enumerate(ex.draw_batch_params(5, n_periods=T, seed=4))]
    pool = WorkerPool(models, 2, np.float32, 128, chunk=2)
    try:
        got = pool.gather(pool.submit(draws))
        pool.submit(draws[:3])  # left for close() to clean up
    finally:
        pool.close()
    assert glob.glob(glob.escape(pool.spill) + "*") == []
    assert len(got) == len(draws)
    for (k, p), g in zip(draws, got):
        ref = pack_problem(models[k], p, np.float32, 128)
        assert "spilled" not in g
        assert np.asarray(g["image"]).tobytes() == ref["image"].tobytes()
        assert g["D0"].tobytes() == ref["D0"].tobytes()
        assert np.array_equal(g["ptr_cls"], ref["ptr_cls"]) and np.array_equal(g["tasks_cat"], ref["tasks_cat"])
