"""Host-side checks of the native solve server (no device): the request files
the lowering workers write parse in the library as the packed request they
came from, malformed files are refused, and the library's tile choice and
task splitting equal the Python autotuner's and server's."""

import ctypes as C
import os
import time

import numpy as np
import pytest

from pylcm_amd import _native as nat
from pylcm_amd import autotune
from pylcm_amd import examples as ex

LCM_E_ARG = -1


def _lib():
    return nat.load_library()


def _inspect(path):
    info = (C.c_int64 * 10)()
    rc = _lib().lcm_req_inspect(str(path).encode(), info, 10)
    return rc, list(info)


def _packed(T=6, g=300, nodes=5, dtype=np.float32):
    from pylcm_amd._batch import pack_problem

    m = ex.retirement_model(T, n_wealth=g, n_consumption=g // 2, income_nodes=nodes)
    p = ex.retirement_params(T, income={"rho": 0.9, "sigma": 0.1, "mu": 0.0} if nodes else None)
    return pack_problem(m, p, dtype, nat.REQ_TILE)


@pytest.mark.parametrize("fetch", ["first", "all", "none"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_request_file_parses_as_written(tmp_path, fetch, dtype):
    from pylcm_amd.serving import write_request_file

    data = _packed(dtype=dtype)
    path = tmp_path / "req"
    fetched = write_request_file(str(path), data, dtype, fetch)
    rc, info = _inspect(path)
    assert rc == 0, nat.host_error()
    outs = data["outs"][0]
    keys = {"first": [k for k in outs if k[0] == 0], "all": sorted(outs), "none": []}[fetch]
    item = np.dtype(dtype).itemsize
    assert [(t, r) for t, r, _, _ in fetched] == keys
    assert [nb for *_, nb in fetched] == [int(np.prod(outs[k].shape)) * item for k in keys]
    code = nat.LCM_DTYPE_F64 if np.dtype(dtype) == np.float64 else nat.LCM_DTYPE_F32
    lds = max(int(_lib().lcm_gs_lds_bytes(code, int(g), int(k))) for g, k in zip(data["lds_g"], data["lds_k"]))
    n_ptr = sum(int(np.asarray(m).sum()) for m in data["masks"].values())
    assert info == [data["n_steps"], data["n_desc"], len(data["tasks_cat"]), len(keys), data["image"].nbytes,
                    data["total_bytes"], sum(nb for *_, nb in fetched), n_ptr, max(data["wdep_by_step"]), lds]
    # the bulk parts sit at the end of the file in the order the placer reads them
    raw = path.read_bytes()
    tail = data["D0"].tobytes() + np.ascontiguousarray(data["tasks_cat"], np.int32).tobytes() + \
        data["image"].tobytes()
    assert raw.endswith(tail)


def test_pointer_masks_flag_exactly_the_relocated_words(tmp_path):
    """Adding a base to the words the file's masks flag gives the records
    the lowering produces for that base."""
    from pylcm_amd._gs_program import DESC_DTYPE, GsProgram
    from pylcm_amd.serving import write_request_file

    data = _packed(T=5, g=200)
    path = tmp_path / "req"
    write_request_file(str(path), data, np.float32, "first")
    raw = path.read_bytes()
    n = data["n_desc"]
    n_steps = data["n_steps"]
    off = 128 + 8 * (4 * n_steps + 2 * sum(1 for k in data["outs"][0] if k[0] == 0))
    masks = np.frombuffer(raw, dtype=np.uint64, count=n, offset=off)
    assert not np.any(masks >> np.uint64(nat.DESC_PTR_WORDS))
    base = 0x7F0000000000
    words = data["D0"].copy().view(np.uint64).reshape(n, -1)
    for w in range(nat.DESC_PTR_WORDS):
        hit = ((masks >> np.uint64(w)) & np.uint64(1)).astype(bool)
        words[hit, w] += np.uint64(base)
    want = GsProgram.records_at(data["D0"], data["masks"], base)
    assert words.tobytes() == want.tobytes()
    assert DESC_DTYPE.itemsize * n == data["D0"].nbytes


def test_malformed_request_files_are_refused(tmp_path):
    from pylcm_amd.serving import write_request_file

    data = _packed(T=4, g=120, nodes=None)
    good = tmp_path / "good"
    write_request_file(str(good), data, np.float32, "first")
    raw = bytearray(good.read_bytes())
    cases = {
        "truncated": bytes(raw[:-7]),
        "magic": b"X" + bytes(raw[1:]),
        "header only": bytes(raw[:100]),
    }
    bad_count = bytearray(raw)
    v = np.frombuffer(bad_count, dtype=np.int64, count=1, offset=128).copy() + 1  # first step's task count
    bad_count[128:136] = v.tobytes()
    cases["task counts"] = bytes(bad_count)
    bad_tile = bytearray(raw)
    bad_tile[72:80] = np.int64(128).tobytes()
    cases["tile"] = bytes(bad_tile)
    for name, blob in cases.items():
        p = tmp_path / name.replace(" ", "_")
        p.write_bytes(blob)
        rc, _ = _inspect(p)
        assert rc == LCM_E_ARG, (name, rc)
        assert nat.host_error(), name
    rc, _ = _inspect(tmp_path / "missing")
    assert rc == LCM_E_ARG and "cannot open" in nat.host_error()


def test_library_tile_choice_equals_the_autotuner():
    table = autotune.load_table()
    rng = np.random.default_rng(0)
    for dts in ("f32", "f64"):
        ent = [e for e in table["entries"] if e["dtype"] == dts]
        u = np.array([e["units"] for e in ent], dtype=np.int64)
        a = np.array([e["n_actions"] for e in ent], dtype=np.int64)
        t = np.array([e["tile"] for e in ent], dtype=np.int32)
        shapes = [(int(x), int(y)) for x, y in zip(rng.integers(1, 1 << 16, 300), rng.integers(1, 1 << 14, 300))]
        shapes += [(e["units"], e["n_actions"]) for e in ent] + [(0, 0), (1, 1)]
        for units, acts in shapes:
            got = _lib().lcm_tile_choose(u.ctypes.data, a.ctypes.data, t.ctypes.data, len(ent), units, acts)
            assert got == autotune.choose(dts, units, acts, table), (dts, units, acts)
    empty = np.zeros(1, dtype=np.int64)
    for units in (1, 256, 257, 2048, 2049, 10 ** 6):
        got = _lib().lcm_tile_choose(empty.ctypes.data, empty.ctypes.data, empty.ctypes.data, 0, units, 7)
        assert got == autotune.fallback_tile(units)


@pytest.mark.parametrize("tile", [64, 128])
def test_library_split_equals_the_server_split(tile):
    from pylcm_amd._gs_program import TASK_DTYPE
    from pylcm_amd.serving import SolveServer

    rng = np.random.default_rng(tile)
    n_w_of_desc = rng.integers(1, 1200, 50).astype(np.int32)
    n = 400
    rec = np.zeros(n, dtype=TASK_DTYPE)
    rec["desc"] = rng.integers(0, 50, n)
    rec["sd"] = rng.integers(0, 9, n)
    rec["w0"] = (rng.integers(0, 5, n) * 256) % np.maximum(n_w_of_desc[rec["desc"]] // 256 * 256, 1)
    want = SolveServer._split(n_w_of_desc, rec, tile)
    out = np.zeros(n * 256 // tile, dtype=TASK_DTYPE)
    m = _lib().lcm_split_tasks(rec.ctypes.data, n, n_w_of_desc.ctypes.data, tile, out.ctypes.data)
    assert m == len(want)
    for f in ("desc", "sd", "w0"):
        assert np.array_equal(out[f][:m], want[f]), f


def test_server_clock_is_the_monotonic_clock():
    """Times the library reports compare with time.perf_counter()."""
    a = time.perf_counter()
    c = _lib().lcm_serve_clock()
    b = time.perf_counter()
    assert a - 1e-3 <= c <= b + 1e-3


def test_native_worker_writes_a_request_file(tmp_path):
    from pylcm_amd.serving import _lower_to_request_file

    m = ex.retirement_model(4, n_wealth=150, n_consumption=80)
    p = ex.retirement_params(4)
    out = _lower_to_request_file((m, p), np.dtype(np.float32).str, str(tmp_path / "r_"), "first")
    assert os.path.exists(out["path"]) and out["worker_s"] > 0
    rc, info = _inspect(out["path"])
    assert rc == 0 and info[3] == len(out["fetched"])


def test_simulation_image_lists_every_pointer(tmp_path):
    """The simulation image built where the solve is lowered: every pointer
    word is listed once under its base, lies inside what that base covers,
    and the EV descriptors point into the solve's descriptor array."""
    from pylcm_amd._gs_program import DESC_DTYPE, GsProgram
    from pylcm_amd._simulate import SimProgram

    m, p = ex.stochastic_workload()
    g = GsProgram.of([(m, p)], np.float32, tile=256, defer=True)
    img = SimProgram(m, p, g).build_image()
    words = img["image"].view(np.uint64)
    seen = set()
    limits = {SimProgram.BASE_TABLES: img["image"].nbytes, SimProgram.BASE_EV: img["ev_bytes"],
              SimProgram.BASE_SOLVE: g.n_desc * DESC_DTYPE.itemsize}
    for kind, offs in img["patch"].items():
        assert np.all(offs % 8 == 0)
        assert seen.isdisjoint(offs.tolist())
        seen |= set(offs.tolist())
        assert np.all(words[offs // 8] < np.uint64(limits[kind])), kind
    n_ev = len(img["ev_keys"])
    assert len(img["patch"][SimProgram.BASE_SOLVE]) == n_ev
    assert np.all(words[img["patch"][SimProgram.BASE_SOLVE] // 8] % np.uint64(DESC_DTYPE.itemsize) == 0)
    assert len(img["patch"][SimProgram.BASE_EV]) >= 3 * n_ev


def test_worker_writes_a_held_request_with_its_simulation(tmp_path):
    import pickle

    from pylcm_amd.serving import _lower_to_request_file

    m = ex.retirement_model(5, n_wealth=120, n_consumption=90, income_nodes=5)
    p = ex.retirement_params(5, income={"rho": 0.9, "sigma": 0.1, "mu": 0.0})
    spec = {"n_agents": 300, "seed": 4, "initial_conditions": ex.serving_initial_conditions}
    out = _lower_to_request_file((m, p, spec), np.dtype(np.float32).str, str(tmp_path / "r_"), "first")
    header = np.frombuffer(open(out["path"], "rb").read(128), dtype=np.int64)
    assert header[10] == nat.REQ_HOLD
    rc, info = _inspect(out["path"])
    assert rc == 0, nat.host_error()
    with open(out["sim_path"], "rb") as f:
        payload = pickle.load(f)
    assert payload["seed"] == 4 and payload["ic"]["_n"] == 300
    assert set(payload["img"]["patch"]) == {0, 1, 2}
    plain = _lower_to_request_file((m, p), np.dtype(np.float32).str, str(tmp_path / "q_"), "first")
    assert "sim_path" not in plain
    assert np.frombuffer(open(plain["path"], "rb").read(128), dtype=np.int64)[10] == 0


def test_simulation_image_same_with_period_sharing(monkeypatch):
    """Periods with equal keys share one simulation lowering; the image and
    its pointer lists are byte-identical to lowering every period."""
    from pylcm_amd._gs_program import GsProgram
    from pylcm_amd._simulate import SimLowering, SimProgram

    m, p = ex.stochastic_workload()
    g = GsProgram.of([(m, p)], np.float32, tile=256, defer=True)
    shared = SimProgram(m, p, g).build_image()
    monkeypatch.setattr(SimLowering, "period_key", lambda self, lay, period: None)
    every = SimProgram(m, p, g).build_image()
    assert shared["image"].tobytes() == every["image"].tobytes()
    assert all(np.array_equal(shared["patch"][k], every["patch"][k]) for k in every["patch"])
    assert shared["lds_max"] == every["lds_max"] and shared["ev_keys"] == every["ev_keys"]
