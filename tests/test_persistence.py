"""Solution files and snapshots on the CPU (the reference's
tests/test_persistence.py covers the same surface through solves; here the
arrays are synthetic, the GPU solves are in test_gpu_persistence.py)."""

import logging
from types import MappingProxyType

import numpy as np
import pytest

from pylcm_amd import examples as ex
from pylcm_amd import persistence as ps
from pylcm_amd.persistence import load_snapshot, load_solution, save_solution


def _solution(dtype=np.float64):
    rng = np.random.default_rng(3)
    V = {}
    for t in range(4):
        V[t] = {"working_life": rng.standard_normal((7, 50)).astype(dtype),
                "retirement": rng.standard_normal(50).astype(dtype)}
    V[4] = {"dead": np.zeros(1, dtype)}
    V[2]["retirement"][3] = -np.inf
    V[1]["working_life"][0, 0] = np.nan
    return MappingProxyType({t: MappingProxyType(V[t]) for t in V})


def _same(a, b):
    assert set(a) == set(b)
    for t in a:
        assert set(a[t]) == set(b[t])
        for r in a[t]:
            x, y = np.asarray(a[t][r]), np.asarray(b[t][r])
            assert x.dtype == y.dtype and x.shape == y.shape
            assert np.array_equal(x, y, equal_nan=True)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_save_and_load_solution_roundtrip(tmp_path, dtype):
    V = _solution(dtype)
    p = save_solution(period_to_regime_to_V_arr=V, path=tmp_path / "solution.h5")
    back = load_solution(path=p)
    _same(V, back)
    assert isinstance(back, MappingProxyType) and isinstance(back[0], MappingProxyType)
    assert list(back) == sorted(back)


def test_save_solution_missing_parent_dir(tmp_path):
    with pytest.raises(FileNotFoundError):
        save_solution(period_to_regime_to_V_arr=_solution(), path=tmp_path / "nope" / "s.h5")


def test_load_solution_rejects_other_files(tmp_path):
    p = tmp_path / "x.h5"
    p.write_bytes(b"not a solution file")
    with pytest.raises(ValueError):
        load_solution(path=p)


def test_hdf5_file_needs_h5py(tmp_path):
    p = tmp_path / "x.h5"
    p.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    if ps._h5py() is None:
        with pytest.raises(ImportError):
            load_solution(path=p)


def test_solve_snapshot_roundtrip_and_retention(tmp_path):
    m, p = ex.deterministic_workload()
    V = _solution()
    dirs = [ps.save_solve_snapshot(model=m, params=p, period_to_regime_to_V_arr=V,
                                   log_path=tmp_path, log_keep_n_latest=2) for _ in range(4)]
    assert [d.name for d in dirs] == [f"solve_snapshot_{k:03d}" for k in (1, 2, 3, 4)]
    left = sorted(d.name for d in tmp_path.iterdir())
    assert left == ["solve_snapshot_003", "solve_snapshot_004"]
    snap = load_snapshot(dirs[-1])
    assert isinstance(snap, ps.SolveSnapshot)
    assert snap.model.n_periods == m.n_periods
    assert list(snap.model.regime_names) == list(m.regime_names)
    assert snap.params == p
    _same(V, snap.period_to_regime_to_V_arr)
    assert (dirs[-1] / "REPRODUCE.md").exists()
    light = load_snapshot(dirs[-1], exclude=["period_to_regime_to_V_arr", "model"])
    assert light.period_to_regime_to_V_arr is None and light.model is None
    assert light.params == p


def test_counter_past_padding_width(tmp_path):
    """Ordering by the parsed counter, not by name (io.py:130-150)."""
    for k in (998, 999):
        (tmp_path / f"solve_snapshot_{k}").mkdir()
    m, p = ex.deterministic_workload()
    d = ps.save_solve_snapshot(model=m, params=p, period_to_regime_to_V_arr=_solution(),
                               log_path=tmp_path, log_keep_n_latest=2)
    assert d.name == "solve_snapshot_1000"
    assert sorted(x.name for x in tmp_path.iterdir()) == ["solve_snapshot_1000", "solve_snapshot_999"]


def test_load_snapshot_warns_on_platform_mismatch(tmp_path, caplog, monkeypatch):
    m, p = ex.deterministic_workload()
    d = ps.save_solve_snapshot(model=m, params=p, period_to_regime_to_V_arr=_solution(),
                               log_path=tmp_path, log_keep_n_latest=3)
    monkeypatch.setattr(ps, "_get_platform", lambda: "fake_arch-FakeOS")
    with caplog.at_level(logging.WARNING, logger="pylcm_amd.persistence"):
        load_snapshot(d)
    assert any("environment may not match" in r.message for r in caplog.records)
