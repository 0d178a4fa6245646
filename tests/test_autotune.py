"""The tile autotuner's table lookup (CPU): nearest measured launch shape of the
same dtype, the fallback without entries, and the committed table's schema."""

import json

import pytest

from pylcm_amd import autotune

TABLE = {"entries": [
    {"dtype": "f32", "units": 4, "n_actions": 400, "tile": 64},
    {"dtype": "f32", "units": 64, "n_actions": 1000, "tile": 128},
    {"dtype": "f32", "units": 4096, "n_actions": 2000, "tile": 256},
    {"dtype": "f64", "units": 4, "n_actions": 400, "tile": 128},
]}


def test_nearest_shape_of_the_same_dtype():
    assert autotune.choose("f32", 3, 500, TABLE) == 64
    assert autotune.choose("f32", 80, 900, TABLE) == 128
    assert autotune.choose("f32", 10_000, 4000, TABLE) == 256
    # f64 only has a small shape measured: it is the nearest for every launch
    assert autotune.choose("f64", 10_000, 4000, TABLE) == 128


def test_fallback_without_entries():
    empty = {"entries": []}
    assert autotune.choose("f32", 10, 100, empty) == autotune.fallback_tile(10) == 64
    assert autotune.choose("f32", 1000, 100, empty) == 128
    assert autotune.choose("f32", 100_000, 100, empty) == 256


def test_committed_table_schema():
    if not autotune.TABLE_PATH.exists():
        pytest.skip("no tuned table committed")
    table = json.loads(autotune.TABLE_PATH.read_text())
    assert table["entries"]
    for e in table["entries"]:
        assert e["dtype"] in ("f32", "f64")
        assert e["tile"] in autotune.TILES
        assert e["units"] >= 1 and e["n_actions"] >= 1
        times = {int(k): v for k, v in e["ms_per_launch"].items()}
        assert set(times) == set(autotune.TILES)
        assert times[e["tile"]] == min(times.values())
