"""Solution files and diagnostic snapshots (`src/lcm/persistence.py:1-220`).

Public surface as in the reference: `save_solution` / `load_solution` for the
`period -> regime -> V` mapping a solve returns (and `simulate` accepts), the
snapshot dataclasses `SolveSnapshot` / `SimulateSnapshot`, and
`load_snapshot` for the directories `Model.solve` / `Model.simulate` write
when `log_path` is set.

Value arrays are stored as datasets `/{period}/{regime}/V_arr`, the layout of
the reference's HDF5 writer (`src/_lcm/persistence/io.py:41-53`). With `h5py`
importable the file is that HDF5 file. Without it the same dataset names go
into an uncompressed NumPy `.npz` archive (a logged warning says so); the
loaders tell the two formats apart by their magic bytes, so a file keeps
loading wherever it was written. Model, params, initial conditions and the
simulation result of a snapshot are cloudpickle files next to the arrays, as
in the reference (`src/_lcm/persistence/snapshots.py:52-146`): load snapshots
only from directories you trust, since unpickling runs code.
"""

from __future__ import annotations

import contextlib
import json
import logging
import pickle
import platform
import shutil
import tempfile
from collections.abc import Mapping, Sequence
from dataclasses import dataclass
from pathlib import Path
from types import MappingProxyType
from typing import Any

import numpy as np

_LOG = logging.getLogger(__name__)

_HDF5_MAGIC = b"\x89HDF\r\n\x1a\n"
_ZIP_MAGIC = b"PK\x03\x04"


def _h5py():
    try:
        import h5py
    except ImportError:
        return None
    return h5py


def _get_platform() -> str:
    """`"{machine}-{system}"`, e.g. `"x86_64-Linux"` (`io.py:22-24`)."""
    return f"{platform.machine()}-{platform.system()}"


def _freeze(tree: dict[int, dict[str, np.ndarray]]):
    return MappingProxyType({t: MappingProxyType(dict(tree[t])) for t in sorted(tree)})


def _dataset_items(period_to_regime_to_V_arr: Mapping):
    for period, regimes in period_to_regime_to_V_arr.items():
        for name, arr in regimes.items():
            yield f"{int(period)}/{name}/V_arr", np.asarray(arr)


def _save_arrays(path: Path, period_to_regime_to_V_arr: Mapping) -> None:
    h5py = _h5py()
    if h5py is not None:
        with h5py.File(path, "w") as fh:
            for key, arr in _dataset_items(period_to_regime_to_V_arr):
                fh.create_dataset(key, data=arr)
        return
    _LOG.warning("h5py is not installed: %s is written as a NumPy .npz archive with the "
                 "HDF5 dataset names", path)
    items = dict(_dataset_items(period_to_regime_to_V_arr))
    tmp = None
    try:
        with tempfile.NamedTemporaryFile(mode="wb", dir=path.parent, delete=False) as fh:
            tmp = Path(fh.name)
            np.savez(fh, **items)
        tmp.replace(path)
        tmp = None
    finally:
        if tmp is not None:
            with contextlib.suppress(OSError):
                tmp.unlink()


def _load_arrays(path: Path):
    with path.open("rb") as fh:
        head = fh.read(8)
    out: dict[int, dict[str, np.ndarray]] = {}
    if head == _HDF5_MAGIC:
        h5py = _h5py()
        if h5py is None:
            raise ImportError(f"{path} is an HDF5 file; reading it needs h5py")
        with h5py.File(path, "r") as fh:
            for pk in fh:
                out[int(pk)] = {name: np.asarray(fh[pk][name]["V_arr"][()]) for name in fh[pk]}
        return _freeze(out)
    if head[:4] == _ZIP_MAGIC:
        with np.load(path, allow_pickle=False) as z:
            for key in z.files:
                pk, name, leaf = key.rsplit("/", 2)
                if leaf != "V_arr":
                    raise ValueError(f"{path}: unexpected dataset {key!r}")
                out.setdefault(int(pk), {})[name] = z[key]
        return _freeze(out)
    raise ValueError(f"{path} is neither an HDF5 file nor a .npz archive")


def save_solution(*, period_to_regime_to_V_arr: Mapping, path: str | Path) -> Path:
    """Write value function arrays (`src/lcm/persistence.py:185-210`).

    Raises FileNotFoundError when the parent directory does not exist.
    """
    p = Path(path)
    if not p.parent.is_dir():
        raise FileNotFoundError(f"Parent directory does not exist: {p.parent}")
    _save_arrays(p, period_to_regime_to_V_arr)
    return p


def load_solution(*, path: str | Path):
    """Read what `save_solution` wrote (`src/lcm/persistence.py:213-220`):
    an immutable `period -> regime -> V` mapping of NumPy arrays."""
    return _load_arrays(Path(path))


# ----------------------------------------------------------------- snapshots
@dataclass(frozen=True)
class SolveSnapshot:
    """A solve run for offline reconstruction (`src/lcm/persistence.py:62-76`)."""

    model: Any
    params: Any
    period_to_regime_to_V_arr: Any
    platform: str


@dataclass(frozen=True)
class SimulateSnapshot:
    """A simulate run for offline reconstruction (`src/lcm/persistence.py:79-99`)."""

    model: Any
    params: Any
    initial_conditions: Any
    period_to_regime_to_V_arr: Any
    result: Any
    platform: str


def _atomic_pickle(obj, path: Path) -> None:
    """All-or-nothing cloudpickle dump (`io.py:163-197`)."""
    import cloudpickle

    tmp = None
    try:
        with tempfile.NamedTemporaryFile(mode="wb", dir=path.parent, delete=False) as fh:
            tmp = Path(fh.name)
            cloudpickle.dump(obj, fh, protocol=pickle.HIGHEST_PROTOCOL)
        tmp.replace(path)
        tmp = None
    finally:
        if tmp is not None:
            with contextlib.suppress(OSError):
                tmp.unlink()


def _counter(entry: Path, prefix: str) -> int:
    try:
        return int(entry.name.removeprefix(f"{prefix}_"))
    except ValueError:
        return -1


def _next_dir(parent: Path, prefix: str) -> Path:
    """`{prefix}_{n:03d}` with n one above the largest existing counter (`io.py:120-127`)."""
    parent.mkdir(parents=True, exist_ok=True)
    counters = [c for e in parent.glob(f"{prefix}_*/") if (c := _counter(e, prefix)) >= 0]
    d = parent / f"{prefix}_{max(counters, default=0) + 1:03d}"
    d.mkdir()
    return d


def _enforce_retention(parent: Path, prefix: str, keep_n_latest: int) -> None:
    """Keep the `keep_n_latest` highest counters (`io.py:130-150`)."""
    dirs = sorted((e for e in parent.glob(f"{prefix}_*/") if _counter(e, prefix) >= 0),
                  key=lambda e: _counter(e, prefix))
    for d in dirs[: max(0, len(dirs) - keep_n_latest)]:
        shutil.rmtree(d)


def _finish(snap: Path, kind: str, fields: list[str]) -> None:
    meta = {"snapshot_type": kind, "platform": _get_platform(), "fields": fields}
    (snap / "metadata.json").write_text(json.dumps(meta, indent=2), encoding="utf-8")
    (snap / "REPRODUCE.md").write_text(
        "# Reproducing this run\n\n```python\nfrom pylcm_amd.persistence import load_snapshot\n"
        f"snapshot = load_snapshot({snap.name!r})\n```\n\nPlatform: {_get_platform()}\n",
        encoding="utf-8")


def save_solve_snapshot(*, model, params, period_to_regime_to_V_arr, log_path, log_keep_n_latest: int) -> Path:
    """Write `solve_snapshot_NNN/` (`src/_lcm/persistence/snapshots.py:52-90`)."""
    parent = Path(log_path)
    snap = _next_dir(parent, "solve_snapshot")
    _atomic_pickle(model, snap / "model.pkl")
    _atomic_pickle(params, snap / "params.pkl")
    _save_arrays(snap / "arrays.h5", period_to_regime_to_V_arr)
    _finish(snap, "solve", ["model", "params"])
    _enforce_retention(parent, "solve_snapshot", log_keep_n_latest)
    return snap


def save_simulate_snapshot(*, model, params, initial_conditions, period_to_regime_to_V_arr, result,
                           log_path, log_keep_n_latest: int) -> Path:
    """Write `simulate_snapshot_NNN/` (`src/_lcm/persistence/snapshots.py:93-140`)."""
    parent = Path(log_path)
    snap = _next_dir(parent, "simulate_snapshot")
    _atomic_pickle(model, snap / "model.pkl")
    _atomic_pickle(params, snap / "params.pkl")
    _atomic_pickle(initial_conditions, snap / "initial_conditions.pkl")
    _atomic_pickle(result, snap / "result.pkl")
    _save_arrays(snap / "arrays.h5", period_to_regime_to_V_arr)
    _finish(snap, "simulate", ["model", "params", "initial_conditions", "result"])
    _enforce_retention(parent, "simulate_snapshot", log_keep_n_latest)
    return snap


def load_snapshot(path, *, exclude: Sequence[str] = ()):
    """Load a snapshot directory (`src/lcm/persistence.py:102-182`).

    `exclude` names fields to skip (set to None), e.g.
    `["period_to_regime_to_V_arr"]`. A platform other than the current one is
    logged as a warning. The pickles are loaded with cloudpickle: only load
    directories you trust.
    """
    import cloudpickle

    path = Path(path)
    meta = json.loads((path / "metadata.json").read_text(encoding="utf-8"))
    saved = meta["platform"]
    if saved != _get_platform():
        _LOG.warning("Snapshot created on %s but loading on %s — environment may not match",
                     saved, _get_platform())
    loaded: dict[str, Any] = {}
    for name in meta["fields"]:
        pkl = path / f"{name}.pkl"
        if name in exclude or not pkl.exists():
            loaded[name] = None
            continue
        with pkl.open("rb") as fh:
            loaded[name] = cloudpickle.load(fh)
    arrays = path / "arrays.h5"
    V = None
    if "period_to_regime_to_V_arr" not in exclude:
        if arrays.exists():
            V = _load_arrays(arrays)
        else:
            _LOG.warning("arrays.h5 not found in %s; period_to_regime_to_V_arr set to None", path)
    kind = meta["snapshot_type"]
    if kind == "solve":
        return SolveSnapshot(model=loaded.get("model"), params=loaded.get("params"),
                             period_to_regime_to_V_arr=V, platform=saved)
    if kind == "simulate":
        return SimulateSnapshot(model=loaded.get("model"), params=loaded.get("params"),
                                initial_conditions=loaded.get("initial_conditions"),
                                period_to_regime_to_V_arr=V, result=loaded.get("result"),
                                platform=saved)
    raise ValueError(f"Unknown snapshot_type: {kind!r}")


__all__ = ["SimulateSnapshot", "SolveSnapshot", "load_snapshot", "load_solution", "save_solution"]
