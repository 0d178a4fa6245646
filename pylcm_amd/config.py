"""Global precision switch.

The reference takes its float dtype from JAX's `jax_enable_x64` flag
(`src/_lcm/dtypes.py:22-31`): float64 when x64 is on, float32 otherwise, and
float32 by default. This module keeps the same rule without JAX:
`pylcm_amd.config.update("enable_x64", True)` selects float64 for every model
solved afterwards. `Model.solve(..., dtype=...)` overrides it per call.
"""

import contextlib

import numpy as np

_state = {"enable_x64": False}


def update(name: str, value: bool) -> None:
    if name not in ("enable_x64", "jax_enable_x64"):
        raise KeyError(f"unknown config option {name!r}")
    _state["enable_x64"] = bool(value)


def x64_enabled() -> bool:
    return _state["enable_x64"]


def canonical_float_dtype() -> np.dtype:
    return np.dtype(np.float64 if _state["enable_x64"] else np.float32)


@contextlib.contextmanager
def precision(bits: int):
    """Temporarily select 32- or 64-bit floats."""
    old = _state["enable_x64"]
    _state["enable_x64"] = bits == 64
    try:
        yield
    finally:
        _state["enable_x64"] = old
