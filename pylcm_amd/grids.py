"""State and action grids, and the `@categorical` decorator.

Mirrors the reference's user-facing grid classes (`src/lcm/grids.py`,
`src/_lcm/grids/continuous.py:111-326`, `src/_lcm/grids/discrete.py:11-64`,
`src/_lcm/grids/categorical.py`). Grid points are computed the way
`jnp.linspace` computes them (`start * (1 - i/div) + stop * (i/div)`, endpoint
appended), so a float32 grid holds the same numbers as the reference's.

Only the host holds these objects; the HIP kernels receive the points (or, for
uniform grids, start/step) through the lowered period descriptors.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass

import numpy as np

from pylcm_amd.exceptions import GridInitializationError

GRID_KIND_LIN = 0
GRID_KIND_LOG = 1
GRID_KIND_IRREG = 2


def categorical(cls=None, *, ordered: bool = False):
    """Turn a class of annotated fields into a category enumeration.

    Each annotated field becomes an integer class attribute, numbered 0, 1, ...
    in declaration order (`src/_lcm/grids/categorical.py`). The decorated class
    keeps the list of names in `__categories__`.
    """

    def wrap(c):
        names = list(getattr(c, "__annotations__", {}).keys())
        if not names:
            raise GridInitializationError(
                f"categorical class {c.__name__} declares no fields"
            )
        for code, name in enumerate(names):
            setattr(c, name, code)
        c.__categories__ = tuple(names)
        c.__ordered__ = bool(ordered)
        return c

    if cls is not None:
        return wrap(cls)
    return wrap


def jax_linspace(start: float, stop: float, n: int, dtype=np.float64) -> np.ndarray:
    """`jnp.linspace(start, stop, n)` evaluated in `dtype` arithmetic."""
    dt = np.dtype(dtype)
    if n == 1:
        return np.asarray([start], dtype=dt)
    div = n - 1
    step = (np.arange(div, dtype=dt) / dt.type(div)).astype(dt)
    s = dt.type(start)
    e = dt.type(stop)
    out = s * (dt.type(1) - step) + e * step
    return np.concatenate([out, np.asarray([e], dtype=dt)]).astype(dt)


class Grid:
    """Base class of every grid."""

    batch_size: int = 0
    distributed: bool = False


@dataclass(frozen=True)
class DiscreteGrid(Grid):
    """A grid over the codes of a `@categorical` class."""

    category_class: type

    def __post_init__(self):
        if not hasattr(self.category_class, "__categories__"):
            raise GridInitializationError(
                "DiscreteGrid needs a class decorated with @categorical"
            )

    @property
    def categories(self) -> tuple[str, ...]:
        return tuple(self.category_class.__categories__)

    @property
    def codes(self) -> tuple[int, ...]:
        return tuple(range(len(self.categories)))

    @property
    def ordered(self) -> bool:
        return bool(self.category_class.__ordered__)

    @property
    def n_points(self) -> int:
        return len(self.categories)

    def to_numpy(self, dtype=np.int32) -> np.ndarray:
        return np.arange(self.n_points, dtype=dtype)


class ContinuousGrid(Grid):
    """A grid whose values are interpolated between points."""

    kind: int

    def to_numpy(self, dtype=np.float64) -> np.ndarray:
        raise NotImplementedError


def _check_uniform(start, stop, n_points):
    if not isinstance(n_points, (int, np.integer)) or n_points < 1:
        raise GridInitializationError("n_points must be a positive integer")
    if not np.isfinite(start) or not np.isfinite(stop):
        raise GridInitializationError("grid bounds must be finite")
    if n_points > 1 and not start < stop:
        raise GridInitializationError("grid start must be smaller than stop")


@dataclass(frozen=True, kw_only=True)
class LinSpacedGrid(ContinuousGrid):
    """Linearly spaced points on [start, stop] (`continuous.py:111-133`)."""

    start: float
    stop: float
    n_points: int
    batch_size: int = 0
    distributed: bool = False
    kind = GRID_KIND_LIN

    def __post_init__(self):
        _check_uniform(self.start, self.stop, self.n_points)

    def to_numpy(self, dtype=np.float64) -> np.ndarray:
        return jax_linspace(self.start, self.stop, self.n_points, dtype)

    def replace(self, **kw) -> "LinSpacedGrid":
        return dataclasses.replace(self, **kw)


@dataclass(frozen=True, kw_only=True)
class LogSpacedGrid(ContinuousGrid):
    """Logarithmically spaced points on [start, stop] (`continuous.py:136-180`).

    Points are `exp(linspace(log start, log stop))` with the endpoints set to
    exactly `start` and `stop` (`src/_lcm/grids/coordinates.py:47-70`).
    """

    start: float
    stop: float
    n_points: int
    batch_size: int = 0
    distributed: bool = False
    kind = GRID_KIND_LOG

    def __post_init__(self):
        _check_uniform(self.start, self.stop, self.n_points)
        if self.start <= 0:
            raise GridInitializationError("LogSpacedGrid needs start > 0")

    def to_numpy(self, dtype=np.float64) -> np.ndarray:
        dt = np.dtype(dtype)
        lin = jax_linspace(np.log(dt.type(self.start)), np.log(dt.type(self.stop)),
                           self.n_points, dt)
        pts = np.exp(lin).astype(dt)
        pts[0] = dt.type(self.start)
        pts[-1] = dt.type(self.stop)
        return pts

    def replace(self, **kw) -> "LogSpacedGrid":
        return dataclasses.replace(self, **kw)


@dataclass(frozen=True, kw_only=True)
class IrregSpacedGrid(ContinuousGrid):
    """Arbitrary increasing points (`continuous.py:226-326`)."""

    points: tuple
    batch_size: int = 0
    distributed: bool = False
    kind = GRID_KIND_IRREG

    def __post_init__(self):
        pts = np.asarray(self.points, dtype=np.float64)
        if pts.ndim != 1 or pts.size < 1:
            raise GridInitializationError("IrregSpacedGrid needs a 1-d points tuple")
        if not np.all(np.isfinite(pts)):
            raise GridInitializationError("IrregSpacedGrid points must be finite")
        if pts.size > 1 and not np.all(np.diff(pts) > 0):
            raise GridInitializationError("IrregSpacedGrid points must increase")
        object.__setattr__(self, "points", tuple(float(p) for p in pts))

    @property
    def n_points(self) -> int:
        return len(self.points)

    @property
    def start(self) -> float:
        return self.points[0]

    @property
    def stop(self) -> float:
        return self.points[-1]

    def to_numpy(self, dtype=np.float64) -> np.ndarray:
        return np.asarray(self.points, dtype=dtype)


def is_continuous(grid) -> bool:
    from pylcm_amd.processes import StochasticProcess

    return isinstance(grid, ContinuousGrid) and not isinstance(grid, StochasticProcess)
