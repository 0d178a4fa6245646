"""Solver configurations a regime can select.

Same names and fields as the reference's shipped solvers
(`src/lcm/solvers.py:42-78`, `src/_lcm/solution/grid_search.py:99-132`,
`src/_lcm/solution/dcegm.py:93-109,151-254`). Here a solver is a plain
configuration object: `Model(...)` lowers each regime according to its solver
into descriptors for the HIP grid-search or DC-EGM kernels.
"""

from __future__ import annotations

from dataclasses import dataclass, field

from pylcm_amd.exceptions import RegimeInitializationError


class Solver:
    """Base class of solver configurations (`src/_lcm/solution/contract.py:549-619`).

    `build_period_kernels(context=...)` returns the regime's `SolutionKernels`
    (`pylcm_amd.contract`); `requires_continuation` says whether the solver
    reads its targets' continuation payloads.
    """

    def build_period_kernels(self, *, context):
        raise NotImplementedError(f"{type(self).__name__} builds no period kernels")

    @property
    def requires_continuation(self) -> bool:
        return False


@dataclass(frozen=True, kw_only=True)
class GridSearch(Solver):
    """Brute-force max over the full action grid (the reference's default)."""

    def build_period_kernels(self, *, context):
        """One HIP grid-search period kernel per active period (`grid_search.py:99-132`)."""
        from pylcm_amd.contract import grid_search_kernels

        return grid_search_kernels(context)


@dataclass(frozen=True, kw_only=True)
class FUESEnvelope:
    """Fast Upper-Envelope Scan (Dobrescu & Shanker 2022) configuration.

    `jump_thresh` is the segment-switch threshold on |dA/dR|;
    `n_points_to_scan=None` scans exhaustively (`dcegm.py:93-109`).
    """

    jump_thresh: float = 2.0
    n_points_to_scan: int | None = None
    scan_unroll: int = 1

    def __post_init__(self):
        if not self.jump_thresh > 0:
            raise RegimeInitializationError("FUESEnvelope.jump_thresh must be positive")
        if self.n_points_to_scan is not None and self.n_points_to_scan < 1:
            raise RegimeInitializationError("FUESEnvelope.n_points_to_scan must be >= 1")


@dataclass(frozen=True, kw_only=True)
class DCEGM(Solver):
    """Discrete-continuous endogenous grid method with an upper envelope.

    Only `FUESEnvelope` is implemented on the device; other envelope configs are
    refused at `Model(...)`.
    """

    savings_grid: object
    envelope: object = field(default_factory=FUESEnvelope)
    refined_grid_factor: float = 2.0
    n_constrained_points: int = 20
    stochastic_node_batch_size: int = 0

    def build_period_kernels(self, *, context):
        """One HIP DC-EGM period kernel per active period (`dcegm.py:359-419`)."""
        from pylcm_amd.contract import dcegm_kernels

        return dcegm_kernels(context)

    @property
    def requires_continuation(self) -> bool:
        return True

    def __post_init__(self):
        if self.refined_grid_factor < 1.0:
            raise RegimeInitializationError("refined_grid_factor must be >= 1")
        if self.n_constrained_points < 1:
            raise RegimeInitializationError("n_constrained_points must be >= 1")


@dataclass(frozen=True, kw_only=True)
class EGM(Solver):
    """Declared for API completeness; refused at `Model(...)`."""

    savings_grid: object = None
