"""MI355X-native backward-induction solver with pylcm's model API.

The public names mirror the reference's `lcm` package (`src/lcm/__init__.py`)
for the part of it this package implements: model and regime specification,
grids, stochastic processes, solver configurations, `Model.solve` and
`Model.simulate`. Solves and simulations run as hand-written HIP kernels for
gfx950 through `liblcm_amd.so`; there is no CPU execution path.
"""

from pylcm_amd import config
from pylcm_amd.ages import AgeGrid
from pylcm_amd.exceptions import (
    GridInitializationError,
    InvalidAdditionalTargetsError,
    InvalidInitialConditionsError,
    InvalidParamsError,
    InvalidValueFunctionError,
    ModelInitializationError,
    NativeKernelError,
    PyLCMError,
    RegimeInitializationError,
    UnsupportedModelError,
)
from pylcm_amd.grids import (
    DiscreteGrid,
    IrregSpacedGrid,
    LinSpacedGrid,
    LogSpacedGrid,
    categorical,
)
from pylcm_amd.model import Model
from pylcm_amd.persistence import (
    SimulateSnapshot,
    SolveSnapshot,
    load_snapshot,
    load_solution,
    save_solution,
)
from pylcm_amd.processes import (
    LogNormalIIDProcess,
    NormalIIDProcess,
    RouwenhorstAR1Process,
    TauchenAR1Process,
    UniformIIDProcess,
)
from pylcm_amd.regime import (
    ConsumptionSavingsRegime,
    ExtremeValueTasteShocks,
    LiquidMargin,
    MarkovTransition,
    Regime,
)
from pylcm_amd.solvers import DCEGM, FUESEnvelope, GridSearch


class LinearAggregator:
    """`U + discount_factor * CE` (`src/lcm/koopmans_aggregation.py:69-98`)."""

    def __call__(self, utility, CE, discount_factor):
        return utility + discount_factor * CE


class LinearExpectation:
    """The linear certainty equivalent (the only one implemented)."""


__all__ = [
    "AgeGrid", "ConsumptionSavingsRegime", "DCEGM", "DiscreteGrid", "ExtremeValueTasteShocks",
    "FUESEnvelope", "GridInitializationError", "GridSearch", "InvalidAdditionalTargetsError",
    "InvalidInitialConditionsError",
    "InvalidParamsError", "InvalidValueFunctionError", "IrregSpacedGrid", "LinSpacedGrid",
    "LinearAggregator", "LinearExpectation", "LiquidMargin", "LogNormalIIDProcess",
    "LogSpacedGrid", "MarkovTransition", "Model", "ModelInitializationError",
    "NativeKernelError", "NormalIIDProcess", "PyLCMError", "RegimeInitializationError",
    "Regime", "RouwenhorstAR1Process", "TauchenAR1Process", "UniformIIDProcess",
    "SimulateSnapshot", "SolveSnapshot", "UnsupportedModelError", "categorical", "config",
    "load_snapshot", "load_solution", "save_solution",
]
