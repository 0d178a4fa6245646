"""Exception hierarchy of the public API.

Names follow the reference's `PyLCMError` family (`src/lcm/exceptions.py:1-110`)
so user code that catches them keeps working. `NativeKernelError` and
`UnsupportedModelError` have no counterpart there: the first maps a nonzero
status of the HIP library, the second is raised at `Model(...)` when a model
falls outside the family the HIP kernels implement (never a silent fallback).
"""


class PyLCMError(Exception):
    """Base class of every error raised by this package."""


class ModelInitializationError(PyLCMError):
    """The model specification is invalid."""


class RegimeInitializationError(ModelInitializationError):
    """A regime specification is invalid (a model initialization error, as in
    `src/lcm/exceptions.py:63`)."""


class GridInitializationError(PyLCMError):
    """A grid specification is invalid."""


class InvalidParamsError(PyLCMError):
    """The `params` passed to solve/simulate do not match the model."""


class InvalidValueFunctionError(PyLCMError):
    """A value function contains NaN (raised in `log_level="debug"`).

    `partial_solution` holds the periods solved before the failure, as in the
    reference (`src/lcm/model.py:552-562`).
    """

    def __init__(self, msg, partial_solution=None):
        super().__init__(msg)
        self.partial_solution = partial_solution


class InvalidInitialConditionsError(PyLCMError):
    """Initial conditions passed to `simulate` are invalid."""


class UnsupportedModelError(ModelInitializationError):
    """The model uses a feature the native kernels do not implement."""


class NativeKernelError(PyLCMError):
    """The HIP library returned a nonzero status or could not be loaded."""

    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


class InvalidAdditionalTargetsError(PyLCMError):
    """`to_dataframe(additional_targets=...)` names a target no regime defines
    (`src/lcm/exceptions.py` InvalidAdditionalTargetsError)."""
