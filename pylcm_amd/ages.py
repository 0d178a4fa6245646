"""Age grid: maps periods to ages (`src/lcm/ages.py`, `src/_lcm/ages.py:21-40`)."""

from __future__ import annotations

import re
from fractions import Fraction

import numpy as np

from pylcm_amd.exceptions import GridInitializationError

_STEP_UNITS = {"Y": Fraction(1), "Q": Fraction(1, 4), "M": Fraction(1, 12)}


def _parse_step(step) -> int | Fraction:
    if isinstance(step, (int, float, Fraction)) and not isinstance(step, bool):
        f = Fraction(step).limit_denominator(10**6)
        return int(f) if f.denominator == 1 else f
    m = re.match(r"^(\d+)?([YMQ])$", str(step), re.IGNORECASE)
    if not m:
        raise GridInitializationError(
            f"Invalid step format: '{step}'. Expected format like 'Y', '2Y', 'M', '3M', 'Q'."
        )
    mult = int(m.group(1)) if m.group(1) else 1
    r = mult * _STEP_UNITS[m.group(2).upper()]
    return int(r) if r.denominator == 1 else r


class AgeGrid:
    """Ages by period: `start..stop` inclusive in `step`, or `exact_values`."""

    def __init__(self, *, start=None, stop=None, step=None, exact_values=None):
        if exact_values is not None:
            vals = tuple(exact_values)
            if len(vals) < 1:
                raise GridInitializationError("AgeGrid needs at least one age")
            self._step = None
        elif start is not None and stop is not None and step is not None:
            st = _parse_step(step)
            if all(isinstance(x, (int, np.integer)) for x in (start, stop, st)):
                # integer ages: plain integer arithmetic (exact, and cheap)
                span = int(stop) - int(start)
                if span < 0 or span % st != 0:
                    raise GridInitializationError("step must divide stop - start evenly")
                vals = tuple(range(int(start), int(stop) + 1, st))
            else:
                span = Fraction(stop) - Fraction(start)
                if span < 0 or span % st != 0:
                    raise GridInitializationError("step must divide stop - start evenly")
                n = int(span // st) + 1
                vals = tuple(Fraction(start) + i * st for i in range(n))
            self._step = st
        else:
            raise GridInitializationError("Must specify 'start/stop/step' or 'exact_values'.")
        self._is_integer = all(isinstance(v, (int, np.integer)) or Fraction(v).denominator == 1 for v in vals)
        if self._is_integer:
            self._exact = tuple(int(v) for v in vals)
        else:
            self._exact = tuple(Fraction(v) for v in vals)
        if any(b <= a for a, b in zip(self._exact, self._exact[1:])):
            raise GridInitializationError("ages must increase")

    @property
    def exact_values(self):
        return self._exact

    @property
    def values(self) -> np.ndarray:
        if self._is_integer:
            return np.asarray(self._exact, dtype=np.int32)
        return np.asarray([float(v) for v in self._exact])

    @property
    def is_integer(self) -> bool:
        return self._is_integer

    @property
    def n_periods(self) -> int:
        return len(self._exact)

    @property
    def step_size(self):
        return None if self._step is None else float(self._step)

    def period_to_age(self, period: int):
        if period < 0 or period >= self.n_periods:
            raise IndexError(f"Period {period} out of bounds for grid with {self.n_periods} periods.")
        v = self._exact[period]
        return int(v) if self._is_integer else float(v)

    def age_to_period(self, age) -> int:
        for i, v in enumerate(self._exact):
            if float(v) == float(age):
                return i
        raise ValueError(f"Age {age} is not a valid grid point.")
