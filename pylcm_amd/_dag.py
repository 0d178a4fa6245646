"""Resolve a regime's user functions by argument name.

The reference composes user functions into a DAG with `dags`
(`src/_lcm/regime_building/Q_and_F.py:1321-1351`): every argument of a function
is a state, an action, `age`/`period`, the output of another function of the
regime, or a parameter. This module does the same resolution on the host, for
evaluating functions on concrete arrays or on the affine tracer of
`_lowering.py`.

Parameters are looked up most-specific first:
`params[regime][function][arg]`, then `params[regime][arg]`, then `params[arg]`,
then the same three levels of the model's `fixed_params`
(`src/lcm/model.py:457-463`).
"""

from __future__ import annotations

import inspect
from functools import lru_cache

import numpy as np

from pylcm_amd.exceptions import InvalidParamsError
from pylcm_amd.regime import MarkovTransition

_MISSING = object()


def unwrap(func):
    return func.func if isinstance(func, MarkovTransition) else func


@lru_cache(maxsize=4096)
def _signature(func):
    sig = inspect.signature(func)
    out = []
    for p in sig.parameters.values():
        if p.kind in (p.VAR_POSITIONAL, p.VAR_KEYWORD):
            continue
        out.append((p.name, p.default is not inspect.Parameter.empty, p.default))
    return tuple(out)


def arg_names(func) -> tuple[str, ...]:
    return tuple(n for n, _, _ in _signature(unwrap(func)))


def lookup_param(params, fixed, regime: str, func_name: str, arg: str):
    for src in (params, fixed):
        if not src:
            continue
        r = src.get(regime)
        if isinstance(r, dict):
            f = r.get(func_name)
            if isinstance(f, dict) and arg in f:
                return f[arg]
            if arg in r and not isinstance(r[arg], dict):
                return r[arg]
        if arg in src and not isinstance(src[arg], dict):
            return src[arg]
    return _MISSING


class Evaluator:
    """Evaluates the functions of one regime on a given variable environment.

    `variables` maps state/action names, `age` and `period` to values (numpy
    arrays, scalars or tracers). Function outputs are memoised per evaluator.
    """

    def __init__(self, *, regime_name, functions, variables, params, fixed_params,
                 param_dtype=None):
        self.regime_name = regime_name
        self.functions = dict(functions)
        self.variables = dict(variables)
        self.params = params or {}
        self.fixed = fixed_params or {}
        self.param_dtype = np.dtype(param_dtype) if param_dtype is not None else None
        self._memo: dict[str, object] = {}

    def _cast(self, v):
        """Float parameters are rounded to the model dtype, as the reference
        casts params to the canonical dtype (`params/processing.py:225`)."""
        if self.param_dtype is None:
            return v
        if isinstance(v, float):
            return float(np.asarray(v, dtype=self.param_dtype))
        if isinstance(v, np.ndarray) and v.dtype.kind == "f":
            return v.astype(self.param_dtype).astype(np.float64)
        return v

    def param(self, func_name: str, arg: str, default=_MISSING):
        v = lookup_param(self.params, self.fixed, self.regime_name, func_name, arg)
        if v is _MISSING:
            if default is not _MISSING:
                return default
            raise InvalidParamsError(
                f"regime {self.regime_name!r}: missing parameter {arg!r} of function {func_name!r}"
            )
        return self._cast(v)

    def call(self, func_name: str, func=None):
        """Evaluate `func` (default: the regime function `func_name`)."""
        if func is None and func_name in self._memo:
            return self._memo[func_name]
        f = unwrap(func if func is not None else self.functions[func_name])
        kwargs = {}
        for name, has_default, default in _signature(f):
            kwargs[name] = self.resolve(name, func_name, has_default, default)
        out = f(**kwargs)
        if func is None:
            self._memo[func_name] = out
        return out

    def resolve(self, name, func_name, has_default=False, default=None):
        if name in self.variables:
            return self.variables[name]
        if name in self.functions and name != func_name:
            return self.call(name)
        v = lookup_param(self.params, self.fixed, self.regime_name, func_name, name)
        if v is not _MISSING:
            return self._cast(v)
        if has_default:
            return default
        raise InvalidParamsError(
            f"regime {self.regime_name!r}: cannot resolve argument {name!r} of {func_name!r}"
            " (not a state, action, function or parameter)"
        )


def param_names_of(*, regime_name, functions, variables, extra=()):
    """Names a function set reads as parameters (for templates and checks)."""
    out = {}
    for fname, f in list(functions.items()) + list(extra):
        for name, has_default, _ in _signature(unwrap(f)):
            if name in variables or name in functions or has_default:
                continue
            out.setdefault(fname, []).append(name)
    return out
