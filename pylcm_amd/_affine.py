"""Affine tracer for partial evaluation of user functions.

The grid-search kernel evaluates, for every state-action pair, the next
continuous state and the constraints. User functions are plain numpy-style
Python, so instead of compiling them the host calls them once per period with
every discrete variable and the continuous action as concrete numpy arrays, and
the continuous state as an `Affine` object `a * w + b`. Arithmetic that stays
affine in `w` (the budget laws `R * (w - c) + y` and constraints `c <= w` of the
reference's examples) comes back as coefficient tables `a`, `b` over the
(discrete state, discrete action, continuous action) grid; anything else raises
`NonAffine` and the caller tabulates that function over the full state grid
instead. The kernel then forms `fma(a, w, b)` per state-action pair.

For forward simulation a process state's value is continuous per agent, so it
is traced as a *feature* symbol too: `Affine.feats` maps a feature id to its
coefficient table. A unary function of an expression that holds features but
no `w` and no action dependence (e.g. `exp(income)`) becomes a new feature,
recorded in a `FeatureSet` with the recipe to compute it per agent. The
simulation kernel then forms `B = B0 + sum_j B_j * phi_j(agent)`.
"""

from __future__ import annotations

import numpy as np


class NonAffine(Exception):
    """The traced expression is not affine in the continuous state."""


# unary functions a feature may be built from, with their device opcodes
FEATURE_OPS = {
    np.exp: 1,
    np.log: 2,
    np.sqrt: 3,
    np.negative: 4,
}


class FeatureSet:
    """Per-agent features: feature k = op(sum_i coef_i * phi_i + const).

    Feature 0.. are the identity features of the traced process states; each
    recipe holds coefficients that may vary with the discrete-state axes (but
    not with actions), kept as broadcast arrays.
    """

    def __init__(self, action_axes=()):
        self.recipes = []  # (op, {fid: coef}, const) ; op 0 = identity of a process state
        self.names = []
        self.action_axes = tuple(action_axes)

    def identity(self, name):
        self.recipes.append((0, {}, None))
        self.names.append(name)
        return len(self.recipes) - 1

    def derive(self, op, feats, const):
        self.recipes.append((op, feats, const))
        self.names.append(None)
        return len(self.recipes) - 1


def _is_affine(x):
    return isinstance(x, Affine)


def _parts(x):
    if isinstance(x, Affine):
        return x.a, x.b, x.feats
    return 0.0, np.asarray(x, dtype=np.float64), {}


def _merge(f1, f2, s1=1.0, s2=1.0):
    out = {k: s1 * v for k, v in f1.items()}
    for k, v in f2.items():
        out[k] = out[k] + s2 * v if k in out else s2 * v
    return out


class Affine:
    """`a * w + sum_j feats[j] * phi_j + b` with numpy-array coefficients."""

    __array_priority__ = 1000

    def __init__(self, a, b, feats=None, fs: FeatureSet | None = None):
        self.a = np.asarray(a, dtype=np.float64)
        self.b = np.asarray(b, dtype=np.float64)
        self.feats = {k: np.asarray(v, dtype=np.float64) for k, v in (feats or {}).items()}
        self.fs = fs

    # -- helpers -------------------------------------------------------------
    @property
    def has_w(self) -> bool:
        return bool(np.any(self.a != 0))

    @property
    def has_feats(self) -> bool:
        return any(np.any(v != 0) for v in self.feats.values())

    @property
    def is_const(self) -> bool:
        return not self.has_w and not self.has_feats

    def const_value(self):
        return self.b

    def _simplify(self):
        return self.b if self.is_const else self

    def _new(self, a, b, feats, other=None):
        fs = self.fs if self.fs is not None else getattr(other, "fs", None)
        return Affine(a, b, feats, fs)

    # -- arithmetic ----------------------------------------------------------
    def __add__(self, o):
        if isinstance(o, Constraint):
            return NotImplemented
        oa, ob, of = _parts(o)
        return self._new(self.a + oa, self.b + ob, _merge(self.feats, of), o)

    __radd__ = __add__

    def __sub__(self, o):
        oa, ob, of = _parts(o)
        return self._new(self.a - oa, self.b - ob, _merge(self.feats, of, 1.0, -1.0), o)

    def __rsub__(self, o):
        oa, ob, of = _parts(o)
        return self._new(oa - self.a, ob - self.b, _merge(of, self.feats, 1.0, -1.0), o)

    def __mul__(self, o):
        if _is_affine(o):
            if o.is_const:
                o = o.b
            elif self.is_const:
                return o * self.b
            else:
                raise NonAffine("product of two state-dependent terms")
        k = np.asarray(o, dtype=np.float64)
        return self._new(self.a * k, self.b * k, {f: v * k for f, v in self.feats.items()})

    __rmul__ = __mul__

    def __truediv__(self, o):
        if _is_affine(o):
            if not o.is_const:
                raise NonAffine("division by a state-dependent term")
            o = o.b
        k = np.asarray(o, dtype=np.float64)
        return self._new(self.a / k, self.b / k, {f: v / k for f, v in self.feats.items()})

    def __rtruediv__(self, o):
        if self.is_const:
            return np.asarray(o, dtype=np.float64) / self.b
        raise NonAffine("division by a state-dependent term")

    def __neg__(self):
        return self._new(-self.a, -self.b, {f: -v for f, v in self.feats.items()})

    def __pos__(self):
        return self

    def __pow__(self, k):
        if self.is_const:
            return self.b**k
        if np.isscalar(k) and k == 1:
            return self
        raise NonAffine("power of a state-dependent term")

    def __rpow__(self, base):
        if self.is_const:
            return np.asarray(base, dtype=np.float64) ** self.b
        raise NonAffine("state-dependent exponent")

    # -- comparisons give constraints ---------------------------------------
    def __le__(self, o):
        return Constraint.make(_as_aff(o) - self, strict=False)

    def __lt__(self, o):
        return Constraint.make(_as_aff(o) - self, strict=True)

    def __ge__(self, o):
        return Constraint.make(self - o, strict=False)

    def __gt__(self, o):
        return Constraint.make(self - o, strict=True)

    def __bool__(self):
        raise NonAffine("truth value of a state-dependent expression")

    # -- features ------------------------------------------------------------
    def _as_feature(self, op):
        """op(self) as a new feature, if self is a feature expression free of
        `w` whose coefficients do not vary with the actions."""
        if self.has_w or self.fs is None:
            raise NonAffine(f"{op} of a state-dependent term")
        for arr in [self.b, *self.feats.values()]:
            for ax in self.fs.action_axes:
                if arr.ndim > ax and arr.shape[ax] > 1:
                    raise NonAffine("feature expression varies with the actions")
        fid = self.fs.derive(FEATURE_OPS[op], dict(self.feats), self.b)
        return Affine(0.0, 0.0, {fid: 1.0}, self.fs)

    # -- numpy protocol ------------------------------------------------------
    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        if method != "__call__" or kwargs.get("out") is not None:
            raise NonAffine(f"unsupported ufunc use {ufunc.__name__}.{method}")
        table = {
            np.add: lambda x, y: _as_aff(x) + y,
            np.subtract: lambda x, y: _as_aff(x) - y,
            np.multiply: lambda x, y: _as_aff(x) * y,
            np.true_divide: lambda x, y: _as_aff(x) / y,
            np.negative: lambda x: -_as_aff(x),
            np.positive: lambda x: x,
            np.less_equal: lambda x, y: _as_aff(x) <= y,
            np.less: lambda x, y: _as_aff(x) < y,
            np.greater_equal: lambda x, y: _as_aff(x) >= y,
            np.greater: lambda x, y: _as_aff(x) > y,
        }
        if ufunc in table:
            return table[ufunc](*inputs)
        if all((not _is_affine(i)) or i.is_const for i in inputs):
            args = [i.b if _is_affine(i) else i for i in inputs]
            return ufunc(*args, **kwargs)
        if len(inputs) == 1 and ufunc in FEATURE_OPS and _is_affine(inputs[0]):
            return inputs[0]._as_feature(ufunc)
        raise NonAffine(f"{ufunc.__name__} of a state-dependent term")

    def __array_function__(self, func, types, args, kwargs):
        if func is np.where and len(args) == 3 and not kwargs:
            cond, x, y = args
            if isinstance(cond, (Affine, Constraint)):
                raise NonAffine("condition depends on the continuous state")
            xa, xb, xf = _parts(x)
            ya, yb, yf = _parts(y)
            keys = set(xf) | set(yf)
            feats = {k: np.where(cond, xf.get(k, 0.0), yf.get(k, 0.0)) for k in keys}
            fs = getattr(x, "fs", None) or getattr(y, "fs", None)
            return Affine(np.where(cond, xa, ya), np.where(cond, xb, yb), feats, fs)._simplify()
        flat = list(args) + list(kwargs.values())
        if all((not _is_affine(i)) or i.is_const for i in flat):
            args2 = [i.b if _is_affine(i) else i for i in args]
            kw2 = {k: (v.b if _is_affine(v) else v) for k, v in kwargs.items()}
            return func(*args2, **kw2)
        raise NonAffine(f"{func.__name__} of a state-dependent term")


def _as_aff(x):
    return x if isinstance(x, Affine) else Affine(0.0, x)


class Constraint:
    """Conjunction of affine inequalities `a_k * w + b_k >= 0` (or `> 0`)."""

    def __init__(self, terms, feats=None):
        self.terms = list(terms)  # [(a, b, strict)] ; a concrete bool mask is (0, 0|-inf, False)
        # per term: {feature id: coefficient} added to b (forward simulation only)
        self.feats = list(feats) if feats is not None else [{} for _ in self.terms]

    @staticmethod
    def make(expr, strict):
        expr = _as_aff(expr)
        if expr.is_const:
            return (expr.b > 0) if strict else (expr.b >= 0)
        if expr.has_feats and not expr.has_w:
            raise NonAffine("constraint on a per-agent process value alone")
        return Constraint([(expr.a, expr.b, strict)], [dict(expr.feats)])

    def __and__(self, o):
        return Constraint(self.terms + _constraint_terms(o), self.feats + _constraint_feats(o))

    __rand__ = __and__

    def __bool__(self):
        raise NonAffine("truth value of a state-dependent constraint")

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        if ufunc in (np.logical_and, np.bitwise_and) and method == "__call__":
            terms, feats = [], []
            for i in inputs:
                terms += _constraint_terms(i)
                feats += _constraint_feats(i)
            return Constraint(terms, feats)
        raise NonAffine(f"{ufunc.__name__} of a state-dependent constraint")

    def __array_function__(self, func, types, args, kwargs):
        raise NonAffine(f"{func.__name__} of a state-dependent constraint")


def _constraint_terms(x):
    if isinstance(x, Constraint):
        return x.terms
    mask = np.asarray(x, dtype=bool)
    return [(np.zeros(()), np.where(mask, 0.0, -np.inf), False)]


def _constraint_feats(x):
    return x.feats if isinstance(x, Constraint) else [{}]
