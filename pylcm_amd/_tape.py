"""Expression tapes: user functions of one continuous variable, per column.

The DC-EGM kernel has to evaluate the regime's utility u(c), its derivative
u'(c) and the inverse marginal utility at points that only exist on the device
(the Euler-inverted consumption, the refined policy row). The user writes those
functions as NumPy code. The host traces each one once with a symbolic input
(`Sym`) and every other argument bound to its per-column value (an array over
the regime's (discrete state, discrete action) columns): arithmetic that does not
touch the input is folded into per-column constants, and each operation on the
input is recorded as one tape instruction. The device interprets the tape in
FP64 with forward-mode derivatives, per lane; the host evaluates the same tape
(NumPy) for tables such as the next resources R'(A) and dR'/dA.

Derivative conventions follow JAX's: d(x**k) = k x**(k-1) dx, and `maximum` /
`minimum` split the derivative evenly at a tie.
"""

from __future__ import annotations

import numpy as np

from pylcm_amd.exceptions import UnsupportedModelError

# opcodes (kept in sync with pylcm_amd/csrc/egm.hip)
OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_POW = 1, 2, 3, 4, 5
OP_NEG, OP_EXP, OP_LOG, OP_SQRT, OP_MAX, OP_MIN = 6, 7, 8, 9, 10, 11
OP_WHERE, OP_LOG1P, OP_EXPM1, OP_ABS = 12, 13, 14, 15

MAX_INSTRUCTIONS = 15  # registers 0..15 on the device (0 holds the input)
HEADER = 4  # shorts: n_ins, output operand, 0, 0


class Sym:
    """A tape register: the traced input or the result of an instruction."""

    __array_priority__ = 2000

    def __init__(self, tape, reg):
        self.tape = tape
        self.reg = reg

    # arithmetic ---------------------------------------------------------------
    def __add__(self, o):
        return self.tape.emit(OP_ADD, self, o)

    def __radd__(self, o):
        return self.tape.emit(OP_ADD, o, self)

    def __sub__(self, o):
        return self.tape.emit(OP_SUB, self, o)

    def __rsub__(self, o):
        return self.tape.emit(OP_SUB, o, self)

    def __mul__(self, o):
        return self.tape.emit(OP_MUL, self, o)

    def __rmul__(self, o):
        return self.tape.emit(OP_MUL, o, self)

    def __truediv__(self, o):
        return self.tape.emit(OP_DIV, self, o)

    def __rtruediv__(self, o):
        return self.tape.emit(OP_DIV, o, self)

    def __pow__(self, o):
        return self.tape.emit(OP_POW, self, o)

    def __rpow__(self, o):
        return self.tape.emit(OP_POW, o, self)

    def __neg__(self):
        return self.tape.emit(OP_NEG, self)

    def __pos__(self):
        return self

    def __abs__(self):
        return self.tape.emit(OP_ABS, self)

    def _no(self, *_):
        raise UnsupportedModelError(
            "a comparison or branch on the traced continuous variable is not supported in a "
            "DC-EGM utility / inverse marginal utility / resources law")

    __lt__ = __le__ = __gt__ = __ge__ = __eq__ = __ne__ = __bool__ = _no

    def __hash__(self):
        return id(self)

    _UNARY = {np.negative: OP_NEG, np.exp: OP_EXP, np.log: OP_LOG, np.sqrt: OP_SQRT,
              np.log1p: OP_LOG1P, np.expm1: OP_EXPM1, np.absolute: OP_ABS}
    _BINARY = {np.add: OP_ADD, np.subtract: OP_SUB, np.multiply: OP_MUL, np.true_divide: OP_DIV,
               np.power: OP_POW, np.float_power: OP_POW, np.maximum: OP_MAX, np.minimum: OP_MIN}

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        if method != "__call__" or kwargs.get("out") is not None:
            raise UnsupportedModelError(f"unsupported use of {ufunc.__name__} on a traced variable")
        if ufunc in self._UNARY and len(inputs) == 1:
            return self.tape.emit(self._UNARY[ufunc], inputs[0])
        if ufunc in self._BINARY and len(inputs) == 2:
            return self.tape.emit(self._BINARY[ufunc], *inputs)
        if ufunc is np.positive:
            return inputs[0]
        if ufunc is np.square:
            return self.tape.emit(OP_MUL, inputs[0], inputs[0])
        if ufunc is np.reciprocal:
            return self.tape.emit(OP_DIV, 1.0, inputs[0])
        raise UnsupportedModelError(f"{ufunc.__name__} of a traced variable is not supported")

    def __array_function__(self, func, types, args, kwargs):
        if func is np.where and len(args) == 3 and not kwargs:
            cond, a, b = args
            if isinstance(cond, Sym):
                self._no()
            return self.tape.emit(OP_WHERE, a, b, mask=np.asarray(cond, dtype=bool))
        if func in (np.asarray, np.asanyarray) and args and args[0] is self:
            return self
        raise UnsupportedModelError(f"{func.__name__} of a traced variable is not supported")


class Tape:
    """One traced function: instructions plus per-column constant slots."""

    def __init__(self, n_cols: int):
        self.n_cols = int(n_cols)
        self.ins = []          # (op, a, b, c) operands: >= 0 register, < 0 constant slot -(k+1)
        self.consts = []       # each (n_cols,) float64
        self._const_key = {}
        self.out = None        # operand of the result
        self.x = Sym(self, 0)

    # -- recording ------------------------------------------------------------
    def _operand(self, v):
        if isinstance(v, Sym):
            if v.tape is not self:
                raise UnsupportedModelError("mixing two traces")
            return v.reg
        arr = np.asarray(v, dtype=np.float64)
        if arr.size != 1 and arr.shape != (self.n_cols,):
            try:
                arr = np.broadcast_to(arr, (self.n_cols,))
            except ValueError as e:
                raise UnsupportedModelError(
                    f"a per-column constant of shape {arr.shape} does not broadcast to "
                    f"{self.n_cols} columns") from e
        arr = np.ascontiguousarray(np.broadcast_to(arr, (self.n_cols,)), dtype=np.float64)
        key = arr.tobytes()
        if key not in self._const_key:
            self._const_key[key] = len(self.consts)
            self.consts.append(arr)
        return -(self._const_key[key] + 1)

    def emit(self, op, a, b=None, mask=None):
        if len(self.ins) >= MAX_INSTRUCTIONS:
            raise UnsupportedModelError(
                f"traced function needs more than {MAX_INSTRUCTIONS} operations on the device")
        if op == OP_WHERE:
            m = np.broadcast_to(mask, (self.n_cols,)).astype(np.float64)
            ops = (self._operand(m), self._operand(a), self._operand(b))
        else:
            ops = (self._operand(a), self._operand(b) if b is not None else 0, 0)
        self.ins.append((op,) + ops)
        return Sym(self, len(self.ins))

    def finish(self, result):
        """Record the traced function's result (a Sym or a constant)."""
        if isinstance(result, Sym):
            self.out = result.reg
        else:
            self.out = self._operand(result)
        return self

    # -- packing for the device -----------------------------------------------
    def program(self) -> np.ndarray:
        """int16 [HEADER + 4 * n_ins]."""
        p = [len(self.ins), self.out, 0, 0]
        for ins in self.ins:
            p += list(ins)
        return np.asarray(p, dtype=np.int16)

    def const_table(self) -> np.ndarray:
        """float64 [n_cols, n_const] (each column's constants contiguous)."""
        if not self.consts:
            return np.zeros((self.n_cols, 1))
        return np.ascontiguousarray(np.stack(self.consts, axis=1))

    # -- host evaluation (same semantics as the device) -----------------------
    def eval(self, x, derivative=False):
        """Evaluate at x of shape (n_cols, ...). Returns value (and d/dx)."""
        x = np.asarray(x, dtype=np.float64)
        ex = (slice(None),) + (None,) * (x.ndim - 1)
        rv = [x]
        rd = [np.ones_like(x)]

        def get(o):
            if o >= 0:
                return rv[o], rd[o]
            c = self.consts[-o - 1][ex]
            return c, np.zeros_like(c)

        with np.errstate(all="ignore"):
            for op, a, b, c in self.ins:
                va, da = get(a)
                if op in (OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_POW, OP_MAX, OP_MIN):
                    vb, db = get(b)
                if op == OP_ADD:
                    v, d = va + vb, da + db
                elif op == OP_SUB:
                    v, d = va - vb, da - db
                elif op == OP_MUL:
                    v, d = va * vb, da * vb + va * db
                elif op == OP_DIV:
                    v = va / vb
                    d = da / vb - va * db / (vb * vb)
                elif op == OP_POW:
                    v = va ** vb
                    d = np.where(da != 0, vb * va ** (vb - 1.0) * da, 0.0) + \
                        np.where(db != 0, v * np.log(va) * db, 0.0)
                elif op == OP_NEG:
                    v, d = -va, -da
                elif op == OP_EXP:
                    v = np.exp(va)
                    d = v * da
                elif op == OP_LOG:
                    v, d = np.log(va), da / va
                elif op == OP_SQRT:
                    v = np.sqrt(va)
                    d = da / (2.0 * v)
                elif op == OP_LOG1P:
                    v, d = np.log1p(va), da / (1.0 + va)
                elif op == OP_EXPM1:
                    v = np.expm1(va)
                    d = (v + 1.0) * da
                elif op == OP_ABS:
                    v, d = np.abs(va), np.sign(va) * da
                elif op in (OP_MAX, OP_MIN):
                    pick_a = (va > vb) if op == OP_MAX else (va < vb)
                    tie = va == vb
                    v = np.where(pick_a | tie, va, vb)
                    v = np.where(np.isnan(va) | np.isnan(vb), np.nan, v)
                    d = np.where(tie, 0.5 * (da + db), np.where(pick_a, da, db))
                elif op == OP_WHERE:
                    mv, _ = get(a)
                    va, da = get(b)
                    vb, db = get(c)
                    v = np.where(mv != 0, va, vb)
                    d = np.where(mv != 0, da, db)
                else:
                    raise AssertionError(op)
                rv.append(np.broadcast_to(v, x.shape))
                rd.append(np.broadcast_to(d, x.shape))
        v, d = get(self.out)
        v = np.broadcast_to(v, x.shape)
        d = np.broadcast_to(d, x.shape)
        return (v, d) if derivative else v


def trace(n_cols: int, call) -> Tape:
    """Trace `call(sym)` (which evaluates the user's function with the symbolic
    input and per-column bound arguments) into a tape."""
    t = Tape(n_cols)
    return t.finish(call(t.x))
