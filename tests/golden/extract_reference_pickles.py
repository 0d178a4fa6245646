"""Write the reference's pickled solution fixtures as `.npz` golden vectors.

Reads (text of the reference's test data, never imported or executed):

* `tests/data/processes/solution_{uniform,normal,lognormal,tauchen,rouwenhorst}.pkl`
  (checked at decimal 5 by the reference's `tests/test_processes.py:33-70`);
* `tests/data/regression_tests/solution.pkl`
  (checked at decimal 5 by `tests/test_regression_test.py:29-66`).

No unpickler runs. The files are decoded as a data format: `pickletools.genops`
yields the opcode stream without executing it, and `_scan` below keeps a stack
of plain Python values (ints, strings, bytes, tuples, dicts) plus inert
`_Call` records. A global reference is kept as its (module, name) strings and
is never imported, resolved or called; a `_Call` of one only records its
arguments. Afterwards `_arrays` turns the records that the allow-list below
recognises (numpy's `_reconstruct` with its `BUILD` state, numpy `dtype`,
jax's `_reconstruct_array` and cloudpickle's `MappingProxyType`) into numpy
arrays by reading the raw bytes with `np.frombuffer`. Any other global, or an
opcode outside the small set these files use, stops the scan with an error.

* `tests/data/regression_tests/{f32,f64}/{mortality,precautionary_savings}_simulation.pkl`
  (pickled simulation DataFrames, checked at atol 1e-5 by
  `tests/test_regression_test.py:69-140`).

Output (committed): `tests/golden/processes/solution_{dist}.npz` and
`tests/golden/regression/solution.npz`, one entry per `"{period}/{regime}"`;
`tests/golden/regression/{f32,f64}/*_simulation.npz`, one entry per column
(categorical columns as strings, "" where missing).

Run from the repository root: `python tests/golden/extract_reference_pickles.py`.
"""

from __future__ import annotations

import pickletools
from pathlib import Path

import numpy as np

REF = Path("/root/reference/tests/data")
OUT = Path(__file__).resolve().parent

_ALLOWED_GLOBALS = {
    ("cloudpickle.cloudpickle", "_builtin_type"),
    ("jax._src.array", "_reconstruct_array"),
    ("numpy._core.multiarray", "_reconstruct"),
    ("numpy._core.numeric", "_frombuffer"),
    ("numpy", "ndarray"),
    ("numpy", "dtype"),
    # names of a pickled pandas DataFrame (kept as strings, never resolved)
    ("pandas", "DataFrame"),
    ("pandas", "Categorical"),
    ("pandas", "CategoricalDtype"),
    ("pandas", "StringDtype"),
    ("pandas", "RangeIndex"),
    ("pandas", "Index"),
    ("pandas.arrays", "ArrowStringArray"),
    ("pandas.core.indexes.base", "_new_Index"),
    ("pandas.core.internals.managers", "BlockManager"),
    ("pandas._libs.internals", "_unpickle_block"),
    ("pandas._libs.arrays", "__pyx_unpickle_NDArrayBacked"),
    ("pyarrow.lib", "_restore_array"),
    ("pyarrow.lib", "type_for_alias"),
    ("pyarrow.lib", "py_buffer"),
    ("builtins", "slice"),
}
_ALLOWED_DTYPES = {"f8", "f4", "i1", "i2", "i4", "i8", "b1"}


class _Global:
    def __init__(self, module, name):
        if (module, name) not in _ALLOWED_GLOBALS:
            raise ValueError(f"global {module}.{name} is not on the allow-list")
        self.module, self.name = module, name


class _Call:
    """An inert record of `func(*args)` (never evaluated), with BUILD state."""

    def __init__(self, func, args):
        self.func, self.args, self.state = func, args, None


class _Mark:
    pass


def _scan(data: bytes):
    """Decode the opcode stream into plain values and `_Call` records."""
    stack, memo = [], {}

    def pop_mark():
        items = []
        while True:
            x = stack.pop()
            if isinstance(x, _Mark):
                return items[::-1]
            items.append(x)

    for op, arg, _pos in pickletools.genops(data):
        name = op.name
        if name in ("PROTO", "FRAME"):
            continue
        if name == "STOP":
            break
        if name in ("SHORT_BINUNICODE", "BINUNICODE", "SHORT_BINBYTES", "BINBYTES",
                    "BININT1", "BININT", "BININT2", "BINFLOAT"):
            stack.append(arg)
        elif name == "BYTEARRAY8":
            stack.append(bytes(arg))
        elif name == "EMPTY_LIST":
            stack.append([])
        elif name == "APPEND":
            v = stack.pop()
            stack[-1].append(v)
        elif name == "APPENDS":
            items = pop_mark()
            stack[-1].extend(items)
        elif name == "NONE":
            stack.append(None)
        elif name == "NEWTRUE":
            stack.append(True)
        elif name == "NEWFALSE":
            stack.append(False)
        elif name == "MEMOIZE":
            memo[len(memo)] = stack[-1]
        elif name in ("BINGET", "LONG_BINGET"):
            stack.append(memo[arg])
        elif name == "MARK":
            stack.append(_Mark())
        elif name == "EMPTY_TUPLE":
            stack.append(())
        elif name in ("TUPLE1", "TUPLE2", "TUPLE3"):
            k = int(name[-1])
            items = tuple(stack[-k:])
            del stack[-k:]
            stack.append(items)
        elif name == "TUPLE":
            stack.append(tuple(pop_mark()))
        elif name == "EMPTY_DICT":
            stack.append({})
        elif name == "SETITEM":
            v = stack.pop()
            k = stack.pop()
            stack[-1][k] = v
        elif name == "SETITEMS":
            items = pop_mark()
            d = stack[-1]
            for k, v in zip(items[::2], items[1::2]):
                d[k] = v
        elif name == "STACK_GLOBAL":
            nm = stack.pop()
            mod = stack.pop()
            stack.append(_Global(mod, nm))
        elif name == "REDUCE":
            args = stack.pop()
            func = stack.pop()
            if not (isinstance(func, _Global) or _is_proxy_type(func)):
                raise ValueError("REDUCE of a value that is not an allow-listed global")
            stack.append(_Call(func, args))
        elif name == "NEWOBJ":
            args = stack.pop()
            cls = stack.pop()
            if not isinstance(cls, _Global):
                raise ValueError("NEWOBJ of a value that is not an allow-listed global")
            stack.append(_Call(cls, args))
        elif name == "BUILD":
            state = stack.pop()
            if not isinstance(stack[-1], _Call):
                raise ValueError("BUILD on a value that is not a call record")
            stack[-1].state = state
        else:
            raise ValueError(f"opcode {name} is outside the decoded subset")
    if len(stack) != 1:
        raise ValueError("malformed stream")
    return stack[0]


def _is(x, module, name):
    return (isinstance(x, _Call) and isinstance(x.func, _Global)
            and x.func.module == module and x.func.name == name)


def _is_proxy_type(x):
    """The record `_builtin_type("MappingProxyType")` (cloudpickle's name for
    `types.MappingProxyType`)."""
    return _is(x, "cloudpickle.cloudpickle", "_builtin_type") and x.args == ("MappingProxyType",)


def _dtype(rec):
    if not _is(rec, "numpy", "dtype"):
        raise ValueError("expected a dtype record")
    code = rec.args[0]
    if code not in _ALLOWED_DTYPES:
        raise ValueError(f"dtype {code!r} is not on the allow-list")
    order = rec.state[1] if rec.state else "<"
    return np.dtype((order if order in "<>" else "") + code)


def _arrays(x):
    """Plain values and recognised records -> dicts and numpy arrays."""
    if isinstance(x, dict):
        return {k: _arrays(v) for k, v in x.items()}
    if isinstance(x, _Call) and isinstance(x.func, _Global):
        f = x.func
        if (f.module, f.name) == ("jax._src.array", "_reconstruct_array"):
            return _arrays(x.args[2])  # (fun, args, arr_state, aval_state)
        if (f.module, f.name) == ("numpy._core.multiarray", "_reconstruct"):
            _ver, shape, dt, fortran, raw = x.state
            if fortran:
                raise ValueError("Fortran-ordered arrays are not expected here")
            dtype = _dtype(dt)
            return np.frombuffer(raw, dtype=dtype).reshape(shape).copy()
    if isinstance(x, tuple) and len(x) == 5 and isinstance(x[4], bytes):
        # the array state tuple itself (jax passes it as `arr_state`)
        _ver, shape, dt, fortran, raw = x
        return np.frombuffer(raw, dtype=_dtype(dt)).reshape(shape).copy()
    raise ValueError(f"unrecognised record {x!r}")


def _proxy_to_dict(x):
    """MappingProxyType(dict) records -> the dict (recursively)."""
    if isinstance(x, _Call) and _is_proxy_type(x.func):
        (inner,) = x.args
        return {k: _proxy_to_dict(v) for k, v in inner.items()}
    if isinstance(x, dict):
        return {k: _proxy_to_dict(v) for k, v in x.items()}
    return x


def decode_solution(path: Path) -> dict:
    """{period: {regime: array}} of a pickled solution mapping."""
    tree = _proxy_to_dict(_scan(path.read_bytes()))
    return {int(t): {str(r): _arrays(a) for r, a in regs.items()} for t, regs in tree.items()}


def _ndarray(rec):
    """`_frombuffer(buf, dtype, shape, order)` -> the array."""
    if not _is(rec, "numpy._core.numeric", "_frombuffer"):
        raise ValueError("expected a _frombuffer record")
    buf, dt, shape, order = rec.args
    return np.frombuffer(buf, dtype=_dtype(dt)).reshape(shape, order=order).copy()


def _strings(rec):
    """An Arrow large-string array (offsets + UTF-8 data, no nulls) -> list of str."""
    if _is(rec, "pandas.core.indexes.base", "_new_Index"):
        rec = rec.args[1]["data"]
    if not _is(rec, "pandas.arrays", "ArrowStringArray"):
        raise ValueError("expected an ArrowStringArray record")
    arr = rec.state["_pa_array"]
    if not _is(arr, "pyarrow.lib", "_restore_array"):
        raise ValueError("expected a pyarrow array record")
    typ, length, null_count, offset, buffers, children, _dict = arr.args[0]
    if not (_is(typ, "pyarrow.lib", "type_for_alias") and typ.args == ("large_string",)):
        raise ValueError("only large_string arrays are decoded")
    if null_count or offset or children:
        raise ValueError("nulls, offsets or children are not expected here")
    _validity, offs, data = buffers
    offs = np.frombuffer(offs.args[0], dtype="<i8")
    data = data.args[0]
    return [data[offs[i]:offs[i + 1]].decode("utf-8") for i in range(length)]


def _column_values(rec):
    """A block's values: a 2-D numeric array (columns x rows), or a
    categorical as a 1 x rows array of strings ("" where missing)."""
    if _is(rec, "numpy._core.numeric", "_frombuffer"):
        return _ndarray(rec)
    if _is(rec, "pandas._libs.arrays", "__pyx_unpickle_NDArrayBacked"):
        cdtype, codes_rec = rec.state
        if not _is(cdtype, "pandas", "CategoricalDtype"):
            raise ValueError("expected a CategoricalDtype record")
        cats = np.array(_strings(cdtype.state["categories"]) + [""])
        codes = _ndarray(codes_rec).astype(np.int64)
        return cats[np.where(codes >= 0, codes, len(cats) - 1)][None, :]
    raise ValueError("unrecognised block values")


def decode_frame(path: Path) -> dict:
    """{column: 1-D array} of a pickled DataFrame with a RangeIndex."""
    root = _scan(path.read_bytes())
    if not _is(root, "pandas", "DataFrame"):
        raise ValueError("expected a DataFrame record")
    mgr = root.state["_mgr"]
    if not _is(mgr, "pandas.core.internals.managers", "BlockManager"):
        raise ValueError("expected a BlockManager record")
    blocks, axes = mgr.args
    columns = _strings(axes[0])
    out = [None] * len(columns)
    for blk in blocks:
        if not _is(blk, "pandas._libs.internals", "_unpickle_block"):
            raise ValueError("expected a block record")
        values_rec, placement, ndim = blk.args
        if ndim != 2:
            raise ValueError("unexpected block layout")
        if _is(placement, "builtins", "slice"):
            cols = list(range(*placement.args))
        else:  # an integer array of column positions
            cols = [int(c) for c in _ndarray(placement)]
        vals = _column_values(values_rec)
        for k, col in enumerate(cols):
            out[col] = vals[k]
    if any(v is None for v in out):
        raise ValueError("a column has no block")
    return dict(zip(columns, out))


def _write(sol: dict, dest: Path) -> None:
    dest.parent.mkdir(parents=True, exist_ok=True)
    flat = {f"{t}/{r}": np.asarray(a) for t, regs in sol.items() for r, a in regs.items()}
    np.savez(dest, **flat)
    print(f"{dest.relative_to(OUT.parent.parent)}: " +
          ", ".join(f"{k} {v.shape} {v.dtype}" for k, v in flat.items()))


def main() -> None:
    for dist in ("uniform", "normal", "lognormal", "tauchen", "rouwenhorst"):
        _write(decode_solution(REF / "processes" / f"solution_{dist}.pkl"),
               OUT / "processes" / f"solution_{dist}.npz")
    _write(decode_solution(REF / "regression_tests" / "solution.pkl"), OUT / "regression" / "solution.npz")
    for prec in ("f32", "f64"):
        for name in ("mortality_simulation", "precautionary_savings_simulation"):
            frame = decode_frame(REF / "regression_tests" / prec / f"{name}.pkl")
            dest = OUT / "regression" / prec / f"{name}.npz"
            dest.parent.mkdir(parents=True, exist_ok=True)
            np.savez(dest, **frame)
            print(f"{dest.relative_to(OUT.parent.parent)}: " +
                  ", ".join(f"{k} {v.dtype}" for k, v in frame.items()))


if __name__ == "__main__":
    main()
