"""ctypes binding of `liblcm_amd.so` (declared in include/pylcm_amd.h).

The library is loaded from the package directory. There is no fallback: if it
is missing or a call fails, `NativeKernelError` is raised with the library's
own message.
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import numpy as np

from pylcm_amd.exceptions import NativeKernelError

LIB_PATH = Path(__file__).resolve().parent / "liblcm_amd.so"

LCM_DTYPE_F32 = 0
LCM_DTYPE_F64 = 1
MAX_TARGETS = 4
MAX_CONSTR = 4
MAX_SCALAR = 4
MAX_TASTE_D = 8
ACT_CHUNK = 256
REQ_MAGIC = 0x31305145524D434C  # "LCMREQ01", the request files of the native solve server
REQ_TILE = 256
REQ_HOLD = 1
DESC_PTR_WORDS = 36
SERVE_LONGEST_FIRST = 1
GS_OCC_DENSE = 0x100  # or'ed into lcm_gs_launch_ex's variant: the float32 kernel's register budget
GS_OCC_SPARSE = 0x200
GS_DENSE_MIN_TASKS = 6144

vp = C.c_void_p


class LcmGsDesc(C.Structure):
    _fields_ = [
        ("V_out", vp),
        ("pol_out", vp),
        ("w_pts", vp),
        ("U", vp),
        ("A", vp),
        ("B", vp),
        ("Fa", vp * MAX_CONSTR),
        ("Fb", vp * MAX_CONSTR),
        ("Vn", vp * MAX_TARGETS),
        ("M", vp * MAX_TARGETS),
        ("Vs", vp * MAX_SCALAR),
        ("ms", vp * MAX_SCALAR),
        ("g_pts", vp),
        ("Flo", vp),
        ("Fhi", vp),
        ("Fchunk", vp),
        ("Af", vp),
        ("Bf", vp),
        ("U_ssd", C.c_int64), ("U_sw", C.c_int64),
        ("A_ssd", C.c_int64), ("A_sw", C.c_int64),
        ("B_ssd", C.c_int64), ("B_sw", C.c_int64),
        ("Fa_ssd", C.c_int64 * MAX_CONSTR), ("Fa_sw", C.c_int64 * MAX_CONSTR),
        ("Fb_ssd", C.c_int64 * MAX_CONSTR), ("Fb_sw", C.c_int64 * MAX_CONSTR),
        ("Fi_ssd", C.c_int64), ("Fc_ssd", C.c_int64),
        ("n_sd_next", C.c_int32 * MAX_TARGETS),
        ("n_sd", C.c_int32), ("n_w", C.c_int32), ("n_d", C.c_int32), ("n_c", C.c_int32),
        ("n_constr", C.c_int32), ("strict_mask", C.c_int32),
        ("n_targets", C.c_int32), ("n_scalar", C.c_int32),
        ("terminal", C.c_int32), ("g_kind", C.c_int32), ("g_n", C.c_int32),
        ("has_taste", C.c_int32),
        ("use_interval", C.c_int32), ("pad0", C.c_int32),
        ("beta", C.c_double), ("taste_scale", C.c_double),
        ("g_start", C.c_double), ("g_stop", C.c_double),
    ]


class LcmGsTask(C.Structure):
    _fields_ = [("desc", C.c_int32), ("sd", C.c_int32), ("w0", C.c_int32), ("pad", C.c_int32)]


SIM_MAX_FEAT = 4
SIM_MAX_REGIMES = 8
SIM_NOALLOC_MAX_ORDER = 1 << 22  # regimes x agents of lcm_sim_period_noalloc


class LcmEvDesc(C.Structure):
    _fields_ = [("gs", vp), ("EV", vp), ("CS", vp),
                ("n_sd", C.c_int32), ("n_d", C.c_int32), ("g_n", C.c_int32), ("pad", C.c_int32),
                ("fin", vp)]


class LcmSimDesc(C.Structure):
    _fields_ = [
        ("U0", vp), ("A", vp), ("B0", vp), ("lo", vp), ("hi", vp),
        ("Uf", vp * SIM_MAX_FEAT), ("Bf", vp * SIM_MAX_FEAT),
        ("fcoef", vp * SIM_MAX_FEAT), ("fconst", vp * SIM_MAX_FEAT),
        ("EV", vp), ("CS", vp), ("sd_base", vp), ("nodes", vp), ("g_pts", vp), ("preg", vp),
        ("Wd", vp * SIM_MAX_REGIMES),
        ("LoB", vp), ("HiB", vp),
        ("LoF", vp * SIM_MAX_FEAT), ("HiF", vp * SIM_MAX_FEAT),
        ("lo_strict", C.c_int32), ("hi_strict", C.c_int32),
        ("n_sdd_next", C.c_int32 * SIM_MAX_REGIMES),
        ("feat_op", C.c_int32 * SIM_MAX_FEAT),
        ("n_feat", C.c_int32), ("n_sdd", C.c_int32), ("n_d", C.c_int32), ("n_c", C.c_int32),
        ("terminal", C.c_int32), ("interp", C.c_int32), ("n_nodes", C.c_int32), ("p_stride", C.c_int32),
        ("g_kind", C.c_int32), ("g_n", C.c_int32), ("n_regimes", C.c_int32), ("regime_code", C.c_int32),
        ("proc_kind", C.c_int32), ("n_rows", C.c_int32),
        ("beta", C.c_double), ("g_start", C.c_double), ("g_stop", C.c_double),
        ("pa", C.c_double), ("pb", C.c_double), ("pc", C.c_double),
        ("taste_scale", C.c_double),
        ("ev_global", C.c_int32), ("pad_", C.c_int32),
        ("ev_fin", vp),
    ]


class LcmAgents(C.Structure):
    _fields_ = [("w", vp), ("z", vp), ("sdd", vp), ("regime", vp), ("start", vp)]


class LcmSimOut(C.Structure):
    _fields_ = [("regime", vp), ("action", vp), ("value", vp), ("w", vp), ("z", vp), ("sdd", vp)]


EGM_MAX_TARGETS = 4
EGM_MAX_ROWS = 8
EGM_MAX_CAND = 4096
EGM_MAX_ROWS_PER_COL = 64


class LcmEgmRow(C.Structure):
    _fields_ = [("grid", vp), ("value", vp), ("marg", vp), ("n_valid", vp), ("len", C.c_int),
                ("pad", C.c_int)]


class LcmEgmCol(C.Structure):
    _fields_ = [
        ("A", vp), ("W", vp), ("u_prog", vp), ("u_c", vp), ("i_prog", vp), ("i_c", vp),
        ("tR", vp * EGM_MAX_TARGETS), ("tG", vp * EGM_MAX_TARGETS), ("tW", vp * EGM_MAX_TARGETS),
        ("tRows", vp * EGM_MAX_TARGETS),
        ("tprob", C.c_double * EGM_MAX_TARGETS), ("tscalar", C.c_double * EGM_MAX_TARGETS),
        ("tscale", C.c_double * EGM_MAX_TARGETS),
        ("tK", C.c_int * EGM_MAX_TARGETS), ("tN", C.c_int * EGM_MAX_TARGETS),
        ("tkind", C.c_int * EGM_MAX_TARGETS),
        ("out_grid", vp), ("out_policy", vp), ("out_value", vp), ("out_marg", vp),
        ("out_counts", vp), ("out_V", vp), ("cand", vp),
        ("beta", C.c_double), ("ratio", C.c_double), ("jump_thresh", C.c_double), ("w_max", C.c_double),
        ("n_s", C.c_int), ("n_w", C.c_int), ("n_con", C.c_int), ("n_pad", C.c_int), ("n_scan", C.c_int),
        ("n_targets", C.c_int), ("pad0", C.c_int), ("pad1", C.c_int),
    ]


class LcmEgmAgg(C.Structure):
    _fields_ = [("cols", vp), ("out", vp), ("scale", C.c_double), ("n_d", C.c_int), ("n_w", C.c_int)]


_SIGS = {
    "lcm_init": ([C.c_int], C.c_int),
    "lcm_last_error": ([], C.c_char_p),
    "lcm_device_info": ([C.POINTER(C.c_int), C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], C.c_int),
    "lcm_ctx_create": ([C.POINTER(vp)], C.c_int),
    "lcm_ctx_destroy": ([vp], C.c_int),
    "lcm_ctx_sync": ([vp], C.c_int),
    "lcm_alloc": ([C.c_size_t, C.POINTER(vp)], C.c_int),
    "lcm_free": ([vp], C.c_int),
    "lcm_host_alloc": ([C.c_size_t, C.POINTER(vp)], C.c_int),
    "lcm_host_free": ([vp], C.c_int),
    "lcm_memcpy_h2d": ([vp, vp, vp, C.c_size_t], C.c_int),
    "lcm_memcpy_d2h": ([vp, vp, vp, C.c_size_t], C.c_int),
    "lcm_memcpy_h2d_async": ([vp, vp, vp, C.c_size_t], C.c_int),
    "lcm_memcpy_d2h_async": ([vp, vp, vp, C.c_size_t], C.c_int),
    "lcm_memcpy_d2d_async": ([vp, vp, vp, C.c_size_t], C.c_int),
    "lcm_memset_async": ([vp, vp, C.c_int, C.c_size_t], C.c_int),
    "lcm_gs_launch": ([vp, C.c_int, vp, vp, C.c_int, C.c_int, C.c_size_t], C.c_int),
    "lcm_gs_launch_ex": ([vp, C.c_int, vp, vp, C.c_int, C.c_int, C.c_size_t, C.c_int], C.c_int),
    "lcm_gs_serve_launch": ([vp, C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_size_t, C.c_int, vp, vp,
                             C.POINTER(vp), C.POINTER(vp), C.POINTER(C.c_size_t), C.c_int, vp], C.c_int),
    "lcm_graph_begin": ([vp], C.c_int),
    "lcm_graph_end": ([vp, C.POINTER(vp)], C.c_int),
    "lcm_graph_launch": ([vp, vp], C.c_int),
    "lcm_graph_destroy": ([vp], C.c_int),
    "lcm_event_create": ([C.POINTER(vp)], C.c_int),
    "lcm_event_destroy": ([vp], C.c_int),
    "lcm_event_record": ([vp, vp], C.c_int),
    "lcm_event_elapsed": ([vp, vp, C.POINTER(C.c_float)], C.c_int),
    "lcm_event_query": ([vp], C.c_int),
    "lcm_event_synchronize": ([vp], C.c_int),
    "lcm_event_elapsed_batch": ([vp, C.POINTER(vp), C.c_int, C.POINTER(C.c_float)], C.c_int),
    "lcm_copy_launch": ([vp, vp, vp, C.c_size_t], C.c_int),
    "lcm_abi_sizes": ([C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], C.c_int),
    "lcm_gs_lds_bytes": ([C.c_int, C.c_int, C.c_int], C.c_size_t),
    "lcm_abi_sizes_sim": ([C.POINTER(C.c_size_t)] * 4, C.c_int),
    "lcm_ev_launch": ([vp, C.c_int, vp, C.c_int, C.c_int], C.c_int),
    "lcm_sim_period_ex": ([vp, C.c_int, vp, C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_uint64,
                           C.c_size_t, vp, vp], C.c_int),
    "lcm_sim_period_noalloc": ([vp, C.c_int, vp, C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_uint64,
                                C.c_size_t], C.c_int),
    "lcm_sim_lds_bytes": ([C.c_int, C.c_int, C.c_int, C.c_int], C.c_size_t),
    "lcm_egm_launch": ([vp, vp, C.c_int, C.c_size_t, C.c_int, C.c_int], C.c_int),
    "lcm_egm_launch_ex": ([vp, vp, C.c_int, C.c_size_t, C.c_int, C.c_int, C.c_int], C.c_int),
    "lcm_egm_aggregate": ([vp, C.c_int, vp, C.c_int, C.c_int], C.c_int),
    "lcm_egm_lds_bytes": ([C.c_int], C.c_size_t),
    "lcm_abi_sizes_egm": ([C.POINTER(C.c_size_t)] * 3, C.c_int),
    "lcm_tape_eval": ([vp, vp, vp, vp, vp, vp, C.c_int], C.c_int),
    "lcm_req_inspect": ([C.c_char_p, C.POINTER(C.c_int64), C.c_int], C.c_int),
    "lcm_tile_choose": ([vp, vp, vp, C.c_int, C.c_int64, C.c_int64], C.c_int),
    "lcm_split_tasks": ([vp, C.c_int64, vp, C.c_int, vp], C.c_int64),
    "lcm_serve_clock": ([], C.c_double),
    "lcm_serve_create": ([C.POINTER(vp), C.c_int, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_size_t, C.c_int,
                          C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_int, C.c_int], C.c_int),
    "lcm_serve_submit": ([vp, C.c_int64, C.c_char_p], C.c_int),
    "lcm_serve_wait": ([vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_int, C.c_int], C.c_int),
    "lcm_serve_fetch": ([vp, C.c_int64, vp, C.c_size_t, C.POINTER(C.c_double), C.c_char_p, C.c_size_t], C.c_int),
    "lcm_serve_held": ([vp, C.c_int64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], C.c_int),
    "lcm_serve_release": ([vp, C.c_int64], C.c_int),
    "lcm_serve_stats": ([vp, C.POINTER(C.c_int64), C.c_int], C.c_int),
    "lcm_serve_busy_ms": ([vp, C.POINTER(C.c_double)], C.c_int),
    "lcm_serve_destroy": ([vp], C.c_int),
    "lcm_philox": ([vp, vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32], C.c_int),
}

_lock = threading.Lock()
_lib = None
_initialised = False


def declared_symbols() -> list[str]:
    return list(_SIGS)


def load_library(path: Path | None = None):
    """Load the shared library and declare every signature (no device call)."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path else LIB_PATH
        if not p.exists():
            raise NativeKernelError(
                f"HIP library {p} is missing; build it with `python -m pylcm_amd.build_native`"
            )
        lib = C.CDLL(str(p))
        for name, (args, res) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        if path is None:
            _lib = lib
        return lib


def _check(status: int, what: str):
    if status != 0:
        msg = _lib.lcm_last_error().decode(errors="replace") if _lib else ""
        raise NativeKernelError(f"{what} failed ({status}): {msg}", status=status)


def lib():
    """The library, with the device initialised (device 0)."""
    global _initialised
    L = load_library()
    if not _initialised:
        with _lock:
            if not _initialised:
                _check(L.lcm_init(0), "lcm_init")
                _initialised = True
    return L


def call(name: str, *args):
    L = lib()
    _check(getattr(L, name)(*args), name)


def host_error() -> str:
    """The library's last error message (host-only entry points)."""
    return load_library().lcm_last_error().decode(errors="replace")


def device_info():
    L = lib()
    n = C.c_int()
    f = C.c_size_t()
    t = C.c_size_t()
    _check(L.lcm_device_info(C.byref(n), C.byref(f), C.byref(t)), "lcm_device_info")
    return n.value, f.value, t.value


class Context:
    """One HIP stream; owns nothing else."""

    def __init__(self):
        L = lib()
        h = vp()
        _check(L.lcm_ctx_create(C.byref(h)), "lcm_ctx_create")
        self.handle = h

    def sync(self):
        call("lcm_ctx_sync", self.handle)

    def close(self):
        if self.handle:
            lib().lcm_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """A device allocation; freed when the object is collected."""

    def __init__(self, nbytes: int):
        L = lib()
        p = vp()
        _check(L.lcm_alloc(C.c_size_t(int(nbytes)), C.byref(p)), "lcm_alloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def free(self):
        if self.ptr:
            lib().lcm_free(vp(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedBuffer:
    """Pinned host memory viewed as a numpy byte array."""

    def __init__(self, nbytes: int):
        L = lib()
        p = vp()
        _check(L.lcm_host_alloc(C.c_size_t(int(nbytes)), C.byref(p)), "lcm_host_alloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)
        self.array = np.ctypeslib.as_array((C.c_uint8 * max(self.nbytes, 1)).from_address(self.ptr))

    def free(self):
        if self.ptr:
            lib().lcm_host_free(vp(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def h2d(ctx: Context, dst: int, arr: np.ndarray):
    a = np.ascontiguousarray(arr)
    call("lcm_memcpy_h2d", ctx.handle, vp(dst), a.ctypes.data_as(vp), C.c_size_t(a.nbytes))


def d2h(ctx: Context, src: int, out: np.ndarray):
    assert out.flags.c_contiguous
    call("lcm_memcpy_d2h", ctx.handle, vp(out.ctypes.data), vp(src), C.c_size_t(out.nbytes))


def d2d(ctx: Context, dst: int, src: int, nbytes: int):
    """Device-to-device copy on ctx's stream (asynchronous)."""
    call("lcm_memcpy_d2d_async", ctx.handle, vp(dst), vp(src), C.c_size_t(int(nbytes)))


class DeviceArray:
    """A device-resident array: `ptr` points into memory that `owner` keeps
    alive. `np.asarray(a)` (or `a.to_numpy()`) copies it to the host after
    the work queued on `ctx` has finished."""

    def __init__(self, ctx: Context, ptr: int, shape, dtype, owner=None):
        self.ctx = ctx
        self.ptr = int(ptr)
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.owner = owner

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64)) if self.shape else 1

    @property
    def nbytes(self) -> int:
        return self.size * self.dtype.itemsize

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.size, dtype=self.dtype)
        d2h(self.ctx, self.ptr, out)
        return out.reshape(self.shape)

    def __array__(self, dtype=None, copy=None):
        a = self.to_numpy()
        return a if dtype is None else a.astype(dtype)

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


class Event:
    def __init__(self):
        h = vp()
        _check(lib().lcm_event_create(C.byref(h)), "lcm_event_create")
        self.handle = h

    def query(self) -> bool:
        st = lib().lcm_event_query(self.handle)
        if st < 0:
            _check(st, "lcm_event_query")
        return st == 1

    def synchronize(self):
        call("lcm_event_synchronize", self.handle)

    def record(self, ctx: Context):
        call("lcm_event_record", ctx.handle, self.handle)

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        _check(lib().lcm_event_elapsed(self.handle, later.handle, C.byref(ms)), "lcm_event_elapsed")
        return float(ms.value)

    def __del__(self):
        try:
            lib().lcm_event_destroy(self.handle)
        except Exception:
            pass


class Graph:
    """An instantiated HIP graph captured from a context's stream."""

    def __init__(self, exec_handle):
        self.handle = exec_handle

    @staticmethod
    def capture(ctx: Context, record):
        call("lcm_graph_begin", ctx.handle)
        try:
            record()
        except BaseException:
            h = vp()
            lib().lcm_graph_end(ctx.handle, C.byref(h))
            if h.value:
                lib().lcm_graph_destroy(h)
            raise
        h = vp()
        call("lcm_graph_end", ctx.handle, C.byref(h))
        return Graph(h)

    def launch(self, ctx: Context):
        call("lcm_graph_launch", ctx.handle, self.handle)

    def __del__(self):
        try:
            if self.handle:
                lib().lcm_graph_destroy(self.handle)
        except Exception:
            pass
