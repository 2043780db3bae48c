"""fastpoly_amd — MI355X (gfx950) build of fastpoly's exact NTT convolution path.

Python mirror of the reference API (fpx::poly::conv / inv / log / exp / sqrt,
NttPlan<M>, release_scratch_memory / scratch_memory_stats) over the C ABI of
libfpxg.so (include/fpxg.h).  Coefficients are the reference's raw Montgomery
limbs (x * 2^32 mod p) as uint32 numpy arrays; results are bit-identical to
the reference CPU library.

Errors map onto the reference's exceptions: DomainError (poly::domain_error),
NttSizeError (ntt_size_error / std::length_error), MemoryError (bad_alloc).
There is no CPU fallback: importing works without a GPU, but every compute
call needs the built library and a device, and raises otherwise.
"""
from __future__ import annotations

import builtins
import ctypes
import os

import numpy as np

__all__ = [
    "DomainError", "NttSizeError", "DeviceError", "UnsupportedModulus", "lib", "lib_path",
    "conv", "inv", "log", "exp", "sqrt", "series", "ntt_forward", "ntt_inverse", "NttPlan",
    "release_scratch_memory", "scratch_memory_stats", "from_int", "to_int", "Mont", "DeviceArray",
    "EXPORTED_SYMBOLS", "Request", "batch", "batch_submit", "wait", "Server", "pow", "divmod", "div", "mod",
]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libfpxg.so"

OK, E_DOMAIN, E_SIZE, E_NOMEM, E_DEVICE, E_UNSUPPORTED, E_ARG, E_BUSY = range(8)
OPS = {"inv": 0, "log": 1, "exp": 2, "sqrt": 3, "conv": 4}
NTT_FORWARD, NTT_FORWARD_PREFIX, NTT_INVERSE, NTT_INVERSE_PREFIX, NTT_INVERSE_HIGHHALF = range(5)
SIZE_MAX = (1 << 64) - 1

# Every extern "C" entry point of include/fpxg.h.
EXPORTED_SYMBOLS = [
    "fpxg_init", "fpxg_shutdown", "fpxg_strerror", "fpxg_last_error", "fpxg_arena_stats",
    "fpxg_release_scratch", "fpxg_conv_u32", "fpxg_series_u32", "fpxg_ntt_u32", "fpxg_conv_dev",
    "fpxg_series_dev", "fpxg_ntt_dev", "fpxg_dev_alloc", "fpxg_dev_free", "fpxg_memcpy_h2d",
    "fpxg_memcpy_d2h", "fpxg_synchronize", "fpxg_stream", "fpxg_event_record", "fpxg_event_elapsed",
    "fpxg_ntt_pass_dev", "fpxg_ntt_plan_info", "fpxg_batch_submit", "fpxg_wait", "fpxg_batch_u32",
    "fpxg_server_create", "fpxg_server_submit", "fpxg_server_wait", "fpxg_server_poll", "fpxg_server_now",
    "fpxg_server_stats_get", "fpxg_server_destroy", "fpxg_autotune", "fpxg_set_option",
    "fpxg_pow_u32", "fpxg_pow_dev", "fpxg_divmod_u32", "fpxg_divmod_dev", "fpxg_product_plan_info",
    "fpxg_conv_dev_noalloc",
]


class FpxgError(RuntimeError):
    code = E_DEVICE


class DomainError(FpxgError, ValueError):
    """poly::domain_error (detail/poly/base.hpp:26-30)."""
    code = E_DOMAIN


class NttSizeError(FpxgError):
    """ntt_size_error (ntt.hpp:59-63) and std::length_error."""
    code = E_SIZE


class DeviceError(FpxgError):
    code = E_DEVICE


class UnsupportedModulus(FpxgError, ValueError):
    code = E_UNSUPPORTED


class ArgumentError(FpxgError, ValueError):
    code = E_ARG


_ERRORS = {E_DOMAIN: DomainError, E_SIZE: NttSizeError, E_DEVICE: DeviceError,
           E_UNSUPPORTED: UnsupportedModulus, E_ARG: ArgumentError}

_u32p = ctypes.POINTER(ctypes.c_uint32)
_size_p = ctypes.POINTER(ctypes.c_size_t)


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_size_t) for n in (
        "cached_bytes", "cached_blocks", "system_allocations", "cache_hits", "in_use_bytes", "capacity_bytes")]


class Req(ctypes.Structure):
    """fpxg_req (include/fpxg.h)."""
    _fields_ = [("op", ctypes.c_int), ("p", ctypes.c_uint32), ("a", ctypes.c_void_p), ("la", ctypes.c_size_t),
                ("b", ctypes.c_void_p), ("lb", ctypes.c_size_t), ("n", ctypes.c_size_t),
                ("out", ctypes.c_void_p), ("out_cap", ctypes.c_size_t), ("out_len", ctypes.c_size_t),
                ("status", ctypes.c_int)]


class Result(ctypes.Structure):
    """fpxg_result (include/fpxg.h)."""
    _fields_ = [("status", ctypes.c_int), ("out_len", ctypes.c_size_t), ("submit_s", ctypes.c_double),
                ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    """fpxg_server_stats (include/fpxg.h)."""
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("iterations", ctypes.c_uint64),
                ("launches", ctypes.c_uint64), ("max_active", ctypes.c_uint64), ("max_queue", ctypes.c_uint64),
                ("peak_in_use_bytes", ctypes.c_size_t)]


def lib_path() -> str:
    return os.path.join(HERE, LIB_NAME)


_lib = None


def lib():
    """The loaded libfpxg.so; raises if it was not built (no fallback path)."""
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `python -m fastpoly_amd.build`")
        L = ctypes.CDLL(path)
        L.fpxg_init.argtypes = [ctypes.c_int, ctypes.c_size_t]
        L.fpxg_strerror.restype = ctypes.c_char_p
        L.fpxg_strerror.argtypes = [ctypes.c_int]
        L.fpxg_last_error.restype = ctypes.c_char_p
        L.fpxg_arena_stats.argtypes = [ctypes.POINTER(Stats)]
        L.fpxg_conv_u32.argtypes = [ctypes.c_uint32, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t,
                                    ctypes.c_size_t, _u32p, ctypes.c_size_t, _size_p]
        L.fpxg_conv_dev.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                    ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, _size_p]
        L.fpxg_conv_dev_noalloc.argtypes = L.fpxg_conv_dev.argtypes
        L.fpxg_series_u32.argtypes = [ctypes.c_int, ctypes.c_uint32, _u32p, ctypes.c_size_t, ctypes.c_size_t, _u32p]
        L.fpxg_series_dev.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.c_size_t, ctypes.c_void_p]
        L.fpxg_ntt_u32.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t,
                                   _u32p]
        L.fpxg_ntt_dev.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t,
                                   ctypes.c_void_p]
        L.fpxg_ntt_pass_dev.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                        ctypes.c_void_p]
        _ip = ctypes.POINTER(ctypes.c_int)
        L.fpxg_ntt_plan_info.argtypes = [ctypes.c_uint32, _ip, _ip, _ip]
        L.fpxg_product_plan_info.argtypes = [ctypes.c_uint32, _ip, _ip]
        L.fpxg_dev_alloc.restype = ctypes.c_void_p
        L.fpxg_dev_alloc.argtypes = [ctypes.c_size_t]
        L.fpxg_dev_free.argtypes = [ctypes.c_void_p]
        L.fpxg_memcpy_h2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.fpxg_memcpy_d2h.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.fpxg_stream.restype = ctypes.c_void_p
        L.fpxg_event_record.argtypes = [ctypes.c_int]
        L.fpxg_event_elapsed.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.fpxg_batch_submit.argtypes = [ctypes.POINTER(Req), ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint64)]
        L.fpxg_wait.argtypes = [ctypes.c_uint64]
        L.fpxg_batch_u32.argtypes = [ctypes.POINTER(Req), ctypes.c_size_t]
        L.fpxg_server_create.restype = ctypes.c_void_p
        L.fpxg_server_create.argtypes = [ctypes.c_size_t, ctypes.c_int]
        L.fpxg_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(Req), ctypes.POINTER(ctypes.c_uint64)]
        L.fpxg_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(Result)]
        L.fpxg_server_poll.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(Result)]
        L.fpxg_server_now.restype = ctypes.c_double
        L.fpxg_server_now.argtypes = [ctypes.c_void_p]
        L.fpxg_server_stats_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServerStats)]
        L.fpxg_autotune.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_float)]
        L.fpxg_set_option.argtypes = [ctypes.c_int, ctypes.c_int64]
        L.fpxg_pow_u32.argtypes = [ctypes.c_uint32, _u32p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_size_t, _u32p]
        L.fpxg_pow_dev.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64,
                                   ctypes.c_size_t, ctypes.c_void_p]
        L.fpxg_divmod_u32.argtypes = [ctypes.c_uint32, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t, _u32p,
                                      ctypes.c_size_t, _size_p, _u32p, ctypes.c_size_t, _size_p]
        L.fpxg_divmod_dev.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                      ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, _size_p,
                                      ctypes.c_void_p, ctypes.c_size_t, _size_p]
        L.fpxg_server_destroy.restype = None
        L.fpxg_server_destroy.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc == OK:
        return
    msg = lib().fpxg_last_error().decode(errors="replace")
    if rc == E_NOMEM:
        raise MemoryError(f"{what}: {msg}")
    raise _ERRORS.get(rc, FpxgError)(f"{what}: {msg}")


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


# ---------------------------------------------------------------------------
# Montgomery helpers on the host (modint.hpp:57-180): conversion of plain
# integers to and from the limb representation.
# ---------------------------------------------------------------------------
def from_int(p: int, x) -> np.ndarray | int:
    if p == 2:
        conv = lambda v: int(v) & 1  # noqa: E731
    else:
        one = (1 << 32) % p
        conv = lambda v: (int(v) % p) * one % p  # noqa: E731
    if np.isscalar(x):
        return conv(x)
    return np.array([conv(v) for v in np.asarray(x).ravel()], dtype=np.uint32)


def to_int(p: int, limbs) -> np.ndarray | int:
    if p == 2:
        conv = lambda v: int(v) & 1  # noqa: E731
    else:
        rinv = builtins.pow(1 << 32, p - 2, p)
        conv = lambda v: int(v) * rinv % p  # noqa: E731
    if np.isscalar(limbs):
        return conv(limbs)
    return np.array([conv(v) for v in np.asarray(limbs).ravel()], dtype=np.uint64)


class Mont:
    """Field descriptor mirroring fpx::Mont<Mod, Root> (root 0 = smallest primitive root)."""

    ALIASES = {998244353: 3, 1004535809: 3, 469762049: 3, 167772161: 3, 754974721: 11, 1224736769: 3}

    def __init__(self, mod: int, root: int = 0):
        self.mod = int(mod)
        self.root = int(root) if root else self.ALIASES.get(self.mod, 0)

    def __repr__(self):
        return f"Mont<{self.mod}, {self.root}>"


# ---------------------------------------------------------------------------
# The reference API.
# ---------------------------------------------------------------------------
def init(device: int = 0, arena_bytes: int = 0) -> None:
    _check(lib().fpxg_init(device, arena_bytes), "fpxg_init")


def conv(p: int, a, b, n: int | None = None) -> np.ndarray:
    """poly::conv(a, b, n) (detail/poly/convolution.hpp:145-147)."""
    a, b = _u32(a), _u32(b)
    lim = SIZE_MAX if n is None else int(n)
    la, lb = min(a.size, lim), min(b.size, lim)
    cap = min(la + lb - 1, lim) if la and lb else 0
    out = np.empty(max(cap, 1), dtype=np.uint32)
    got = ctypes.c_size_t(0)
    same = a is b
    rc = lib().fpxg_conv_u32(p, _ptr(a), a.size, _ptr(a) if same else _ptr(b), b.size, lim, _ptr(out),
                             out.size, ctypes.byref(got))
    _check(rc, "conv")
    return out[: got.value].copy()


def series(op: str, p: int, a, n: int) -> np.ndarray:
    a = _u32(a)
    out = np.empty(max(int(n), 1), dtype=np.uint32)
    rc = lib().fpxg_series_u32(OPS[op], p, _ptr(a), a.size, int(n), _ptr(out))
    _check(rc, op)
    return out[: int(n)].copy()


def inv(p: int, a, n: int) -> np.ndarray:
    """poly::inv (detail/poly/inverse.hpp:163)."""
    return series("inv", p, a, n)


def log(p: int, a, n: int) -> np.ndarray:
    """poly::log (detail/poly/exponential.hpp:85)."""
    return series("log", p, a, n)


def exp(p: int, a, n: int) -> np.ndarray:
    """poly::exp (detail/poly/exponential.hpp:107)."""
    return series("exp", p, a, n)


def sqrt(p: int, a, n: int) -> np.ndarray:
    """poly::sqrt (detail/poly/square_root.hpp:12)."""
    return series("sqrt", p, a, n)


def _ntt(p: int, root: int, data, variant: int, count: int = 0) -> np.ndarray:
    x = _u32(data).copy()
    _check(lib().fpxg_ntt_u32(p, root, x.size, variant, count, _ptr(x)), "ntt")
    return x


def ntt_forward(p: int, data, root: int = 0) -> np.ndarray:
    """NttPlan<M>::forward (ntt.hpp:142): canonical limbs, bit-reversed order."""
    return _ntt(p, root, data, NTT_FORWARD)


def ntt_inverse(p: int, data, root: int = 0) -> np.ndarray:
    """NttPlan<M>::inverse (ntt.hpp:179), normalised by 1/n."""
    return _ntt(p, root, data, NTT_INVERSE)


class NttPlan:
    """NttPlan<Mont<mod, root>>::get(n) (ntt.hpp:130-192) backed by the device."""

    def __init__(self, field: Mont, n: int):
        self.field, self.n = field, int(n)

    def forward(self, a):
        return _ntt(self.field.mod, self.field.root, a, NTT_FORWARD)

    forward_lazy = forward  # canonical values satisfy the [0, 2p) contract

    def forward_prefix(self, a, count: int):
        return _ntt(self.field.mod, self.field.root, a, NTT_FORWARD_PREFIX, count)

    def inverse(self, a):
        return _ntt(self.field.mod, self.field.root, a, NTT_INVERSE)

    def inverse_prefix(self, a, count: int):
        return _ntt(self.field.mod, self.field.root, a, NTT_INVERSE_PREFIX, count)

    def inverse_highhalf(self, a):
        return _ntt(self.field.mod, self.field.root, a, NTT_INVERSE_HIGHHALF)


def release_scratch_memory() -> None:
    """release_scratch_memory (memory.hpp:251-255): return idle arena memory."""
    _check(lib().fpxg_release_scratch(), "release_scratch")


def scratch_memory_stats() -> dict:
    """scratch_memory_stats (memory.hpp:256-258) for the device arena."""
    s = Stats()
    _check(lib().fpxg_arena_stats(ctypes.byref(s)), "arena_stats")
    return {name: getattr(s, name) for name, _ in Stats._fields_}


# ---------------------------------------------------------------------------
# Device-resident buffers for the *_dev entry points.
# ---------------------------------------------------------------------------
class DeviceArray:
    """uint32 buffer in the library's device arena."""

    def __init__(self, n: int):
        self.n = int(n)
        self.ptr = lib().fpxg_dev_alloc(max(self.n, 1) * 4)
        if not self.ptr:
            _check(E_NOMEM, "dev_alloc")

    @classmethod
    def from_host(cls, a) -> "DeviceArray":
        a = _u32(a)
        d = cls(a.size)
        if a.size:
            _check(lib().fpxg_memcpy_h2d(d.ptr, a.ctypes.data, a.size * 4), "h2d")
        return d

    def to_host(self, n: int | None = None) -> np.ndarray:
        n = self.n if n is None else int(n)
        out = np.empty(max(n, 1), dtype=np.uint32)
        if n:
            _check(lib().fpxg_memcpy_d2h(out.ctypes.data, self.ptr, n * 4), "d2h")
        return out[:n]

    def free(self) -> None:
        if self.ptr:
            lib().fpxg_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def conv_dev(p: int, a: DeviceArray, la: int, b: DeviceArray, lb: int, out: DeviceArray,
             n: int | None = None) -> int:
    got = ctypes.c_size_t(0)
    lim = SIZE_MAX if n is None else int(n)
    _check(lib().fpxg_conv_dev(p, a.ptr, la, b.ptr, lb, lim, out.ptr, out.n, ctypes.byref(got)), "conv_dev")
    return got.value


def conv_dev_noalloc(p: int, a: DeviceArray, la: int, b: DeviceArray, lb: int, out: DeviceArray,
                     n: int | None = None) -> int:
    """conv_dev through fpxg_conv_dev_noalloc: scratch from the library's static
    device array, no device allocation."""
    got = ctypes.c_size_t(0)
    lim = SIZE_MAX if n is None else int(n)
    _check(lib().fpxg_conv_dev_noalloc(p, a.ptr, la, b.ptr, lb, lim, out.ptr, out.n, ctypes.byref(got)),
           "conv_dev_noalloc")
    return got.value


def series_dev(op: str, p: int, a: DeviceArray, la: int, n: int, out: DeviceArray) -> None:
    _check(lib().fpxg_series_dev(OPS[op], p, a.ptr, la, n, out.ptr), op + "_dev")


def ntt_dev(p: int, data: DeviceArray, variant: int = NTT_FORWARD, count: int = 0, root: int = 0) -> None:
    _check(lib().fpxg_ntt_dev(p, root, data.n, variant, count, data.ptr), "ntt_dev")


def synchronize() -> None:
    _check(lib().fpxg_synchronize(), "synchronize")


def event_record(slot: int) -> None:
    _check(lib().fpxg_event_record(slot), "event_record")


def event_elapsed_ms(a: int, b: int) -> float:
    ms = ctypes.c_float(0)
    _check(lib().fpxg_event_elapsed(a, b, ctypes.byref(ms)), "event_elapsed")
    return float(ms.value)


def ntt_pass_dev(p: int, data: DeviceArray, pass_index: int, root: int = 0) -> None:
    _check(lib().fpxg_ntt_pass_dev(p, root, data.n, pass_index, data.ptr), "ntt_pass_dev")


def ntt_plan_info(n: int) -> dict:
    r, c, t = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    _check(lib().fpxg_ntt_plan_info(n, ctypes.byref(r), ctypes.byref(c), ctypes.byref(t)), "plan_info")
    return {"row_log": r.value, "col_levels": c.value, "tile_cols_log": t.value}


def product_plan_info(n: int) -> dict:
    """Split of an n-point fused product (column levels, row words)."""
    r, c = ctypes.c_int(0), ctypes.c_int(0)
    _check(lib().fpxg_product_plan_info(n, ctypes.byref(r), ctypes.byref(c)), "product_plan_info")
    return {"row_log": r.value, "col_levels": c.value}


# ---------------------------------------------------------------------------
# Batches: many conv / series calls merged into shared launches.
# ---------------------------------------------------------------------------
class Request:
    """One call of a batch: ("conv", p, a, b[, n]) or (op, p, a, n) for op in inv/log/exp/sqrt."""

    def __init__(self, op: str, p: int, a, b=None, n: int | None = None):
        self.op, self.p = op, int(p)
        self.a = _u32(a)
        self.b = None if b is None else _u32(b)
        if op == "conv":
            self.n = SIZE_MAX if n is None else int(n)
        else:
            if n is None:
                raise ValueError("series requests need n")
            self.n = int(n)
        self.out = None
        self.status = None

    def out_capacity(self) -> int:
        if self.op == "conv":
            la, lb = min(self.a.size, self.n), min(self.b.size, self.n)
            return 0 if la == 0 or lb == 0 else min(la + lb - 1, self.n)
        return self.n


def batch(requests) -> list:
    """Runs host-array requests as one merged batch (fpxg_batch_u32).  Returns
    each request's result array, or the exception the single call would raise."""
    requests = list(requests)
    k = len(requests)
    arr = (Req * max(k, 1))()
    outs = []
    for i, r in enumerate(requests):
        cap = r.out_capacity()
        out = np.zeros(max(cap, 1), dtype=np.uint32)
        outs.append(out)
        q = arr[i]
        q.op, q.p = OPS[r.op], r.p
        q.a, q.la = (r.a.ctypes.data if r.a.size else None), r.a.size
        if r.b is not None:
            q.b, q.lb = (r.b.ctypes.data if r.b.size else None), r.b.size
        q.n, q.out, q.out_cap = r.n, out.ctypes.data, cap
    _check(lib().fpxg_batch_u32(arr, k), "batch")
    results = []
    for i, r in enumerate(requests):
        st = arr[i].status
        r.status = st
        if st == OK:
            r.out = outs[i][: arr[i].out_len].copy()
            results.append(r.out)
        else:
            msg = fpxg_strerror(st)
            results.append(MemoryError(msg) if st == E_NOMEM else _ERRORS.get(st, FpxgError)(msg))
    return results


def fpxg_strerror(code: int) -> str:
    return lib().fpxg_strerror(code).decode()


def batch_submit(reqs) -> int:
    """Device-pointer batch: reqs is a ctypes array of Req (device addresses).
    Returns the ticket; statuses and out_len are filled before return."""
    t = ctypes.c_uint64(0)
    _check(lib().fpxg_batch_submit(reqs, len(reqs), ctypes.byref(t)), "batch_submit")
    return t.value


def wait(ticket: int) -> None:
    _check(lib().fpxg_wait(ticket), "wait")


# ---------------------------------------------------------------------------
# Serving: continuous batching over host-pointer requests.
# ---------------------------------------------------------------------------
class Server:
    """fpxg_server: submit() returns an id at once; wait(id) returns
    (result array or exception, timings dict) once the request completes."""

    def __init__(self, hbm_budget_bytes: int = 0, levels_per_iteration: int = 0):
        self._h = lib().fpxg_server_create(hbm_budget_bytes, levels_per_iteration)
        if not self._h:
            _check(E_DEVICE, "server_create")
        self._live = {}

    def now(self) -> float:
        return lib().fpxg_server_now(self._h)

    def submit(self, r: "Request") -> int:
        cap = r.out_capacity()
        out = np.zeros(max(cap, 1), dtype=np.uint32)
        q = Req()
        q.op, q.p = OPS[r.op], r.p
        q.a, q.la = (r.a.ctypes.data if r.a.size else None), r.a.size
        if r.b is not None:
            q.b, q.lb = (r.b.ctypes.data if r.b.size else None), r.b.size
        q.n, q.out, q.out_cap = r.n, out.ctypes.data, cap
        rid = ctypes.c_uint64(0)
        _check(lib().fpxg_server_submit(self._h, ctypes.byref(q), ctypes.byref(rid)), "server_submit")
        self._live[rid.value] = (r, q, out)  # keeps every buffer alive until completion
        return rid.value

    def _collect(self, rid: int, res: Result):
        r, q, out = self._live.pop(rid)
        t = {"submit_s": res.submit_s, "admit_s": res.admit_s, "done_s": res.done_s,
             "latency_s": res.done_s - res.submit_s}
        if res.status == OK:
            return out[: res.out_len].copy(), t
        msg = fpxg_strerror(res.status)
        return (MemoryError(msg) if res.status == E_NOMEM else _ERRORS.get(res.status, FpxgError)(msg)), t

    def wait(self, rid: int):
        res = Result()
        _check(lib().fpxg_server_wait(self._h, rid, ctypes.byref(res)), "server_wait")
        return self._collect(rid, res)

    def poll(self, rid: int):
        """None while pending, else as wait()."""
        res = Result()
        rc = lib().fpxg_server_poll(self._h, rid, ctypes.byref(res))
        if rc == E_BUSY:
            return None
        _check(rc, "server_poll")
        return self._collect(rid, res)

    def stats(self) -> dict:
        s = ServerStats()
        _check(lib().fpxg_server_stats_get(self._h, ctypes.byref(s)), "server_stats")
        return {name: getattr(s, name) for name, _ in ServerStats._fields_}

    def close(self) -> None:
        if self._h:
            lib().fpxg_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(min_log: int = 13, max_log: int = 24, reps: int = 5) -> dict:
    """Picks the fastest pass split per transform size (fpxg_autotune);
    returns {log2 n: (column levels, microseconds)}."""
    k = max_log - min_log + 1
    cols = (ctypes.c_int * k)()
    us = (ctypes.c_float * k)()
    _check(lib().fpxg_autotune(min_log, max_log, reps, cols, us), "autotune")
    return {min_log + i: (cols[i], float(us[i])) for i in range(k)}


OPT_GRAPHS = 0
OPT_ROW_DMA = 1
OPT_CRT_FUSED_LOG = 2
OPT_STEP_LOG_MAX = 3


def set_graphs(enabled: bool) -> None:
    """Replay repeated single series calls from captured hipGraphs (default on)."""
    _check(lib().fpxg_set_option(OPT_GRAPHS, 1 if enabled else 0), "set_option")


def set_row_dma(enabled: bool) -> None:
    """Forward row passes over many 2^14-word rows: one persistent prefetching
    workgroup per CU (default) or one workgroup per row (same results)."""
    _check(lib().fpxg_set_option(OPT_ROW_DMA, 1 if enabled else 0), "set_option")


def set_crt_fused_log(log2_points: int) -> None:
    """Three-prime products of at least 2^log2_points points reconstruct by
    Garner inside their inverse column pass; 0 restores the default (same
    results)."""
    _check(lib().fpxg_set_option(OPT_CRT_FUSED_LOG, int(log2_points)), "set_option")


def set_step_log_max(log2_points: int) -> None:
    """Newton steps of 2^10 .. 2^log2_points points (9 <= log2_points <= 13)
    over a prime with its own transform run as one workgroup's launch, larger
    ones as four-step products; 0 restores the default, 12 (same results)."""
    _check(lib().fpxg_set_option(OPT_STEP_LOG_MAX, int(log2_points)), "set_option")


def pow(p: int, a, k: int, n: int) -> np.ndarray:  # noqa: A001 (the reference's name)
    """poly::pow(a, k, n) (detail/poly/power.hpp:17-87)."""
    a = _u32(a)
    out = np.empty(max(n, 1), dtype=np.uint32)
    _check(lib().fpxg_pow_u32(p, _ptr(a), a.size, int(k), n, _ptr(out)), "poly::pow")
    return out[:n].copy()


def divmod(p: int, a, b):  # noqa: A001
    """poly::divmod(a, b) (detail/poly/division.hpp:12-104) -> (q, r), both trimmed."""
    a, b = _u32(a), _u32(b)
    cap = max(a.size, 1)
    q = np.empty(cap, dtype=np.uint32)
    r = np.empty(cap, dtype=np.uint32)
    lq, lr = ctypes.c_size_t(0), ctypes.c_size_t(0)
    _check(lib().fpxg_divmod_u32(p, _ptr(a), a.size, _ptr(b), b.size, _ptr(q), cap, ctypes.byref(lq), _ptr(r), cap,
                                 ctypes.byref(lr)), "poly::divmod")
    return q[: lq.value].copy(), r[: lr.value].copy()


def div(p: int, a, b) -> np.ndarray:
    """poly::div (division.hpp:86-87)."""
    return divmod(p, a, b)[0]


def mod(p: int, a, b) -> np.ndarray:
    """poly::mod (division.hpp:89-100)."""
    return divmod(p, a, b)[1]


def pow_dev(p: int, a: DeviceArray, la: int, k: int, n: int, out: DeviceArray) -> None:
    _check(lib().fpxg_pow_dev(p, a.ptr, la, int(k), n, out.ptr), "pow_dev")


def divmod_dev(p: int, a: DeviceArray, la: int, b: DeviceArray, lb: int, q: DeviceArray, r: DeviceArray):
    """Device divmod; returns (len q, len r)."""
    lq, lr = ctypes.c_size_t(0), ctypes.c_size_t(0)
    _check(lib().fpxg_divmod_dev(p, a.ptr, la, b.ptr, lb, q.ptr, q.n, ctypes.byref(lq), r.ptr, r.n,
                                 ctypes.byref(lr)), "divmod_dev")
    return lq.value, lr.value
