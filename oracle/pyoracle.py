"""ctypes front end of the CPU oracle (oracle/liboracle.so) and, where it was
built, of the reference shim (oracle/_ref/libfpref_*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the fastpoly_amd package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

OPS = {"inv": 0, "log": 1, "exp": 2, "sqrt": 3}
E_OK, E_DOMAIN, E_SIZE, E_NOMEM, E_ARG = 0, 1, 2, 3, 4

_u32p = ctypes.POINTER(ctypes.c_uint32)


class Field(ctypes.Structure):
    _fields_ = [
        ("p", ctypes.c_uint32),
        ("ninv", ctypes.c_uint32),
        ("one", ctypes.c_uint32),
        ("r2", ctypes.c_uint32),
        ("root", ctypes.c_uint32),
        ("maxlog", ctypes.c_int),
    ]


class MT64(ctypes.Structure):
    _fields_ = [("mt", ctypes.c_uint64 * 312), ("idx", ctypes.c_int)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        FP = ctypes.POINTER(Field)
        L.fpo_field_init.argtypes = [FP, ctypes.c_uint32, ctypes.c_uint32]
        L.fpo_alias_root.argtypes = [ctypes.c_uint32]
        L.fpo_alias_root.restype = ctypes.c_uint32
        for name in ("fpo_reduce", "fpo_from_int"):
            getattr(L, name).argtypes = [FP, ctypes.c_uint64]
            getattr(L, name).restype = ctypes.c_uint32
        for name in ("fpo_mul", "fpo_add", "fpo_sub"):
            getattr(L, name).argtypes = [FP, ctypes.c_uint32, ctypes.c_uint32]
            getattr(L, name).restype = ctypes.c_uint32
        L.fpo_val.argtypes = [FP, ctypes.c_uint32]
        L.fpo_val.restype = ctypes.c_uint32
        L.fpo_pow.argtypes = [FP, ctypes.c_uint32, ctypes.c_uint64]
        L.fpo_pow.restype = ctypes.c_uint32
        L.fpo_minv.argtypes = [FP, ctypes.c_uint32]
        L.fpo_minv.restype = ctypes.c_uint32
        L.fpo_sqrt.argtypes = [FP, ctypes.c_uint32, _u32p]
        L.fpo_primitive_root.argtypes = [ctypes.c_uint32]
        L.fpo_primitive_root.restype = ctypes.c_uint32
        MP = ctypes.POINTER(MT64)
        L.fpo_mt64_seed.argtypes = [MP, ctypes.c_uint64]
        L.fpo_mt64_next.argtypes = [MP]
        L.fpo_mt64_next.restype = ctypes.c_uint64
        L.fpo_rnd.argtypes = [FP, MP, _u32p, ctypes.c_size_t]
        L.fpo_checksum.argtypes = [_u32p, ctypes.c_size_t]
        L.fpo_checksum.restype = ctypes.c_uint64
        L.fpo_ntt_forward.argtypes = [FP, _u32p, ctypes.c_uint32]
        L.fpo_ntt_inverse.argtypes = [FP, _u32p, ctypes.c_uint32]
        L.fpo_conv.argtypes = [FP, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t,
                               ctypes.c_size_t, _u32p, ctypes.POINTER(ctypes.c_size_t)]
        L.fpo_conv_crt.argtypes = [FP, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t,
                                   ctypes.c_size_t, ctypes.c_size_t, _u32p]
        L.fpo_conv_naive.argtypes = [FP, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t,
                                     ctypes.c_size_t, _u32p]
        L.fpo_crt_prime_count.argtypes = [ctypes.c_uint32, ctypes.c_size_t]
        for name in ("fpo_inv", "fpo_log", "fpo_exp", "fpo_sqrt_series"):
            getattr(L, name).argtypes = [FP, _u32p, ctypes.c_size_t, ctypes.c_size_t, _u32p]
        L.fpo_poly_pow.argtypes = [FP, _u32p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_size_t, _u32p]
        _szp = ctypes.POINTER(ctypes.c_size_t)
        L.fpo_divmod.argtypes = [FP, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t, _u32p, _szp, _u32p, _szp]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: oracle error {code}")
        self.code = code


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def field(p: int, root: int | None = None) -> Field:
    f = Field()
    if root is None:
        root = lib().fpo_alias_root(p)
    rc = lib().fpo_field_init(ctypes.byref(f), p, root)
    if rc:
        raise OracleError(rc, "field")
    return f


class MT:
    """std::mt19937_64."""

    def __init__(self, seed: int):
        self.g = MT64()
        lib().fpo_mt64_seed(ctypes.byref(self.g), seed)

    def next(self) -> int:
        return lib().fpo_mt64_next(ctypes.byref(self.g))

    def rnd(self, p: int, n: int) -> np.ndarray:
        """bench.cpp rnd(): Mont::from_int(rng() % p) limbs."""
        out = np.empty(n, dtype=np.uint32)
        f = field(p)
        lib().fpo_rnd(ctypes.byref(f), ctypes.byref(self.g), _ptr(out), n)
        return out


def checksum(a) -> int:
    a = _u32(a)
    return int(lib().fpo_checksum(_ptr(a), a.size))


def from_int(p: int, x: int) -> int:
    f = field(p)
    return int(lib().fpo_from_int(ctypes.byref(f), x))


def to_int(p: int, limb: int) -> int:
    f = field(p)
    return int(lib().fpo_val(ctypes.byref(f), limb))


def mont_mul(p: int, a: int, b: int) -> int:
    f = field(p)
    return int(lib().fpo_mul(ctypes.byref(f), a, b))


def mont_pow(p: int, a: int, e: int) -> int:
    f = field(p)
    return int(lib().fpo_pow(ctypes.byref(f), a, e))


def mont_sqrt(p: int, a: int):
    f = field(p)
    out = ctypes.c_uint32(0)
    ok = lib().fpo_sqrt(ctypes.byref(f), a, ctypes.byref(out))
    return int(out.value) if ok else None


def ntt(p: int, a, inverse: bool = False, root: int | None = None) -> np.ndarray:
    f = field(p, root)
    x = _u32(a).copy()
    fn = lib().fpo_ntt_inverse if inverse else lib().fpo_ntt_forward
    rc = fn(ctypes.byref(f), _ptr(x), x.size)
    if rc:
        raise OracleError(rc, "ntt")
    return x


def conv(p: int, a, b, n: int | None = None) -> np.ndarray:
    f = field(p)
    a, b = _u32(a), _u32(b)
    lim = (1 << 64) - 1 if n is None else n
    cap = max(a.size + b.size, 1)
    out = np.zeros(cap, dtype=np.uint32)
    got = ctypes.c_size_t(0)
    rc = lib().fpo_conv(ctypes.byref(f), _ptr(a), a.size, _ptr(b), b.size, lim, _ptr(out),
                        ctypes.byref(got))
    if rc:
        raise OracleError(rc, "conv")
    return out[: got.value].copy()


def conv_crt(p: int, a, b, out_count: int, block: int = 1 << 23) -> np.ndarray:
    f = field(p)
    a, b = _u32(a), _u32(b)
    out = np.zeros(max(out_count, 1), dtype=np.uint32)
    rc = lib().fpo_conv_crt(ctypes.byref(f), _ptr(a), a.size, _ptr(b), b.size, out_count,
                            block, _ptr(out))
    if rc:
        raise OracleError(rc, "conv_crt")
    return out[:out_count].copy()


def conv_naive(p: int, a, b, out_count: int | None = None) -> np.ndarray:
    f = field(p)
    a, b = _u32(a), _u32(b)
    if out_count is None:
        out_count = a.size + b.size - 1 if a.size and b.size else 0
    out = np.zeros(max(out_count, 1), dtype=np.uint32)
    lib().fpo_conv_naive(ctypes.byref(f), _ptr(a), a.size, _ptr(b), b.size, out_count, _ptr(out))
    return out[:out_count].copy()


def crt_prime_count(p: int, terms: int) -> int:
    return int(lib().fpo_crt_prime_count(p, terms))


def series(op: str, p: int, a, n: int) -> np.ndarray:
    f = field(p)
    a = _u32(a)
    out = np.zeros(max(n, 1), dtype=np.uint32)
    fn = {"inv": lib().fpo_inv, "log": lib().fpo_log, "exp": lib().fpo_exp,
          "sqrt": lib().fpo_sqrt_series}[op]
    rc = fn(ctypes.byref(f), _ptr(a), a.size, n, _ptr(out))
    if rc:
        raise OracleError(rc, op)
    return out[:n].copy()


def poly_pow(p: int, a, k: int, n: int) -> np.ndarray:
    """poly::pow(a, k, n) (power.hpp:17-87)."""
    f = field(p)
    a = _u32(a)
    out = np.zeros(max(n, 1), dtype=np.uint32)
    rc = lib().fpo_poly_pow(ctypes.byref(f), _ptr(a), a.size, k, n, _ptr(out))
    if rc:
        raise OracleError(rc, "pow")
    return out[:n].copy()


def divmod(p: int, a, b):
    """poly::divmod(a, b) (division.hpp:12-104) -> (q, r), both trimmed."""
    f = field(p)
    a, b = _u32(a), _u32(b)
    q = np.zeros(max(a.size, 1), dtype=np.uint32)
    r = np.zeros(max(a.size, 1), dtype=np.uint32)
    lq, lr = ctypes.c_size_t(0), ctypes.c_size_t(0)
    rc = lib().fpo_divmod(ctypes.byref(f), _ptr(a), a.size, _ptr(b), b.size, _ptr(q), ctypes.byref(lq),
                          _ptr(r), ctypes.byref(lr))
    if rc:
        raise OracleError(rc, "divmod")
    return q[: lq.value].copy(), r[: lr.value].copy()


# --------------------------------------------------------------------------
# Reference shim (oracle/_ref), built here from /root/reference/include.
# --------------------------------------------------------------------------

_ref = None


def _cpu_flags() -> set:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def ref_path() -> str | None:
    flags = _cpu_flags()
    names = []
    if {"avx512f", "avx512vl"} <= flags:
        names.append("libfpref_avx512.so")
    if "avx2" in flags:
        names.append("libfpref_avx2.so")
    for nm in names:
        path = os.path.join(REF_DIR, nm)
        if os.path.exists(path):
            return path
    return None


def ref():
    """The compiled reference shim, or None where it was not built."""
    global _ref
    if _ref is None:
        path = ref_path()
        if path is None:
            return None
        L = ctypes.CDLL(path)
        L.fpref_backend.restype = ctypes.c_char_p
        L.fpref_conv.argtypes = [ctypes.c_uint32, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t,
                                 ctypes.c_size_t, _u32p, ctypes.POINTER(ctypes.c_size_t)]
        L.fpref_series.argtypes = [ctypes.c_int, ctypes.c_uint32, _u32p, ctypes.c_size_t,
                                   ctypes.c_size_t, _u32p]
        L.fpref_ntt.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32, _u32p]
        L.fpref_msqrt.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p]
        if hasattr(L, "fpref_pow"):
            _szp = ctypes.POINTER(ctypes.c_size_t)
            L.fpref_pow.argtypes = [ctypes.c_uint32, _u32p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_size_t, _u32p]
            L.fpref_divmod.argtypes = [ctypes.c_uint32, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t, _u32p, _szp,
                                       _u32p, _szp]
        _ref = L
    return _ref


def ref_conv(p: int, a, b, n: int | None = None) -> np.ndarray:
    L = ref()
    a, b = _u32(a), _u32(b)
    lim = (1 << 64) - 1 if n is None else n
    out = np.zeros(max(a.size + b.size, 1), dtype=np.uint32)
    got = ctypes.c_size_t(0)
    rc = L.fpref_conv(p, _ptr(a), a.size, _ptr(b), b.size, lim, _ptr(out), ctypes.byref(got))
    if rc:
        raise OracleError(rc, "ref_conv")
    return out[: got.value].copy()


def ref_pow(p: int, a, k: int, n: int) -> np.ndarray:
    L = ref()
    a = _u32(a)
    out = np.zeros(max(n, 1), dtype=np.uint32)
    rc = L.fpref_pow(p, _ptr(a), a.size, k, n, _ptr(out))
    if rc:
        raise OracleError(rc, "ref_pow")
    return out[:n].copy()


def ref_divmod(p: int, a, b):
    L = ref()
    a, b = _u32(a), _u32(b)
    q = np.zeros(max(a.size, 1), dtype=np.uint32)
    r = np.zeros(max(a.size, 1), dtype=np.uint32)
    lq, lr = ctypes.c_size_t(0), ctypes.c_size_t(0)
    rc = L.fpref_divmod(p, _ptr(a), a.size, _ptr(b), b.size, _ptr(q), ctypes.byref(lq), _ptr(r), ctypes.byref(lr))
    if rc:
        raise OracleError(rc, "ref_divmod")
    return q[: lq.value].copy(), r[: lr.value].copy()


def ref_series(op: str, p: int, a, n: int) -> np.ndarray:
    L = ref()
    a = _u32(a)
    out = np.zeros(max(n, 1), dtype=np.uint32)
    rc = L.fpref_series(OPS[op], p, _ptr(a), a.size, n, _ptr(out))
    if rc:
        raise OracleError(rc, "ref_" + op)
    return out[:n].copy()


def ref_ntt(p: int, a, inverse: bool = False) -> np.ndarray:
    L = ref()
    x = _u32(a).copy()
    rc = L.fpref_ntt(p, 1 if inverse else 0, x.size, _ptr(x))
    if rc:
        raise OracleError(rc, "ref_ntt")
    return x


def ref_msqrt(p: int, limb: int):
    out = ctypes.c_uint32(0)
    rc = ref().fpref_msqrt(p, limb, ctypes.byref(out))
    return None if rc else int(out.value)
