"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle restates the reference's CPU multiply (joelodom/pi
crates/bignum/src/ntt.rs, integer.rs:445-635) in C; see gl_oracle.c.  Only
tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may use
this module, and only as the checker or the timed CPU baseline.

Also wraps two independent references:
* GMP's documented low-level ``mpn_mul``/``mpn_sqr`` (exported from
  libgmp.so.10 as ``__gmpn_mul``/``__gmpn_sqr``), called through its public
  API only;
* CPython ``int`` arithmetic.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

P = 0xFFFFFFFF00000001
U64 = ctypes.c_uint64
SZ = ctypes.c_size_t
PU64 = ctypes.POINTER(ctypes.c_uint64)

_lib = None


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name in ("canonicalize",):
            getattr(L, "orc_" + name).restype = U64
            getattr(L, "orc_" + name).argtypes = [U64]
        for name in ("add", "sub", "mul", "reduce128", "pow"):
            getattr(L, "orc_" + name).restype = U64
            getattr(L, "orc_" + name).argtypes = [U64, U64]
        for name in ("inv", "primitive_root_of_unity"):
            getattr(L, "orc_" + name).restype = U64
            getattr(L, "orc_" + name).argtypes = [U64]
        for name in ("bit_reverse", "ntt_forward", "ntt_inverse", "four_step_forward",
                     "four_step_inverse"):
            getattr(L, "orc_" + name).restype = None
            getattr(L, "orc_" + name).argtypes = [PU64, SZ]
        for name in ("build_twiddles_serial", "build_twiddles_parallel"):
            getattr(L, "orc_" + name).restype = None
            getattr(L, "orc_" + name).argtypes = [U64, SZ, PU64]
        L.orc_twiddle_cache_clear.restype = None
        L.orc_twiddle_cache_entries.restype = SZ
        L.orc_pack.restype = None
        L.orc_pack.argtypes = [PU64, SZ, PU64, SZ]
        for name in ("unpack", "unpack_serial", "unpack_parallel"):
            getattr(L, "orc_" + name).restype = SZ
            getattr(L, "orc_" + name).argtypes = [PU64, SZ, PU64, SZ]
        for name in ("mul_mag_ntt", "mul_schoolbook", "mul_karatsuba", "mul_mag"):
            getattr(L, "orc_" + name).restype = SZ
            getattr(L, "orc_" + name).argtypes = [PU64, SZ, PU64, SZ, PU64]
        L.orc_mul_mag_dispatch.restype = SZ
        L.orc_mul_mag_dispatch.argtypes = [PU64, SZ, PU64, SZ, PU64, SZ, SZ, SZ]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        for name in ("set_four_step_threshold", "set_target_task_size", "set_ntt_threshold",
                     "set_karatsuba_threshold", "set_parallel_karatsuba_threshold"):
            getattr(L, "orc_" + name).restype = None
            getattr(L, "orc_" + name).argtypes = [SZ]
        L.orc_get_four_step_threshold.restype = SZ
        L.orc_apply_host_knobs.argtypes = [ctypes.c_int]
        L.orc_random_vec.argtypes = [SZ, U64, PU64]
        L.orc_make_int.argtypes = [U64, SZ, PU64]
        L.orc_splitmix64_fill.argtypes = [U64, SZ, PU64]
        L.orc_mod_small.restype = U64
        L.orc_mod_small.argtypes = [PU64, SZ, U64]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(PU64)


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


# ---- field ----
def add(a, b): return lib().orc_add(a, b)
def sub(a, b): return lib().orc_sub(a, b)
def mul(a, b): return lib().orc_mul(a, b)
def pow_(a, e): return lib().orc_pow(a, e)
def inv(a): return lib().orc_inv(a)
def canonicalize(x): return lib().orc_canonicalize(x)
def primitive_root_of_unity(n): return lib().orc_primitive_root_of_unity(n)


def reduce128(x: int) -> int:
    return lib().orc_reduce128(x & (2**64 - 1), x >> 64)


# ---- transforms (in place on a copy; returns the result) ----
def _inplace(fn, a):
    x = _u64(a).copy()
    getattr(lib(), fn)(_ptr(x), x.size)
    return x


def bit_reverse(a): return _inplace("orc_bit_reverse", a)
def ntt_forward(a): return _inplace("orc_ntt_forward", a)
def ntt_inverse(a): return _inplace("orc_ntt_inverse", a)
def four_step_forward(a): return _inplace("orc_four_step_forward", a)
def four_step_inverse(a): return _inplace("orc_four_step_inverse", a)


def build_twiddles(omega: int, count: int, parallel: bool) -> np.ndarray:
    out = np.zeros(count, dtype=np.uint64)
    fn = lib().orc_build_twiddles_parallel if parallel else lib().orc_build_twiddles_serial
    fn(omega, count, _ptr(out))
    return out


def pack(limbs, ncoeffs: int) -> np.ndarray:
    l = _u64(limbs)
    out = np.full(ncoeffs, 0xDEAD, dtype=np.uint64)  # stale contents, as a pooled buffer
    lib().orc_pack(_ptr(l), l.size, _ptr(out), ncoeffs)
    return out


def _unpack(fn, coeffs) -> np.ndarray:
    c = _u64(coeffs)
    cap = (c.size + 3) // 4 + 4
    out = np.zeros(cap, dtype=np.uint64)
    n = getattr(lib(), fn)(_ptr(c), c.size, _ptr(out), cap)
    return out[:n].copy()


def unpack(c): return _unpack("orc_unpack", c)
def unpack_serial(c): return _unpack("orc_unpack_serial", c)
def unpack_parallel(c): return _unpack("orc_unpack_parallel", c)


# ---- multiplies ----
def _mul2(fn, a, b, extra=1) -> np.ndarray:
    x, y = _u64(a), _u64(b)
    out = np.zeros(x.size + y.size + extra, dtype=np.uint64)
    n = getattr(lib(), fn)(_ptr(x), x.size, _ptr(y), y.size, _ptr(out))
    return out[:n].copy()


def mul_mag_ntt(a, b): return _mul2("orc_mul_mag_ntt", a, b)
def schoolbook(a, b): return _mul2("orc_mul_schoolbook", a, b)
def karatsuba(a, b): return _mul2("orc_mul_karatsuba", a, b)
def mul_mag(a, b): return _mul2("orc_mul_mag", a, b)


def mul_mag_dispatch(a, b, ntt_max_coeffs: int, ntt_threshold: int, karatsuba_threshold: int):
    x, y = _u64(a), _u64(b)
    out = np.zeros(x.size + y.size + 1, dtype=np.uint64)
    n = lib().orc_mul_mag_dispatch(_ptr(x), x.size, _ptr(y), y.size, _ptr(out),
                                   ntt_max_coeffs, ntt_threshold, karatsuba_threshold)
    return out[:n].copy()


def set_threads(n: int): lib().orc_set_threads(n)
def get_threads() -> int: return lib().orc_get_threads()
def set_four_step_threshold(n: int): lib().orc_set_four_step_threshold(n)
def get_four_step_threshold() -> int: return lib().orc_get_four_step_threshold()
def apply_host_knobs(cores: int): lib().orc_apply_host_knobs(cores)


# ---- generators ----
def random_vec(n: int, seed: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_random_vec(n, seed & (2**64 - 1), _ptr(out))
    return out


def make_int(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_make_int(seed & (2**64 - 1), n, _ptr(out))
    return out


def splitmix64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_splitmix64_fill(seed & (2**64 - 1), n, _ptr(out))
    return out


def mod_small(limbs, q: int) -> int:
    x = _u64(limbs)
    return lib().orc_mod_small(_ptr(x), x.size, q)


# ---- independent references ----
def to_int(limbs) -> int:
    x = _u64(limbs)
    return int.from_bytes(x.astype("<u8").tobytes(), "little")


def from_int(v: int) -> np.ndarray:
    if v == 0:
        return np.zeros(0, dtype=np.uint64)
    nbytes = (v.bit_length() + 63) // 64 * 8
    return np.frombuffer(v.to_bytes(nbytes, "little"), dtype="<u8").astype(np.uint64)


def trim(x: np.ndarray) -> np.ndarray:
    n = x.size
    while n > 0 and x[n - 1] == 0:
        n -= 1
    return x[:n]


_gmp = None


def gmp():
    """libgmp.so.10 through its documented mpn API, or None when absent."""
    global _gmp
    if _gmp is None:
        try:
            g = ctypes.CDLL("libgmp.so.10")
            g.__gmpn_mul.restype = U64
            g.__gmpn_mul.argtypes = [PU64, PU64, ctypes.c_long, PU64, ctypes.c_long]
            g.__gmpn_sqr.restype = None
            g.__gmpn_sqr.argtypes = [PU64, PU64, ctypes.c_long]
            _gmp = g
        except OSError:
            _gmp = False
    return _gmp or None


def gmp_mul(a, b) -> np.ndarray:
    """Product by GMP mpn_mul (requires libgmp); trimmed like the reference."""
    g = gmp()
    x, y = _u64(a), _u64(b)
    if x.size == 0 or y.size == 0:
        return np.zeros(0, dtype=np.uint64)
    if x.size < y.size:
        x, y = y, x
    out = np.zeros(x.size + y.size, dtype=np.uint64)
    g.__gmpn_mul(_ptr(out), _ptr(x), x.size, _ptr(y), y.size)
    return trim(out).copy()


def gmp_sqr(a) -> np.ndarray:
    g = gmp()
    x = _u64(a)
    if x.size == 0:
        return np.zeros(0, dtype=np.uint64)
    out = np.zeros(2 * x.size, dtype=np.uint64)
    g.__gmpn_sqr(_ptr(out), _ptr(x), x.size)
    return trim(out).copy()
