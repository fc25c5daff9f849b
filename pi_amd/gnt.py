"""Python binding of libgnt.so through its C ABI (include/gnt.h).

Mirrors the reference's multiply surface (joelodom/pi crates/bignum):
``mul_mag_ntt(a, b)`` (ntt.rs:806) and ``mul_mag(a, b)`` (integer.rs:445)
take little-endian u64 limb arrays and return the trimmed product.  There is
no CPU fallback: without the built library or a GPU these raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgnt.so")

U64 = ctypes.c_uint64
SZ = ctypes.c_size_t
PU64 = ctypes.POINTER(ctypes.c_uint64)
PSZ = ctypes.POINTER(ctypes.c_size_t)

GNT_OK = 0
GNT_E_ARG = -1
GNT_E_TOO_LARGE = -2
GNT_E_NOMEM = -3
GNT_E_DEVICE = -4
GNT_E_TIMEOUT = -5
GNT_E_STATE = -6

# Every entry point include/gnt.h declares (checked by the ABI test).
EXPORTS = (
    "gnt_version", "gnt_last_error", "gnt_init", "gnt_shutdown", "gnt_mul", "gnt_sqr", "gnt_mul_mag",
    "gnt_mul_mag_dispatch", "gnt_mul_noalloc",
    "gnt_buf_upload", "gnt_buf_download", "gnt_buf_len", "gnt_buf_release", "gnt_mul_buf", "gnt_sqr_buf",
    "gnt_sync", "gnt_mul_batch", "gnt_mul_batch_serial", "gnt_plan_info", "gnt_set_timing", "gnt_get_timing",
    "gnt_batch_create", "gnt_batch_upload", "gnt_batch_run", "gnt_batch_download", "gnt_batch_info",
    "gnt_batch_destroy", "gnt_serve_start", "gnt_serve_stop", "gnt_submit", "gnt_poll", "gnt_wait",
    "gnt_serve_stats", "gnt_set_plan", "gnt_load_plans", "gnt_set_pass_tables", "gnt_tune",
    "gnt_set_staging",
)


class GntError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gnt error {code}: {msg}")
        self.code = code


class GntConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("hbm_cap_bytes", SZ), ("plan_cache_path", ctypes.c_char_p),
                ("serve_split_log_n", SZ), ("serve_large_inflight", SZ), ("serve_device_limbs", SZ),
                ("reserved", SZ * 1)]

SERVE_DEVICE_OFF = 2**64 - 1  # gnt_config.serve_device_limbs: no device-resident loop


class GntReq(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int), ("a", PU64), ("na", SZ), ("b", PU64), ("nb", SZ),
                ("a_buf", ctypes.c_void_p), ("b_buf", ctypes.c_void_p)]


class GntJob(ctypes.Structure):
    _fields_ = [("a", PU64), ("na", SZ), ("b", PU64), ("nb", SZ), ("out", PU64), ("out_cap", SZ),
                ("out_len", SZ), ("is_sqr", ctypes.c_int), ("status", ctypes.c_int)]


_lib = None


def load() -> ctypes.CDLL:
    """Load libgnt.so (raises if it was never built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GntError(GNT_E_STATE, f"{LIB_PATH} is not built; run python -m pi_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    L.gnt_version.restype = ctypes.c_char_p
    L.gnt_last_error.restype = ctypes.c_char_p
    L.gnt_init.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(GntConfig)]
    L.gnt_shutdown.argtypes = [ctypes.c_void_p]
    L.gnt_mul.argtypes = [ctypes.c_void_p, PU64, SZ, PU64, SZ, PU64, SZ, PSZ]
    L.gnt_sqr.argtypes = [ctypes.c_void_p, PU64, SZ, PU64, SZ, PSZ]
    L.gnt_mul_mag.argtypes = [ctypes.c_void_p, PU64, SZ, PU64, SZ, PU64, SZ, PSZ, SZ, SZ]
    L.gnt_mul_mag_dispatch.argtypes = [ctypes.c_void_p, PU64, SZ, PU64, SZ, PU64, SZ, PSZ, SZ, SZ, SZ]
    L.gnt_mul_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, SZ, ctypes.c_void_p, SZ, ctypes.c_void_p, SZ,
                                  ctypes.c_void_p, ctypes.c_void_p]
    L.gnt_buf_upload.argtypes = [ctypes.c_void_p, PU64, SZ, ctypes.POINTER(ctypes.c_void_p)]
    L.gnt_buf_download.argtypes = [ctypes.c_void_p, ctypes.c_void_p, PU64, SZ, PSZ]
    L.gnt_buf_len.argtypes = [ctypes.c_void_p, ctypes.c_void_p, PSZ]
    L.gnt_buf_release.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.gnt_mul_buf.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    L.gnt_sqr_buf.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    L.gnt_sync.argtypes = [ctypes.c_void_p]
    L.gnt_mul_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(GntJob), SZ]
    L.gnt_mul_batch_serial.argtypes = [ctypes.c_void_p, ctypes.POINTER(GntJob), SZ]
    L.gnt_plan_info.argtypes = [SZ, SZ, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_int)]
    L.gnt_batch_create.argtypes = [ctypes.c_void_p, PSZ, PSZ, ctypes.POINTER(ctypes.c_int), SZ, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_void_p)]
    L.gnt_batch_upload.argtypes = [ctypes.c_void_p, ctypes.POINTER(GntJob)]
    L.gnt_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.gnt_batch_download.argtypes = [ctypes.c_void_p, ctypes.POINTER(GntJob)]
    L.gnt_batch_info.argtypes = [ctypes.c_void_p, PSZ, PSZ, PSZ]
    L.gnt_batch_destroy.argtypes = [ctypes.c_void_p]
    L.gnt_serve_start.argtypes = [ctypes.c_void_p]
    L.gnt_serve_stop.argtypes = [ctypes.c_void_p]
    L.gnt_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(GntReq), ctypes.POINTER(ctypes.c_uint64)]
    L.gnt_poll.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]
    L.gnt_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p),
                           ctypes.POINTER(ctypes.c_double)]
    L.gnt_serve_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]
    L.gnt_set_plan.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.gnt_set_pass_tables.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.gnt_load_plans.argtypes = [ctypes.c_char_p]
    L.gnt_set_staging.argtypes = [ctypes.c_int, SZ, SZ]
    L.gnt_tune.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p]
    L.gnt_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.gnt_get_timing.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_uint64)]
    _lib = L
    return L


def _check(rc: int):
    if rc != GNT_OK:
        raise GntError(rc, load().gnt_last_error().decode(errors="replace"))


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def _p(a: np.ndarray):
    return a.ctypes.data_as(PU64)


def plan_info(na: int, nb: int):
    L = load()
    log_n, npass = ctypes.c_int(), ctypes.c_int()
    lr = (ctypes.c_int * 4)()
    _check(L.gnt_plan_info(na, nb, ctypes.byref(log_n), ctypes.byref(npass), lr))
    return log_n.value, [lr[i] for i in range(npass.value)]


class DeviceBuf:
    """A limb array resident in HBM (gnt_buf)."""

    def __init__(self, ctx: "Context", handle: ctypes.c_void_p):
        self.ctx = ctx
        self.handle = handle

    def __len__(self) -> int:
        n = SZ()
        _check(load().gnt_buf_len(self.ctx.h, self.handle, ctypes.byref(n)))
        return n.value

    def download(self) -> np.ndarray:
        n = len(self)
        out = np.zeros(max(n, 1), dtype=np.uint64)
        got = SZ()
        _check(load().gnt_buf_download(self.ctx.h, self.handle, _p(out), out.size, ctypes.byref(got)))
        return out[:got.value].copy()

    def release(self):
        if self.handle:
            _check(load().gnt_buf_release(self.ctx.h, self.handle))
            self.handle = None


class Context:
    """One device context (gnt_ctx)."""

    def __init__(self, device: int = 0, hbm_cap_bytes: int = 0, plan_cache: Optional[str] = None,
                 serve_split_log_n: int = 0, serve_large_inflight: int = 0, serve_device_limbs: int = 0):
        L = load()
        cfg = GntConfig(device=device, hbm_cap_bytes=hbm_cap_bytes,
                        plan_cache_path=plan_cache.encode() if plan_cache else None,
                        serve_split_log_n=serve_split_log_n, serve_large_inflight=serve_large_inflight,
                        serve_device_limbs=serve_device_limbs)
        h = ctypes.c_void_p()
        _check(L.gnt_init(ctypes.byref(h), ctypes.byref(cfg)))
        self.h = h

    def close(self):
        if self.h:
            load().gnt_shutdown(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- the reference's surface ----
    def mul_mag_ntt(self, a, b) -> np.ndarray:
        """ntt::mul_mag_ntt (ntt.rs:806): trimmed |a|*|b|."""
        x, y = _u64(a), _u64(b)
        out = np.zeros(max(x.size + y.size, 1), dtype=np.uint64)
        n = SZ()
        _check(load().gnt_mul(self.h, _p(x), x.size, _p(y), y.size, _p(out), out.size, ctypes.byref(n)))
        return out[:n.value].copy()

    def sqr(self, a) -> np.ndarray:
        x = _u64(a)
        out = np.zeros(max(2 * x.size, 1), dtype=np.uint64)
        n = SZ()
        _check(load().gnt_sqr(self.h, _p(x), x.size, _p(out), out.size, ctypes.byref(n)))
        return out[:n.value].copy()

    def mul_mag(self, a, b, ntt_threshold: int = 0, karatsuba_threshold: int = 0,
                ntt_max_coeffs: int = 0) -> np.ndarray:
        """mul_mag / mul_mag_dispatch (integer.rs:445-486) with the GPU leaf;
        ntt_max_coeffs (0: the library cap 2^32) is the dispatch's transform
        cap, which the reference's tests lower to force Karatsuba above it."""
        x, y = _u64(a), _u64(b)
        out = np.zeros(x.size + y.size + 1, dtype=np.uint64)
        n = SZ()
        _check(load().gnt_mul_mag_dispatch(self.h, _p(x), x.size, _p(y), y.size, _p(out), out.size, ctypes.byref(n),
                                           ntt_max_coeffs, ntt_threshold, karatsuba_threshold))
        return out[:n.value].copy()

    # ---- device-resident ----
    def upload(self, a) -> DeviceBuf:
        x = _u64(a)
        h = ctypes.c_void_p()
        _check(load().gnt_buf_upload(self.h, _p(x), x.size, ctypes.byref(h)))
        return DeviceBuf(self, h)

    def mul_buf(self, a: DeviceBuf, b: DeviceBuf) -> DeviceBuf:
        h = ctypes.c_void_p()
        _check(load().gnt_mul_buf(self.h, a.handle, b.handle, ctypes.byref(h)))
        return DeviceBuf(self, h)

    def sqr_buf(self, a: DeviceBuf) -> DeviceBuf:
        h = ctypes.c_void_p()
        _check(load().gnt_sqr_buf(self.h, a.handle, ctypes.byref(h)))
        return DeviceBuf(self, h)

    def sync(self):
        _check(load().gnt_sync(self.h))

    # ---- batches ----
    def _jobs(self, pairs: Sequence):
        keep, outs = [], []
        jobs = (GntJob * len(pairs))()
        for i, pr in enumerate(pairs):
            a = _u64(pr[0])
            sq = pr[1] is None
            b = a if sq else _u64(pr[1])
            out = np.zeros(max(a.size + (a.size if sq else b.size), 1), dtype=np.uint64)
            keep += [a, b]
            outs.append(out)
            jobs[i] = GntJob(_p(a), a.size, _p(b), b.size, _p(out), out.size, 0, int(sq), 0)
        return jobs, keep, outs

    def mul_batch(self, pairs: Sequence, serial: bool = False):
        """Products of (a, b) pairs; b None squares.  Returns trimmed results."""
        jobs, keep, outs = self._jobs(pairs)
        fn = load().gnt_mul_batch_serial if serial else load().gnt_mul_batch
        rc = fn(self.h, jobs, len(pairs))
        for j in jobs:
            if j.status != GNT_OK:
                raise GntError(j.status, load().gnt_last_error().decode(errors="replace"))
        _check(rc)
        return [o[:j.out_len].copy() for o, j in zip(outs, jobs)]

    def batch_plan(self, pairs: Sequence, graph: bool = True) -> "BatchPlan":
        return BatchPlan(self, pairs, graph)

    # ---- timing ----
    def set_timing(self, on: bool):
        _check(load().gnt_set_timing(self.h, int(on)))

    def get_timing(self):
        ms = (ctypes.c_double * 5)()
        n = ctypes.c_uint64()
        _check(load().gnt_get_timing(self.h, ms, 5, ctypes.byref(n)))
        return list(ms), n.value


class BatchPlan:
    """gnt_batch: fixed product shapes, device-resident buffers, optional hipGraph."""

    def __init__(self, ctx: Context, pairs: Sequence, graph: bool = True):
        self.ctx = ctx
        self.jobs, self._keep, self.outs = ctx._jobs(pairs)
        n = len(pairs)
        na = (SZ * n)(*[j.na for j in self.jobs])
        nb = (SZ * n)(*[j.nb for j in self.jobs])
        sq = (ctypes.c_int * n)(*[j.is_sqr for j in self.jobs])
        h = ctypes.c_void_p()
        _check(load().gnt_batch_create(ctx.h, na, nb, sq, n, int(graph), ctypes.byref(h)))
        self.h = h
        _check(load().gnt_batch_upload(self.h, self.jobs))

    def run(self, serial: bool = False):
        _check(load().gnt_batch_run(self.h, 1 if serial else 0))

    def results(self):
        _check(load().gnt_batch_download(self.h, self.jobs))
        return [o[:j.out_len].copy() for o, j in zip(self.outs, self.jobs)]

    def info(self):
        a, b, c = SZ(), SZ(), SZ()
        _check(load().gnt_batch_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"launches_per_run": a.value, "iterations": b.value, "device_bytes": c.value}

    def close(self):
        if self.h:
            load().gnt_batch_destroy(self.h)
            self.h = None


class Server:
    """The serving loop (continuous batching) of a context."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self._keep = {}
        _check(load().gnt_serve_start(ctx.h))

    OPS = {"mul": 0, "sqr": 1, "add": 2, "sub": 3}

    def submit(self, a=None, b=None, square: bool = False, a_buf: Optional[DeviceBuf] = None,
               b_buf: Optional[DeviceBuf] = None, op: Optional[str] = None) -> int:
        """Queue a*b (or a*a with square=True); op="add"/"sub" queues a+b / a-b (a >= b)."""
        req = GntReq()
        if op is None:
            op = "sqr" if square else "mul"
        req.op = self.OPS[op]
        square = op == "sqr"
        keep = []
        if a_buf is not None:
            req.a_buf = a_buf.handle
        else:
            x = _u64(a)
            keep.append(x)
            req.a, req.na = _p(x), x.size
        if not square:
            if b_buf is not None:
                req.b_buf = b_buf.handle
            else:
                y = _u64(b)
                keep.append(y)
                req.b, req.nb = _p(y), y.size
        t = ctypes.c_uint64()
        _check(load().gnt_submit(self.ctx.h, ctypes.byref(req), ctypes.byref(t)))
        self._keep[t.value] = keep  # host operands live until the request completes
        return t.value

    def done(self, ticket: int) -> bool:
        d = ctypes.c_int()
        _check(load().gnt_poll(self.ctx.h, ticket, ctypes.byref(d)))
        return bool(d.value)

    def wait(self, ticket: int, timeout_us: int = -1):
        h = ctypes.c_void_p()
        lat = ctypes.c_double()
        _check(load().gnt_wait(self.ctx.h, ticket, timeout_us, ctypes.byref(h), ctypes.byref(lat)))
        self._keep.pop(ticket, None)
        return DeviceBuf(self.ctx, h), lat.value

    def stats(self):
        v = (ctypes.c_uint64 * 10)()
        _check(load().gnt_serve_stats(self.ctx.h, v, 10))
        return dict(zip(["iterations", "launches", "completed", "peak_admitted", "peak_bytes", "bytes_in_flight",
                         "small_lane_iterations", "large_lane_iterations", "device_loop_requests",
                         "device_loop_launches"], list(v)))

    def stop(self):
        _check(load().gnt_serve_stop(self.ctx.h))


def set_pass_tables(max_log_m: int) -> int:
    """Largest pass block size (log2) whose twiddles come from a table; -1
    disables tables.  Returns the previous setting (process-wide)."""
    prev = ctypes.c_int()
    _check(load().gnt_set_pass_tables(max_log_m, ctypes.byref(prev)))
    return prev.value


def set_staging(threads: int = 4, chunk_bytes: int = 8 << 20, min_bytes: int = 4 << 20):
    """Host <-> device copies of at least min_bytes go through pinned staging
    chunks copied by `threads` host threads; threads=0 copies directly.
    Process-wide; the arguments' defaults are the library's."""
    _check(load().gnt_set_staging(threads, chunk_bytes, min_bytes))


def set_plan(log_n: int, log_r=None):
    """Install a pass split for transforms of length 2^log_n (None: built-in rule)."""
    arr = (ctypes.c_int * 4)(*(list(log_r or []) + [0] * (4 - len(log_r or []))))
    _check(load().gnt_set_plan(log_n, len(log_r or []), arr))


def load_plans(path: str):
    _check(load().gnt_load_plans(path.encode()))


def tune(ctx: "Context", min_log_n: int, max_log_n: int, reps: int = 3, path: Optional[str] = None):
    """Time every split of each length on the device, install the fastest, write the cache."""
    _check(load().gnt_tune(ctx.h, min_log_n, max_log_n, reps, path.encode() if path else None))


def plan(log_n: int):
    """(npass, radices) the library uses for a length-2^log_n transform."""
    na = 1 << (log_n - 3) if log_n >= 3 else 1
    ln, np_, lr = ctypes.c_int(), ctypes.c_int(), (ctypes.c_int * 4)()
    _check(load().gnt_plan_info(na, na, ctypes.byref(ln), ctypes.byref(np_), lr))
    return np_.value, list(lr)[:np_.value]


_default: Optional[Context] = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default


def mul_mag_ntt(a, b) -> np.ndarray:
    return default_context().mul_mag_ntt(a, b)


def mul_mag(a, b) -> np.ndarray:
    return default_context().mul_mag(a, b)


def sqr(a) -> np.ndarray:
    return default_context().sqr(a)
