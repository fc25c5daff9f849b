"""gnt_mul_noalloc through the C ABI against gnt_mul_buf (the entry point whose
kernels it shares) and GMP, bit for bit: two calls in a row, the second with
a smaller input, for products and squares, on the context's stream and on a
caller's stream; operands, outputs and length words are the caller's device
memory (hipMalloc through the HIP runtime, before the calls).  Sizes above the
scratch limit are refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    from pi_amd import gnt
    gnt.load()  # the runtime libgnt.so already loaded
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


class Dev:
    """A caller-owned device array (hipMalloc), filled from / read into numpy."""

    def __init__(self, hip, nbytes, fill=None):
        self.hip, self.n = hip, nbytes
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(nbytes, 8)) == 0
        if fill is not None:
            src = np.ascontiguousarray(fill)
            assert src.nbytes == nbytes
            assert hip.hipMemcpy(self.p, src.ctypes.data, nbytes, H2D) == 0

    def read(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(ctx, hip, a, b, stream):
    """One gnt_mul_noalloc call (b None squares); returns the trimmed product."""
    from pi_amd import gnt
    L = gnt.load()
    na = a.size
    nb = na if b is None else b.size
    da = Dev(hip, a.nbytes, a)
    db = None if b is None else Dev(hip, b.nbytes, b)
    garbage = np.full(na + nb, 0xA5A5A5A5A5A5A5A5, dtype=np.uint64)   # stale output contents
    do = Dev(hip, garbage.nbytes, garbage)
    dl = Dev(hip, 4, np.array([0xFFFFFFFF], dtype=np.uint32))
    try:
        rc = L.gnt_mul_noalloc(ctx.h, da.p, na, None if db is None else db.p, nb, do.p, na + nb, dl.p, stream)
        assert rc == gnt.GNT_OK, L.gnt_last_error()
        if stream:
            assert hip.hipStreamSynchronize(stream) == 0
        else:
            ctx.sync()
        n = int(dl.read(np.uint32, 1)[0])
        assert n <= na + nb
        return do.read(np.uint64, na + nb)[:n]
    finally:
        for d in (da, db, do, dl):
            if d is not None:
                d.free()


def _mul_buf(ctx, a, b):
    da = ctx.upload(a)
    db = da if b is None else ctx.upload(b)
    r = ctx.sqr_buf(da) if b is None else ctx.mul_buf(da, db)
    out = r.download()
    for h in {id(x): x for x in (da, db, r)}.values():
        h.release()
    return out


@pytest.mark.parametrize("use_stream", [False, True])
def test_noalloc_matches_mul_buf(ctx, orc, use_stream):
    hip = _hip()
    stream = ctypes.c_void_p()
    if use_stream:
        assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        # two calls in a row, the second smaller; then squares the same way;
        # then all-ones (maximal carries) at the size limit 4(na+nb) = 2^22
        cases = [((1 << 19) - 3, (1 << 19) - 700), (5000, 3001), ((1 << 19) - 1, None), (777, None),
                 ((1 << 19) + 100, (1 << 19) - 100), (1, 1)]
        for i, (na, nb) in enumerate(cases):
            print(f"noalloc case {i}: {na} x {nb} stream={use_stream}", flush=True)
            if i == 4:
                a = np.full(na, 2**64 - 1, dtype=np.uint64)
                b = np.full(nb, 2**64 - 1, dtype=np.uint64)
            else:
                a = orc.splitmix64(40 + i, na)
                b = None if nb is None else orc.splitmix64(60 + i, nb)
            got = _noalloc(ctx, hip, a, b, stream.value)
            assert np.array_equal(got, _mul_buf(ctx, a, b)), i
            want = orc.gmp_sqr(a) if b is None else orc.gmp_mul(a, b)
            assert np.array_equal(got, want), i
    finally:
        if use_stream:
            hip.hipStreamDestroy(stream)


def test_noalloc_refuses_beyond_its_scratch(ctx):
    from pi_amd import gnt
    L = gnt.load()
    hip = _hip()
    # 4(na+nb) = 2^22 + 4: one limb past the limit; nothing is launched
    d = Dev(hip, 8)
    try:
        rc = L.gnt_mul_noalloc(ctx.h, d.p, (1 << 19) + 1, d.p, 1 << 19, d.p, 1 << 21, d.p, None)
        assert rc == gnt.GNT_E_TOO_LARGE
        assert b"2^22" in L.gnt_last_error()
        # empty operand: length 0, no product
        dl = Dev(hip, 4, np.array([7], dtype=np.uint32))
        assert L.gnt_mul_noalloc(ctx.h, d.p, 0, d.p, 5, d.p, 5, dl.p, None) == gnt.GNT_OK
        ctx.sync()
        assert int(dl.read(np.uint32, 1)[0]) == 0
        dl.free()
    finally:
        d.free()
