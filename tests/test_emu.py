"""CPU checks of the GPU kernels' arithmetic, plans and index math.

tests/emu/libgnt_emu.so runs the same per-task code as the HIP kernels
(pi_amd/csrc/gnt_pass.hpp, gl_field.hpp) one task at a time on the CPU.  Its
field operations are checked against exact integer arithmetic, its forward
transform against the closed-form DFT, and its whole products against the
oracle and GMP, so index or twiddle mistakes show up before any GPU run.
"""
import ctypes
import os
import random
import subprocess

import numpy as np
import pytest

P = 0xFFFFFFFF00000001
HERE = os.path.dirname(os.path.abspath(__file__))
U = ctypes.c_uint64
PU = ctypes.POINTER(U)


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "emu")], check=True)
    E = ctypes.CDLL(os.path.join(HERE, "emu", "libgnt_emu.so"))
    for f in ("emu_add", "emu_sub", "emu_mul_mod", "emu_reduce128"):
        getattr(E, f).restype = U
        getattr(E, f).argtypes = [U, U]
    E.emu_mul_pow2.restype = U
    E.emu_mul_pow2.argtypes = [U, ctypes.c_int]
    E.emu_root.restype = U
    E.emu_root.argtypes = [ctypes.c_int]
    E.emu_forward.argtypes = [PU, ctypes.c_int]
    E.emu_inverse.argtypes = [PU, ctypes.c_int]
    E.emu_mul.restype = ctypes.c_size_t
    E.emu_mul.argtypes = [PU, ctypes.c_size_t, PU, ctypes.c_size_t, PU]
    E.emu_plan.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3
    return E


def emu_mul(E, a, b=None):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    nb = a.size if b is None else len(b)
    out = np.zeros(a.size + nb, dtype=np.uint64)
    pb = None if b is None else np.ascontiguousarray(b, dtype=np.uint64)
    n = E.emu_mul(a.ctypes.data_as(PU), a.size, None if pb is None else pb.ctypes.data_as(PU), nb,
                  out.ctypes.data_as(PU))
    return out[:n]


def test_field_ops_exact(emu):
    rnd = random.Random(5)
    vals = [0, 1, 2, P - 1, P - 2, 2**32 - 1, 2**32, P // 2, (P + 1) // 2] + [rnd.randrange(P) for _ in range(200)]
    for a in vals:
        for b in vals[:30]:
            assert emu.emu_add(a, b) == (a + b) % P
            assert emu.emu_sub(a, b) == (a - b) % P
            assert emu.emu_mul_mod(a, b) == a * b % P
    for x in [0, 1, 2**128 - 1, P * P, P * P - 1] + [rnd.randrange(2**128) for _ in range(500)]:
        assert emu.emu_reduce128(x & (2**64 - 1), x >> 64) == x % P


def test_reduce128_borrow_and_carry_edges(emu):
    """The reduction's corner cases: lo equal to or just below the top word hh
    (the borrow of lo - hh), hl = 0 and hl = 2^32 - 1 (the carry of
    t + hl * EPS), sums landing on p and just above it."""
    rnd = random.Random(128)
    words = [0, 1, 2, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFE, 0xFFFFFFFF]
    his = [(h << 32) | l for h in words for l in words]
    cases = []
    for hi in his:
        hh = hi >> 32
        for lo in {0, 1, hh, (hh - 1) % 2**64, hh + 1, P - 1, P, P + 1, 2**64 - 1, 2**64 - 2**32, 2**32 - 1}:
            cases.append((lo % 2**64, hi))
    for _ in range(20000):
        hi = rnd.randrange(2**64)
        k = rnd.randrange(4)
        lo = [hi >> 32, (hi >> 32) - 1, rnd.randrange(2**64), rnd.randrange(2**32)][k] % 2**64
        cases.append((lo, hi))
    for lo, hi in cases:
        assert emu.emu_reduce128(lo, hi) == ((hi << 64) | lo) % P, (hex(lo), hex(hi))


def test_shift_multiplies(emu):
    """x * 2^e for every exponent used by the register DFTs (2^192 = 1, 2^96 = -1)."""
    assert pow(2, 96, P) == P - 1 and pow(2, 192, P) == 1
    rnd = random.Random(6)
    for x in [0, 1, P - 1, 2**63, 2**32 + 5] + [rnd.randrange(P) for _ in range(40)]:
        for e in range(192):
            assert emu.emu_mul_pow2(x, e) == x * pow(2, e, P) % P


def test_shift_by_72_edge_cases(emu):
    """x * 2^72 has its own one-reduction path (x*2^72 = xl*2^40 - (xl + xh)*2^8):
    words at their extremes drive every borrow and wrap in it."""
    rnd = random.Random(72)
    words = [0, 1, 2, 0xFF, 0xFFFFFF, 0x1000000, 0x7FFFFFFF, 0x80000000, 0xFFFFFF00, 0xFFFFFFFE, 0xFFFFFFFF]
    xs = [(h << 32) | l for h in words for l in words]
    xs += [rnd.randrange(P) for _ in range(20000)]
    xs += [rnd.randrange(2**64) for _ in range(2000)]          # non-canonical inputs too
    for x in xs:
        assert emu.emu_mul_pow2(x, 72) == x * pow(2, 72, P) % P, hex(x)
        assert emu.emu_mul_pow2(x, 168) == x * pow(2, 168, P) % P, hex(x)


def test_library_root(emu):
    for log_n in range(1, 31):
        w = emu.emu_root(log_n)
        n = 1 << log_n
        assert pow(w, n, P) == 1 and pow(w, n // 2, P) == P - 1
        if log_n <= 6:
            assert w == pow(2, 192 // n, P)  # small-order roots are powers of two


def test_plans(emu):
    for log_n in range(1, 31):
        npass = ctypes.c_int()
        lr, lc = (ctypes.c_int * 4)(), (ctypes.c_int * 4)()
        emu.emu_plan(log_n, ctypes.byref(npass), lr, lc)
        assert sum(lr[i] for i in range(npass.value)) == log_n
        m = log_n
        for p in range(npass.value):
            s = m - lr[p]
            assert lr[p] + lc[p] <= 13                     # tile fits 64 KiB of LDS
            if s > 0:
                assert lc[p] <= s and lc[p] >= 3           # >= 64-byte contiguous runs
                assert lr[p] <= 10
            m = s


@pytest.mark.parametrize("log_n", list(range(1, 10)))
def test_forward_is_dft_in_bit_reversed_order(emu, log_n):
    n = 1 << log_n
    rnd = random.Random(log_n)
    x = np.array([rnd.randrange(P) for _ in range(n)], dtype=np.uint64)
    y = x.copy()
    emu.emu_forward(y.ctypes.data_as(PU), log_n)
    w = emu.emu_root(log_n)
    xs = [int(v) for v in x]
    for pos in rnd.sample(range(n), min(n, 12)):
        k = int(format(pos, f"0{log_n}b")[::-1], 2)
        assert int(y[pos]) == sum(xs[i] * pow(w, i * k, P) for i in range(n)) % P
    z = y.copy()
    emu.emu_inverse(z.ctypes.data_as(PU), log_n)
    ninv = pow(n, P - 2, P)
    assert [int(v) * ninv % P for v in z] == xs


@pytest.mark.parametrize("log_n", [14, 17, 21])
def test_round_trip_multi_pass(emu, log_n):
    n = 1 << log_n
    x = np.random.default_rng(log_n).integers(0, 2**63, size=n, dtype=np.uint64)
    y = x.copy()
    emu.emu_forward(y.ctypes.data_as(PU), log_n)
    emu.emu_inverse(y.ctypes.data_as(PU), log_n)
    # a plan of two or more passes carries the 1/N in the inverse twiddles of
    # pass 0, so the round trip is the identity
    assert all(int(y[i]) == int(x[i]) for i in range(0, n, 997))


@pytest.mark.parametrize("la,lb", [(1, 1), (1, 4), (8, 5), (200, 200), (1000, 3), (2048, 2048), (5000, 3000),
                                   (1 << 15, 1 << 15), (1 << 18, 1 << 17)])
def test_products_match_oracle(emu, orc, la, lb):
    a, b = orc.splitmix64(la, la), orc.splitmix64(lb + 7, lb)
    assert np.array_equal(emu_mul(emu, a, b), orc.mul_mag_ntt(a, b))
    assert np.array_equal(emu_mul(emu, a), orc.mul_mag_ntt(a, a))


@pytest.mark.parametrize("n", [1, 5, 1000, 20000])
def test_all_ones(emu, orc, n):
    a = np.full(n, 2**64 - 1, dtype=np.uint64)
    want = np.zeros(2 * n, dtype=np.uint64)
    want[0] = 1
    want[n] = np.uint64(2**64 - 2)
    want[n + 1:] = np.uint64(2**64 - 1)
    assert np.array_equal(emu_mul(emu, a, a), want)
    assert np.array_equal(emu_mul(emu, a), want)


def _splits(log_n):
    import itertools
    out = []
    for np_ in (2, 3, 4):
        for c in itertools.product(range(7, 11), repeat=np_):
            if sum(c) == log_n:
                out.append(list(c))
    return out


@pytest.mark.parametrize("log_n", [14, 17, 21])
def test_every_split_gives_the_same_product(emu, orc, log_n):
    """The autotuner may pick any split of a length into radix-2^7..2^10 passes:
    every one must give the exact product (forward and inverse both follow it)."""
    E = emu
    E.emu_set_plan.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    na = (1 << (log_n - 3)) - 3
    a = orc.random_vec(na, log_n)
    b = orc.random_vec(na - 5, log_n + 1)
    want = orc.gmp_mul(a, b)
    try:
        for sp in _splits(log_n):
            arr = (ctypes.c_int * 4)(*(sp + [0] * (4 - len(sp))))
            assert E.emu_set_plan(log_n, len(sp), arr) == 0
            npass, lr, lc = ctypes.c_int(), (ctypes.c_int * 4)(), (ctypes.c_int * 4)()
            E.emu_plan(log_n, ctypes.byref(npass), lr, lc)
            assert list(lr)[:npass.value] == sp
            assert np.array_equal(emu_mul(E, a, b), want), sp
            assert np.array_equal(emu_mul(E, a), orc.gmp_sqr(a)), sp
    finally:
        E.emu_set_plan(log_n, 0, None)
    # an uncovered split is refused
    bad = (ctypes.c_int * 4)(6, log_n - 6, 0, 0)
    assert E.emu_set_plan(log_n, 2, bad) != 0


def test_pass_table_entries_are_the_stepped_twiddles(emu):
    """Entry n*s + j of a pass's table is omega_M^(j * bitrev_R(n)) (inverse:
    its reciprocal), the twiddle the stepping sweep produces for row n,
    column j of a block."""
    E = emu
    E.emu_pass_table_entry.restype = U
    E.emu_pass_table_entry.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    rnd = random.Random(11)
    for log_m, log_r in [(14, 7), (19, 9), (22, 10), (17, 8)]:
        w = emu.emu_root(log_m)
        s = 1 << (log_m - log_r)
        for _ in range(40):
            n, j = rnd.randrange(1 << log_r), rnd.randrange(s)
            kc = int(format(n, f"0{log_r}b")[::-1], 2)
            want = pow(w, j * kc, P)
            assert E.emu_pass_table_entry(n * s + j, log_m, log_r, 0) == want
            assert E.emu_pass_table_entry(n * s + j, log_m, log_r, 1) * want % P == 1


@pytest.mark.parametrize("max_log", [-1, 14, 30])
def test_table_and_stepped_twiddles_give_the_same_products(emu, orc, max_log):
    """Passes with a table multiply their twiddles from it (inverse ones in the
    store of the stage before); passes without one step them.  Every mix of
    the two must give the exact product and the exact transform."""
    E = emu
    E.emu_set_table_max_log.argtypes = [ctypes.c_int]
    try:
        E.emu_set_table_max_log(max_log)
        for la, lb in [(2048, 2048), (5000, 3000), (1 << 15, 1 << 15), (1 << 18, 1 << 17)]:
            a, b = orc.splitmix64(la, la), orc.splitmix64(lb + 7, lb)
            assert np.array_equal(emu_mul(E, a, b), orc.mul_mag_ntt(a, b)), (la, lb)
            assert np.array_equal(emu_mul(E, a), orc.mul_mag_ntt(a, a)), la
        # 2^17: two passes; 2^21: three passes, where with max_log 14 pass 0
        # steps its twiddles and pass 1 takes them from its table
        for log_n in (17, 21) if max_log == 14 else (17,):
            n = 1 << log_n
            x = np.random.default_rng(3).integers(0, 2**63, size=n, dtype=np.uint64)
            y = x.copy()
            E.emu_forward(y.ctypes.data_as(PU), log_n)
            E.emu_inverse(y.ctypes.data_as(PU), log_n)
            assert all(int(y[i]) == int(x[i]) for i in range(0, n, 991)), log_n
    finally:
        E.emu_set_table_max_log(22)


def test_fused_tile_layout(emu):
    """The fused multiply's half-width tiles: their LDS positions are one to
    one within the tile's words (padded 4-column tiles leave 2 words after
    every 8 rows and 32 after every 128), every column-DFT task of a padded
    tile addresses its rows as one base plus compile-time offsets, and two
    A+B tile pairs fit a CU's 160 KiB of LDS."""
    emu.emu_fused_tile_layout_check.restype = ctypes.c_long
    assert emu.emu_fused_tile_layout_check() == 0
