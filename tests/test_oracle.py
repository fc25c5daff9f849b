"""Pins of the CPU oracle (oracle/gl_oracle.c), which the GPU tests use as checker.

Known-answer tests are the values in the reference's own tests
(crates/bignum/src/ntt.rs:1114-2101, integer.rs:1391-1410); the rest are
closed forms, brute force (naive DFT, naive convolution, schoolbook) and the
independent references GMP mpn_mul and CPython int, chosen so that a dropped
term, a wrong sign, a transposed operand or a wrong twiddle fails one of them.
"""
import hashlib
import json
import os

import numpy as np
import pytest

P = 0xFFFFFFFF00000001
EPS = 0xFFFFFFFF
HERE = os.path.dirname(os.path.abspath(__file__))


def golden():
    with open(os.path.join(HERE, "golden", "products.json")) as f:
        return json.load(f)


# ---- field: ntt.rs:1114-1212 ----

def test_canonicalize_add_sub_kats(orc):
    assert [orc.canonicalize(x) for x in (0, 1, P - 1, P, P + 1)] == [0, 1, P - 1, 0, 1]
    assert orc.add(0, 0) == 0 and orc.add(1, 2) == 3 and orc.add(P - 1, 1) == 0 and orc.add(P - 1, P - 1) == P - 2
    assert orc.sub(0, 0) == 0 and orc.sub(5, 3) == 2 and orc.sub(0, 1) == P - 1 and orc.sub(0, P - 1) == 1


def test_mul_kats(orc):
    assert orc.mul(0, 12345) == 0 and orc.mul(1, 12345) == 12345 and orc.mul(2, 3) == 6
    assert orc.mul(P - 1, P - 1) == 1
    assert orc.mul((P + 1) // 2, 2) == 1


def test_reduce128_battery(orc):
    cases = [0, 1, P - 1, P, P + 1, 2 * P, 2 * P - 1, P * P - 1, P * (P - 1), 2**128 - 1, 2**128 - 2,
             0x123456789ABCDEF01122334455667788, 0xFFFF0000FFFF0000FFFF0000FFFF0000,
             0xAAAABBBBCCCCDDDDEEEEFFFF00001111]
    for x in cases:
        assert orc.reduce128(x) == x % P, hex(x)


def test_mul_matches_naive_mod(orc):
    xs = [0, 1, 2, 100, EPS, EPS + 1, P - 1, P // 2, P // 3, 0xDEADBEEFCAFEBABE & (P - 1),
          0x123456789ABCDEF0 & (P - 1)]
    for a in xs:
        for b in xs:
            assert orc.mul(a, b) == a * b % P


def test_pow_inv(orc):
    for a in (2, 3, 7, 12345, P - 1, P // 2):
        assert orc.pow_(a, P - 1) == 1
    for a in (1, 2, 3, 7, 12345, P - 1, P // 2):
        assert orc.mul(a, orc.inv(a)) == 1


def test_root_order(orc):
    # ntt.rs:1424-1436: omega_n has order exactly n
    for log_n in (1, 2, 4, 8, 16, 20, 32):
        n = 1 << log_n
        w = orc.primitive_root_of_unity(n)
        assert pow(w, n, P) == 1 and pow(w, n // 2, P) == P - 1
    # 7 generates the group: 7^((P-1)/q) != 1 for each prime q | P-1
    for q in (2, 3, 5, 17, 257, 65537):
        assert (P - 1) % q == 0 and pow(7, (P - 1) // q, P) != 1


# ---- transforms ----

def test_bit_reverse_kat(orc):
    assert list(orc.bit_reverse(np.arange(8))) == [0, 4, 2, 6, 1, 5, 3, 7]
    x = (np.arange(1 << 18, dtype=np.uint64) * np.uint64(0xDEADBEEF))
    assert np.array_equal(orc.bit_reverse(orc.bit_reverse(x)), x)  # parallel path, involution


def test_forward_is_naive_dft(orc):
    """ntt_forward output is A[k] = sum a_i w^(ik) in natural order (closed form)."""
    for log_n in (1, 2, 3, 5, 7):
        n = 1 << log_n
        a = orc.random_vec(n, log_n)
        w = orc.primitive_root_of_unity(n)
        want = [sum(int(a[i]) * pow(w, i * k, P) for i in range(n)) % P for k in range(n)]
        assert [int(x) for x in orc.ntt_forward(a)] == want


@pytest.mark.parametrize("log_n", [1, 2, 3, 4, 8, 12])
def test_round_trip(orc, log_n):
    a = orc.random_vec(1 << log_n, log_n)
    assert np.array_equal(orc.ntt_inverse(orc.ntt_forward(a)), a)


def conv_via(orc, fwd, inv, a, b, n):
    x = np.zeros(n, dtype=np.uint64); x[:len(a)] = a
    y = np.zeros(n, dtype=np.uint64); y[:len(b)] = b
    X, Y = fwd(x), fwd(y)
    Z = np.array([orc.mul(int(p), int(q)) for p, q in zip(X, Y)], dtype=np.uint64)
    return inv(Z)


def naive_conv(a, b):
    out = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] = (out[i + j] + int(x) * int(y)) % P
    return out


def test_convolution_kats(orc):
    r = conv_via(orc, orc.ntt_forward, orc.ntt_inverse, [1, 2], [3, 4], 4)
    assert list(r[:3]) == [3, 10, 8]
    r = conv_via(orc, orc.four_step_forward, orc.four_step_inverse, [1, 2, 3], [4, 5, 6], 16)
    assert list(r[:5]) == [4, 13, 28, 27, 18] and not r[5:].any()
    r = conv_via(orc, orc.four_step_forward, orc.four_step_inverse, [7, 11, 13], [2, 3, 5], 8)
    assert list(r[:5]) == [14, 43, 94, 94, 65]


@pytest.mark.parametrize("in_len,seed", [(7, 1), (15, 2), (31, 3), (63, 4), (127, 5)])
def test_convolution_random(orc, in_len, seed):
    a = orc.random_vec(in_len, seed) & np.uint64(0xFFFF)
    b = orc.random_vec(in_len, seed ^ 0xFF) & np.uint64(0xFFFF)
    n = 1 << (2 * in_len - 1 - 1).bit_length()
    r = conv_via(orc, orc.ntt_forward, orc.ntt_inverse, a, b, n)
    assert [int(x) for x in r[:2 * in_len - 1]] == naive_conv(a, b)


@pytest.mark.parametrize("n", [4, 8, 16, 32, 128, 256, 512, 1024, 2048])
def test_four_step_round_trip(orc, n):
    a = orc.random_vec(n, 0xCAFEF00D + n)
    assert np.array_equal(orc.four_step_inverse(orc.four_step_forward(a)), a)


@pytest.mark.parametrize("n,s1,s2", [(64, 0x11112222, 0x33334444), (256, 0xAAAABBBB, 0xCCCCDDDD),
                                     (32, 0x10102020, 0x30304040), (512, 0x50506060, 0x70708080)])
def test_four_step_convolution(orc, n, s1, s2):
    half = n // 2
    a = orc.random_vec(half, s1) & np.uint64(0xFFFF)
    b = orc.random_vec(half, s2) & np.uint64(0xFFFF)
    r = conv_via(orc, orc.four_step_forward, orc.four_step_inverse, a, b, n)
    assert [int(x) for x in r[:n - 1]] == naive_conv(a, b)


def test_twiddles_parallel_matches_serial(orc):
    w = orc.primitive_root_of_unity(1 << 20)
    for count in (1 << 19, 8192, 8192 * 3 + 17):
        s = orc.build_twiddles(w, count, parallel=False)
        p = orc.build_twiddles(w, count, parallel=True)
        assert np.array_equal(s, p)
        assert int(s[5]) == pow(w, 5, P)


# ---- pack / unpack: ntt.rs:1321-1356, 1729-1768 ----

def test_pack_unpack_round_trips(orc):
    for limbs in ([0xDEADBEEFCAFEBABE],
                  [0xDEADBEEFCAFEBABE, 0x123456789ABCDEF0, 0x55555555AAAAAAAA, 2**64 - 1, 1]):
        c = orc.pack(limbs, 4 * len(limbs) + 5)
        assert not c[4 * len(limbs):].any()
        assert [int(x) for x in orc.unpack(c)] == limbs
    r = orc.random_vec(40, 12345)
    assert np.array_equal(orc.unpack(orc.pack(r, 170)), orc.trim(r))


def test_unpack_parallel_matches_serial(orc):
    n = (1 << 20) + 4321
    i = np.arange(n, dtype=np.uint64)
    c = ((i * np.uint64(0x123456789ABCDEF0)) >> np.uint64(8)) & np.uint64(0xFFFFFFFF)
    assert np.array_equal(orc.unpack_serial(c), orc.unpack_parallel(c))
    n = (1 << 20) * 4
    i = np.arange(n, dtype=np.uint64)
    c = np.where(i % np.uint64(7) == 0, np.uint64(0xFFFFFFFF), (i * np.uint64(0xDEADBEEF)) & np.uint64(0xFFFFFFFF))
    s = orc.unpack_serial(c)
    assert np.array_equal(s, orc.unpack_parallel(c))
    # the low 1024 bits only depend on the first 64 coefficients (and carries from them)
    assert orc.to_int(s[:16]) == sum(int(x) << (16 * k) for k, x in enumerate(c[:64])) % (1 << 1024)


def test_unpack_exact_sum_small(orc):
    rng = np.random.default_rng(3)
    c = rng.integers(0, 2**58, size=777, dtype=np.uint64)
    assert orc.to_int(orc.unpack(c)) == sum(int(x) << (16 * k) for k, x in enumerate(c))


# ---- products ----

def test_mul_kats(orc):
    assert list(orc.mul_mag_ntt([3], [7])) == [21]
    m = [2**64 - 1]
    assert orc.to_int(orc.mul_mag_ntt(m, m)) == (2**64 - 1) ** 2
    assert list(orc.mul_mag([12345], [67890])) == [838102050]  # integer.rs:1391-1395
    assert orc.to_int(orc.mul_mag([0, 1], [0, 1])) == 2**128     # integer.rs:1397-1410


def test_golden_small_products(orc):
    g = golden()
    assert [hex(int(x)) for x in orc.splitmix64(1, 4)] == g["generators"]["splitmix64_seed1_first4"]
    assert [hex(int(x)) for x in orc.random_vec(4, 7)] == g["generators"]["random_vec_7_seed7_first4"]
    for case in g["small"]:
        a = orc.random_vec(case["la"], case["seed_a"])
        b = orc.random_vec(case["lb"], case["seed_b"])
        want = [int(x, 16) for x in case["product"]]
        assert [int(x) for x in orc.mul_mag_ntt(a, b)] == want
        assert [int(x) for x in orc.schoolbook(a, b)] == want


def _sha(limbs):
    return hashlib.sha256(np.ascontiguousarray(limbs, dtype="<u8").tobytes()).hexdigest()


def test_golden_hashed_products(orc):
    """C1: the 2^16 x 2^16 product (N = 2^19) against GMP/CPython-made hashes."""
    for case in golden()["hashed"]:
        a = orc.splitmix64(case["seed_a"], case["la"])
        b = orc.splitmix64(case["seed_b"], case["lb"])
        if a[-1] == 0:
            a[-1] = 1
        if b[-1] == 0:
            b[-1] = 1
        c = orc.mul_mag_ntt(a, b)
        assert c.size == case["len"] and _sha(c) == case["sha256"]


def test_four_step_product_routes(orc):
    """ntt.rs:2055-2101: mul_mag_ntt through four-step (square and non-square N)."""
    prev = orc.get_four_step_threshold()
    try:
        for n, sa, sb in ((300, 0x1357, 0x2468), (600, 0xAB, 0xCD)):
            a, b = orc.random_vec(n, sa), orc.random_vec(n, sb)
            want = orc.schoolbook(a, b)
            orc.set_four_step_threshold(1)
            assert np.array_equal(orc.mul_mag_ntt(a, b), want)
            orc.set_four_step_threshold(1 << 25)
            assert np.array_equal(orc.mul_mag_ntt(a, b), want)
    finally:
        orc.set_four_step_threshold(prev)


def test_karatsuba_and_dispatch(orc):
    a, b = orc.make_int(5, 3000), orc.make_int(6, 2000)
    want = orc.to_int(a) * orc.to_int(b)
    assert orc.to_int(orc.karatsuba(a, b)) == want
    assert np.array_equal(orc.karatsuba(a, b), orc.schoolbook(a, b))
    # Karatsuba above the NTT cap (integer.rs:1727-1754): a tiny injected cap
    a, b = orc.make_int(7, 20000), orc.make_int(8, 9000)
    want = orc.to_int(a) * orc.to_int(b)
    assert orc.to_int(orc.mul_mag_dispatch(a, b, 1 << 14, 8192, 32)) == want
    assert orc.to_int(orc.mul_mag_dispatch(a, b, 1 << 32, 8192, 32)) == want
    # guard formula: combined limbs above 2^30 (4(na+nb) > 2^32) do not fit one transform
    lim = 1 << 32
    assert (lim // 4 + 1) * 4 > lim and (lim // 4 - 1) * 4 <= lim


def test_gmp_agrees_where_present(orc):
    if orc.gmp() is None:
        pytest.skip("libgmp.so.10 not present")
    a, b = orc.splitmix64(31, 50000), orc.splitmix64(32, 40000)
    assert np.array_equal(orc.mul_mag_ntt(a, b), orc.gmp_mul(a, b))
    assert np.array_equal(orc.gmp_sqr(a), orc.gmp_mul(a, a))


def test_fingerprint_detects_a_planted_error(orc):
    """The residue check used at 2^24 limbs catches a single wrong limb."""
    a, b = orc.splitmix64(41, 5000), orc.splitmix64(42, 5000)
    c = orc.mul_mag_ntt(a, b)
    q = (1 << 61) - 1
    assert orc.mod_small(c, q) == orc.mod_small(a, q) * orc.mod_small(b, q) % q
    bad = c.copy(); bad[2500] ^= np.uint64(1 << 17)
    assert orc.mod_small(bad, q) != orc.mod_small(a, q) * orc.mod_small(b, q) % q
