"""GPU parity: the HIP multiply (through the C ABI) against the oracle.

Bit-exact everywhere: the product of two integers is unique.  Cases follow the
reference's own multiply tests (crates/bignum/src/ntt.rs:1384-1487,
1527-1567, 1712-1722, 2055-2101) with its `random_vec` inputs, plus GMP
mpn_mul/mpn_sqr and CPython int as independent references, closed forms for
all-ones operands, and residue fingerprints at the 2^24-limb size.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M61 = (1 << 61) - 1
FP_PRIMES = (M61, 0x1FFFFFFFFFFFFFC7, 0x0FFFFFFFFFFFFFC5)  # 2^61-1 and two primes below 2^61, 2^60


def log(msg):
    print(msg, flush=True)


def check(orc, got, a, b=None):
    want = orc.gmp_mul(a, b) if b is not None else orc.gmp_sqr(a)
    assert got.dtype == np.uint64
    assert np.array_equal(got, want), f"mismatch: got {got.size} limbs, want {want.size}"


# ---- the reference's mul_mag_ntt tests (ntt.rs:1384-1487) ----

def test_tiny_3x7(ctx, orc):
    log("3*7")
    assert list(ctx.mul_mag_ntt([3], [7])) == [21]


@pytest.mark.parametrize("la,lb,seed", [(1, 1, 1), (1, 4, 2), (4, 4, 3), (8, 5, 4), (32, 16, 5), (60, 60, 6),
                                        (200, 200, 7)])
def test_small_random_vs_schoolbook(ctx, orc, la, lb, seed):
    log(f"random_vec {la}x{lb} seed {seed}")
    a = orc.random_vec(la, seed)
    b = orc.random_vec(lb, seed ^ 0xFF)
    assert np.array_equal(ctx.mul_mag_ntt(a, b), orc.schoolbook(a, b))


def test_u64_max_squared(ctx, orc):
    m = np.array([2**64 - 1], dtype=np.uint64)
    got = ctx.mul_mag_ntt(m, m)
    assert orc.to_int(got) == (2**64 - 1) ** 2
    assert np.array_equal(got, orc.schoolbook(m, m))


def test_medium_500(ctx, orc):
    a = orc.random_vec(500, 0xDEADBEEFCAFEBABE)
    b = orc.random_vec(500, 0xDEADBEEFCAFEBABE ^ 0xFF)
    assert np.array_equal(ctx.mul_mag_ntt(a, b), orc.schoolbook(a, b))


@pytest.mark.parametrize("n,sa,sb", [(200, 0x123456789ABCDEF0, 0x0FEDCBA987654321), (300, 0xCAFE0001, 0xBEEF0002),
                                     (300, 0x1357, 0x2468), (400, 0x1111222233334444, 0x5555666677778888),
                                     (600, 0xAB, 0xCD)])
def test_reference_seeds(ctx, orc, n, sa, sb):
    a, b = orc.random_vec(n, sa), orc.random_vec(n, sb)
    want = orc.schoolbook(a, b)
    assert np.array_equal(ctx.mul_mag_ntt(a, b), want)
    assert np.array_equal(ctx.mul_mag_ntt(a, b), want)  # pooled buffers reused


def test_32768_limbs(ctx, orc):
    a, b = orc.random_vec(32768, 0xAA), orc.random_vec(32768, 0xBB)
    assert np.array_equal(ctx.mul_mag_ntt(a, b), orc.mul_mag_ntt(a, b))


# ---- edge cases ----

def test_empty_and_zero(ctx, orc):
    e = np.zeros(0, dtype=np.uint64)
    assert ctx.mul_mag_ntt(e, [5]).size == 0
    assert ctx.mul_mag_ntt([5], e).size == 0
    assert ctx.sqr(e).size == 0
    assert ctx.mul_mag_ntt([0], [7]).size == 0          # zero, not normalised
    assert ctx.mul_mag_ntt([0, 0, 0], [1, 2]).size == 0


def test_unnormalised_inputs_trimmed(ctx, orc):
    a = np.array([5, 0, 0, 0], dtype=np.uint64)
    b = np.array([7, 9, 0], dtype=np.uint64)
    got = ctx.mul_mag_ntt(a, b)
    assert list(got) == [35, 45]
    assert np.array_equal(got, orc.mul_mag_ntt(a, b))


@pytest.mark.parametrize("top", [0, 2046, 2047, 2048, 4095, 6000, 16383])
def test_zero_top_limbs_trimmed(ctx, orc, top):
    """Products whose highest nonzero limb `top` lies below the last 2048-limb
    carry tile (every operand padded with thousands of zero limbs): the
    trimmed length is top + 1, also at and around tile boundaries."""
    log(f"zero top limbs, top={top}")
    ia = top // 3
    a = np.zeros(ia + 9000, dtype=np.uint64); a[ia] = 3; a[:ia] = orc.splitmix64(top + 1, ia)
    b = np.zeros(top - ia + 7000, dtype=np.uint64); b[top - ia] = 5
    got = ctx.mul_mag_ntt(a, b)
    assert got.size == top + 1   # limb `top` of the product is a[ia] * 5 plus a carry below 5: nonzero, no overflow
    assert np.array_equal(got, orc.mul_mag_ntt(a, b))
    assert orc.to_int(got) == orc.to_int(a) * orc.to_int(b)


@pytest.mark.parametrize("la,lb", [(1000, 3), (1, 100000), (3, 70000), (12345, 54321), (8191, 8193)])
def test_unbalanced(ctx, orc, la, lb):
    log(f"unbalanced {la}x{lb}")
    a, b = orc.splitmix64(la, la), orc.splitmix64(lb + 1, lb)
    check(orc, ctx.mul_mag_ntt(a, b), a, b)


def test_every_transform_length(ctx, orc):
    """One product per transform length 2^3 .. 2^24 under the tuned plans that
    bench.py and pi_compute load (their splits asserted)."""
    from pi_amd import gnt
    from conftest import tuned_plans
    for log_n, sp in tuned_plans().items():
        assert gnt.plan(log_n) == (len(sp), sp), log_n
    _every_length(ctx, orc)


def test_every_transform_length_builtin_plans(ctx, orc, builtin_plans):
    """The same with the library's built-in splits (no plan cache)."""
    from pi_amd import gnt
    assert gnt.plan(27) == (3, [9, 9, 9])
    _every_length(ctx, orc)


def _every_length(ctx, orc):
    for log_n in range(3, 25):
        # na + nb = 2^(log_n-2) gives n_out = 2^log_n - 1
        tot = 1 << (log_n - 2) if log_n >= 3 else 2
        na = max(1, tot // 2 - 3)
        nb = max(1, tot - na)
        a, b = orc.splitmix64(100 + log_n, na), orc.splitmix64(200 + log_n, nb)
        log(f"log_n={log_n} na={na} nb={nb}")
        check(orc, ctx.mul_mag_ntt(a, b), a, b)
        check(orc, ctx.sqr(a), a)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 1000, 4096, 65537, 1 << 18])
def test_all_ones_closed_form(ctx, orc, n):
    """(B^n - 1)^2 = [1, 0 x (n-1), 2^64-2, (2^64-1) x (n-1)]: maximal carries."""
    log(f"all-ones n={n}")
    a = np.full(n, 2**64 - 1, dtype=np.uint64)
    want = np.zeros(2 * n, dtype=np.uint64)
    want[0] = 1
    want[n] = np.uint64(2**64 - 2)
    want[n + 1:] = np.uint64(2**64 - 1)
    assert np.array_equal(ctx.mul_mag_ntt(a, a), want)
    assert np.array_equal(ctx.sqr(a), want)


def test_all_ones_rectangular(ctx, orc):
    """(B^m - 1)(B^n - 1) = B^(m+n) - B^m - B^n + 1."""
    m, n = 3000, 777
    a = np.full(m, 2**64 - 1, dtype=np.uint64)
    b = np.full(n, 2**64 - 1, dtype=np.uint64)
    B = 1 << 64
    assert orc.to_int(ctx.mul_mag_ntt(a, b)) == B ** (m + n) - B**m - B**n + 1


def test_powers_of_two(ctx, orc):
    a = np.zeros(5000, dtype=np.uint64); a[-1] = 1 << 63
    b = np.zeros(3000, dtype=np.uint64); b[-1] = 1 << 7
    assert orc.to_int(ctx.mul_mag_ntt(a, b)) == (1 << (64 * 4999 + 63)) * (1 << (64 * 2999 + 7))


def test_carry_chain_across_tiles(ctx, orc):
    """A long propagate run fed by one generate: x = B^n - 1 times (B^k + 1)."""
    n, k = 10000, 1
    a = np.full(n, 2**64 - 1, dtype=np.uint64)
    b = np.zeros(k + 1, dtype=np.uint64); b[0] = 1; b[k] = 1
    B = 1 << 64
    assert orc.to_int(ctx.mul_mag_ntt(a, b)) == (B**n - 1) * (B**k + 1)


def test_cpython_int_cross_check(ctx, orc):
    a, b = orc.splitmix64(11, 3333), orc.splitmix64(12, 4444)
    assert orc.to_int(ctx.mul_mag_ntt(a, b)) == orc.to_int(a) * orc.to_int(b)


# ---- workload distributions at 2^20 limbs (exact vs GMP) ----

def _dist(orc, kind, n, seed):
    if kind == "uniform":
        x = orc.splitmix64(seed, n)
    elif kind == "allones":
        x = np.full(n, 2**64 - 1, dtype=np.uint64)
    else:  # sparse: 1% nonzero limbs
        r = orc.splitmix64(seed, 2 * n)
        keep = (r[:n] % np.uint64(100)) == 0
        x = np.where(keep, r[n:], np.uint64(0)).astype(np.uint64)
    if x[-1] == 0:
        x[-1] = 1
    return x


@pytest.mark.parametrize("kind", ["uniform", "allones", "sparse"])
def test_2p20_distributions(ctx, orc, kind):
    n = 1 << 20
    log(f"2^20 {kind}")
    a, b = _dist(orc, kind, n, 3), _dist(orc, kind, n, 4)
    check(orc, ctx.mul_mag_ntt(a, b), a, b)
    check(orc, ctx.sqr(a), a)


def _fingerprints_ok(orc, c, a, b):
    for q in FP_PRIMES:
        if orc.mod_small(c, q) != (orc.mod_small(a, q) * orc.mod_small(b, q)) % q:
            return False
    return True


@pytest.mark.parametrize("kind", ["uniform", "allones", "sparse"])
def test_2p24_distributions(ctx, orc, kind):
    """2^24 x 2^24 (N = 2^27) and the square of a: GMP when present, residue
    fingerprints and the low limb always; the closed form for all-ones."""
    from pi_amd import gnt
    from conftest import tuned_plans
    n = 1 << 24
    # the benched configuration: the tuned split of N = 2^27
    assert gnt.plan(27) == (3, tuned_plans()[27]) == (3, [8, 9, 10])
    log(f"2^24 {kind}")
    a, b = _dist(orc, kind, n, 5), _dist(orc, kind, n, 6)
    c = ctx.mul_mag_ntt(a, b)
    assert c.size in (2 * n, 2 * n - 1)
    assert int(c[0]) == (int(a[0]) * int(b[0])) % (1 << 64)
    assert _fingerprints_ok(orc, c, a, b)
    if kind == "allones":
        assert c[0] == 1 and c[n] == 2**64 - 2
        assert not c[1:n].any() and (c[n + 1:] == np.uint64(2**64 - 1)).all()
    elif orc.gmp() is not None:
        assert np.array_equal(c, orc.gmp_mul(a, b))
    s = ctx.sqr(a)
    assert _fingerprints_ok(orc, s, a, a)
    if kind == "allones":
        assert np.array_equal(s, c)  # a == b: the square is the closed form above
    elif orc.gmp() is not None:
        assert np.array_equal(s, orc.gmp_sqr(a))


def _ones_product_ok(c, m, n):
    """(B^m - 1)(B^n - 1) = B^(m+n) - B^m - B^n + 1 for m >= n."""
    return (c.size == m + n and c[0] == 1 and not c[1:n].any()
            and (c[n:m] == np.uint64(2**64 - 1)).all() and c[m] == np.uint64(2**64 - 2)
            and (c[m + 1:] == np.uint64(2**64 - 1)).all())


def test_transform_lengths_up_to_the_cap(ctx, orc):
    """One product and one square per transform length 2^25 .. 2^30, each with
    4(na+nb) - 1 just under the power of two: length, low limb and residue
    fingerprints (GMP at these sizes would take minutes).  Then 2^31 and the cap
    2^32 (4(na+nb) <= 2^32, as ntt_fits, integer.rs:476-478), four-pass plans,
    by the all-ones closed form, and one limb more is refused with
    GNT_E_TOO_LARGE."""
    from pi_amd import gnt
    for log_n in range(25, 31):
        tot = 1 << (log_n - 2)
        na = tot // 2 + 12345
        nb = tot - na
        log(f"log_n={log_n} na={na} nb={nb}")
        a, b = orc.splitmix64(300 + log_n, na), orc.splitmix64(400 + log_n, nb)
        c = ctx.mul_mag_ntt(a, b)
        assert c.size in (na + nb, na + nb - 1)
        assert int(c[0]) == (int(a[0]) * int(b[0])) % (1 << 64)
        assert _fingerprints_ok(orc, c, a, b), log_n
        del c
        s = ctx.sqr(a[: tot // 2])
        assert _fingerprints_ok(orc, s, a[: tot // 2], a[: tot // 2]), log_n
        del s, a, b
    for log_n in (31, 32):
        m, n = (1 << (log_n - 3)) + 5, (1 << (log_n - 3)) - 5
        assert gnt.plan_info(m, n)[0] == log_n and len(gnt.plan_info(m, n)[1]) == 4
        log(f"all-ones {m} x {n} (log_n={log_n})")
        ones = np.full(m, np.uint64(2**64 - 1))
        c = ctx.mul_mag_ntt(ones, ones[:n])
        assert _ones_product_ok(c, m, n), log_n
        del c
    with pytest.raises(gnt.GntError) as e:
        ctx.mul_mag_ntt(ones, ones[: n + 1])
    assert e.value.code == gnt.GNT_E_TOO_LARGE
    del ones


# ---- device-resident operands, batches, dispatch ----

def test_device_chain_reuse(ctx, orc):
    """A result kept in HBM feeds the next product: ((a*b)^2) * a."""
    a, b = orc.splitmix64(21, 7000), orc.splitmix64(22, 5000)
    da, db = ctx.upload(a), ctx.upload(b)
    p = ctx.mul_buf(da, db)
    s = ctx.sqr_buf(p)
    t = ctx.mul_buf(s, da)
    A, Bv = orc.to_int(a), orc.to_int(b)
    assert orc.to_int(t.download()) == (A * Bv) ** 2 * A
    assert len(p) == orc.gmp_mul(a, b).size
    for h in (da, db, p, s, t):
        h.release()


def test_batch_matches_single(ctx, orc):
    rng = np.random.default_rng(7)
    pairs = []
    for i in range(40):
        n = int(round(2 ** rng.uniform(8, 16)))
        a = orc.splitmix64(1000 + i, n)
        pairs.append((a, None if i % 3 == 0 else orc.splitmix64(2000 + i, n)))
    got = ctx.mul_batch(pairs)
    ser = ctx.mul_batch(pairs, serial=True)
    for (a, b), g, s in zip(pairs, got, ser):
        want = orc.gmp_sqr(a) if b is None else orc.gmp_mul(a, b)
        assert np.array_equal(g, want) and np.array_equal(s, want)


@pytest.mark.parametrize("la,lb", [(5, 7), (40, 100), (3000, 2500), (9000, 9000), (20000, 9000), (100, 30000)])
def test_mul_mag_dispatch(ctx, orc, la, lb):
    a, b = orc.splitmix64(la, la), orc.splitmix64(lb + 3, lb)
    check(orc, ctx.mul_mag(a, b), a, b)
    # a low NTT threshold routes the Karatsuba leaves to the GPU too
    check(orc, ctx.mul_mag(a, b, ntt_threshold=64), a, b)


def test_karatsuba_above_the_transform_cap(ctx, orc):
    """The reference's karatsuba_above_ntt_routes_correctly_on_oversize
    (integer.rs:1727-1754) with its inputs: a cap of 100 coefficients forces
    Karatsuba on 200-limb operands, a large cap takes the transform; both equal
    schoolbook.  Then GPU leaves under an injected cap: 2^17 x 2^17 limbs with
    4(na+nb) <= 2^17 per leaf split into Karatsuba levels whose concurrent
    sub-products (>= 512 limbs) are GPU transforms, unbalanced operands, and the
    same with the serving loop batching the concurrent leaves."""
    from pi_amd import gnt
    M = 1 << 64
    a = np.array([(0xDEADBEEFCAFE0001 * (i + 1)) % M for i in range(200)], dtype=np.uint64)
    b = np.array([(0xFACEF00D12345678 * (i + 7)) % M for i in range(200)], dtype=np.uint64)
    want = orc.schoolbook(a, b)
    assert np.array_equal(ctx.mul_mag(a, b, ntt_threshold=50, karatsuba_threshold=32, ntt_max_coeffs=100), want)
    assert np.array_equal(ctx.mul_mag(a, b, ntt_threshold=50, karatsuba_threshold=32, ntt_max_coeffs=1 << 32), want)
    cases = [(1 << 17, 1 << 17, 1 << 17), ((1 << 17) + 999, 70001, 1 << 16), (3 << 15, 5000, 1 << 15)]
    for la, lb, cap in cases:
        x, y = orc.splitmix64(la + 5, la), orc.splitmix64(lb + 6, lb)
        log(f"karatsuba above cap {cap}: {la} x {lb}")
        check(orc, ctx.mul_mag(x, y, ntt_threshold=1024, ntt_max_coeffs=cap), x, y)
    c2 = gnt.Context(0)
    srv = gnt.Server(c2)
    try:
        x, y = orc.splitmix64(77, 1 << 17), orc.splitmix64(78, (1 << 17) - 3)
        check(orc, c2.mul_mag(x, y, ntt_threshold=1024, ntt_max_coeffs=1 << 16), x, y)
        assert srv.stats()["completed"] >= 9          # the leaves went through the loop
    finally:
        srv.stop()
        c2.close()


# ---- batch plans: level-grouped launches, hipGraph capture and replay ----

def _mixed_pairs(orc, n, lo, hi, seed, sqr_every=4):
    rng = np.random.default_rng(seed)
    pairs = []
    for i in range(n):
        m = int(round(2 ** rng.uniform(lo, hi)))
        if i % 11 == 5:
            a = np.full(m, 2**64 - 1, dtype=np.uint64)   # maximal carries inside a batch
        else:
            a = orc.splitmix64(seed * 1000 + i, m)
        b = None if i % sqr_every == 0 else orc.splitmix64(seed * 1000 + 500 + i, int(round(2 ** rng.uniform(lo, hi))))
        pairs.append((a, b))
    return pairs


def _want(orc, a, b):
    return orc.gmp_sqr(a) if b is None else orc.gmp_mul(a, b)


@pytest.mark.parametrize("graph", [True, False])
def test_batch_plan_replay(ctx, orc, graph):
    from pi_amd import gnt
    pairs = _mixed_pairs(orc, 48, 1, 15, 3)
    log(f"batch plan graph={graph} n={len(pairs)}")
    plan = ctx.batch_plan(pairs, graph=graph)
    info = plan.info()
    assert info["launches_per_run"] < 6 * len(pairs)   # levels grouped, not one sequence per product
    plan.run()
    for (a, b), got in zip(pairs, plan.results()):
        assert np.array_equal(got, _want(orc, a, b))
    # same shapes, new operands: upload and replay
    pairs2 = [(orc.splitmix64(77 + i, a.size), None if b is None else orc.splitmix64(99 + i, b.size))
              for i, (a, b) in enumerate(pairs)]
    jobs, keep, outs = ctx._jobs(pairs2)
    gnt._check(gnt.load().gnt_batch_upload(plan.h, jobs))
    plan.run()
    plan.run()   # a second replay on the same inputs gives the same results
    for (a, b), got in zip(pairs2, plan.results()):
        assert np.array_equal(got, _want(orc, a, b))
    plan.run(serial=True)  # one launch sequence per product, same buffers
    for (a, b), got in zip(pairs2, plan.results()):
        assert np.array_equal(got, _want(orc, a, b))
    plan.close()


def test_batch_plan_uneven_large(ctx, orc):
    """Products from 2^8 to 2^20 limbs in one plan (1-, 2- and 3-pass plans together)."""
    pairs = _mixed_pairs(orc, 24, 8, 20, 5, sqr_every=3)
    plan = ctx.batch_plan(pairs, graph=True)
    plan.run()
    for (a, b), got in zip(pairs, plan.results()):
        assert np.array_equal(got, _want(orc, a, b))
    plan.close()


# ---- serving loop: continuous batching ----

def test_serving_mix_equals_alone(ctx, orc):
    """Requests submitted together (work overlapping on the device) give the
    same bytes as the same request alone, and the exact product."""
    from pi_amd import gnt
    rng = np.random.default_rng(11)
    reqs = []
    for i in range(40):
        n = int(round(2 ** rng.uniform(4, 17)))
        a = orc.splitmix64(3000 + i, n)
        sq = i % 3 == 0
        b = None if sq else orc.splitmix64(4000 + i, int(round(2 ** rng.uniform(4, 17))))
        reqs.append((a, b))
    srv = gnt.Server(ctx)
    try:
        tickets = [srv.submit(a, b, square=b is None) for a, b in reqs]
        results = []
        for t in tickets:
            buf, lat = srv.wait(t)
            assert lat > 0
            results.append(buf.download())
            buf.release()
        st = srv.stats()
        assert st["completed"] >= len(reqs) and st["peak_admitted"] > 1   # they were in flight together
        # a result kept on the device as the next request's operand (no host round trip)
        a, b = reqs[1]
        t = srv.submit(a, b)
        prod, _ = srv.wait(t)
        t2 = srv.submit(a_buf=prod, square=True)
        sq, _ = srv.wait(t2)
        A, B = orc.to_int(a), orc.to_int(b)
        assert orc.to_int(sq.download()) == (A * B) ** 2
        prod.release(); sq.release()
    finally:
        srv.stop()
    for (a, b), got in zip(reqs, results):
        alone = ctx.sqr(a) if b is None else ctx.mul_mag_ntt(a, b)
        assert np.array_equal(got, alone)
        assert np.array_equal(got, orc.gmp_sqr(a) if b is None else orc.gmp_mul(a, b))


def test_serving_admission_cap(orc):
    """With a tiny HBM cap the loop admits one request at a time and still completes all."""
    from pi_amd import gnt
    c = gnt.Context(0, hbm_cap_bytes=1 << 20)
    srv = gnt.Server(c)
    try:
        pairs = [(orc.splitmix64(50 + i, 3000), orc.splitmix64(60 + i, 2000)) for i in range(6)]
        ts = [srv.submit(a, b) for a, b in pairs]
        outs = []
        for t in ts:
            buf, _ = srv.wait(t)
            outs.append(buf.download())
            buf.release()
        assert srv.stats()["peak_bytes"] <= (1 << 20) or srv.stats()["peak_admitted"] == 1
        for (a, b), o in zip(pairs, outs):
            assert np.array_equal(o, orc.gmp_mul(a, b))
    finally:
        srv.stop()
        c.close()


def test_serving_two_lanes(orc):
    """Both lanes at once (a split at 2^14 sends larger transforms to the
    large lane), a small-lane result taken as a large-lane operand, and a
    cap that makes one lane wait for bytes the other frees: every result
    exact and the same as alone."""
    from pi_amd import gnt
    rng = np.random.default_rng(21)
    reqs = []
    for i in range(30):
        a = orc.splitmix64(7000 + i, int(round(2 ** rng.uniform(3, 13))))
        b = None if i % 4 == 0 else orc.splitmix64(8000 + i, int(round(2 ** rng.uniform(3, 13))))
        reqs.append((a, b))
    for cap in (0, 3 << 20):
        print(f"cap {cap}", flush=True)
        c = gnt.Context(0, hbm_cap_bytes=cap, serve_split_log_n=14)
        srv = gnt.Server(c)
        try:
            ts = [srv.submit(a, b, square=b is None) for a, b in reqs]
            outs = []
            for t in ts:
                buf, _ = srv.wait(t)
                outs.append(buf.download())
                buf.release()
            # small lane (N = 2^11) feeding the large lane (N = 2^15)
            x, y = orc.splitmix64(91, 150), orc.splitmix64(92, 200)
            z = orc.splitmix64(93, 5000)
            t = srv.submit(x, y)
            xy, _ = srv.wait(t)
            t2 = srv.submit(b=z, a_buf=xy)
            xyz, _ = srv.wait(t2)
            assert orc.to_int(xyz.download()) == orc.to_int(x) * orc.to_int(y) * orc.to_int(z)
            xy.release()
            xyz.release()
            st = srv.stats()
            assert st["small_lane_iterations"] > 0 and st["large_lane_iterations"] > 0, st
            if cap:
                assert st["peak_bytes"] <= cap or st["peak_admitted"] == 1, st
        finally:
            srv.stop()
            c.close()
        for (a, b), got in zip(reqs, outs):
            assert np.array_equal(got, orc.gmp_sqr(a) if b is None else orc.gmp_mul(a, b))


@pytest.mark.parametrize("inflight", [0, 3, (1 << 64) - 1])
def test_serving_late_b_operand(orc, inflight):
    """Multiplies whose host B (2 MiB and more) is still uploading when A's
    forward passes start (host A and host B, and a device A with host B),
    several at once, with the large lane advancing 1 (default), 3 or any
    number of products: every product exact and the same as alone."""
    from pi_amd import gnt
    c = gnt.Context(0, serve_split_log_n=20, serve_large_inflight=inflight)
    srv = gnt.Server(c)
    try:
        n = 1 << 18
        ops = [(orc.splitmix64(500 + i, n - 7 * i), orc.splitmix64(600 + i, n + 5 * i)) for i in range(4)]
        ts = [srv.submit(a, b) for a, b in ops]
        dev_a = c.upload(ops[0][0])
        t_dev = srv.submit(b=ops[1][1], a_buf=dev_a)
        outs = []
        for t in ts:
            buf, _ = srv.wait(t)
            outs.append(buf.download())
            buf.release()
        buf, _ = srv.wait(t_dev)
        got_dev = buf.download()
        buf.release()
        dev_a.release()
        st = srv.stats()
        assert st["large_lane_iterations"] > 0, st
    finally:
        srv.stop()
        c.close()
    for (a, b), got in zip(ops, outs):
        assert np.array_equal(got, orc.gmp_mul(a, b))
    assert np.array_equal(got_dev, orc.gmp_mul(ops[0][0], ops[1][1]))


# ---- serving loop: sums and differences (binary splitting's t = x +- y) ----

def _rand_int(rng, n):
    return int.from_bytes(rng.bytes(8 * n), "little") | (1 << (64 * n - 1))


def _limbs(x):
    from oracle import oracle
    return oracle.from_int(x)


def test_serving_add_sub(ctx, orc):
    """a + b and a - b (a >= b) through the serving loop, exact against CPython
    int, with carries and borrows that cross carry tiles (2048 limbs) and
    differences whose top limbs cancel."""
    from pi_amd import gnt
    rng = np.random.default_rng(5)
    ones = (1 << (64 * 5000)) - 1
    X = _rand_int(rng, 6000)
    cases = [
        ("add", _rand_int(rng, 3000), _rand_int(rng, 3000)),
        ("add", _rand_int(rng, 5000), _rand_int(rng, 17)),
        ("add", _rand_int(rng, 3), _rand_int(rng, 4100)),
        ("add", ones, 1),                                   # carry through 5000 limbs
        ("add", ones, ones),
        ("sub", _rand_int(rng, 5000), _rand_int(rng, 4999)),
        ("sub", _rand_int(rng, 5000), _rand_int(rng, 12)),
        ("sub", 1 << (64 * 4500), 1),                       # borrow through 4500 limbs
        ("sub", X + 12345, X),                              # top 6000 limbs cancel
        ("sub", X + (1 << (64 * 2500)), X),                 # cancellation above a carry tile
        ("sub", X, X),                                      # zero
        ("add", 5, 7),
        ("sub", 7, 5),
        # limb 0 of a - b is 2^64 - 1 and carries out (the two's complement's
        # +1): it generates and must not propagate as well
        ("sub", (1 << 128) - 1, 1 << 64),
        ("sub", (2**64 - 1) | (0xDEADBEEF << 64), 0x1234 << 64),
        ("sub", (2**64 - 1) | (X << 64), (X - 1) << 64),
    ]
    srv = gnt.Server(ctx)
    try:
        ts = [srv.submit(_limbs(a), _limbs(b), op=op) for op, a, b in cases]
        for i, ((op, a, b), t) in enumerate(zip(cases, ts)):
            buf, _ = srv.wait(t)
            want = a + b if op == "add" else a - b
            got = buf.download()
            assert orc.to_int(got) == want, f"case {i}: {op} {a.bit_length()} {b.bit_length()}"
            assert len(buf) == (want.bit_length() + 63) // 64, f"case {i}: length"
            buf.release()
        # device-resident chain: (x*y) - (z*w) and (x*y) + (z*w) from product results
        x, y, z, w = (_rand_int(rng, n) for n in (4000, 3900, 2000, 1500))
        tp = srv.submit(_limbs(x), _limbs(y))
        tq = srv.submit(_limbs(z), _limbs(w))
        p, _ = srv.wait(tp)
        q, _ = srv.wait(tq)
        td = srv.submit(a_buf=p, b_buf=q, op="sub")
        ts_ = srv.submit(a_buf=p, b_buf=q, op="add")
        d, _ = srv.wait(td)
        s, _ = srv.wait(ts_)
        assert orc.to_int(d.download()) == x * y - z * w
        assert orc.to_int(s.download()) == x * y + z * w
        for b_ in (p, q, d, s):
            b_.release()
    finally:
        srv.stop()


def test_concurrent_host_products_join_the_serving_loop(orc):
    """gnt_mul from several host threads while the serving loop runs: every
    call goes through the loop (continuous batching) and is exact."""
    import threading
    from pi_amd import gnt
    c = gnt.Context(0)
    srv = gnt.Server(c)
    try:
        rng = np.random.default_rng(21)
        jobs = [(orc.splitmix64(700 + i, int(2 ** rng.uniform(10, 17))), orc.splitmix64(800 + i, int(2 ** rng.uniform(10, 17))))
                for i in range(24)]
        out = [None] * len(jobs)

        def work(k):
            for i in range(k, len(jobs), 6):
                a, b = jobs[i]
                out[i] = c.mul_mag_ntt(a, b) if i % 4 else c.sqr(a)

        ts = [threading.Thread(target=work, args=(k,)) for k in range(6)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        st = srv.stats()
        assert st["completed"] >= len(jobs)            # all of them went through the loop
        for i, (a, b) in enumerate(jobs):
            want = orc.gmp_mul(a, b) if i % 4 else orc.gmp_sqr(a)
            assert np.array_equal(out[i], want), i
    finally:
        srv.stop()
        c.close()
