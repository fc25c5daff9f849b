"""Host <-> device copies through pinned staging (gnt_set_staging): uploads
and downloads are exact across chunk boundaries, copy-team sizes (growing
and shrinking) and the direct path; products through every host entry point
(single, batch, serving loop) equal GMP's with staging on and off; and
threads copying through one context at once each get their own data."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHUNK = 64 << 10                    # 8192 limbs per staging chunk
SIZES = [1, 8191, 8192, 8193, 3 * 8192 + 5, 4 * 8192, 5 * 8192 + 1, 200_003, (1 << 20) + 3]


@pytest.fixture
def staging():
    from pi_amd import gnt
    yield gnt.set_staging
    gnt.set_staging()  # the library's defaults


def _limbs(orc, seed, n):
    a = orc.splitmix64(seed, n)
    a[-1] |= 1
    return a


@pytest.mark.parametrize("threads", [0, 1, 2, 4, 8, 3])
def test_round_trip_across_chunks(ctx, orc, staging, threads):
    staging(threads, CHUNK, 0)
    for n in SIZES:
        print(f"threads {threads} n {n}", flush=True)
        a = _limbs(orc, n, n)
        buf = ctx.upload(a)
        got = buf.download()
        buf.release()
        assert np.array_equal(got, a), n


def test_products_staged_and_direct(ctx, orc, staging):
    from pi_amd import gnt
    a = _limbs(orc, 5, (1 << 19) + 11)
    b = _limbs(orc, 6, (1 << 19) - 3)
    want, want_sq = orc.gmp_mul(a, b), orc.gmp_sqr(a)
    pairs = [(a, b), (b[:70_001], a[:90_000]), (a[:5000], b[:3])]
    want_batch = [orc.gmp_mul(x, y) for x, y in pairs]
    for cfg in [(0, 8 << 20, 0), (4, 1 << 20, 0), (4, 8 << 20, 4 << 20), (2, 192 << 10, 0), (8, CHUNK, 1 << 16)]:
        print(f"staging {cfg}", flush=True)
        staging(*cfg)
        assert np.array_equal(ctx.mul_mag_ntt(a, b), want), cfg
        assert np.array_equal(ctx.sqr(a), want_sq), cfg
        for got, w in zip(ctx.mul_batch(pairs), want_batch):
            assert np.array_equal(got, w), cfg
        srv = gnt.Server(ctx)
        try:
            t1 = srv.submit(a, b)
            t2 = srv.submit(a, square=True)
            r1, _ = srv.wait(t1)
            r2, _ = srv.wait(t2)
            assert np.array_equal(r1.download(), want), cfg
            assert np.array_equal(r2.download(), want_sq), cfg
            r1.release()
            r2.release()
        finally:
            srv.stop()


def test_concurrent_copies(ctx, orc, staging):
    staging(4, CHUNK, 0)
    errors = []

    def worker(k):
        try:
            for i in range(6):
                n = 50_000 + 9_973 * k + 1_001 * i
                a = _limbs(orc, 1000 * k + i, n)
                buf = ctx.upload(a)
                got = buf.download()
                buf.release()
                if not np.array_equal(got, a):
                    errors.append((k, i))
        except Exception as e:  # reported below
            errors.append((k, repr(e)))

    ths = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not errors


def test_batch_packed_transfers(ctx, orc, staging):
    """Many uneven pairs and squares in one batch: the packed operand and
    result transfers cross chunk boundaries inside and between pieces."""
    import random
    rnd = random.Random(12)
    pairs, want = [], []
    for k in range(300):
        na = int(2 ** rnd.uniform(0, 16))
        nb = int(2 ** rnd.uniform(0, 16))
        a = _limbs(orc, 3 * k + 1, na)
        if k % 7 == 0:
            pairs.append((a, None))
            want.append(orc.gmp_sqr(a))
        else:
            b = _limbs(orc, 3 * k + 2, nb)
            pairs.append((a, b))
            want.append(orc.gmp_mul(a, b))
    for cfg in [(0, 8 << 20, 0), (4, CHUNK, 0), (3, 192 << 10, 0), (4, 8 << 20, 4 << 20)]:
        print(f"staging {cfg}", flush=True)
        staging(*cfg)
        got = ctx.mul_batch(pairs)
        for i, (g, w) in enumerate(zip(got, want)):
            assert np.array_equal(g, w), (cfg, i)
        bp = ctx.batch_plan(pairs, graph=False)
        try:
            bp.run()
            for i, (g, w) in enumerate(zip(bp.results(), want)):
                assert np.array_equal(g, w), (cfg, "plan", i)
        finally:
            bp.close()
