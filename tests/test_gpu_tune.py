"""Autotuned plans on the device: every split the tuner may choose gives the
exact product, and a tuned plan cache round-trips through gnt_init."""
import itertools
import json

import numpy as np
import pytest

from conftest import PLAN_CACHE, tuned_plans

pytestmark = pytest.mark.gpu


def _splits(log_n):
    out = []
    for np_ in (2, 3, 4):
        for c in itertools.product(range(7, 11), repeat=np_):
            if sum(c) == log_n:
                out.append(list(c))
    return out


@pytest.mark.parametrize("log_n", [14, 19, 21, 24])
def test_every_split_exact(ctx, orc, log_n):
    from pi_amd import gnt
    na = (1 << (log_n - 3)) - 1
    a = orc.splitmix64(log_n, na)
    b = orc.splitmix64(log_n + 100, na - 7)
    want, want_sq = orc.gmp_mul(a, b), orc.gmp_sqr(a)
    try:
        for sp in _splits(log_n):
            print(f"log_n {log_n} split {sp}", flush=True)
            gnt.set_plan(log_n, sp)
            assert gnt.plan(log_n) == (len(sp), sp)
            assert np.array_equal(ctx.mul_mag_ntt(a, b), want), sp
            assert np.array_equal(ctx.sqr(a), want_sq), sp
    finally:
        gnt.load_plans(PLAN_CACHE)


@pytest.mark.parametrize("log_n", range(11, 29))
def test_tuned_lengths_exact(ctx, orc, log_n):
    """Every length of the committed plan cache under its tuned split, as
    bench.py and pi_compute run it: one product with 4(na+nb) - 1 just under
    2^log_n (unbalanced) and one square, exact against GMP."""
    from pi_amd import gnt
    sp = tuned_plans()[log_n]
    assert gnt.plan(log_n) == (len(sp), sp)
    tot = 1 << (log_n - 2)
    na = tot // 2 + tot // 8
    a, b = orc.splitmix64(900 + log_n, na), orc.splitmix64(950 + log_n, tot - na)
    assert gnt.plan_info(na, tot - na) == (log_n, sp)
    print(f"log_n {log_n} split {sp} na {na} nb {tot - na}", flush=True)
    assert np.array_equal(ctx.mul_mag_ntt(a, b), orc.gmp_mul(a, b)), sp
    s = a[: tot // 2]
    assert np.array_equal(ctx.sqr(s), orc.gmp_sqr(s)), sp


def test_tune_writes_loadable_cache(ctx, orc, tmp_path):
    from pi_amd import gnt
    path = tmp_path / "plans.json"
    try:
        gnt.tune(ctx, 12, 18, reps=2, path=str(path))
        cache = json.loads(path.read_text())
        assert set(cache["plans"]) == {str(n) for n in range(12, 19)}
        for n, sp in cache["plans"].items():
            assert gnt.plan(int(n)) == (len(sp), sp)          # installed
            k = ",".join(map(str, sp))
            timed, timed_sq = cache["us"][n], cache["us_sqr"][n]
            assert set(timed) == set(timed_sq) and k in timed
            # the fastest multiply plus square among the timed candidates
            assert timed[k] + timed_sq[k] <= min(timed[x] + timed_sq[x] for x in timed) + 0.02
        for n in range(12, 19):
            gnt.set_plan(n)
        c2 = gnt.Context(0, plan_cache=str(path))             # gnt_init loads it
        try:
            for n, sp in cache["plans"].items():
                assert gnt.plan(int(n)) == (len(sp), sp)
            a = orc.splitmix64(7, 1 << 14)
            assert np.array_equal(c2.mul_mag_ntt(a, a), orc.gmp_sqr(a))
        finally:
            c2.close()
    finally:
        gnt.load_plans(PLAN_CACHE)


@pytest.mark.parametrize("max_log", [-1, 14, 30])
def test_table_and_stepped_twiddles_exact(ctx, orc, max_log):
    """Passes up to 2^max_log take twiddles from tables (the inverse ones applied
    by the store of the stage before), larger ones step them: every mix gives
    the exact product, single and batched."""
    from pi_amd import gnt
    prev = gnt.set_pass_tables(max_log)
    try:
        for la, lb in [(3000, 2000), (1 << 15, (1 << 15) - 9), (1 << 18, 1 << 17), (1 << 21, 1 << 20)]:
            print(f"max_log {max_log} sizes {la} {lb}", flush=True)
            a, b = orc.splitmix64(la, la), orc.splitmix64(lb + 3, lb)
            assert np.array_equal(ctx.mul_mag_ntt(a, b), orc.gmp_mul(a, b)), (la, lb)
            assert np.array_equal(ctx.sqr(a), orc.gmp_sqr(a)), la
        pairs = [(orc.splitmix64(n, n), orc.splitmix64(n + 1, n // 3 + 1)) for n in (700, 9000, 70000, 300000)]
        got = ctx.mul_batch(pairs)
        for (a, b), r in zip(pairs, got):
            assert np.array_equal(r, orc.gmp_mul(a, b))
    finally:
        gnt.set_pass_tables(prev)
