"""The serving loop's device-resident path: requests of at most 2048 result
limbs (products) or sums/differences of operands up to 2047 limbs are claimed
by the persistent kernel's workgroups from a ring in pinned host memory and
computed whole in LDS.  Every result is checked against CPython int (exact),
and against the same request run on the lanes (device loop off)."""
import threading
import time

import numpy as np
import pytest

from pi_amd import gnt

pytestmark = pytest.mark.gpu

MASK = (1 << 64) - 1


def limbs(x):
    from oracle import oracle
    return oracle.from_int(x)


def val(v):
    from oracle import oracle
    return oracle.to_int(v)


def rand_int(rng, n):
    return int.from_bytes(rng.bytes(8 * n), "little") | (1 << (64 * n - 1))


def cases():
    """(op, a, b) with the device loop's edges: one-limb operands, the largest
    product (na + nb = 2048) and sum (2047-limb operands), all-ones limbs
    (maximal carries), unbalanced shapes, a difference to zero and a
    difference whose limb 0 is 2^64 - 1 with a carry out."""
    rng = np.random.default_rng(77)
    ones = lambda n: (1 << (64 * n)) - 1  # noqa: E731
    out = [("mul", 1, 1), ("mul", MASK, MASK), ("sqr", MASK, None), ("mul", 3, rand_int(rng, 7)),
           ("mul", rand_int(rng, 1024), rand_int(rng, 1024)), ("sqr", rand_int(rng, 1024), None),
           ("mul", ones(1024), ones(1024)), ("sqr", ones(1024), None), ("mul", rand_int(rng, 2047), 12345),
           ("mul", rand_int(rng, 1), rand_int(rng, 2047)), ("mul", ones(1000), ones(48)),
           ("add", ones(2047), 1), ("add", rand_int(rng, 2047), rand_int(rng, 2047)), ("add", 5, 7),
           ("sub", 1 << (64 * 2046), 1), ("sub", rand_int(rng, 2048), rand_int(rng, 2000)),
           ("sub", (1 << 128) - 1, 1 << 64), ("sub", 12345, 12345)]
    for i in range(60):
        na = int(2 ** rng.uniform(0, 10))
        nb = int(2 ** rng.uniform(0, 10))
        op = ["mul", "sqr", "add", "sub"][i % 4]
        a, b = rand_int(rng, na), rand_int(rng, nb)
        if op == "sub" and b > a:
            a, b = b, a
        out.append((op, a, None if op == "sqr" else b))
    return out


def want(op, a, b):
    return {"mul": lambda: a * b, "sqr": lambda: a * a, "add": lambda: a + b, "sub": lambda: a - b}[op]()


def run_all(srv, cs):
    ts = [srv.submit(limbs(a), None if b is None else limbs(b), op=op) for op, a, b in cs]
    res = []
    for t in ts:
        buf, _ = srv.wait(t)
        res.append((buf.download(), len(buf)))
        buf.release()
    return res


def test_device_loop_exact_and_same_as_lanes():
    cs = cases()
    c = gnt.Context(0)
    srv = gnt.Server(c)
    try:
        print("device loop on", flush=True)
        got = run_all(srv, cs)
        st = srv.stats()
        assert st["device_loop_requests"] == len(cs), st
        assert st["device_loop_launches"] >= 1 and st["small_lane_iterations"] == 0, st
    finally:
        srv.stop()
        c.close()
    for i, ((op, a, b), (v, n)) in enumerate(zip(cs, got)):
        w = want(op, a, b)
        assert val(v) == w, f"case {i}: {op} {a.bit_length()} {0 if b is None else b.bit_length()}"
        assert n == (w.bit_length() + 63) // 64, f"case {i}: trimmed length"
    c = gnt.Context(0, serve_device_limbs=gnt.SERVE_DEVICE_OFF)
    srv = gnt.Server(c)
    try:
        print("device loop off", flush=True)
        lanes = run_all(srv, cs)
        assert srv.stats()["device_loop_requests"] == 0
    finally:
        srv.stop()
        c.close()
    for (v, n), (u, m) in zip(got, lanes):
        assert np.array_equal(v, u) and n == m


def test_device_loop_chains_with_the_lanes():
    """Results kept on the device as operands: device loop -> device loop,
    device loop -> lane (a product too large for the loop), lane -> lane."""
    rng = np.random.default_rng(8)
    x, y, z = rand_int(rng, 300), rand_int(rng, 200), rand_int(rng, 5000)
    c = gnt.Context(0)
    srv = gnt.Server(c)
    try:
        xy, _ = srv.wait(srv.submit(limbs(x), limbs(y)))           # loop
        xy2, _ = srv.wait(srv.submit(a_buf=xy, square=True))        # loop -> loop (1000 limbs)
        big, _ = srv.wait(srv.submit(b=limbs(z), a_buf=xy2))        # loop -> lane (6000 limbs)
        s1, _ = srv.wait(srv.submit(limbs(z + 1), limbs(z)))        # lane
        s2, _ = srv.wait(srv.submit(limbs(z), limbs(z)))            # lane
        d, _ = srv.wait(srv.submit(a_buf=s1, b_buf=s2, op="sub"))   # lane -> lane
        small, _ = srv.wait(srv.submit(limbs(x), limbs(7)))         # loop
        t, _ = srv.wait(srv.submit(a_buf=small, b_buf=xy, op="add"))  # loop -> loop
        assert val(xy2.download()) == (x * y) ** 2
        assert val(big.download()) == (x * y) ** 2 * z
        assert val(d.download()) == z
        assert val(t.download()) == 7 * x + x * y
        st = srv.stats()
        assert st["device_loop_requests"] == 4 and st["iterations"] > 0, st
        for b in (xy, xy2, big, s1, s2, d, small, t):
            b.release()
    finally:
        srv.stop()
        c.close()


def test_device_loop_relaunches_after_idle():
    """The persistent kernel leaves after 100 us without requests; requests
    published later start a new launch and still complete."""
    rng = np.random.default_rng(9)
    c = gnt.Context(0)
    srv = gnt.Server(c)
    try:
        for k in range(4):
            a, b = rand_int(rng, 100 + k), rand_int(rng, 50)
            buf, lat = srv.wait(srv.submit(limbs(a), limbs(b)))
            assert val(buf.download()) == a * b
            buf.release()
            time.sleep(0.02)
        st = srv.stats()
        assert st["device_loop_requests"] == 4 and st["device_loop_launches"] >= 2, st
    finally:
        srv.stop()
        c.close()


def test_device_loop_many_submitters():
    """Eight host threads submitting small requests at once (more than the
    128-slot ring holds: the overflow goes to the lanes), every result exact."""
    rng = np.random.default_rng(10)
    jobs = [(rand_int(rng, int(2 ** rng.uniform(0, 9))), rand_int(rng, int(2 ** rng.uniform(0, 9))))
            for _ in range(400)]
    out = [None] * len(jobs)
    c = gnt.Context(0)
    srv = gnt.Server(c)
    try:
        def work(k):
            ts = [(i, srv.submit(limbs(jobs[i][0]), limbs(jobs[i][1]))) for i in range(k, len(jobs), 8)]
            for i, t in ts:
                buf, _ = srv.wait(t)
                out[i] = buf.download()
                buf.release()

        th = [threading.Thread(target=work, args=(k,)) for k in range(8)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        st = srv.stats()
        assert st["device_loop_requests"] > 0, st
    finally:
        srv.stop()
        c.close()
    for (a, b), v in zip(jobs, out):
        assert val(v) == a * b


@pytest.mark.parametrize("graph", [True, False])
def test_batch_of_small_products_is_one_persistent_launch(graph):
    """A batch plan whose products all fit one workgroup (at most 2048 result
    limbs) runs as one persistent launch (and the kernel that zeroes its claim
    counter); replayed three times with new operands and eager work between
    the replays, every replay exact.  A plan mixing small and large products
    takes the persistent launch beside the level-grouped ones."""
    from oracle import oracle as orc
    rng = np.random.default_rng(12)
    shapes = [(1, 1), (1, 2047), (1024, 1024), (700, None), (1024, None), (3, 5)]
    shapes += [(int(2 ** rng.uniform(0, 10)), None if i % 3 == 0 else int(2 ** rng.uniform(0, 10)))
               for i in range(90)]

    def draw(seed):
        return [(orc.splitmix64(seed + 2 * i, na), None if nb is None else orc.splitmix64(seed + 2 * i + 1, nb))
                for i, (na, nb) in enumerate(shapes)]

    c = gnt.Context(0)
    try:
        pairs = draw(1000)
        plan = c.batch_plan(pairs, graph=graph)
        assert plan.info()["launches_per_run"] == 2, plan.info()
        for rep in range(3):
            print(f"replay {rep}", flush=True)
            if rep:
                pairs = draw(1000 + 500 * rep)
                jobs, keep, outs = c._jobs(pairs)
                gnt._check(gnt.load().gnt_batch_upload(plan.h, jobs))
            plan.run()
            got = plan.results()
            x, y = orc.splitmix64(rep, 3000), orc.splitmix64(rep + 9, 2000)
            assert np.array_equal(c.mul_mag_ntt(x, y), orc.gmp_mul(x, y))   # eager work between replays
            for (a, b), g in zip(pairs, got):
                assert np.array_equal(g, orc.gmp_sqr(a) if b is None else orc.gmp_mul(a, b))
        plan.close()
        mixed = draw(7) + [(orc.splitmix64(5, 3000), orc.splitmix64(6, 5000)), (orc.splitmix64(8, 40000), None)]
        plan = c.batch_plan(mixed, graph=graph)
        plan.run()
        for (a, b), g in zip(mixed, plan.results()):
            assert np.array_equal(g, orc.gmp_sqr(a) if b is None else orc.gmp_mul(a, b))
        assert plan.info()["launches_per_run"] > 2
        plan.close()
    finally:
        c.close()
