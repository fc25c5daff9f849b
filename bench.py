"""Benchmark: Goldilocks-NTT multiply throughput on one MI355X.

Workload (default): one 2^24 x 2^24-limb product per step (uniform random
64-bit limbs, SplitMix64 seeds 3/4, top limb nonzero), transform length
N = 2^27, operands resident in HBM when the timed region starts.  A step is
one full product: pack + two forward transforms + pointwise + inverse + carry.

Prints ONE JSON line.  `value` is Mlimbs/s = (na + nb) / t / 1e6.
`roofline.achieved` is the algorithmic byte count of a product,
B_mul = 16 (na + nb) + 80 N, divided by the measured product time (HIP events
on the library's stream), against the 8 TB/s HBM peak; `roofline.valu` puts
the same launch against VALU instruction issue, the bound the transform
kernels actually meet (instruction counts from rocprofv3).  `cpu_baseline` times
the CPU restatement of the reference multiply on a bounded sample of the same
workload on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ROOT = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PLANS = os.path.join(ROOT, "pi_amd", "plans_gfx950.json")


def make_ctx(args, **kw):
    """Context with the tuned plan cache when one is present."""
    from pi_amd import gnt
    return gnt.Context(0, plan_cache=args.plans or None, **kw)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """SplitMix64 stream (vectorised numpy; same sequence as the oracle's)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def operand(seed: int, n: int) -> np.ndarray:
    x = splitmix64(seed, n)
    if x[-1] == 0:
        x[-1] = 1
    return x


def operand_dist(dist: str, seed: int, n: int) -> np.ndarray:
    """uniform limbs, all-ones limbs (the longest carry chains), or sparse:
    each limb nonzero with probability 1/100 (decided by one stream, values
    from another), top limb nonzero."""
    if dist == "uniform":
        return operand(seed, n)
    if dist == "ones":
        return np.full(n, np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
    pick = splitmix64(seed, n) % np.uint64(100) == 0
    x = np.where(pick, splitmix64(seed + 1000, n), np.uint64(0)).astype(np.uint64)
    if x[-1] == 0:
        x[-1] = 1
    return x


# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (KiB per dispatch; FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0, the
# factors a plain 1 GiB device copy reads), keyed by (op, log2 N, radices).
# VALU wave instructions per launch of the dominant (fused) kernel, measured
# with rocprofv3 --pmc SQ_INSTS_VALU, keyed like HBM_TRAFFIC.  The secondary
# bound: 1024 SIMDs issue one wave64 instruction each per 4 cycles at 2.4 GHz
# for everything but plain 32-bit add/sub/xor/mov (2 cycles).
# Both tables hold counter readings of the kernels built from the sources
# whose digest is PROFILED_SOURCES (kernel_sources_digest()); with any other
# kernel source the line reports them as null rather than stale figures.
PROFILED_SOURCES = "4848b4a455e4e859"
VALU_INSTS = {
    ("mul", 27, (8, 9, 10)): 821559296,
    ("sqr", 27, (8, 9, 10)): 647495680,
}
VALU_ISSUE_PEAK = 1024 * 2.4e9 / 4  # wave64 instructions per second at 4 cycles each

HBM_TRAFFIC = {
    ("mul", 27, (8, 9, 10)): {"fused": 3227483886, "product": 14614631803},
}

KERNEL_SOURCES = ("gl_field.hpp", "gnt_plan.hpp", "gnt_pass.hpp", "gnt_carry.hpp", "gnt_kernels.hip",
                  "gnt_kernels.hpp", "gnt_api.hip")


def kernel_sources_digest() -> str:
    """First 16 hex digits of the SHA-256 over the sources that define the
    single-product launches (field arithmetic, plans, tile bodies, carry,
    kernels, their host launch sequence)."""
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "pi_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


# Three primes below 2^31 (2^31 - 1 and the next two): a wrong product passes
# all three residue checks with probability about 2^-93.
FP_PRIMES = (2147483647, 2147483629, 2147483587)
_POW32 = {}  # q -> [2^(32k) mod q for k < len]


def _pow32(q: int, n: int) -> np.ndarray:
    """2^(32k) mod q for k < n (extended by doubling: w[m:2m] = w[:m] * 2^(32m))."""
    w = _POW32.get(q)
    if w is None or w.size < n:
        m = 1 if w is None else w.size
        if w is None:
            w = np.ones(1, dtype=np.uint64)
        while m < n:
            step = np.uint64(pow(2, 32 * m, q))
            w = np.concatenate([w, (w * step) % np.uint64(q)])  # < 2^62: exact
            m *= 2
        _POW32[q] = w
    return w[:n]


_CHUNK = 1 << 15


_SEG = 1 << 20
_POOL = None


def _segment_residue(h32, w) -> int:
    """sum h_k w_k mod q over one segment (q is folded in by the caller): w_k
    is split at bit 16 so that h_k * (w_k mod 2^16) < 2^48 and
    h_k * (w_k >> 16) < 2^47, and each partial sum runs over at most 2^15
    terms (< 2^63)."""
    h = h32.astype(np.uint64)
    pad = (-h.size) % _CHUNK
    if pad:
        h = np.concatenate([h, np.zeros(pad, dtype=np.uint64)])
        w = np.concatenate([w, np.zeros(pad, dtype=np.uint64)])
    lo = (h * (w & np.uint64(0xFFFF))).reshape(-1, _CHUNK).sum(axis=1, dtype=np.uint64)
    hi = (h * (w >> np.uint64(16))).reshape(-1, _CHUNK).sum(axis=1, dtype=np.uint64)
    return sum(int(v) for v in lo) + (sum(int(v) for v in hi) << 16)


def residues(x, primes=FP_PRIMES):
    """The integer with little-endian limbs x, mod each prime: its 32-bit
    halves h_k weighted by w_k = 2^(32k) mod q, in segments on host threads."""
    global _POOL
    from concurrent.futures import ThreadPoolExecutor
    if _POOL is None:
        _POOL = ThreadPoolExecutor(max_workers=host_threads())
    h = np.ascontiguousarray(x, dtype="<u8").view("<u4")
    out = []
    for q in primes:
        w = _pow32(q, h.size)
        parts = _POOL.map(lambda k: _segment_residue(h[k:k + _SEG], w[k:k + _SEG]), range(0, h.size, _SEG))
        out.append(sum(parts) % q)
    return tuple(out)


def residues_many(arrays, primes=FP_PRIMES):
    """residues() of many arrays, one array per host thread."""
    residues(np.zeros(1, dtype=np.uint64))  # starts the pool
    big = max((x.size for x in arrays), default=0)
    for q in primes:
        _pow32(q, 2 * big)

    def one(x):
        h = np.ascontiguousarray(x, dtype="<u8").view("<u4")
        return tuple(sum(_segment_residue(h[k:k + _SEG], _pow32(q, h.size)[k:k + _SEG])
                         for k in range(0, h.size, _SEG)) % q for q in primes)

    return list(_POOL.map(one, arrays))


def residues_match(ra, rb, rc, primes=FP_PRIMES) -> bool:
    return all(z == x * y % q for x, y, z, q in zip(ra, rb, rc, primes))


def product_ok(c, a, b) -> bool:
    """c == a * b by residues mod three primes (independent of the
    transform), plus the low limb and the length bounds."""
    ra, rb, rc = residues(a), residues(b), residues(c)
    ok = all(z == x * y % q for x, y, z, q in zip(ra, rb, rc, FP_PRIMES))
    ok &= c.size in (a.size + b.size, a.size + b.size - 1)
    return bool(ok and int(c[0]) == (int(a[0]) * int(b[0])) % (1 << 64))


_GMP = None


def gmp_product(a, b):
    """GMP's mpn_mul (libgmp's public API), or None where libgmp is absent."""
    global _GMP
    import ctypes
    if _GMP is None:
        try:
            _GMP = ctypes.CDLL("libgmp.so.10")
        except OSError:
            _GMP = False
    if not _GMP:
        return None
    if a.size < b.size:
        a, b = b, a
    a, b = np.ascontiguousarray(a, dtype=np.uint64), np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros(a.size + b.size, dtype=np.uint64)
    P = ctypes.c_void_p
    _GMP.__gmpn_mul(P(out.ctypes.data), P(a.ctypes.data), ctypes.c_long(a.size), P(b.ctypes.data),
                    ctypes.c_long(b.size))
    n = out.size
    while n and out[n - 1] == 0:
        n -= 1
    return out[:n]


def valu_block(insts, ms):
    """The launch against VALU issue: measured wave instructions / time, over the
    chip's issue rate at 4 cycles per instruction (the kernels' real bound)."""
    if not insts:
        return None
    rate = insts / (ms * 1e-3)
    return {"wave_instructions_per_launch": insts, "achieved": round(rate / 1e9, 1),
            "peak": round(VALU_ISSUE_PEAK / 1e9, 1), "unit": "G wave64 instructions/s (4-cycle issue)",
            "frac": round(rate / VALU_ISSUE_PEAK, 4)}


def b_mul(na: int, nb: int, log_n: int) -> float:
    return 16.0 * (na + nb) + 80.0 * (1 << log_n)


def b_sqr(n: int, log_n: int) -> float:
    return 24.0 * n + 48.0 * (1 << log_n)


def cpu_baseline(log_size: int, op: str, threads: int, timed: int = 3, budget_s: float = 40.0) -> dict:
    """The oracle (CPU restatement of ntt.rs::mul_mag_ntt, with the host knobs
    config_gen.rs:56-99 picks for `threads` cores) on the metric's own
    operands: one warm-up product, then the median of `timed` products (fewer
    if the time budget runs out first)."""
    from oracle import oracle as orc

    orc.set_threads(threads)
    orc.apply_host_knobs(threads)
    four_step = orc.get_four_step_threshold()   # the threshold the timed runs use
    n = 1 << log_size
    a, b = operand(3, n), operand(4, n)
    if op == "sqr":
        b = a
    t_start = time.perf_counter()
    orc.mul_mag_ntt(a, b)                        # warm-up: pool buffers and twiddle tables
    times = []
    while len(times) < timed and (time.perf_counter() - t_start) < budget_s:
        t0 = time.perf_counter()
        orc.mul_mag_ntt(a, b)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    # one thread on a 2^(log_size-4) sample, and GMP's mpn_mul (one thread) beside it
    n1 = n >> 4
    a1, b1 = a[:n1].copy(), (a if op == "sqr" else b)[:n1].copy()
    orc.set_threads(1)
    orc.apply_host_knobs(1)
    orc.mul_mag_ntt(a1, b1)
    t0 = time.perf_counter()
    orc.mul_mag_ntt(a1, b1)
    t1 = time.perf_counter() - t0
    orc.set_threads(threads)
    orc.apply_host_knobs(threads)
    gmp = None
    if orc.gmp() is not None:
        t0 = time.perf_counter()
        orc.gmp_mul(a, b) if op == "mul" else orc.gmp_sqr(a)
        gmp = round(2 * n / (time.perf_counter() - t0) / 1e6, 3)
    return {
        "value": round(2 * n / t / 1e6, 3),
        "unit": "Mlimbs/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{op} 2^{log_size} x 2^{log_size} limbs (the metric's operands): 1 warm-up, median of "
                  f"{len(times)} timed ({', '.join(f'{x:.2f}' for x in times)} s); four-step threshold "
                  f"{four_step} (host knobs for {threads} cores)",
        "cpu_model": cpu_model(),
        "one_thread": {"value": round(2 * n1 / t1 / 1e6, 3),
                       "sample": f"1 x {op} 2^{log_size - 4} limbs after 1 warm-up"},
        "gmp_mpn_1thread": {"value": gmp, "sample": f"1 x {op} 2^{log_size} limbs"} if gmp else None,
    }


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log-size", type=int, default=24, help="log2 limbs per operand")
    ap.add_argument("--op", choices=["mul", "sqr"], default="mul")
    ap.add_argument("--cpu-log-size", type=int, default=24, help="log2 limbs of the CPU-baseline sample")
    ap.add_argument("--cpu-timed", type=int, default=3, help="timed CPU-baseline products after the warm-up")
    ap.add_argument("--no-batched", action="store_true", help="single: skip the 1024-pair batched run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["single", "batched", "serve", "chudnovsky", "c2", "copies", "small"],
                    default="single",
                    help="single: the metric (one 2^24 product per step); batched: 1024 uneven pairs per step; "
                         "serve: open-loop Poisson request stream through the serving loop; "
                         "chudnovsky: pi to --digits digits end to end; "
                         "c2: mul and sqr at 2^20 and 2^24 limbs over three operand distributions; "
                         "copies: host <-> device copy rates and the PCIe-inclusive host-buffer multiply; "
                         "small: closed-loop latency of small served products, device-resident loop against the lanes")
    ap.add_argument("--rates", default="100,200,400,800,1600", help="serve: arrival rates (requests/s) to sweep")
    ap.add_argument("--requests", type=int, default=300, help="serve: requests per rate")
    ap.add_argument("--pairs", type=int, default=1024)
    ap.add_argument("--plans", default=DEFAULT_PLANS if os.path.exists(DEFAULT_PLANS) else "",
                    help="plan cache written by the autotuner ('' = built-in plans)")
    ap.add_argument("--tune", default="", help="LO:HI: autotune transform lengths 2^LO..2^HI, write --tune-out, exit")
    ap.add_argument("--tune-out", default=DEFAULT_PLANS)
    ap.add_argument("--digits", type=int, default=100_000_000, help="chudnovsky: decimal digits")
    ap.add_argument("--latency-rate", type=float, default=400.0,
                    help="single: arrival rate (requests/s) of the serving run whose p50/p99 the line reports")
    ap.add_argument("--latency-requests", type=int, default=1000)
    ap.add_argument("--no-latency", action="store_true", help="single: skip the serving latency run")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
              "independent products shard across GPUs with no exchange, see the design notes.", file=sys.stderr)
        sys.exit(2)

    from pi_amd import gnt

    if args.tune:
        lo, hi = (int(x) for x in args.tune.split(":"))
        ctx = gnt.Context(0)  # tuned from the built-in plans
        t0 = time.perf_counter()
        gnt.tune(ctx, lo, hi, reps=5, path=args.tune_out)
        ctx.close()
        print(json.dumps({"tuned": [lo, hi], "path": args.tune_out, "s": round(time.perf_counter() - t0, 2)}))
        return
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "chudnovsky":
        return bench_chudnovsky(args)
    if args.workload == "c2":
        return bench_c2(args)
    if args.workload == "copies":
        return bench_copies(args)
    if args.workload == "small":
        return bench_small(args)

    n = 1 << args.log_size
    a, b = operand(3, n), operand(4, n)
    ctx = make_ctx(args)
    log_n, radices = gnt.plan_info(n, n)
    da = ctx.upload(a)
    db = ctx.upload(b)

    def step():
        r = ctx.sqr_buf(da) if args.op == "sqr" else ctx.mul_buf(da, db)
        r.release()  # stream-ordered: the buffer is reused by the next step

    for _ in range(args.warmup):
        step()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3

    # device time per product with HIP events on the library's stream
    ctx.set_timing(True)
    for _ in range(max(3, min(args.steps, 10))):
        step()
    ctx.sync()
    phases, prods = ctx.get_timing()
    ctx.set_timing(False)
    t_prod_ms = phases[4] / prods
    limbs = 2 * n
    value = limbs / (ms_per_step * 1e-3) / 1e6
    bytes_alg = b_sqr(n, log_n) if args.op == "sqr" else b_mul(n, n, log_n)
    achieved = bytes_alg / (t_prod_ms * 1e-3) / 1e9
    # the dominant single launch: the fused kernel reads A's and B's arrays and
    # writes A's (a square reads and writes one array): 24N (16N) bytes
    fused_ms = phases[1] / prods
    fused_bytes = (16 if args.op == "sqr" else 24) * (1 << log_n)
    fused_gbs = fused_bytes / (fused_ms * 1e-3) / 1e9
    digest = kernel_sources_digest()
    fresh = digest == PROFILED_SOURCES
    key = (args.op, log_n, tuple(radices))
    measured = HBM_TRAFFIC.get(key) if fresh else None
    fused_traffic = measured["fused"] if measured else None
    product_traffic = measured["product"] if measured else None

    # the timed path's result, whole: residues mod three primes, low limb, length
    r = ctx.sqr_buf(da) if args.op == "sqr" else ctx.mul_buf(da, db)
    out = r.download()
    r.release()
    ok = product_ok(out, a, a if args.op == "sqr" else b)
    del out

    line = {
        "metric": "Goldilocks-NTT multiply Mlimbs/s at 2^24 limbs; batched p50/p99 latency",
        "value": round(value, 3),
        "unit": "Mlimbs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (SplitMix64 uniform limbs, seeds 3/4)",
        "config": {
            "workload": f"single GPU {args.op} of 2^{args.log_size} x 2^{args.log_size} limbs",
            "limbs_a": n, "limbs_b": n, "log_n": log_n, "radices_log2": radices,
            "inputs": "resident in HBM", "result_exact_by_residues": ok,
            "plans": os.path.relpath(args.plans, ROOT) if args.plans else "built-in",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": ("k_last_fused_ab (A's and B's last forward passes + pointwise + first inverse pass)"
                       if args.op == "mul" else "k_last_fused (last forward pass + pointwise square + first inverse pass)"),
            "achieved": round(fused_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(fused_gbs / HBM_PEAK_GBS, 4),
            "traffic": fused_traffic,
            "algorithmic_bytes_per_launch": fused_bytes,
            "ms_per_launch": round(fused_ms, 4),
            "valu": valu_block(VALU_INSTS.get(key) if fresh else None, fused_ms),
            "counters_from": {"kernel_sources": digest, "profiled_sources": PROFILED_SOURCES,
                              "current": fresh, "profiles": "profiles/mul_2p24_hbm_traffic.md, "
                                                            "profiles/transform_valu_2p24.md"},
            "whole_product": {
                "achieved": round(achieved, 2), "frac": round(achieved / HBM_PEAK_GBS, 4),
                "bytes_model": "B_mul = 16(na+nb)+80N" if args.op == "mul" else "B_sqr = 24n+48N",
                "traffic": product_traffic,
                "device_ms_per_product": round(t_prod_ms, 4),
                "phase_ms_per_product": {k: round(v / prods, 4) for k, v in
                                         zip(["forward", "last_fused", "inverse", "carry"], phases[:4])},
            },
        },
    }
    ctx.close()
    if not args.no_batched:
        # configs' batched case beside the single product: 1024 uneven pairs
        # as one captured plan, against one launch sequence per product
        line["batched"] = batched_run(args, graph_steps=max(5, args.steps // 2), serial_steps=3)
    if not args.no_latency:
        # the metric's second half: p50/p99 latency of mixed-size requests
        # through the continuously batching serving loop
        sctx, srv, src = serve_open(args)
        pt = serve_rate(srv, src, np.random.default_rng(8), args.latency_rate, args.latency_requests)
        srv.stop()
        sctx.close()
        line["latency"] = {"p50_ms": pt["p50_ms"], "p99_ms": pt["p99_ms"], "unit": "ms",
                           "workload": SERVE_WORKLOAD + f"; {args.latency_rate:g} req/s offered",
                           "measured_from": "scheduled arrival, operands in host memory (upload included), "
                                            "to result in HBM",
                           "run": pt}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(min(args.cpu_log_size, args.log_size), args.op, host_threads(),
                                            timed=args.cpu_timed)
    print(json.dumps(line), flush=True)


def batched_pairs(n_pairs: int, seed: int = 7):
    """1024 pairs, n = round(2^U[8,20]) limbs (same n for both operands), uniform limbs."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(2.0 ** rng.uniform(8, 20, size=n_pairs)).astype(np.int64)
    pairs = []
    for i, m in enumerate(sizes):
        a, b = operand(seed * 100000 + 2 * i, int(m)), operand(seed * 100000 + 2 * i + 1, int(m))
        pairs.append((a, b))
    return pairs


def batched_modes(ctx, pairs, modes, steps: dict, warmup: int):
    """Time each (label, graph, serial) mode of one batch plan over `pairs`
    and check its results: every pair by residues mod three primes (with the
    low limb and length), and every 128th pair limb for limb against GMP."""
    limbs = sum(a.size + b.size for a, b in pairs)
    ra = residues_many([a for a, _ in pairs])
    rb = residues_many([b for _, b in pairs])
    res = {}
    for label, graph, serial in modes:
        plan = ctx.batch_plan(pairs, graph=graph)
        for _ in range(warmup):
            plan.run(serial=serial)
        ctx.sync()
        k = steps[label]
        t0 = time.perf_counter()
        for _ in range(k):
            plan.run(serial=serial)
        ctx.sync()
        dt = (time.perf_counter() - t0) / k
        info = plan.info()
        outs = plan.results()
        plan.close()
        rc = residues_many(outs)
        ok = all(residues_match(x, y, z) and c.size in (a.size + b.size, a.size + b.size - 1)
                 and int(c[0]) == (int(a[0]) * int(b[0])) % (1 << 64)
                 for (a, b), c, x, y, z in zip(pairs, outs, ra, rb, rc))
        gmp_ok = None
        for i in range(0, len(pairs), 128):
            want = gmp_product(*pairs[i])
            if want is None:
                break
            gmp_ok = (gmp_ok is not False) and bool(np.array_equal(outs[i], want))
        del outs
        res[label] = {"ms": round(dt * 1e3, 3), "Mlimbs_per_s": round(limbs / dt / 1e6, 1), "runs": k,
                      "launches_per_run": info["launches_per_run"] if not serial else None,
                      "all_pairs_exact_by_residues": bool(ok), "sampled_pairs_equal_gmp": gmp_ok}
    return res


def batched_run(args, graph_steps: int, serial_steps: int) -> dict:
    """The batched key of the default line: 1024 uneven pairs (n = round(2^U[8,20])
    limbs, seed 7) as one captured plan and one launch sequence per product."""
    pairs = batched_pairs(args.pairs)
    limbs = sum(a.size + b.size for a, b in pairs)
    ctx = make_ctx(args)
    res = batched_modes(ctx, pairs, (("batched_graph", True, False), ("per_product", False, True)),
                        {"batched_graph": graph_steps, "per_product": serial_steps}, warmup=2)
    ctx.close()
    return {"workload": f"{len(pairs)} pairs, n = round(2^U[8,20]) limbs per operand (seed 7), uniform limbs, "
                        "operands resident in HBM",
            "total_operand_limbs": limbs, "unit": "Mlimbs/s",
            "value": res["batched_graph"]["Mlimbs_per_s"],
            "batched_over_per_product": round(res["batched_graph"]["Mlimbs_per_s"]
                                              / res["per_product"]["Mlimbs_per_s"], 3),
            "modes": res}


def bench_batched(args):
    """Config: batched uneven products.  Batched plan (graph replay, and direct
    level launches) against one launch sequence per product, same buffers."""
    pairs = batched_pairs(args.pairs)
    ctx = make_ctx(args)
    res = batched_modes(ctx, pairs, (("batched_graph", True, False), ("batched_launches", False, False),
                                     ("per_product", False, True)),
                        {"batched_graph": args.steps, "batched_launches": args.steps, "per_product": args.steps},
                        warmup=args.warmup)
    limbs = sum(a.size + b.size for a, b in pairs)
    line = {
        "metric": "batched uneven products Mlimbs/s",
        "value": res["batched_graph"]["Mlimbs_per_s"], "unit": "Mlimbs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["batched_graph"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (SplitMix64 uniform limbs, sizes seed 7)",
        "config": {"workload": f"{len(pairs)} pairs, n = round(2^U[8,20]) limbs per operand",
                   "plans": os.path.relpath(args.plans, ROOT) if args.plans else "built-in",
                   "total_operand_limbs": limbs, "modes": res},
    }
    ctx.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_batched(pairs, host_threads())
    print(json.dumps(line), flush=True)


def host_threads() -> int:
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return min(threads, int(os.environ.get("OMP_NUM_THREADS", threads)))


def cpu_baseline_batched(pairs, threads: int, sample: int = 128) -> dict:
    """The oracle's mul_mag routing (schoolbook / Karatsuba / NTT as
    integer.rs:460-486 picks them) on the first `sample` pairs, run
    concurrently on a pool of host threads as binary splitting's sibling
    subtrees run under rayon."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as orc

    orc.set_threads(1)
    orc.apply_host_knobs(threads)
    sub = pairs[:sample]
    limbs = sum(a.size + b.size for a, b in sub)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(lambda p: orc.mul_mag(p[0], p[1]), sub))
    t = time.perf_counter() - t0
    orc.set_threads(threads)
    return {"value": round(limbs / t / 1e6, 3), "unit": "Mlimbs/s", "cores": threads, "kind": "port",
            "sample": f"first {len(sub)} of the {len(pairs)} pairs ({limbs} operand limbs), one pair per thread",
            "cpu_model": cpu_model()}


SERVE_POOL_LIMBS = (1 << 25) + (1 << 24)
Q61 = (1 << 61) - 1


def _residue(x) -> int:
    return int.from_bytes(np.ascontiguousarray(x, dtype="<u8").tobytes(), "little") % Q61


def serve_open(args):
    """A serving loop with the HBM admission cap, its operand pool (slices of
    one random buffer, seed 8) and every size class allocated once."""
    from pi_amd import gnt

    src = operand(8, SERVE_POOL_LIMBS)
    ctx = make_ctx(args, hbm_cap_bytes=200 * 10**9)
    srv = gnt.Server(ctx)
    for e in range(8, 25, 2):  # warm-up: first-use allocations of every size class
        n = 1 << e
        t = srv.submit(src[:n], src[n:2 * n])
        buf, _ = srv.wait(t)
        buf.release()
    return ctx, srv, src


def _latency_by_size(lat_ms, sizes) -> dict:
    """p50/p99 (ms, from scheduled arrival) per band of the larger operand's
    limbs: 2^[8,12), 2^[12,16), 2^[16,20), 2^[20,25)."""
    out = {}
    for lo, hi in ((8, 12), (12, 16), (16, 20), (20, 25)):
        sel = lat_ms[(sizes >= (1 << lo)) & (sizes < (1 << hi))]
        if sel.size:
            out[f"2^{lo}-2^{min(hi, 24)}"] = {"requests": int(sel.size), "p50_ms": round(float(np.percentile(sel, 50)), 3),
                                                "p99_ms": round(float(np.percentile(sel, 99)), 3)}
    return out


def serve_rate(srv, src, rng, rate: float, requests: int) -> dict:
    """One open-loop Poisson run at `rate` requests/s: n = round(2^U[8,24])
    limbs, 70% multiply / 30% square, 20% of requests take a finished result
    that is still on the device as operand.  p50/p99 run from each request's
    scheduled arrival to its completion (the library's submit-to-completion
    time plus any delay of the submit call itself, so a stalled submitter does
    not hide queueing); the library's own submit-to-completion figures are
    reported beside them.  Every 25th request with fresh operands is checked
    exactly (residue mod 2^61-1) after the run."""
    import threading

    # the request schedule, drawn before the clock starts
    sched = []
    gap = rng.exponential(1.0 / rate, size=requests)
    for i in range(requests):
        n = int(round(2 ** rng.uniform(8, 24)))
        sched.append(dict(at=float(gap[:i + 1].sum()), n=n, sq=bool(rng.uniform() < 0.3),
                          reuse=bool(rng.uniform() < 0.2), keep=bool(rng.uniform() < 0.5),
                          oa=int(rng.integers(0, SERVE_POOL_LIMBS - n)),
                          ob=int(rng.integers(0, SERVE_POOL_LIMBS - n))))
    lock = threading.Lock()
    ready = []                      # finished results kept on the device for reuse
    consumed, sampled = [], []
    lat, lat_open, limbs_done = [], [], [0]
    tickets, meta, late, size_of, sizes, small = [], {}, {}, {}, [], []
    done_evt = threading.Event()

    def waiter():
        for k in range(requests):
            while k >= len(tickets):
                time.sleep(0.0001)
            t = tickets[k]
            buf, us = srv.wait(t)
            lat.append(us)
            lat_open.append(us + late[t])
            sizes.append(size_of[t])
            n_in, check, keep = meta[t]
            small.append(n_in <= DEVICE_LOOP_LIMBS)
            limbs_done[0] += n_in
            if check is not None:
                sampled.append((buf, check))
            elif keep and n_in <= (1 << 24):
                with lock:
                    ready.append((buf, n_in))
                    while len(ready) > 8:
                        ready.pop(0)[0].release()
            else:
                buf.release()
        done_evt.set()

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    for i, r in enumerate(sched):
        wait_s = t0 + r["at"] - time.perf_counter()
        if wait_s > 0:
            time.sleep(wait_s)
        n = r["n"]
        a = src[r["oa"]:r["oa"] + n]
        b = src[r["ob"]:r["ob"] + n]
        a_buf = None
        if r["reuse"]:
            with lock:
                if ready:
                    a_buf, n_a = ready.pop(int(rng.integers(0, len(ready))))
        if a_buf is None:
            n_a = n
        t_call = time.perf_counter()
        check = None
        if a_buf is None and i % 25 == 0:
            check = ("sq", r["oa"], n) if r["sq"] else ("mul", r["oa"], r["ob"], n)
        if r["sq"]:
            t = srv.submit(None if a_buf is not None else a, square=True, a_buf=a_buf)
            n_in = 2 * n_a
        else:
            t = srv.submit(None if a_buf is not None else a, b, a_buf=a_buf)
            n_in = n_a + n
        meta[t] = (n_in, check, r["keep"])
        # the band of a request is its larger operand: a reused device result
        # can be far larger than the drawn size n
        size_of[t] = n_a if r["sq"] else max(n_a, n)
        # time between the scheduled arrival and the submit call, so that a
        # submit delayed by an earlier one still counts from its arrival
        late[t] = max(0.0, t_call - (t0 + r["at"])) * 1e6
        if a_buf is not None:
            consumed.append(a_buf)
        tickets.append(t)
    done_evt.wait()
    th.join()
    wall = time.perf_counter() - t0
    # exactness of the sampled requests, off the clock
    ok = 0
    for buf, chk in sampled:
        c = _residue(buf.download())
        if chk[0] == "sq":
            want = _residue(src[chk[1]:chk[1] + chk[2]]) ** 2 % Q61
        else:
            want = _residue(src[chk[1]:chk[1] + chk[3]]) * _residue(src[chk[2]:chk[2] + chk[3]]) % Q61
        ok += int(c == want)
        buf.release()
    for buf, _ in ready:
        buf.release()
    for buf in consumed:
        buf.release()
    lat_ms = np.array(lat_open) / 1e3
    sub_ms = np.array(lat) / 1e3
    sm = np.array(small, dtype=bool)
    loop = {"requests": int(sm.sum())}
    if sm.any():
        loop.update(p50_ms=round(float(np.percentile(lat_ms[sm], 50)), 3),
                    p99_ms=round(float(np.percentile(lat_ms[sm], 99)), 3))
    return {"rate_req_per_s": rate, "requests": requests,
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
            "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
            "submit_to_done_p50_ms": round(float(np.percentile(sub_ms, 50)), 3),
            "submit_to_done_p99_ms": round(float(np.percentile(sub_ms, 99)), 3),
            "by_size": _latency_by_size(lat_ms, np.array(sizes)),
            # requests whose result has at most 2048 limbs (the device-resident loop's)
            "result_le_2048_limbs": loop,
            "throughput_Mlimbs_per_s": round(limbs_done[0] / wall / 1e6, 1),
            "achieved_req_per_s": round(requests / wall, 1),
            "exact_checks_passed": f"{ok}/{len(sampled)}"}


# Largest result (limbs) of the serving loop's device-resident path (the
# library default of gnt_config.serve_device_limbs).
DEVICE_LOOP_LIMBS = 2048

SERVE_WORKLOAD = ("open-loop Poisson; n = round(2^U[8,24]); 70/30 mul/sqr; 20% reuse of device results; "
                  "HBM cap 200 GB")


def bench_serve(args):
    """Config: serving.  The arrival-rate sweep of serve_rate through one
    serving loop; `value` is the best throughput over the sweep."""
    rng = np.random.default_rng(8)
    ctx, srv, src = serve_open(args)
    points = []
    for rate in [float(r) for r in args.rates.split(",")]:
        points.append(serve_rate(srv, src, rng, rate, args.requests))
        print(json.dumps({"serve_point": points[-1]}), file=sys.stderr, flush=True)
    stats = srv.stats()
    srv.stop()
    ctx.close()
    best = max(points, key=lambda p: p["throughput_Mlimbs_per_s"])
    line = {
        "metric": "served multiply/square p50/p99 latency and throughput",
        "value": best["throughput_Mlimbs_per_s"], "unit": "Mlimbs/s", "n_gpus": 1,
        "steps": args.requests, "warmup": 9, "ms_per_step": best["p50_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (SplitMix64 limbs, Poisson arrivals, seed 8)",
        "config": {"plans": os.path.relpath(args.plans, ROOT) if args.plans else "built-in",
                   "workload": SERVE_WORKLOAD, "sweep": points, "server": stats},
    }
    print(json.dumps(line), flush=True)


def bench_small(args):
    """Closed-loop latency of small served products, one request at a time:
    n x n-limb multiplies (n = 64, 512, 1024; results of at most 2048 limbs)
    through the device-resident loop and through the lanes (the loop off),
    back to back (the loop stays alive) and 5 ms apart (each request finds the
    loop's kernel gone and starts a launch).  Every result checked exactly
    against Python integers."""
    from pi_amd import gnt
    reps, warm = args.requests, 10
    rows = []
    for n in (64, 512, 1024):
        a, b = operand(11, n), operand(12, n)
        want = _to_int(a) * _to_int(b)
        for dev in (True, False):
            for gap in (0.0, 0.005):
                ctx = make_ctx(args, serve_device_limbs=0 if dev else gnt.SERVE_DEVICE_OFF)
                srv = gnt.Server(ctx)
                lat, rt, ok = [], [], True
                for i in range(reps + warm):
                    t0 = time.perf_counter()
                    buf, us = srv.wait(srv.submit(a, b))
                    t1 = time.perf_counter()
                    if i == 0:
                        ok = _to_int(buf.download()) == want
                    buf.release()
                    if i >= warm:
                        lat.append(us)
                        rt.append((t1 - t0) * 1e6)
                    if gap:
                        time.sleep(gap)
                st = srv.stats()
                srv.stop()
                ctx.close()
                rows.append({"n": n, "path": "device loop" if dev else "lanes",
                             "arrivals": "back to back" if not gap else "5 ms apart",
                             "submit_to_done_p50_us": round(float(np.percentile(lat, 50)), 1),
                             "submit_to_done_p99_us": round(float(np.percentile(lat, 99)), 1),
                             "python_round_trip_p50_us": round(float(np.percentile(rt, 50)), 1),
                             "device_loop_launches": st["device_loop_launches"], "exact": ok})
                print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
    best = min(r["submit_to_done_p50_us"] for r in rows if r["path"] == "device loop" and r["n"] == 1024)
    line = {"metric": "small served product latency, device-resident loop vs lanes (us, submit to done)",
            "value": best, "unit": "us", "n_gpus": 1, "steps": reps, "warmup": warm, "ms_per_step": best / 1e3,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
            "data": "synthetic (SplitMix64 limbs, seeds 11/12)",
            "config": {"workload": "closed loop, one n x n multiply at a time, n = 64, 512, 1024",
                       "value_is": "p50 for n = 1024 through the device loop", "rows": rows}}
    print(json.dumps(line))


def _to_int(v) -> int:
    return int.from_bytes(np.ascontiguousarray(v, dtype="<u8").tobytes(), "little")


def bench_chudnovsky(args):
    """pi to --digits decimal digits with pi_compute (binary-splitting merges
    through the serving loop, final assembly and decimal conversion over GPU
    products); digits checked against the pinned SHA-256 when one exists."""
    import hashlib
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "pi_amd", "pi_compute")
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "pi_digits.json")))["digits"].get(str(args.digits))
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "pi.txt")
        cmd = [exe, str(args.digits), out]
        if args.plans:
            cmd += ["--plans", args.plans]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, capture_output=True, text=True)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            sys.stderr.write(r.stderr)
            sys.exit(r.returncode)
        st = json.loads(r.stdout.strip().splitlines()[-1])
        h = hashlib.sha256()
        with open(out, "rb") as f:
            for chunk in iter(lambda: f.read(1 << 24), b""):
                h.update(chunk)
    merge_s = st.get("merge_busy_s", sum(L["s"] for L in st["levels"]))  # GPU merge levels or waves
    line = {
        "metric": f"Chudnovsky pi to {args.digits} digits, wall time",
        "value": round(wall, 3),
        "unit": "s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(wall * 1e3, 1),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "the Chudnovsky series (no input data)",
        "config": {"workload": f"pi to {args.digits} digits", "n_terms": st["n_terms"], "prec_bits": st["prec_bits"],
                   "gpu_min_limbs": st["gpu_min_limbs"], "fa_min_limbs": st.get("fa_min_limbs"),
                   "dec_min_limbs": st.get("dec_min_limbs"),
                   "host_threads": st["threads"],
                   "plans": os.path.relpath(args.plans, ROOT) if args.plans else "built-in"},
        "digits_sha256": h.hexdigest(),
        "byte_identical_to_pinned": (h.hexdigest() == gold["sha256"]) if gold else None,
        "in_process_s": st["t_total"],
        "phases_s": {k: st[k] for k in ("t_init", "t_host_bs", "t_upload", "t_merge", "t_download", "t_serve_stop",
                                         "t_final_assembly", "t_decimal", "t_write", "t_shutdown") if k in st},
        "merge_order": st.get("merge_order"),
        "merge_busy_s": st.get("merge_busy_s"),
        "gpu_merge_share": round(merge_s / wall, 4),
        "levels": st["levels"],
        "decimal_levels": st.get("decimal_levels"),
    }
    print(json.dumps(line), flush=True)



def bench_copies(args):
    """Host <-> device copies of pageable host memory (128 MiB), straight
    through the runtime and through pinned staging, and the PCIe-inclusive
    rate of the host-buffer multiply gnt_mul (upload two 2^24-limb operands,
    product, download) with each.  Every copy and product is checked."""
    import ctypes

    from pi_amd import gnt
    L = gnt.load()
    ctx = make_ctx(args)
    n = 16 << 20
    a = operand(3, n)
    out = np.ones(n, dtype=np.uint64)
    m = 1 << 24
    x, y = operand(3, m), operand(4, m)
    o = np.ones(2 * m, dtype=np.uint64)
    modes = {}
    ref_low = None
    for name, cfg in (("direct", (0, 8 << 20, 4 << 20)), ("staged", (4, 8 << 20, 4 << 20))):
        gnt.set_staging(*cfg)
        up, dn, mul = [], [], []
        for _ in range(max(3, args.steps // 4) + 1):
            h = ctypes.c_void_p()
            t0 = time.perf_counter()
            gnt._check(L.gnt_buf_upload(ctx.h, gnt._p(a), n, ctypes.byref(h)))
            t1 = time.perf_counter()
            got = gnt.SZ()
            gnt._check(L.gnt_buf_download(ctx.h, h, gnt._p(out), n, ctypes.byref(got)))
            t2 = time.perf_counter()
            gnt._check(L.gnt_buf_release(ctx.h, h))
            up.append(t1 - t0)
            dn.append(t2 - t1)
            ok_copy = got.value == n and np.array_equal(out, a)
            out[:] = 1
            ln = gnt.SZ()
            t0 = time.perf_counter()
            gnt._check(L.gnt_mul(ctx.h, gnt._p(x), m, gnt._p(y), m, gnt._p(o), 2 * m, ctypes.byref(ln)))
            mul.append(time.perf_counter() - t0)
            # every run gives the same product: low limb exact, whole result by a fingerprint
            low = (int(x[0]) * int(y[0])) % (1 << 64)
            fp = int(np.bitwise_xor.reduce(o[:ln.value] * np.uint64(0x9E3779B97F4A7C15)))
            ref_low = ref_low or (ln.value, fp)
            ok_mul = int(o[0]) == low and (ln.value, fp) == ref_low
            o[:] = 1
            if not (ok_copy and ok_mul):
                raise SystemExit(f"copies: check failed in mode {name}")
        t_up, t_dn, t_mul = (float(np.median(v[1:])) for v in (up, dn, mul))
        modes[name] = {"threads": cfg[0], "chunk_bytes": cfg[1], "min_bytes": cfg[2],
                       "upload_GBs": round(n * 8 / t_up / 1e9, 2), "download_GBs": round(n * 8 / t_dn / 1e9, 2),
                       "gnt_mul_2p24_ms": round(t_mul * 1e3, 3),
                       "gnt_mul_2p24_Mlimbs_per_s": round(2 * m / t_mul / 1e6, 1), "checks_ok": True}
        print(json.dumps({name: modes[name]}), file=sys.stderr, flush=True)
    gnt.set_staging()
    ctx.close()
    line = {
        "metric": "host-buffer multiply Mlimbs/s at 2^24 limbs (PCIe-inclusive)",
        "value": modes["staged"]["gnt_mul_2p24_Mlimbs_per_s"], "unit": "Mlimbs/s", "n_gpus": 1,
        "steps": max(3, args.steps // 4), "warmup": 1,
        "ms_per_step": modes["staged"]["gnt_mul_2p24_ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic (SplitMix64 uniform limbs, seeds 3/4)",
        "config": {"workload": "128 MiB pageable copies each way; gnt_mul of 2^24 x 2^24 limbs from and to "
                               "pageable host memory", "modes": modes},
    }
    print(json.dumps(line), flush=True)


def bench_c2(args):
    """Single products at 2^20 and 2^24 limbs, mul and sqr, uniform / all-ones /
    sparse operands: device time per product (HIP events), Mlimbs/s and the
    byte-model GB/s; every result checked (residues mod three primes, low limb,
    length)."""
    from pi_amd import gnt
    ctx = make_ctx(args)
    rows = []
    for log_size in (20, 24):
        n = 1 << log_size
        for dist in ("uniform", "ones", "sparse"):
            a, b = operand_dist(dist, 3 if dist != "sparse" else 5, n), operand_dist(dist, 4 if dist != "sparse" else 6, n)
            da, db = ctx.upload(a), ctx.upload(b)
            for op in ("mul", "sqr"):
                log_n, _ = gnt.plan_info(n, n)
                step = (lambda: ctx.sqr_buf(da).release()) if op == "sqr" else (lambda: ctx.mul_buf(da, db).release())
                for _ in range(2):
                    step()
                ctx.sync()
                reps = 20 if log_size == 20 else 6
                ctx.set_timing(True)
                for _ in range(reps):
                    step()
                ctx.sync()
                phases, prods = ctx.get_timing()
                ctx.set_timing(False)
                t_ms = phases[4] / prods
                r = ctx.sqr_buf(da) if op == "sqr" else ctx.mul_buf(da, db)
                out = r.download()
                r.release()
                ok = product_ok(out, a, a if op == "sqr" else b)  # residues, low limb, length
                del out
                byt = b_sqr(n, log_n) if op == "sqr" else b_mul(n, n, log_n)
                rows.append({"limbs": n, "op": op, "dist": dist, "device_ms": round(t_ms, 4),
                             "Mlimbs_per_s": round(2 * n / (t_ms * 1e-3) / 1e6, 1),
                             "model_GBps": round(byt / (t_ms * 1e-3) / 1e9, 1), "check": ok})
            da.release()
            db.release()
    ctx.close()
    top = [r for r in rows if r["limbs"] == 1 << 24 and r["op"] == "mul" and r["dist"] == "uniform"][0]
    print(json.dumps({
        "metric": "Goldilocks-NTT single products at 2^20 and 2^24 limbs (three distributions)",
        "value": top["Mlimbs_per_s"], "unit": "Mlimbs/s", "n_gpus": 1, "steps": 6, "warmup": 2,
        "ms_per_step": top["device_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic (SplitMix64: uniform seeds 3/4, sparse seeds 5/6, all-ones)",
        "config": {"workload": "mul and sqr, 2^20 and 2^24 limbs per operand, inputs resident in HBM",
                   "plans": os.path.relpath(args.plans, ROOT) if args.plans else "built-in", "rows": rows},
    }), flush=True)


if __name__ == "__main__":
    main()
