"""Write tests/golden/products.json: reference products from GMP and CPython.

Independent of the oracle and of the GPU path: every product below is computed
by CPython int multiplication and, when libgmp.so.10 is present, confirmed by
GMP's mpn_mul.  Inputs come from generators whose definitions are restated in
oracle/gl_oracle.c and pinned by the first entries (their first limbs).

Run: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as orc  # noqa: E402

MASK = (1 << 64) - 1


def py_splitmix64(seed, n):
    out, x = [], seed
    for _ in range(n):
        x = (x + 0x9E3779B97F4A7C15) & MASK
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        out.append(z ^ (z >> 31))
    return out


def py_random_vec(n, seed):  # ntt.rs:1216-1226
    P = 0xFFFFFFFF00000001
    s = (seed * 0x9E3779B97F4A7C15 + 1) & MASK
    out = []
    for _ in range(n):
        s = (s * 6364136223846793005 + 1442695040888963407) & MASK
        out.append(s % P)
    return out


def to_int(limbs):
    return sum(int(x) << (64 * i) for i, x in enumerate(limbs))


def limbs_of(v):
    out = []
    while v:
        out.append(v & MASK)
        v >>= 64
    return out


def sha(limbs):
    return hashlib.sha256(b"".join(int(x).to_bytes(8, "little") for x in limbs)).hexdigest()


def main():
    g = {"generators": {}, "small": [], "hashed": []}
    g["generators"]["splitmix64_seed1_first4"] = [hex(x) for x in py_splitmix64(1, 4)]
    g["generators"]["random_vec_7_seed7_first4"] = [hex(x) for x in py_random_vec(4, 7)]
    # the reference's test operand shapes (ntt.rs:1394-1411), full products
    for la, lb, seed in [(1, 1, 1), (1, 4, 2), (4, 4, 3), (8, 5, 4), (32, 16, 5), (60, 60, 6), (200, 200, 7)]:
        a, b = py_random_vec(la, seed), py_random_vec(lb, seed ^ 0xFF)
        c = limbs_of(to_int(a) * to_int(b))
        if orc.gmp() is not None:
            assert [int(x) for x in orc.gmp_mul(a, b)] == c
        g["small"].append({"gen": "random_vec", "la": la, "lb": lb, "seed_a": seed, "seed_b": seed ^ 0xFF,
                           "product": [hex(x) for x in c]})
    # C1: 2^16 x 2^16 SplitMix64 seeds 1/2 (top limb forced nonzero), and larger hashed cases
    for la, lb, sa, sb in [(1 << 16, 1 << 16, 1, 2), (1 << 12, 3000, 11, 12), (100000, 7, 13, 14)]:
        a, b = py_splitmix64(sa, la), py_splitmix64(sb, lb)
        a[-1] = a[-1] or 1
        b[-1] = b[-1] or 1
        c = limbs_of(to_int(a) * to_int(b))
        if orc.gmp() is not None:
            assert [int(x) for x in orc.gmp_mul(a, b)] == c
        g["hashed"].append({"gen": "splitmix64", "la": la, "lb": lb, "seed_a": sa, "seed_b": sb,
                            "len": len(c), "sha256": sha(c), "low": hex(c[0]), "high": hex(c[-1])})
    with open(os.path.join(HERE, "products.json"), "w") as f:
        json.dump(g, f, indent=1)
    print("wrote", os.path.join(HERE, "products.json"))


if __name__ == "__main__":
    main()
