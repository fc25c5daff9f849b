"""CPU restatement of the engine's three carry steps (k_ctiles, k_cscan,
k_cripple in pi_amd/csrc/gnt_engine.hip), tile for tile, checked against
CPython int for products' limb sums, sums and differences.  It pins the
tile-level rules the kernels rely on: (generate, propagate) aggregation, the
carry-in scan over tiles, the sparse ripple, and the trimmed length when a
difference's top limbs cancel (a carry that leaves the top clears every
all-ones tile it crosses)."""
import random

import pytest

TILE = 1024
MASK = (1 << 64) - 1
SUMS, ADD, SUB = 0, 1, 2


def _limb_values(mode, n_limbs, A, B, j):
    """(s, generate) of limb j as k_ctiles forms it, before in-tile carries."""
    if j >= n_limbs:
        return 0, 0, True            # padding: propagates (never written)
    if mode == SUMS:
        L, Hprev = A[j], (B[j - 1] if j > 0 else 0)
        s = (L + Hprev) & MASK
        return s, int(s < L), s == MASK
    a = A[j] if j < len(A) else 0
    b = B[j] if j < len(B) else 0
    if mode == SUB:
        b = ~b & MASK
    s = (a + b) & MASK
    g = int(s < a)
    if mode == SUB and j == 0:
        s = (s + 1) & MASK
        g |= int(s == 0)
    return s, g, s == MASK


def carry_tiles(mode, n_limbs, A, B):
    nt = (n_limbs + TILE - 1) // TILE
    out = [0] * n_limbs
    gp, hi = [], []
    for k in range(nt):
        carry, h = 0, 0
        G, P = 0, 1
        for j in range(k * TILE, (k + 1) * TILE):
            s, g, p = _limb_values(mode, n_limbs, A, B, j)
            z = (s + carry) & MASK
            carry = g | (int(p) & carry)
            G, P = g | (int(p) & G), P & int(p)   # aggregate of the limbs so far
            if j < n_limbs:
                out[j] = z
                if z:
                    h = j + 1
        gp.append((G, P))
        hi.append(h)
    # k_cscan: carry into each tile, length excluding all-ones tiles that a carry clears
    cin, carry, maxhi = [], 0, 0
    for k in range(nt):
        cin.append(carry)
        G, P = gp[k]
        if not (carry and P):
            maxhi = max(maxhi, hi[k])
        carry = G | (P & carry)
    # k_cripple
    for k in range(nt):
        if not cin[k]:
            continue
        for j in range(k * TILE, min((k + 1) * TILE, n_limbs)):
            out[j] = (out[j] + 1) & MASK
            if out[j]:
                maxhi = max(maxhi, j + 1)
                break
    return out, maxhi


def to_limbs(x, n=None):
    v = []
    while x:
        v.append(x & MASK)
        x >>= 64
    if n is not None:
        v += [0] * (n - len(v))
    return v


def value(v):
    return sum(x << (64 * i) for i, x in enumerate(v))


def run_add(a, b, sub):
    A, B = to_limbs(a), to_limbs(b)
    n = len(A) if sub else max(len(A), len(B)) + 1
    out, length = carry_tiles(SUB if sub else ADD, n, A, B)
    return value(out), length


def _cases():
    rnd = random.Random(3)
    X = rnd.getrandbits(64 * 3000) | (1 << (64 * 3000 - 1))
    ones = (1 << (64 * 2500)) - 1
    return [
        ("add", rnd.getrandbits(64 * 2100), rnd.getrandbits(64 * 2100)),
        ("add", ones, 1),
        ("add", ones, ones),
        ("sub", 1 << (64 * 2200), 1),
        ("sub", X + 12345, X),                 # top 3000 limbs cancel (a carry leaves the top)
        ("sub", X + (1 << (64 * 1500)), X),
        ("sub", X, X),
        ("sub", X, rnd.getrandbits(64 * 40)),
        ("sub", 7, 5),
    ]


@pytest.mark.parametrize("i", range(9))
def test_add_sub_tiles(i):
    op, a, b = _cases()[i]
    got, length = run_add(a, b, op == "sub")
    want = a + b if op == "add" else a - b
    assert got == want
    assert length == len(to_limbs(want))


def test_product_limb_sums():
    """SUMS mode: limb j = L_j + H_(j-1), as the last inverse pass leaves them."""
    rnd = random.Random(9)
    for n in (5, 1023, 1024, 1025, 3000):
        L = [rnd.getrandbits(64) for _ in range(n)]
        H = [rnd.getrandbits(20) for _ in range(n - 1)] + [0]
        run = range(n // 2, min(n - 2, n // 2 + 1500))
        for j in run:
            L[j] = MASK                   # long propagate run
        L[n - 1] = rnd.getrandbits(32)    # the value fits n limbs, as a product's does
        out, length = carry_tiles(SUMS, n, L, H)
        want = sum((L[j] + (H[j] << 64)) << (64 * j) for j in range(n))
        assert value(out) == want
        assert length == len(to_limbs(want))
