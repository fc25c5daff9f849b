"""CPU restatement of the carry kernels' tile routines (carry_tile and
carry_ripple_tile in pi_amd/csrc/gnt_carry.hpp, with the tile scan of
k_cscan in pi_amd/csrc/gnt_engine.hip), wave for wave and row for row,
checked against CPython int for products' limb sums, sums and differences.
It pins the rules the kernels rely on: the 64 carries of a row from
((G|P) + G + cin) ^ P, the per-lane (m0, pre) bits that re-apply a wave's
carry-in, the (generate, propagate) aggregation over rows, waves and tiles,
the ripple of a tile's carry-in, and the trimmed length when a difference's
top limbs cancel (a carry that leaves the top clears every all-ones tile it
crosses)."""
import random

import pytest

WAVES, ROWS = 4, 8
TILE = WAVES * ROWS * 64
MASK = (1 << 64) - 1
SUMS, ADD, SUB = 0, 1, 2


def _limb_values(mode, n_limbs, A, B, j):
    """(y, generate, propagate) of limb j before carries (the Src functors)."""
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
    return s, g, s == MASK and not g


def row_carries(G, P, cin):
    """Carries into the 64 lanes of a row and the carry out (row_carries)."""
    A = G | P
    t = (A + G) & MASK
    S = (t + cin) & MASK
    return S ^ P, int(t < A) | int(S < t)


def _bits(flags):
    return sum(1 << i for i, f in enumerate(flags) if f)


def carry_tile(mode, n_limbs, A, B, tile, out, nwaves=WAVES):
    """One tile as the kernel works it (nwaves waves of 64 lanes): returns
    ((generate, propagate), hi)."""
    tile_limbs = nwaves * ROWS * 64
    waves = []
    for w in range(nwaves):
        jw = tile * tile_limbs + w * ROWS * 64
        rows = []
        for r in range(ROWS):
            vals = [_limb_values(mode, n_limbs, A, B, jw + r * 64 + lane) for lane in range(64)]
            G = _bits(v[1] for v in vals)
            P = _bits(v[2] for v in vals)
            M, co = row_carries(G, P, 0)
            m0 = [(M >> lane) & 1 for lane in range(64)]
            pre = [int((~P & ((1 << lane) - 1)) == 0) for lane in range(64)]
            rows.append(([v[0] for v in vals], m0, pre, co, int(P == MASK)))
        c = 0
        for _, _, _, co, rp in rows:
            c = co | (rp & c)
        waves.append((rows, c, int(all(rw[4] for rw in rows))))
    # second pass: every wave with its carry-in
    cw, agg_g, agg_p, hi = 0, 0, 1, 0
    for w, (rows, g, p) in enumerate(waves):
        jw = tile * tile_limbs + w * ROWS * 64
        c = cw
        for r, (y, m0, pre, co, rp) in enumerate(rows):
            for lane in range(64):
                j = jw + r * 64 + lane
                z = (y[lane] + (m0[lane] | (c & pre[lane]))) & MASK
                if j < n_limbs:
                    out[j] = z
                    if z:
                        hi = max(hi, j + 1)
            c = co | (rp & c)
        agg_g, agg_p = g | (p & agg_g), agg_p & p
        cw = g | (p & cw)
    return (agg_g, agg_p), hi


def carry_ripple_tile(out, n_limbs, tile):
    """+1 into a tile, a row of 64 limbs at a time; 1 + the incremented limb."""
    j1 = min(n_limbs, (tile + 1) * TILE)
    for j0 in range(tile * TILE, j1, 64):
        lanes = [j for j in range(j0, j0 + 64) if j < j1]
        stop = [j for j in lanes if out[j] != MASK]
        first = stop[0] if stop else j0 + 64
        for j in lanes:
            if j < first:
                out[j] = 0
            elif j == first:
                out[j] += 1
        if stop:
            return first + 1
    return 0


def carry_tiles(mode, n_limbs, A, B):
    nt = (n_limbs + TILE - 1) // TILE
    out = [0] * n_limbs
    gp, hi = [], []
    for k in range(nt):
        g, h = carry_tile(mode, n_limbs, A, B, k, out)
        gp.append(g)
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
        if cin[k]:
            maxhi = max(maxhi, carry_ripple_tile(out, n_limbs, k))
    return out, maxhi


def test_row_carries_is_the_carry_chain():
    """Brute force: the closed form equals c_i = g_(i-1) | (p_(i-1) & c_(i-1))."""
    rnd = random.Random(5)
    for _ in range(3000):
        kind = rnd.random()
        g = [0] * 64
        p = [0] * 64
        for i in range(64):
            u = rnd.random()
            if kind < 0.3:           # mostly propagate: long runs
                p[i] = int(u < 0.9)
                g[i] = int(not p[i] and u < 0.95)
            else:
                g[i] = int(u < 0.3)
                p[i] = int(not g[i] and u < 0.7)
        cin = rnd.getrandbits(1)
        M, cout = row_carries(_bits(g), _bits(p), cin)
        c = cin
        for i in range(64):
            assert (M >> i) & 1 == c
            c = g[i] | (p[i] & c)
        assert cout == c


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
        # limb 0 of a - b is 2^64 - 1 with a carry out (the two's complement's
        # +1): generating, so it must not also propagate
        ("sub", (1 << 128) - 1, 1 << 64),
        ("sub", MASK | (0xDEADBEEF << 64), 0x1234 << 64),
        ("sub", MASK | (X << 64), (X - 1) << 64),
    ]


@pytest.mark.parametrize("i", range(12))
def test_one_tile_of_sixteen_waves(i):
    """The device-resident loop's workgroup (16 waves) takes a whole request in
    one carry tile of 8192 limbs (no tile scan, no ripple): the same cases,
    cut to that size."""
    op, a, b = _cases()[i]
    a &= (1 << (64 * 8191)) - 1
    b &= (1 << (64 * 8191)) - 1
    if op == "sub" and b > a:
        a, b = b, a
    A, B = to_limbs(a), to_limbs(b)
    n = max(len(A), 1) if op == "sub" else max(len(A), len(B)) + 1
    out = [0] * n
    (g, p), hi = carry_tile(SUB if op == "sub" else ADD, n, A, B, 0, out, nwaves=16)
    want = (a - b) if op == "sub" else (a + b)
    assert value(out) == want
    assert hi == len(to_limbs(want))



@pytest.mark.parametrize("i", range(12))
def test_add_sub_tiles(i):
    op, a, b = _cases()[i]
    got, length = run_add(a, b, op == "sub")
    want = a + b if op == "add" else a - b
    assert got == want
    assert length == len(to_limbs(want))


def test_product_limb_sums():
    """SUMS mode: limb j = L_j + H_(j-1), as the last inverse pass leaves them."""
    rnd = random.Random(9)
    for n in (5, 1023, 2047, 2048, 2049, 4100):
        L = [rnd.getrandbits(64) for _ in range(n)]
        H = [rnd.getrandbits(20) for _ in range(n - 1)] + [0]
        run = range(n // 2, min(n - 2, n // 2 + 2100))
        for j in run:
            L[j] = MASK                   # long propagate run
        L[n - 1] = rnd.getrandbits(32)    # the value fits n limbs, as a product's does
        out, length = carry_tiles(SUMS, n, L, H)
        want = sum((L[j] + (H[j] << 64)) << (64 * j) for j in range(n))
        assert value(out) == want
        assert length == len(to_limbs(want))
