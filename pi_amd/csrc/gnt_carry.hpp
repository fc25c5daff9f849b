// Tile routines of the exact carry (limb values -> limbs), shared by the
// single-product kernels (gnt_kernels.hip) and the stage engine
// (gnt_engine.hip).  Reference: the carry loop of crates/bignum/src/ntt.rs:669-791.
//
// Every limb j has a value y_j (a limb sum L_j + H_{j-1}, or a + b) and a
// generate bit g_j (that sum overflowed); it propagates an incoming carry
// when y_j = 2^64 - 1 and g_j = 0 (p_j), so g_j and p_j are never both set.  Limb j of the
// result is y_j + c_j with c_j = g_{j-1} | (p_{j-1} & c_{j-1}).
//
// A carry tile is CARRY_TILE consecutive limbs, worked by one workgroup of
// CARRY_WAVES waves; wave w owns CARRY_ROWS rows of 64 consecutive limbs
// (lane = limb within the row), so every load and store is a contiguous
// 512-byte wave access.  The 64 carries of a row come from two ballots and
// one scalar add (row_carries); rows chain through one carry bit, waves
// through an LDS word each.  A tile's (generate, propagate) pair goes to the
// tile scan; tiles that receive a carry are fixed by carry_ripple_tile.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gnt_pass.hpp"

namespace gnt {

constexpr int CARRY_WAVES = 4;
constexpr int CARRY_ROWS = 8;
constexpr int CARRY_BLOCK = 64 * CARRY_WAVES;
constexpr int CARRY_TILE = CARRY_BLOCK * CARRY_ROWS;  // 2048 limbs
constexpr int CARRY_RIPPLE_BLOCK = 64;                 // one wave per tile

// Carries into the 64 lanes of a row whose lanes have generate bits G and
// propagate bits P (disjoint), with carry cin entering lane 0.  The chain
// c_i = g_{i-1} | (p_{i-1} & c_{i-1}) is the carry chain of the binary sum
// (G|P) + G + cin (lane i: a_i = g_i|p_i, b_i = g_i, so a_i & b_i = g_i and
// a_i ^ b_i = p_i), so the carry into lane i is bit i of that sum xor P, and
// the carry out of lane 63 is the sum's overflow.
__device__ __forceinline__ uint64_t row_carries(uint64_t G, uint64_t P, uint32_t cin, uint32_t &cout) {
    const uint64_t A = G | P;
    const uint64_t t = A + G;
    const uint64_t S = t + cin;
    cout = (uint32_t)(t < A) | (uint32_t)(S < t);
    return S ^ P;
}

// Limb values of a product from its limb sums: y_j = L_j + H_{j-1}.
struct SumsSrc {
    const uint64_t *Ls, *Hs;
    __device__ __forceinline__ void operator()(uint64_t j, uint64_t &y, uint32_t &g) const {
        const uint64_t L = Ls[j], Hm = j ? Hs[j - 1] : 0;
        y = L + Hm;
        g = (uint32_t)(y < L);
    }
};

// Limb values of a sum or difference: a + b, or a + ~b + 1 over n_limbs = na
// limbs (b zero-extended before the complement).
struct LinearSrc {
    const uint64_t *a, *b;
    uint64_t na, nb;
    bool sub;
    __device__ __forceinline__ void operator()(uint64_t j, uint64_t &y, uint32_t &g) const {
        const uint64_t x = j < na ? a[j] : 0;
        uint64_t v = j < nb ? b[j] : 0;
        if (sub) v = ~v;
        y = x + v;
        g = (uint32_t)(y < x);
        if (sub && j == 0) {  // the +1 of the two's complement
            y += 1;
            g |= (uint32_t)(y == 0);
        }
    }
};

// Phase 1 of carry tile `tile`: limb values from src(j, y, g) for j < n_limbs,
// carries resolved inside the tile as if none entered it, limbs written to
// out.  Returns, identical in every thread, the tile's (generate, propagate)
// pair (gp_combine form) and 1 + the index of its highest nonzero limb (0 if
// none).  Called by all 64 * WAVES threads of the workgroup; a tile is
// 64 * WAVES * CARRY_ROWS limbs (CARRY_TILE for the carry kernels' WAVES =
// CARRY_WAVES); `sh` is 2*WAVES words of LDS.
template <class Src, int WAVES = CARRY_WAVES>
__device__ __forceinline__ void carry_tile(const Src &src, uint64_t n_limbs, uint64_t tile,
                                           uint64_t *__restrict__ out, uint32_t *sh, uint32_t &gp_tile,
                                           uint32_t &hi_tile) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t jw = tile * (uint64_t)(64 * WAVES * CARRY_ROWS) + (uint64_t)w * (CARRY_ROWS * 64);
    uint64_t y[CARRY_ROWS];
    uint32_t gb = 0, pb = 0;
#pragma unroll
    for (int r = 0; r < CARRY_ROWS; r++) {
        const uint64_t j = jw + r * 64 + lane;
        uint64_t v = 0;
        uint32_t g = 0;
        if (j < n_limbs) src(j, v, g);
        y[r] = v;
        gb |= g << r;
        // limbs past the end propagate, so a tile is all-ones exactly when
        // its limbs inside the result are.  A generating limb never
        // propagates (its carry out is 1 whatever comes in): a difference's
        // limb 0 can be 2^64 - 1 with the +1 of the two's complement's carry,
        // and marking both would hand lane 1 a carry twice.
        pb |= (uint32_t)(j >= n_limbs || (v == ~0ull && !g)) << r;
    }
    // Rows with no carry into the wave: per lane, the carry it receives (m0)
    // and whether every lane below it propagates (pre); per row, its carry
    // out (rg) and whether all its lanes propagate (rp).  With a carry c
    // into the row, lane i receives m0_i | (c & pre_i).
    uint32_t m0 = 0, pre = 0, rg = 0, rp = 0, c = 0;
    const uint64_t below = (1ull << lane) - 1;
#pragma unroll
    for (int r = 0; r < CARRY_ROWS; r++) {
        const uint64_t G = __ballot((gb >> r) & 1), P = __ballot((pb >> r) & 1);
        uint32_t co;
        const uint64_t M = row_carries(G, P, 0, co);
        m0 |= (uint32_t)((M >> lane) & 1) << r;
        pre |= (uint32_t)((~P & below) == 0) << r;
        rg |= co << r;
        rp |= (uint32_t)(P == ~0ull) << r;
        c = co | ((uint32_t)(P == ~0ull) & c);
    }
    if (lane == 0) sh[w] = c | ((uint32_t)(rp == (1u << CARRY_ROWS) - 1) << 1);
    __syncthreads();
    uint32_t cw = 0, agg = 2;
#pragma unroll
    for (int v = 0; v < WAVES; v++) {
        const uint32_t gp = sh[v];
        if (v < w) cw = (gp & 1) | ((gp >> 1) & cw);
        agg = gp_combine(agg, gp);
    }
    // the rows again, with the carry into this wave
    uint64_t hi = 0;
    c = cw;
#pragma unroll
    for (int r = 0; r < CARRY_ROWS; r++) {
        const uint32_t in_lane = ((m0 >> r) & 1) | (c & (pre >> r) & 1);
        c = ((rg >> r) & 1) | ((rp >> r) & c);
        const uint64_t j = jw + r * 64 + lane;
        const uint64_t z = y[r] + in_lane;
        const bool in = j < n_limbs;
        if (in) out[j] = z;
        const uint64_t nz = __ballot(in && z != 0);
        if (nz) hi = jw + r * 64 + (63 - __clzll((long long)nz)) + 1;
    }
    if (lane == 0) sh[WAVES + w] = (uint32_t)hi;
    __syncthreads();
    uint32_t h = 0;
#pragma unroll
    for (int v = 0; v < WAVES; v++) h = max(h, sh[WAVES + v]);
    gp_tile = agg;
    hi_tile = h;
}

// Phase 3 of a tile that receives a carry: the +1 clears the run of all-ones
// limbs at the bottom of the tile and increments the first other limb (a
// carry through the whole tile leaves it).  One wave; returns 1 + the index
// of the incremented limb, 0 if the carry left the tile or the result.
__device__ __forceinline__ uint32_t carry_ripple_tile(uint64_t *__restrict__ out, uint64_t n_limbs,
                                                      uint64_t tile) {
    const int lane = threadIdx.x & 63;
    const uint64_t j1 = min(n_limbs, (tile + 1) * CARRY_TILE);
    for (uint64_t j0 = tile * CARRY_TILE; j0 < j1; j0 += 64) {
        const uint64_t j = j0 + lane;
        const bool in = j < j1;
        const uint64_t z = in ? out[j] : 0;
        const uint64_t stop = __ballot(in && z != ~0ull);
        const int first = stop ? __ffsll((long long)stop) - 1 : 64;
        if (in && lane <= first) out[j] = lane == first ? z + 1 : 0;
        if (stop) return (uint32_t)(j0 + first + 1);
    }
    return 0;
}

}  // namespace gnt
