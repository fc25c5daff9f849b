// Tile-level bodies of the transform passes and of the carry, shared by the
// HIP kernels and by the CPU emulation used in tests.
//
// Every routine is written against an executor X: on the device X is the
// workgroup (tasks strided by the block size, X::sync() is a barrier); in the
// emulation X runs every task in turn and sync() does nothing.  The same
// source therefore defines what the GPU computes and what the tests check on
// the CPU.  See gnt_plan.hpp for the pass structure.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "gl_field.hpp"
#include "gnt_plan.hpp"

namespace gnt {

using gl::u32;
using gl::u64;

#if defined(__HIPCC__) || defined(__HIP__)
template <int B>
struct DevExec {
    static constexpr int step = B;
    __device__ static int first() { return threadIdx.x; }
    __device__ static void sync() { __syncthreads(); }
};
#endif

struct HostExec {
    static constexpr int step = 1;
    static int first() { return 0; }
    static void sync() {}
};

#define GNT_FOR(X, L, n) for (int L = X::first(); L < (n); L += X::step)

// omega_{2^32}^e from the two-level table.
GL_HD u64 root_pow(const u64 *__restrict__ wlo, const u64 *__restrict__ whi, u32 e) {
    return gl::mul(wlo[e & ((1u << TW_SPLIT) - 1)], whi[e >> TW_SPLIT]);
}

template <int LOGR>
struct Sub {
    static constexpr int count = (LOGR + 3) / 4;
    static constexpr int log(int i) { return LOGR / count + (i < LOGR % count ? 1 : 0); }
    // block size (log) seen by sub-step i inside the R-point column DFT
    static constexpr int log_m(int i) { return i == 0 ? LOGR : log_m(i - 1) - log(i - 1); }
};

// Global address of tile element (row n, tile column c).
GL_HD size_t tile_addr(size_t cg0, int n, int c, int log_m, int log_s) {
    size_t cg = cg0 + (size_t)c;
    size_t b = cg >> log_s;
    size_t j = cg & (((size_t)1 << log_s) - 1);
    return (b << log_m) + ((size_t)n << log_s) + j;
}

// Tile element index -> (row, column): columns fastest when columns are
// adjacent in memory (log_s > 0), rows fastest in the contiguous last pass.
GL_HD void tile_rc(int idx, int log_r, int log_c, int log_s, int &n, int &c) {
    if (log_s > 0) { c = idx & ((1 << log_c) - 1); n = idx >> log_c; }
    else { n = idx & ((1 << log_r) - 1); c = idx >> log_r; }
}

// Coefficient i of an operand packed 16 bits per coefficient (ntt.rs:640-665).
GL_HD u64 coeff_from_limbs(const u64 *__restrict__ limbs, size_t n_limbs, size_t i) {
    size_t l = i >> 2;
    if (l >= n_limbs) return 0;
    return (limbs[l] >> (16 * (i & 3))) & 0xFFFFull;
}

// Sub-step I of the forward R-point column DFT for task L (DIF, Bailey form):
// an r-point DFT over rows base + u*t, then the twiddle omega_{2^LM}^(jj*k1),
// or, after the last sub-step, the pass twiddle omega_M^(j*bitrev_R(p)).
template <int LOGR, int I>
GL_HD void fwd_task(u64 *tile, int L, int log_c, size_t cg0, int log_m, int log_s, const u64 *__restrict__ wlo,
                    const u64 *__restrict__ whi) {
    constexpr int A = Sub<LOGR>::log(I);
    constexpr int R = 1 << A;
    constexpr int LM = Sub<LOGR>::log_m(I);
    constexpr int LT = LM - A;
    constexpr bool LAST = (I == Sub<LOGR>::count - 1);
    const int c = L & ((1 << log_c) - 1);
    const int rest = L >> log_c;
    const int jj = rest & ((1 << LT) - 1);
    const int blk = rest >> LT;
    const int base = (blk << LM) + jj;
    u64 v[R];
#pragma unroll
    for (int u = 0; u < R; u++) v[u] = tile[((base + (u << LT)) << log_c) + c];
    gl::dft_dif<R>(v);
    if (!LAST) {
        // omega_{2^LM}^e = whi[e << (16 - LM)] (LM <= 13): one lookup
#pragma unroll
        for (int q = 1; q < R; q++) {
            u32 k1 = gl::bitrev((u32)q, A);
            v[q] = gl::mul(v[q], whi[((u32)jj * k1) << (TW_SPLIT - LM)]);
        }
    } else if (log_s > 0) {
        size_t j = (cg0 + (size_t)c) & (((size_t)1 << log_s) - 1);
#pragma unroll
        for (int q = 0; q < R; q++) {
            u32 p = (u32)(base + (q << LT));
            u32 kc = gl::bitrev(p, LOGR);
            u64 ex = ((u64)j * kc) << (ROOT_LOG - log_m);
            v[q] = gl::mul(v[q], root_pow(wlo, whi, (u32)ex));
        }
    }
#pragma unroll
    for (int q = 0; q < R; q++) tile[((base + (q << LT)) << log_c) + c] = v[q];
}

// Inverse of fwd_task<LOGR, I>: undo the twiddle, then the inverse r-point
// DFT (DIT, bit-reversed input, natural output).  Unscaled.
template <int LOGR, int I>
GL_HD void inv_task(u64 *tile, int L, int log_c, size_t cg0, int log_m, int log_s, const u64 *__restrict__ wlo,
                    const u64 *__restrict__ whi) {
    constexpr int A = Sub<LOGR>::log(I);
    constexpr int R = 1 << A;
    constexpr int LM = Sub<LOGR>::log_m(I);
    constexpr int LT = LM - A;
    constexpr bool LAST = (I == Sub<LOGR>::count - 1);
    const int c = L & ((1 << log_c) - 1);
    const int rest = L >> log_c;
    const int jj = rest & ((1 << LT) - 1);
    const int blk = rest >> LT;
    const int base = (blk << LM) + jj;
    u64 v[R];
#pragma unroll
    for (int q = 0; q < R; q++) v[q] = tile[((base + (q << LT)) << log_c) + c];
    if (!LAST) {
#pragma unroll
        for (int q = 1; q < R; q++) {
            u32 k1 = gl::bitrev((u32)q, A);
            u32 e = ((u32)jj * k1) & ((1u << LM) - 1);
            // omega^-e = omega^(2^LM - e)
            v[q] = gl::mul(v[q], whi[(((1u << LM) - e) & ((1u << LM) - 1)) << (TW_SPLIT - LM)]);
        }
    } else if (log_s > 0) {
        size_t j = (cg0 + (size_t)c) & (((size_t)1 << log_s) - 1);
#pragma unroll
        for (int q = 0; q < R; q++) {
            u32 p = (u32)(base + (q << LT));
            u32 kc = gl::bitrev(p, LOGR);
            u64 ex = ((u64)j * kc) << (ROOT_LOG - log_m);
            v[q] = gl::mul(v[q], root_pow(wlo, whi, 0u - (u32)ex));
        }
    }
    gl::dft_dit_inv<R>(v);
#pragma unroll
    for (int u = 0; u < R; u++) tile[((base + (u << LT)) << log_c) + c] = v[u];
}

// ---- tile-level routines ----

struct TileGeom {
    int log_m, log_c, log_s, tile_n;
    size_t cg0;
};

template <int LOGR>
GL_HD TileGeom tile_geom(const PassDesc &d, size_t tile_index) {
    TileGeom g;
    g.log_m = d.log_m;
    g.log_c = d.log_c;
    g.log_s = d.log_m - LOGR;
    g.tile_n = 1 << (LOGR + d.log_c);
    g.cg0 = tile_index << d.log_c;
    return g;
}

// Load a tile from x, or (limbs != nullptr) pack it from the operand's limbs.
template <class X, int LOGR>
GL_HD void tile_load(u64 *tile, const TileGeom &g, const u64 *__restrict__ x, const u64 *__restrict__ limbs,
                     size_t n_limbs) {
    GNT_FOR(X, idx, g.tile_n) {
        int n, c;
        tile_rc(idx, LOGR, g.log_c, g.log_s, n, c);
        size_t a = tile_addr(g.cg0, n, c, g.log_m, g.log_s);
        tile[(n << g.log_c) + c] = limbs ? coeff_from_limbs(limbs, n_limbs, a) : x[a];
    }
    X::sync();
}

template <class X, int LOGR>
GL_HD void tile_store(const u64 *tile, const TileGeom &g, u64 *__restrict__ x) {
    GNT_FOR(X, idx, g.tile_n) {
        int n, c;
        tile_rc(idx, LOGR, g.log_c, g.log_s, n, c);
        x[tile_addr(g.cg0, n, c, g.log_m, g.log_s)] = tile[(n << g.log_c) + c];
    }
}

template <class X, int LOGR, int I>
GL_HD void tile_fwd(u64 *tile, const TileGeom &g, const u64 *wlo, const u64 *whi) {
    if constexpr (I < Sub<LOGR>::count) {
        const int ntask = g.tile_n >> Sub<LOGR>::log(I);
        GNT_FOR(X, L, ntask) fwd_task<LOGR, I>(tile, L, g.log_c, g.cg0, g.log_m, g.log_s, wlo, whi);
        X::sync();
        tile_fwd<X, LOGR, I + 1>(tile, g, wlo, whi);
    }
}

// Inverse sub-steps I, I-1, ..., 0.
template <class X, int LOGR, int I>
GL_HD void tile_inv(u64 *tile, const TileGeom &g, const u64 *wlo, const u64 *whi) {
    if constexpr (I >= 0) {
        const int ntask = g.tile_n >> Sub<LOGR>::log(I);
        GNT_FOR(X, L, ntask) inv_task<LOGR, I>(tile, L, g.log_c, g.cg0, g.log_m, g.log_s, wlo, whi);
        X::sync();
        tile_inv<X, LOGR, I - 1>(tile, g, wlo, whi);
    }
}

// tile *= other (at the same global positions, or the tile itself) * 2^(192 - log_n):
// the pointwise product with the 1/N of the inverse (ntt.rs:845-853, 376-391).
template <class X, int LOGR>
GL_HD void tile_pointwise(u64 *tile, const TileGeom &g, const u64 *__restrict__ other, int log_n) {
    const int e = 192 - log_n;
    GNT_FOR(X, idx, g.tile_n) {
        int n, c;
        tile_rc(idx, LOGR, g.log_c, g.log_s, n, c);
        const int t = (n << g.log_c) + c;
        u64 v = tile[t];
        u64 w = other ? other[tile_addr(g.cg0, n, c, g.log_m, g.log_s)] : v;
        tile[t] = gl::mul_pow2(gl::mul(v, w), e);
    }
    X::sync();
}

// Limb sums of the finished coefficients of a tile of the LAST inverse pass
// (the pass over blocks of size N, so tile addresses are coefficient
// indices): S_j = sum_t c[4j+t] 2^(16t) = L_j + 2^64 H_j for every limb whose
// four coefficients lie in this tile, j < n_limbs.  Coefficients at or past
// n_coeffs are zero by construction of the product; they are masked anyway.
template <class X, int LOGR>
GL_HD void tile_limb_sums(const u64 *tile, const TileGeom &g, size_t n_coeffs, size_t n_limbs, u64 *__restrict__ Ls,
                          u64 *__restrict__ Hs) {
    const int nq = g.tile_n >> 2;
    GNT_FOR(X, q, nq) {
        // four consecutive coefficients: along columns (log_s > 0) or rows
        int n, c;
        tile_rc(q << 2, LOGR, g.log_c, g.log_s, n, c);
        size_t a0 = tile_addr(g.cg0, n, c, g.log_m, g.log_s);
        size_t j = a0 >> 2;
        if (j < n_limbs) {
            const int t0 = (n << g.log_c) + c;
            const int step = g.log_s > 0 ? 1 : (1 << g.log_c);
            u64 cc[4];
#pragma unroll
            for (int k = 0; k < 4; k++) cc[k] = (a0 + k < n_coeffs) ? tile[t0 + k * step] : 0;
            u64 lo = cc[0], hi = 0, t;
            t = cc[1] << 16; lo += t; hi += (cc[1] >> 48) + (lo < t);
            t = cc[2] << 32; lo += t; hi += (cc[2] >> 32) + (lo < t);
            t = cc[3] << 48; lo += t; hi += (cc[3] >> 16) + (lo < t);
            Ls[j] = lo;
            Hs[j] = hi;
        }
    }
}

// Carry-lookahead pairs: bit 0 = generate, bit 1 = propagate.  `lo` covers
// the lower limbs, `hi` the limbs just above; identity is 2 (g=0, p=1).
GL_HD u32 gp_combine(u32 lo, u32 hi) {
    u32 g = (hi & 1) | ((hi >> 1) & lo & 1);
    u32 p = (hi >> 1) & (lo >> 1) & 1;
    return g | (p << 1);
}

}  // namespace gnt
