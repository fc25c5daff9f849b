// Per-task bodies of the transform passes and of the carry, shared by the HIP
// kernels (one task per thread between barriers) and by the CPU emulation
// used in tests (tasks run one after another).  See gnt_plan.hpp for the
// pass structure.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "gl_field.hpp"
#include "gnt_plan.hpp"

namespace gnt {

using gl::u32;
using gl::u64;

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
// an r-point DFT over rows base + u*t, then the twiddle omega_{m}^(jj*k1),
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
#pragma unroll
        for (int q = 1; q < R; q++) {
            u32 k1 = gl::bitrev((u32)q, A);
            u32 e = ((u32)jj * k1) << (ROOT_LOG - LM);
            v[q] = gl::mul(v[q], root_pow(wlo, whi, e));
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
            u32 e = ((u32)jj * k1) << (ROOT_LOG - LM);
            v[q] = gl::mul(v[q], root_pow(wlo, whi, 0u - e));
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

// ---- carry helpers ----

// S_j = sum_t c[4j+t] * 2^(16t) = L + 2^64 H for limb j (coefficients past
// n_coeffs are zero).
GL_HD void limb_sum(const u64 *__restrict__ c, size_t n_coeffs, size_t j, u64 &L, u64 &H) {
    size_t i0 = 4 * j;
    u64 c0 = i0 + 0 < n_coeffs ? c[i0 + 0] : 0;
    u64 c1 = i0 + 1 < n_coeffs ? c[i0 + 1] : 0;
    u64 c2 = i0 + 2 < n_coeffs ? c[i0 + 2] : 0;
    u64 c3 = i0 + 3 < n_coeffs ? c[i0 + 3] : 0;
    u64 lo = c0, hi = 0, t;
    t = c1 << 16; lo += t; hi += (c1 >> 48) + (lo < t);
    t = c2 << 32; lo += t; hi += (c2 >> 32) + (lo < t);
    t = c3 << 48; lo += t; hi += (c3 >> 16) + (lo < t);
    L = lo;
    H = hi;
}

// Carry-lookahead pairs: bit 0 = generate, bit 1 = propagate.  `lo` covers
// the lower limbs, `hi` the limbs just above; identity is 2 (g=0, p=1).
GL_HD u32 gp_combine(u32 lo, u32 hi) {
    u32 g = (hi & 1) | ((hi >> 1) & lo & 1);
    u32 p = (hi >> 1) & (lo >> 1) & 1;
    return g | (p << 1);
}

}  // namespace gnt
