// Goldilocks field arithmetic, p = 2^64 - 2^32 + 1, for gfx950 (and the host,
// so the same code is unit-tested on the CPU).
//
// Elements are canonical u64 in [0, p).  Products are built from 32x32->64
// pieces (v_mad_u64_u32 on CDNA) and reduced with 2^64 = 2^32 - 1 (mod p) and
// 2^96 = -1 (mod p).  The reduction identity is the one the reference's
// `reduce128` uses (crates/bignum/src/ntt.rs:218-252); the instruction
// sequence here is written for a 32-bit VALU with carry flags.
//
// Roots of unity: 2 has multiplicative order 192 (2^96 = -1), so 2^(192/m) is
// a primitive m-th root for m | 64.  This library picks its 2^32-th root r
// with r^(2^26) = 8, so that every transform's small-order roots are powers
// of two and their multiplications are shifts.  (The reference uses
// 7^((p-1)/n); products are independent of that choice.)
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define GL_HD __host__ __device__ __forceinline__
#else
#define GL_HD static inline
#endif

namespace gl {

typedef uint64_t u64;
typedef uint32_t u32;

constexpr u64 P = 0xFFFFFFFF00000001ull;
constexpr u64 EPS = 0xFFFFFFFFull;  // 2^64 mod p

// (a + b) mod p for a, b in [0, p).  s = a + b; when the add carried or
// s >= p the answer is s - p = s + EPS (mod 2^64).  The correction is a 32-bit
// mask (EPS < 2^32), so it costs one select and one add, not a 64-bit select.
GL_HD u64 add(u64 a, u64 b) {
    u64 s;
    bool c = __builtin_add_overflow(a, b, &s);
    u32 m = (c | (s >= P)) ? (u32)EPS : 0u;
    return s + (u64)m;
}

// 64-bit a - b as two 32-bit subtracts with borrow (v_sub_co_u32 /
// v_subb_co_u32); *borrow receives the borrow out of bit 63.  Written with
// the carry builtins because a 64-bit __builtin_sub_overflow is lowered to the
// subtract plus a 64-bit compare that recomputes the borrow.
GL_HD u64 sub_borrow(u64 a, u64 b, u32 *borrow) {
    unsigned b1, b2;
    const unsigned lo = __builtin_subc((unsigned)a, (unsigned)b, 0u, &b1);
    const unsigned hi = __builtin_subc((unsigned)(a >> 32), (unsigned)(b >> 32), b1, &b2);
    *borrow = b2;
    return ((u64)hi << 32) | lo;
}

// The device form of the reduction takes the borrow of lo - hh from
// v_subbrev_co_u32 and the carry of t + hl * EPS from v_mad_u64_u32's
// carry-out, as lane masks.  The compiler expresses neither: it recomputes a
// borrow or a carry with a separate 64-bit compare, and forms hl * EPS and the
// 64-bit add in two instructions (a general multiply goes from 21 to 18
// instructions).  Each asm block keeps its masks to itself, declares SCC
// clobbered where it writes it, and puts every VALU write of a mask two wait
// states (s_nop 1) ahead of the VALU instruction that reads it.  The host form
// below the #else is the reference the CPU tests check; the device and host
// forms are compared on edge values by gnt_calib's field check, and whole
// products bit for bit by the GPU tests.
#if defined(__HIP_DEVICE_COMPILE__)
#define GL_DEVICE_ASM 1
#endif

// (a - b) mod p for a, b in [0, p).  On borrow add p, i.e. subtract EPS
// (cannot underflow again: a - b + 2^64 >= 2^64 - p + 1 > EPS).
#ifdef GL_DEVICE_ASM
// d - B * EPS for the borrow B of d = a - b, as (d_hi - B) 2^32 + (d_lo + B):
// the low word takes B as its carry-in, and the high word loses B unless
// that add carried out of the low word (then d_lo + 1 wrapped to 0 and the
// 2^32 it carried cancels the -2^32), i.e. it loses B & ~carry, formed on
// the scalar unit.  Four vector instructions instead of the five of a
// selected EPS subtracted with borrow.
__device__ __forceinline__ u64 sub(u64 a, u64 b) {
    u32 lo, hi;
    u64 b1, bw, c, m;
    asm("v_sub_co_u32_e64 %0, %2, %6, %8\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %1, %3, %7, %9, %2\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %0, %4, %0, 0, %3\n\t"
        "s_andn2_b64 %5, %3, %4\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32_e64 %1, %2, %1, 0, %5"
        : "=&v"(lo), "=v"(hi), "=&s"(b1), "=&s"(bw), "=&s"(c), "=&s"(m)
        : "v"((u32)a), "v"((u32)(a >> 32)), "v"((u32)b), "v"((u32)(b >> 32))
        : "scc");
    return ((u64)hi << 32) | lo;
}
#else
GL_HD u64 sub(u64 a, u64 b) {
    u32 bw, bw2;
    const u64 d = sub_borrow(a, b, &bw);
    return sub_borrow(d, bw ? EPS : 0u, &bw2);
}
#endif

GL_HD u64 neg(u64 a) { return sub(0, a); }

#ifdef GL_DEVICE_ASM
// (t + hl * EPS) mod p for any t < 2^64, hl < 2^32: s = t + hl * EPS with its
// carry c (< 2^65 - 2^33); when c, s + EPS < p; otherwise s >= p wraps to
// s - p by the same + EPS.
__device__ __forceinline__ u64 add_hl_eps(u64 t, u32 hl) {
    // s = t + hl * EPS (carry c); g = c or s >= p; s += g * EPS mod 2^64.
    // The scalar OR of the two masks writes SCC, which the compiler may hold
    // live across the block (its own selects in single-lane regions), so SCC
    // is declared clobbered.  The first instruction reads every vector input,
    // so nothing is early clobbered but the masks.
    u64 s, c, g, junk;
    u32 f;
    asm("v_mad_u64_u32 %0, %1, %5, -1, %6\n\t"
        "v_cmp_lt_u64_e64 %2, %7, %0\n\t"
        "s_or_b64 %2, %2, %1\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32_e64 %4, 0, 1, %2\n\t"
        "v_mad_u64_u32 %0, %3, %4, -1, %0"
        : "=v"(s), "=&s"(c), "=&s"(g), "=&s"(junk), "=v"(f)
        : "v"(hl), "v"(t), "s"(P - 1)
        : "scc");
    return s;
}

// lo - hh as a value in [0, 2^64) congruent mod p: on borrow subtract EPS
// (lo - hh + 2^64 - EPS = lo - hh + p, in [p - hh, p)), the EPS taken off as
// in sub: the borrow carried into the low word, B & ~carry out of the high.
__device__ __forceinline__ u64 sub_hh(u64 lo, u32 hh) {
    u32 r0, r1;
    u64 m, c, m2;
    asm("v_sub_co_u32_e64 %0, %2, %5, %7\n\t"
        "s_nop 1\n\t"
        "v_subbrev_co_u32_e64 %1, %2, 0, %6, %2\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32_e64 %0, %3, %0, 0, %2\n\t"
        "s_andn2_b64 %4, %2, %3\n\t"
        "s_nop 1\n\t"
        "v_subbrev_co_u32_e64 %1, %3, 0, %1, %4"
        : "=&v"(r0), "=v"(r1), "=&s"(m), "=&s"(c), "=&s"(m2)
        : "v"((u32)lo), "v"((u32)(lo >> 32)), "v"(hh)
        : "scc");
    return ((u64)r1 << 32) | r0;
}
#endif

// Reduce hi*2^64 + lo (any 128-bit value) to [0, p):
//   x = lo + (hh*2^32 + hl)*2^64 = lo - hh + hl*(2^32 - 1)  (mod p)
GL_HD u64 reduce128(u64 lo, u64 hi) {
#ifdef GL_DEVICE_ASM
    return add_hl_eps(sub_hh(lo, (u32)(hi >> 32)), (u32)hi);
#else
    u64 hh = hi >> 32;
    u64 hl = hi & 0xFFFFFFFFull;
    u32 bw, bw2;
    u64 t0 = sub_borrow(lo, hh, &bw);             // on borrow: lo - hh + p, no second borrow
    t0 = sub_borrow(t0, bw ? EPS : 0u, &bw2);
    u64 t1 = (hl << 32) - hl;                   // hl * EPS < p
    u64 s;
    bool c = __builtin_add_overflow(t0, t1, &s);    // wrapped: true value s + 2^64 = s + EPS (mod p)
    u32 m = (c | (s >= P)) ? (u32)EPS : 0u;
    return s + (u64)m;                          // lands in [0, p) in both cases
#endif
}

// Full 64x64 -> 128 product from four 32x32 -> 64 pieces.
GL_HD void mul_wide(u64 a, u64 b, u64 &lo, u64 &hi) {
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 p00 = (u64)a0 * b0;
    u64 mid = (u64)a0 * b1 + (p00 >> 32);       // < 2^64
    u64 mid2 = (u64)a1 * b0 + (u32)mid;          // < 2^64
    lo = (mid2 << 32) | (u32)p00;
    hi = (u64)a1 * b1 + (mid >> 32) + (mid2 >> 32);
}

GL_HD u64 mul(u64 a, u64 b) {
    u64 lo, hi;
    mul_wide(a, b, lo, hi);
    return reduce128(lo, hi);
}

// x * 2^f mod p for 0 < f < 64: the shifted value is hi*2^64 + lo, hi < 2^f.
GL_HD u64 shl_reduce(u64 x, int f) {
#ifdef GL_DEVICE_ASM
    // f <= 32: the high word is below 2^32, so nothing is subtracted
    if (f <= 32) return add_hl_eps(x << f, (u32)(x >> (64 - f)));
#endif
    return reduce128(x << f, x >> (64 - f));
}

// x * 2^72 mod p with one reduction instead of two: with x = xh*2^32 + xl,
// 2^96 = -1 and 2^64 = 2^32 - 1 give x*2^72 = xl*2^40 - (xl + xh)*2^8, and
// with y = xl*2^8 = yh*2^32 + yl that is A*2^32 - B for A = yl + yh (33 bits)
// and B = yh + (xl + xh)*2^8 (42 bits); A's bit 32 is 2^64 = EPS.
GL_HD u64 mul_2_72(u64 x) {
    const u32 xl = (u32)x, xh = (u32)(x >> 32);
    const u32 yl = xl << 8, yh = xl >> 24;
    unsigned ca;
    const unsigned al = __builtin_addc(yl, yh, 0u, &ca);
    const u64 B = ((((u64)xl + xh) << 8)) + yh;
    const u64 V0 = ((u64)al << 32) | (ca ? (u32)EPS : 0u);
#ifdef GL_DEVICE_ASM
    const u64 v = sub(V0, B);               // V0 - B, or on borrow V0 - B + p (V0 - B > -p)
    const u32 g = v >= P;
    return v + (u64)g * EPS;                // v >= p wraps to v - p
#else
    u32 bw, bw2;
    u64 v = sub_borrow(V0, B, &bw);         // on borrow add p: subtract EPS (V0 - B > -p)
    v = sub_borrow(v, bw ? EPS : 0u, &bw2);
    return v >= P ? v + EPS : v;            // wraps to v - p
#endif
}

// 2^e mod p for 0 <= e < 192, as a plain constant.
GL_HD u64 pow2_const(int e) {
    // 2^e for e < 64 is exact; beyond, reduce by doubling.
    u64 r = 1;
    for (int i = 0; i < e; i++) r = add(r, r);
    return r;
}

// x * 2^e mod p, 0 <= e < 192.  With e a compile-time constant (after
// unrolling) every branch folds away.
GL_HD u64 mul_pow2(u64 x, int e) {
    bool negate = e >= 96;
    int f = negate ? e - 96 : e;
    u64 r;
    if (f == 0) r = x;
    else if (f == 72) r = mul_2_72(x);
    else if (f < 64) r = shl_reduce(x, f);
    else r = shl_reduce(shl_reduce(x, f - 48), 48);
    return negate ? neg(r) : r;
}

GL_HD u64 pow(u64 base, u64 exp) {
    u64 acc = 1;
    while (exp) {
        if (exp & 1) acc = mul(acc, base);
        exp >>= 1;
        if (exp) base = mul(base, base);
    }
    return acc;
}

GL_HD u64 inv(u64 a) { return pow(a, P - 2); }

GL_HD u32 bitrev(u32 x, int bits) {
    u32 r = 0;
    for (int i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

// ---------------------------------------------------------------------------
// Register-resident small DFTs.  R in {2, 4, 8, 16, 32, 64}; omega_R = 2^(192/R).
// ---------------------------------------------------------------------------

// The butterflies are unrolled by template recursion (butterfly K of the
// level of blocks of M: block b = (K / h) * M, offset j = K % h, h = M / 2), so
// every register index and twiddle exponent is a compile-time constant
// whatever the unroller makes of the inline-assembly reduction.

// One level of the forward DIF: blocks of M, butterflies K .. R/2 - 1.
template <int R, int M, int K = 0>
GL_HD void dif_level(u64 *v) {
    if constexpr (K < R / 2) {
        constexpr int h = M / 2, b = (K / h) * M, j = K % h;
        const u64 u = v[b + j], w = v[b + j + h];
        v[b + j] = add(u, w);
        v[b + j + h] = mul_pow2(sub(u, w), (192 / M) * j);
        dif_level<R, M, K + 1>(v);
    }
}

// Forward DIF: natural-order input, output register q holds frequency
// bitrev_R(q).  Sum_i v[i] * omega_R^(i*k).  M0 < R starts at the level of
// blocks of M0 (the levels above done by the caller).
template <int R, int M0 = R>
GL_HD void dft_dif(u64 *v) {
    if constexpr (M0 >= 2) {
        dif_level<R, M0>(v);
        dft_dif<R, M0 / 2>(v);
    }
}

// v[j + R/2] = v[j] * omega_R^j for j >= J.
template <int R, int J = 0>
GL_HD void upper_from_lower(u64 *v) {
    if constexpr (J < R / 2) {
        v[J + R / 2] = mul_pow2(v[J], (192 / R) * J);
        upper_from_lower<R, J + 1>(v);
    }
}

// dft_dif<R> of v whose upper half v[R/2 ..] is zero: the first level's
// butterflies reduce to v[j + R/2] = v[j] * omega_R^j (v[R/2 ..] not read).
template <int R>
GL_HD void dft_dif_upper_zero(u64 *v) {
    upper_from_lower<R>(v);
    dft_dif<R, R / 2>(v);
}

// One level of the inverse DIT: blocks of M, butterflies K .. R/2 - 1.
template <int R, int M, int K = 0>
GL_HD void dit_level(u64 *v) {
    if constexpr (K < R / 2) {
        constexpr int h = M / 2, b = (K / h) * M, j = K % h;
        constexpr int e = (192 - (192 / M) * j) % 192;
        const u64 u = v[b + j];
        // 2^e = -2^(e-96) for e >= 96: multiply by 2^(e-96) and swap the
        // butterfly's add and subtract instead of negating
        if constexpr (e >= 96) {
            const u64 w = mul_pow2(v[b + j + h], e - 96);
            v[b + j + h] = add(u, w);  // the add first, as below: fewer live registers
            v[b + j] = sub(u, w);
        } else {
            const u64 w = mul_pow2(v[b + j + h], e);
            v[b + j] = add(u, w);
            v[b + j + h] = sub(u, w);
        }
        dit_level<R, M, K + 1>(v);
    }
}

// Inverse DIT (unscaled): register q holds frequency bitrev_R(q); output is
// natural order, v[i] = Sum_k V[k] * omega_R^(-i*k)  (R times the inverse DFT).
template <int R, int M = 2>
GL_HD void dft_dit_inv(u64 *v) {
    if constexpr (M <= R) {
        dit_level<R, M>(v);
        dft_dit_inv<R, M * 2>(v);
    }
}

}  // namespace gl
