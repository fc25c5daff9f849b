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

// Twiddle tables (device or host copies):
//   wlo[i] = r^i, whi[i] = r^(i 2^16) for i < 2^16   (r the library's 2^32-th root)
//   wsm[2^LM + e] = omega_{2^LM}^e for e < 2^LM, LM <= TILE_LOG   (compact, per size)
struct Tw {
    const u64 *wlo, *whi, *wsm;
};
constexpr int WSM_SIZE = 2 << TILE_LOG;

// omega_{2^32}^e from the two-level table.
GL_HD u64 root_pow(const Tw &tw, u32 e) { return gl::mul(tw.wlo[e & ((1u << TW_SPLIT) - 1)], tw.whi[e >> TW_SPLIT]); }

// Entry i = n*s + j of the twiddle table of a pass with block size
// M = 2^log_m and radix 2^log_r (row n, column j of a block, s = M / R):
// omega_M^(j * bitrev_R(n)), the twiddle tile_store_step steps to, or its
// inverse, times 2^scale_e.  The table is indexed by an element's global
// position mod M.  The inverse of pass 0 (M = N) of a plan of two or more
// passes carries the inverse transform's 1/N = 2^(192 - log N) this way.
GL_HD u64 pass_table_entry(const Tw &tw, u32 i, int log_m, int log_r, bool inv, int scale_e = 0) {
    const int log_s = log_m - log_r;
    const u32 n = i >> log_s, j = i & ((1u << log_s) - 1);
    u32 e = (u32)(((u64)j * gl::bitrev(n, log_r)) << (ROOT_LOG - log_m));
    if (inv) e = 0u - e;
    const u64 w = root_pow(tw, e);
    return scale_e ? gl::mul_pow2(w, scale_e) : w;
}

// Sub-steps of an R-point column DFT: radix <= 8 each (8 values in registers
// keeps a thread within 64 VGPRs), as even as possible, larger ones first.
template <int LOGR, bool L16 = false>
struct Sub {
    // L16 with LOGR = 10 or 13: the column DFT ends in one radix-16 sub-step
    // instead of two radix-4 ones, which removes the twiddle stage between
    // them; the last sub-step has no twiddle after it, so its 16 values fit
    // the register budget where the kernel holds little else (with LOGR = 7
    // the fused multiply kernel spills two registers, so 7 keeps 3+2+2).
    static constexpr bool last16 = L16 && LOGR >= 10 && LOGR % 3 == 1;
    static constexpr int count = last16 ? (LOGR - 1) / 3 : (LOGR + 2) / 3;
    static constexpr int log(int i) {
        return last16 ? (i == count - 1 ? 4 : 3) : LOGR / count + (i < LOGR % count ? 1 : 0);
    }
    // block size (log) seen by sub-step i inside the R-point column DFT
    static constexpr int log_m(int i) { return i == 0 ? LOGR : log_m(i - 1) - log(i - 1); }
    // The shift stage: the one sub-step (not the last) whose block has at most
    // 64 elements, so its twiddles omega_{2^LM}^e = 2^(e 192/2^LM) are powers
    // of two; -1 when there is none.
    static constexpr int shift_stage(int i = 0) {
        return i >= count - 1 ? -1 : (log_m(i) <= 6 ? i : shift_stage(i + 1));
    }
};

constexpr int bitrev_c(int x, int bits) { return bits == 0 ? 0 : ((x & 1) << (bits - 1)) | bitrev_c(x >> 1, bits - 1); }

// Tile shape, fixed at compile time: R = 2^LOGR rows, 2^LOGC columns;
// STRIDED: columns are adjacent in memory (a pass with column stride > 1);
// otherwise the pass is the contiguous last one (rows adjacent).  PAD: a
// contiguous tile of 4 columns (the fused multiply's half-width tiles) whose
// rows are padded in LDS (tix); the tile then takes tile_words words of LDS.
template <int LOGR_, int LOGC_, bool STRIDED_, bool L16_ = false, bool PAD_ = false>
struct Geom {
    using S = Sub<LOGR_, L16_>;  // sub-steps of the column DFT
    static constexpr int LOGR = LOGR_;
    static constexpr int log_c = LOGC_;
    static constexpr bool strided = STRIDED_;
    static constexpr int tile_n = 1 << (LOGR_ + LOGC_);
    static constexpr bool pad4 = PAD_ && !STRIDED_ && LOGC_ == 2;                 // 1024 x 4 (radix 2^10)
    static constexpr bool pad8 = PAD_ && !STRIDED_ && LOGC_ == 3 && LOGR_ == 9;  // 512 x 8
    static constexpr bool padded = pad4 || pad8;
    // Words of padding before row n (tix): rows of 4 words, 2 after every 8
    // rows and 32 more after every 128; rows of 8 words, 3 after every 4 rows
    // and 16 more after every 64.
    static constexpr int row_pad(int n) {
        return pad4 ? 2 * (n >> 3) + 32 * (n >> 7) : pad8 ? 3 * (n >> 2) + 16 * (n >> 6) : 0;
    }
    static constexpr int tile_words = tile_n + row_pad((1 << LOGR_) - 1);
    int log_m, log_s;  // pass block size, column stride (logs)
    size_t cg0;        // first column of the tile
};

// Shift stage of a tile shape's column DFT (Sub::shift_stage).  Its
// twiddle omega_{2^LM}^(jj * k) depends on the task's offset jj inside its
// block of 2^LM rows and on the register k; when every wave's tasks share one
// jj the twiddles are shifts by compile-time amounts in each of 2^LT branches
// taken by whole waves (a general multiply from a table otherwise).  Tasks of
// that sub-step are therefore numbered column fastest, then block, then jj,
// so that a wave's 64 tasks cover 2^log_c columns of 2^(6 - log_c) blocks.
// Strided passes with at least 4 columns per tile only: measured on a 2^24-limb
// product, it takes 5% of the VALU instructions out of those launches; in the
// contiguous last pass (rows of 4 or 8 columns, XOR-swizzled) the same change
// raised LDS bank conflicts by 60% and gained nothing, and a row swizzle that
// removes the conflicts cost more address arithmetic than the shifts saved.
template <class G>
struct ShiftTw {
    static constexpr int I = G::S::shift_stage();
    static constexpr int LM = I >= 0 ? G::S::log_m(I) : 0;
    static constexpr int A = I >= 0 ? G::S::log(I) : 0;
    static constexpr int LT = LM - A;
    static constexpr int LOG_NBLK = G::LOGR - LM;
    static constexpr bool on = I >= 0 && G::strided && G::log_c >= 2 && G::log_c + LOG_NBLK >= 6;
};

template <class G>
GL_HD G make_geom(const PassDesc &d, size_t tile_index) {
    G g;
    g.log_m = d.log_m;
    g.log_s = d.log_m - G::LOGR;
    g.cg0 = tile_index << G::log_c;
    return g;
}

// LDS position of tile element (row n, column c).  A strided pass only walks
// its tile along rows (loads, stores and the column DFT tasks all put
// consecutive lanes on consecutive columns), so it is plain row-major and a
// task's R addresses are one base plus constant offsets.  The contiguous last
// pass also walks down columns (its rows are adjacent in memory), so its
// columns are XOR-swizzled by the row.  With rows of 32 columns or more both
// walks then hit distinct banks.  With 8 columns or fewer a row covers only
// part of the banks and a column permutation cannot separate rows that share
// them: the down-column walks and the last column-DFT sub-steps conflict.
//
// gfx950's LDS, measured with SQ_LDS_BANK_CONFLICT on single-pattern
// kernels: a ds_read_b64 is served 32 lanes at a time from 32 banks of 8
// bytes, a ds_write_b64 16 lanes at a time from 16 banks of 8 bytes (a wave's
// lanes at the same bank and different addresses serialise).  The fused
// multiply's half-width tiles of 4 and of 8 columns are therefore padded
// instead (Geom::row_pad).  Rows of 4 words (1024 rows): 2 unused words after
// every 8 rows put the next 8 rows 2 banks over, so a down-column walk has at
// most 2 lanes per bank and every column-DFT walk is conflict-free; 32 more
// words after every 128 rows (a multiple of the bank count, so no bank
// changes) make rows 128 apart 576 = 9 x 64 words apart, which the paired
// ds_read2st64_b64 / ds_write2st64_b64 accesses of the first column sub-step
// need.  Rows of 8 words (512 rows): 3 words after every 4 rows make the
// down-column reads conflict-free and halve the conflicts of the tile's walks
// in a model of the banks above, and 16 more after every 64 rows keep rows 64
// apart 576 words apart for the same pairing.  A layout
// that also permutes the low row bits by higher ones removes every conflict
// but its addresses are no longer one base plus constants: the compiler
// drops the paired LDS accesses and adds address arithmetic, which on a
// 2^24-limb product cost more than the conflicts did.
template <class G>
GL_HD int tix(const G &, int n, int c) {
    if constexpr (G::strided) return (n << G::log_c) + c;
    else if constexpr (G::padded) return (n << G::log_c) + c + G::row_pad(n);
    else return (n << G::log_c) + (c ^ (n & ((1 << G::log_c) - 1)));
}

// Offset of row base + k from row base in a padded tile, for the rows of a
// column-DFT task: base = (blk << LM) + jj with jj < 2^LT and k = u << LT (u <
// 2^(LM - LT)), so base and k have no bits in common and, for every shift s,
// (base + k) >> s = (base >> s) + (k >> s) (when LM <= s both sides are in
// one aligned 2^s block).  The task then addresses its rows as one base plus
// compile-time constants.
template <class G>
GL_HD constexpr int prow_off(int k) {
    return (k << G::log_c) + G::row_pad(k);
}

// Global index of tile element (row n, tile column c).  Transforms hold at
// most 2^MAX_LOG_N = 2^32 elements, so indices fit in 32 bits.  A pass whose
// block is the whole 2^32-element array has block index b = 0; its shift count
// is taken mod 32 (what the hardware shift does anyway) so that the 32 is not
// an undefined shift.
GL_HD u32 tile_addr(size_t cg0, int n, int c, int log_m, int log_s) {
    u32 cg = (u32)cg0 + (u32)c;
    u32 b = cg >> log_s;
    u32 j = cg & ((1u << log_s) - 1);
    return (b << (log_m & 31)) + ((u32)n << log_s) + j;
}
static_assert(MAX_LOG_N <= 32, "32-bit element indices");

// Tile element index -> (row, column): columns fastest when columns are
// adjacent in memory (strided passes), rows fastest in the contiguous last pass.
template <class G>
GL_HD void tile_rc(int idx, int &n, int &c) {
    if (G::strided) { c = idx & ((1 << G::log_c) - 1); n = idx >> G::log_c; }
    else { n = idx & ((1 << G::LOGR) - 1); c = idx >> G::LOGR; }
}

// Coefficient i of an operand packed 16 bits per coefficient (ntt.rs:640-665).
GL_HD u64 coeff_from_limbs(const u64 *__restrict__ limbs, size_t n_limbs, size_t i) {
    size_t l = i >> 2;
    if (l >= n_limbs) return 0;
    return (limbs[l] >> (16 * (i & 3))) & 0xFFFFull;
}

// A value every lane of the wave holds alike, moved to a scalar register so
// that branches on it are taken by the whole wave.
GL_HD int wave_uniform_int(int x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_readfirstlane(x);
#else
    return x;
#endif
}

// v[q] *= 2^(sign * SH * J * bitrev_A(q)) for q = 1 .. 2^A - 1 with J fixed at
// compile time: shifts, negations and one-reduction multiplies only.
template <int SH, int A, int J, bool INV>
GL_HD void shift_twiddles(u64 *v) {
#pragma unroll
    for (int q = 1; q < (1 << A); q++) {
        const int e = (SH * J * bitrev_c(q, A)) % 192;
        v[q] = gl::mul_pow2(v[q], INV ? (192 - e) % 192 : e);
    }
}

// The same with J = jj, a wave-uniform value below NJ, one branch per value.
template <int SH, int A, int NJ, bool INV, int J = 0>
GL_HD void shift_twiddles_by(u64 *v, int jj) {
    if constexpr (J < NJ) {
        if (jj == J) shift_twiddles<SH, A, J, INV>(v);
        else shift_twiddles_by<SH, A, NJ, INV, J + 1>(v, jj);
    }
}

// Task L of sub-step I: column c, offset jj inside its block of 2^LM rows,
// block blk.  Column fastest, then jj, then block; for the shift stage of a
// shape with ShiftTw on, column, then block, then jj (whole waves share jj).
template <class G, int I>
GL_HD void task_coords(int L, int &c, int &jj, int &blk) {
    constexpr int LM = G::S::log_m(I);
    constexpr int LT = LM - G::S::log(I);
    c = L & ((1 << G::log_c) - 1);
    const int rest = L >> G::log_c;
    if constexpr (ShiftTw<G>::on && I == ShiftTw<G>::I) {
        blk = rest & ((1 << ShiftTw<G>::LOG_NBLK) - 1);
        jj = rest >> ShiftTw<G>::LOG_NBLK;
    } else {
        jj = rest & ((1 << LT) - 1);
        blk = rest >> LT;
    }
}

// Sub-step I of the forward R-point column DFT for task L (DIF): an r-point
// DFT over rows base + u*t of column c, then the twiddle omega_{2^LM}^(jj*k1)
// (none after the last sub-step; the pass twiddle is applied on the way out).
// UPPER_ZERO (sub-step 0 only): the tile's rows R/2 .. R-1 are zero and not
// read (a packing pass whose operand fills at most half the transform).
template <class G, int I, bool UPPER_ZERO = false>
GL_HD void fwd_task(u64 *tile, const G &g, int L, const Tw &tw) {
    constexpr int LOGR = G::LOGR;
    constexpr int A = G::S::log(I);
    constexpr int R = 1 << A;
    constexpr int LM = G::S::log_m(I);
    constexpr int LT = LM - A;
    constexpr bool LAST = (I == G::S::count - 1);
    static_assert(!UPPER_ZERO || (I == 0 && R >= 2), "upper-zero input: first sub-step only");
    int c, jj, blk;
    task_coords<G, I>(L, c, jj, blk);
    const int base = (blk << LM) + jj;
    u64 v[R];
    if constexpr (G::padded) {
        static_assert(!UPPER_ZERO, "padded tiles belong to contiguous passes, which never pack");
        const u64 *p = tile + tix(g, base, c);
#pragma unroll
        for (int u = 0; u < R; u++) v[u] = p[prow_off<G>(u << LT)];
        gl::dft_dif<R>(v);
    } else if constexpr (UPPER_ZERO) {
#pragma unroll
        for (int u = 0; u < R / 2; u++) v[u] = tile[tix(g, base + (u << LT), c)];
        gl::dft_dif_upper_zero<R>(v);
    } else {
#pragma unroll
        for (int u = 0; u < R; u++) v[u] = tile[tix(g, base + (u << LT), c)];
        gl::dft_dif<R>(v);
    }
    if constexpr (ShiftTw<G>::on && I == ShiftTw<G>::I) {
        shift_twiddles_by<192 / (1 << LM), A, (1 << LT), false>(v, wave_uniform_int(jj));
    } else if constexpr (!LAST) {
        const u64 *w = tw.wsm + (1 << LM);
#pragma unroll
        for (int q = 1; q < R; q++) v[q] = gl::mul(v[q], w[(u32)jj * gl::bitrev((u32)q, A)]);
    }
    if constexpr (G::padded) {
        u64 *p = tile + tix(g, base, c);
#pragma unroll
        for (int q = 0; q < R; q++) p[prow_off<G>(q << LT)] = v[q];
    } else {
#pragma unroll
        for (int q = 0; q < R; q++) tile[tix(g, base + (q << LT), c)] = v[q];
    }
}

// Inverse of fwd_task<LOGR, I>: undo the twiddle, then the inverse r-point
// DFT (DIT, bit-reversed input, natural output).  Unscaled.
template <class G, int I>
GL_HD void inv_task(u64 *tile, const G &g, int L, const Tw &tw) {
    constexpr int LOGR = G::LOGR;
    constexpr int A = G::S::log(I);
    constexpr int R = 1 << A;
    constexpr int LM = G::S::log_m(I);
    constexpr int LT = LM - A;
    constexpr bool LAST = (I == G::S::count - 1);
    int c, jj, blk;
    task_coords<G, I>(L, c, jj, blk);
    const int base = (blk << LM) + jj;
    u64 *p = tile + tix(g, base, c);
    u64 v[R];
#pragma unroll
    for (int q = 0; q < R; q++) v[q] = G::padded ? p[prow_off<G>(q << LT)] : tile[tix(g, base + (q << LT), c)];
    if constexpr (ShiftTw<G>::on && I == ShiftTw<G>::I) {
        shift_twiddles_by<192 / (1 << LM), A, (1 << LT), true>(v, wave_uniform_int(jj));
    } else if constexpr (!LAST) {
        const u64 *w = tw.wsm + (1 << LM);
#pragma unroll
        for (int q = 1; q < R; q++) {
            u32 e = ((u32)jj * gl::bitrev((u32)q, A)) & ((1u << LM) - 1);
            v[q] = gl::mul(v[q], w[((1u << LM) - e) & ((1u << LM) - 1)]);  // omega^-e
        }
    }
    gl::dft_dit_inv<R>(v);
#pragma unroll
    for (int u = 0; u < R; u++) {
        if constexpr (G::padded) p[prow_off<G>(u << LT)] = v[u];
        else tile[tix(g, base + (u << LT), c)] = v[u];
    }
}

// ---- tile-level routines ----

// Two adjacent tile columns of one row, as one 16-byte access in HBM and in
// LDS (a strided tile is plain row-major, and an even column's global index
// and tile index are even, so the pair is 16-byte aligned in both).  With
// two 8-byte LDS accesses per lane instead, consecutive lanes are 16 bytes
// apart and every access conflicts 2-way.
struct alignas(16) U64Pair {
    u64 v[2];
};

// Load a tile from x, or (limbs != nullptr) pack it from the operand's limbs.
template <class X, class G>
GL_HD void tile_load(u64 *tile, const G &g, const u64 *__restrict__ x, const u64 *__restrict__ limbs,
                     size_t n_limbs) {
    if constexpr (G::strided && G::log_c >= 1) {
        if (!limbs) {
            GNT_FOR(X, k, G::tile_n / 2) {
                const int n = (2 * k) >> G::log_c, c = (2 * k) & ((1 << G::log_c) - 1);
                *reinterpret_cast<U64Pair *>(tile + tix(g, n, c)) =
                    *reinterpret_cast<const U64Pair *>(x + tile_addr(g.cg0, n, c, g.log_m, g.log_s));
            }
            X::sync();
            return;
        }
    }
    GNT_FOR(X, idx, G::tile_n) {
        int n, c;
        tile_rc<G>(idx, n, c);
        u32 a = tile_addr(g.cg0, n, c, g.log_m, g.log_s);
        tile[tix(g, n, c)] = limbs ? coeff_from_limbs(limbs, n_limbs, a) : x[a];
    }
    X::sync();
}

template <class X, class G>
GL_HD void tile_store(const u64 *tile, const G &g, u64 *__restrict__ x) {
    if constexpr (G::strided && G::log_c >= 1) {
        GNT_FOR(X, k, G::tile_n / 2) {
            const int n = (2 * k) >> G::log_c, c = (2 * k) & ((1 << G::log_c) - 1);
            *reinterpret_cast<U64Pair *>(x + tile_addr(g.cg0, n, c, g.log_m, g.log_s)) =
                *reinterpret_cast<const U64Pair *>(tile + tix(g, n, c));
        }
        return;
    }
    GNT_FOR(X, idx, G::tile_n) {
        int n, c;
        tile_rc<G>(idx, n, c);
        x[tile_addr(g.cg0, n, c, g.log_m, g.log_s)] = tile[tix(g, n, c)];
    }
}

// Store a tile with every element multiplied by T[a mod 2^log_t], a its
// global position: a forward pass's own table, or the table of the inverse
// pass that reads this output next (applied here so that pass loads plain
// data).  One multiply per element instead of the two of stepped twiddles.
template <class X, class G>
GL_HD void tile_store_tw(const u64 *tile, const G &g, u64 *__restrict__ x, const u64 *__restrict__ T, int log_t) {
    const u32 mt = (1u << log_t) - 1;
    if constexpr (G::strided && G::log_c >= 1) {
        GNT_FOR(X, k, G::tile_n / 2) {
            const int n = (2 * k) >> G::log_c, c = (2 * k) & ((1 << G::log_c) - 1);
            const u32 a = tile_addr(g.cg0, n, c, g.log_m, g.log_s);
            const U64Pair w = *reinterpret_cast<const U64Pair *>(T + (a & mt));
            U64Pair t = *reinterpret_cast<const U64Pair *>(tile + tix(g, n, c));
            t.v[0] = gl::mul(t.v[0], w.v[0]);
            t.v[1] = gl::mul(t.v[1], w.v[1]);
            *reinterpret_cast<U64Pair *>(x + a) = t;
        }
        return;
    }
    GNT_FOR(X, idx, G::tile_n) {
        int n, c;
        tile_rc<G>(idx, n, c);
        const u32 a = tile_addr(g.cg0, n, c, g.log_m, g.log_s);
        x[a] = gl::mul(tile[tix(g, n, c)], T[a & mt]);
    }
}

// The table a tile's store applies (T null: a plain store).
struct StoreTw {
    const u64 *T;
    int log_t;
};

template <class X, class G>
GL_HD void tile_store_any(const u64 *tile, const G &g, u64 *__restrict__ x, const StoreTw &st) {
    if (st.T) tile_store_tw<X, G>(tile, g, x, st.T, st.log_t);
    else tile_store<X, G>(tile, g, x);
}

// Pass twiddles of a strided pass without a table, stepped in registers on
// the way out (forward) or in (inverse).  Element (row n, column j of the
// block) takes omega_M^(j * kc), kc = bitrev_R(n).  A task owns one column
// and the STEP_ROWS rows whose kc run over [STEP_ROWS * grp, +STEP_ROWS), so
// the twiddle advances by omega_M^j from one to the next: two table lookups
// and one multiply per element besides the element's own.  Consecutive tasks
// take consecutive columns, so every LDS access walks a row and every global
// access is a contiguous run of 2^log_c elements.
constexpr int STEP_ROWS = 8;

template <class G>
GL_HD void step_task(const G &g, int L, int &c, u32 &kc0, u32 &j) {
    c = L & ((1 << G::log_c) - 1);
    kc0 = (u32)(L >> G::log_c) * STEP_ROWS;
    j = (u32)((g.cg0 + (size_t)c) & (((size_t)1 << g.log_s) - 1));
}

template <class X, class G>
GL_HD void tile_store_step(const u64 *tile, const G &g, u64 *__restrict__ x, const Tw &tw) {
    static_assert(G::strided && (1 << G::LOGR) % STEP_ROWS == 0, "stepped twiddles: strided passes only");
    const int sh = ROOT_LOG - g.log_m;
    GNT_FOR(X, L, G::tile_n / STEP_ROWS) {
        int c;
        u32 kc0, j;
        step_task(g, L, c, kc0, j);
        u64 w = root_pow(tw, (j * kc0) << sh);
        const u64 u = root_pow(tw, j << sh);
#pragma unroll
        for (int k = 0; k < STEP_ROWS; k++) {
            const int n = (int)gl::bitrev(kc0 + k, G::LOGR);
            x[tile_addr(g.cg0, n, c, g.log_m, g.log_s)] = gl::mul(tile[tix(g, n, c)], w);
            w = gl::mul(w, u);
        }
    }
}

// Inverse: the loaded values times omega_M^(-j * kc) * 2^scale_e, into the
// tile (scale_e: the 1/N the inverse pass 0 carries, or 0).
template <class X, class G>
GL_HD void tile_load_step(u64 *tile, const G &g, const u64 *__restrict__ x, const Tw &tw, int scale_e) {
    static_assert(G::strided && (1 << G::LOGR) % STEP_ROWS == 0, "stepped twiddles: strided passes only");
    const int sh = ROOT_LOG - g.log_m;
    GNT_FOR(X, L, G::tile_n / STEP_ROWS) {
        int c;
        u32 kc0, j;
        step_task(g, L, c, kc0, j);
        u64 v[STEP_ROWS];
#pragma unroll
        for (int k = 0; k < STEP_ROWS; k++)
            v[k] = x[tile_addr(g.cg0, (int)gl::bitrev(kc0 + k, G::LOGR), c, g.log_m, g.log_s)];
        u64 w = root_pow(tw, 0u - ((j * kc0) << sh));
        if (scale_e) w = gl::mul_pow2(w, scale_e);
        const u64 u = root_pow(tw, 0u - (j << sh));
#pragma unroll
        for (int k = 0; k < STEP_ROWS; k++) {
            tile[tix(g, (int)gl::bitrev(kc0 + k, G::LOGR), c)] = gl::mul(v[k], w);
            w = gl::mul(w, u);
        }
    }
    X::sync();
}

template <class X, class G, int I, bool UPPER_ZERO = false>
GL_HD void tile_fwd(u64 *tile, const G &g, const Tw &tw) {
    if constexpr (I < G::S::count) {
        const int ntask = G::tile_n >> G::S::log(I);
        GNT_FOR(X, L, ntask) fwd_task<G, I, UPPER_ZERO>(tile, g, L, tw);
        X::sync();
        tile_fwd<X, G, I + 1>(tile, g, tw);
    }
}

// Pack a tile of a strided first pass (block N, so element a = n 2^log_s + j
// is coefficient a) from the operand's limbs, 16 bits per coefficient
// (ntt.rs:640-665): one task per limb, which loads it once and writes its four
// coefficients to the row; rows at or beyond `rows` are left unwritten.
template <class X, class G>
GL_HD void tile_pack(u64 *tile, const G &g, const u64 *__restrict__ limbs, size_t n_limbs, int rows) {
    static_assert(G::strided && G::log_c >= 2, "a row holds whole limbs");
    constexpr int QL = G::log_c - 2;  // limbs per row (log)
    GNT_FOR(X, k, rows << QL) {
        const int n = k >> QL, q = (k & ((1 << QL) - 1)) << 2;
        const size_t l = tile_addr(g.cg0, n, q, g.log_m, g.log_s) >> 2;
        const u64 v = l < n_limbs ? limbs[l] : 0;
        // the limb's four coefficients are adjacent words of the row (q is a
        // multiple of 4): two 16-byte LDS writes
        U64Pair lo, hi;
        lo.v[0] = v & 0xFFFFull;
        lo.v[1] = (v >> 16) & 0xFFFFull;
        hi.v[0] = (v >> 32) & 0xFFFFull;
        hi.v[1] = v >> 48;
        U64Pair *dst = reinterpret_cast<U64Pair *>(tile + tix(g, n, q));
        dst[0] = lo;
        dst[1] = hi;
    }
    X::sync();
}

// Inverse sub-steps I, I-1, ..., 0.
template <class X, class G, int I>
GL_HD void tile_inv(u64 *tile, const G &g, const Tw &tw) {
    if constexpr (I >= 0) {
        const int ntask = G::tile_n >> G::S::log(I);
        GNT_FOR(X, L, ntask) inv_task<G, I>(tile, g, L, tw);
        X::sync();
        tile_inv<X, G, I - 1>(tile, g, tw);
    }
}

// The exponent e of the 1/N = 2^e of the inverse transform (ntt.rs:376-391),
// applied by the pointwise product of a single-pass plan; a plan of two or
// more passes folds it into the inverse twiddles of pass 0 (its block is N):
// the stepped twiddles' start value or that pass's table.
GL_HD int inv_scale_e(int log_n) { return 192 - log_n; }

// tile *= other (at the same global positions, or the tile itself), times
// 2^inv_scale_e(log_n) when the tile's pass is the plan's only one
// (ntt.rs:845-853).
template <class X, class G>
GL_HD void tile_pointwise(u64 *tile, const G &g, const u64 *__restrict__ other, int log_n) {
    const int e = g.log_m == log_n ? inv_scale_e(log_n) : 0;
    GNT_FOR(X, idx, G::tile_n) {
        int n, c;
        tile_rc<G>(idx, n, c);
        const int t = tix(g, n, c);
        u64 v = tile[t];
        u64 w = other ? other[tile_addr(g.cg0, n, c, g.log_m, g.log_s)] : v;
        v = gl::mul(v, w);
        tile[t] = e ? gl::mul_pow2(v, e) : v;
    }
    X::sync();
}

// Limb sums of the finished coefficients of a tile of the LAST inverse pass
// (the pass over blocks of size N, so tile addresses are coefficient
// indices): S_j = sum_t c[4j+t] 2^(16t) = L_j + 2^64 H_j for every limb whose
// four coefficients lie in this tile, j < n_limbs.  Coefficients at or past
// n_coeffs are zero by construction of the product; they are masked anyway.
template <class X, class G>
GL_HD void tile_limb_sums(const u64 *tile, const G &g, size_t n_coeffs, size_t n_limbs, u64 *__restrict__ Ls,
                          u64 *__restrict__ Hs) {
    const int nq = G::tile_n >> 2;
    GNT_FOR(X, q, nq) {
        // four consecutive coefficients: along a row (log_s > 0) or down a column
        int n, c;
        tile_rc<G>(q << 2, n, c);
        u32 a0 = tile_addr(g.cg0, n, c, g.log_m, g.log_s);
        u32 j = a0 >> 2;
        if (j < n_limbs) {
            u64 cc[4];
            if constexpr (G::strided) {
                // four adjacent words of the row: two 16-byte LDS reads
                const U64Pair *src = reinterpret_cast<const U64Pair *>(tile + tix(g, n, c));
                const U64Pair lo = src[0], hi = src[1];
                cc[0] = lo.v[0];
                cc[1] = lo.v[1];
                cc[2] = hi.v[0];
                cc[3] = hi.v[1];
#pragma unroll
                for (int k = 0; k < 4; k++) cc[k] = (a0 + k < n_coeffs) ? cc[k] : 0;
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++) cc[k] = (a0 + k < n_coeffs) ? tile[tix(g, n + k, c)] : 0;
            }
            u64 lo = cc[0], hi = 0, t;
            t = cc[1] << 16; lo += t; hi += (cc[1] >> 48) + (lo < t);
            t = cc[2] << 32; lo += t; hi += (cc[2] >> 32) + (lo < t);
            t = cc[3] << 48; lo += t; hi += (cc[3] >> 16) + (lo < t);
            Ls[j] = lo;
            Hs[j] = hi;
        }
    }
}

// ---- pass bodies: what each kernel runs on one tile ----

// Forward pass.  With `limbs` set the tile is packed from the operand's limbs.
// A strided pass multiplies its twiddles on the way out, from its table T or
// stepped (T null).
template <class X, class G>
GL_HD void body_fwd(u64 *tile, const G &g, u64 *x, const u64 *limbs, size_t n_limbs, const Tw &tw, const u64 *T) {
    if constexpr (G::strided && G::log_c >= 2) {
        if (limbs) {
            // first pass of a strided plan (block N): rows R/2 .. R-1 hold
            // coefficients from N/2 on, zero when the operand fills at most
            // half the transform (every balanced product)
            const bool upper_zero = 4 * (u64)n_limbs <= ((u64)1 << (g.log_m - 1));
            tile_pack<X, G>(tile, g, limbs, n_limbs, upper_zero ? (1 << G::LOGR) / 2 : 1 << G::LOGR);
            if (upper_zero) tile_fwd<X, G, 0, true>(tile, g, tw);
            else tile_fwd<X, G, 0>(tile, g, tw);
        } else {
            tile_load<X, G>(tile, g, x, nullptr, 0);
            tile_fwd<X, G, 0>(tile, g, tw);
        }
    } else {
        tile_load<X, G>(tile, g, x, limbs, n_limbs);
        tile_fwd<X, G, 0>(tile, g, tw);
    }
    if constexpr (G::strided) {
        if (T) tile_store_tw<X, G>(tile, g, x, T, g.log_m);
        else tile_store_step<X, G>(tile, g, x, tw);
    } else {
        tile_store<X, G>(tile, g, x);
    }
}

// Inverse pass; SUMS: it is the last one and writes limb sums.  STEP: a
// strided pass whose inverse twiddles were not applied by the producer of its
// input (it has no table) steps them as it loads.  TSTORE: the store
// applies the table of the next inverse pass (st).  scale_e: the 1/N the
// stepped twiddles of pass 0 carry (0 for other passes).  Compile-time choices, so
// each kernel carries only the phases it runs (a kernel holding both forms
// of a phase needs registers it then spills).
template <class X, class G, bool SUMS, bool STEP, bool TSTORE>
GL_HD void body_inv(u64 *tile, const G &g, u64 *x, const Tw &tw, u64 *Ls, u64 *Hs, size_t n_coeffs,
                    size_t n_limbs, int scale_e, const StoreTw &st) {
    if constexpr (G::strided && STEP) tile_load_step<X, G>(tile, g, x, tw, scale_e);
    else tile_load<X, G>(tile, g, x, nullptr, 0);
    tile_inv<X, G, G::S::count - 1>(tile, g, tw);
    if constexpr (SUMS) tile_limb_sums<X, G>(tile, g, n_coeffs, n_limbs, Ls, Hs);
    else if constexpr (TSTORE) tile_store_tw<X, G>(tile, g, x, st.T, st.log_t);
    else tile_store<X, G>(tile, g, x);
}

// Which inverse-pass variant a pass runs (InvForm::of).
enum InvForm : int { INV_STEP = 0, INV_PLAIN = 1, INV_TSTORE = 2 };

// Last forward pass of B (from xb, or packed from limbs), pointwise with A
// (a_other at the same positions; null squares), 1/N, first inverse pass;
// result to xa, or limb sums when this is the only pass.  The last pass has
// column stride 1, so it carries no pass twiddle.
template <class X, class G>
GL_HD void body_fused(u64 *tile, const G &g, const u64 *xb, const u64 *limbs, size_t n_limbs_in, u64 *xa,
                      const u64 *a_other, int log_n, const Tw &tw, u64 *Ls, u64 *Hs, size_t n_coeffs,
                      size_t n_limbs, const StoreTw &st) {
    tile_load<X, G>(tile, g, xb, limbs, n_limbs_in);
    tile_fwd<X, G, 0>(tile, g, tw);
    tile_pointwise<X, G>(tile, g, a_other, log_n);
    tile_inv<X, G, G::S::count - 1>(tile, g, tw);
    if (Ls) tile_limb_sums<X, G>(tile, g, n_coeffs, n_limbs, Ls, Hs);
    else tile_store_any<X, G>(tile, g, xa, st);
}

// Last forward pass of A and of B (both read after their earlier passes), the
// pointwise product, 1/N and the first inverse pass, result to xa: one tile
// round trip instead of a separate pass over A.  G is a half-width tile
// (2^(TILE_LOG-1) elements) of the plan's contiguous last pass, so A's and
// B's tiles share the LDS of about one full tile: tile[0, w) holds A's,
// tile[w, 2w) B's, w = G::tile_words.  Plans of two or more passes (no pass
// twiddle on the last pass).
template <class X, class G, int I>
GL_HD void tile_fwd_pair(u64 *tile, const G &g, const Tw &tw) {
    if constexpr (I < G::S::count) {
        const int ntask = G::tile_n >> G::S::log(I);
        GNT_FOR(X, L, 2 * ntask) {
            const bool b = L >= ntask;
            fwd_task<G, I>(tile + (b ? G::tile_words : 0), g, b ? L - ntask : L, tw);
        }
        X::sync();
        tile_fwd_pair<X, G, I + 1>(tile, g, tw);
    }
}

template <class X, class G>
GL_HD void body_fused_ab(u64 *tile, const G &g, u64 *xa, const u64 *xb, int log_n, const Tw &tw,
                         const StoreTw &st) {
    u64 *ta = tile, *tb = tile + G::tile_words;
    GNT_FOR(X, idx, 2 * G::tile_n) {
        const bool b = idx >= G::tile_n;
        const int k = b ? idx - G::tile_n : idx;
        int n, c;
        tile_rc<G>(k, n, c);
        const u32 a = tile_addr(g.cg0, n, c, g.log_m, g.log_s);
        (b ? tb : ta)[tix(g, n, c)] = b ? xb[a] : xa[a];
    }
    X::sync();
    tile_fwd_pair<X, G, 0>(tile, g, tw);
    // the 1/N rides on the inverse twiddles of pass 0 (two or more passes)
    GNT_FOR(X, idx, G::tile_n) {
        int n, c;
        tile_rc<G>(idx, n, c);
        const int t = tix(g, n, c);
        ta[t] = gl::mul(ta[t], tb[t]);
    }
    X::sync();
    tile_inv<X, G, G::S::count - 1>(ta, g, tw);
    tile_store_any<X, G>(ta, g, xa, st);
}

// Half-width geometry of a full-tile contiguous pass shape (the fused A+B
// tiles); FullG itself when the shape cannot be halved (never launched then).
template <class FullG>
struct HalfGeom {
    static constexpr bool ok = !FullG::strided && FullG::log_c >= 1 && FullG::LOGR + FullG::log_c == TILE_LOG;
    // the fused multiply kernel holds few registers besides a task's values:
    // a radix-16 last sub-step fits it without spilling
    using type = Geom<FullG::LOGR, ok ? FullG::log_c - 1 : FullG::log_c, false, true, true>;
};

// LDS bytes of the fused A+B kernel for the last-pass shape FullG.
template <class FullG>
constexpr size_t fused_ab_lds_bytes() {
    return 2 * (size_t)HalfGeom<FullG>::type::tile_words * sizeof(u64);
}

// Carry-lookahead pairs: bit 0 = generate, bit 1 = propagate.  `lo` covers
// the lower limbs, `hi` the limbs just above; identity is 2 (g=0, p=1).
GL_HD u32 gp_combine(u32 lo, u32 hi) {
    u32 g = (hi & 1) | ((hi >> 1) & lo & 1);
    u32 p = (hi >> 1) & (lo >> 1) & 1;
    return g | (p << 1);
}

}  // namespace gnt
