// HIP kernels of the Goldilocks-NTT multiply, written for gfx950 (wave64,
// 64 KiB LDS tiles).  Integer-only: no MFMA, no floating point.
//
// One product (reference: crates/bignum/src/ntt.rs:806-865):
//   k_fwd_pass       forward passes of A and B (pack fused into the first)
//                    except their last passes
//   k_last_fused_ab  A's and B's last forward passes, the pointwise product
//                    and the first inverse pass, in one tile round trip
//                    (k_last_fused: the same for a square, or for a
//                    single-pass plan after A's whole transform, which also
//                    applies the 1/N; with more passes the 1/N rides on
//                    pass 0's inverse twiddles)
//   k_inv_pass       the remaining inverse passes; the last one emits limb
//                    sums L_j + 2^64 H_j instead of coefficients
//   k_carry_*        exact carry: tile scan -> tile-carry scan -> sparse ripple
// A square runs A through k_last_fused with itself.
#include <hip/hip_runtime.h>

#include "gnt_carry.hpp"
#include "gnt_kernels.hpp"
#include "gnt_pass.hpp"

namespace gnt {

// Transform kernels: one 64 KiB tile per workgroup of TBLOCK threads; two
// tiles per CU give 8 waves per SIMD to hide LDS and HBM latency, which
// needs <= 64 VGPRs per thread (512-thread tiles at 4 waves per SIMD and 128
// VGPRs measured 12% slower, even with radix-16 sub-steps).
constexpr int TBLOCK = 1024;
constexpr int TWAVES = TBLOCK / 64 * 2 / 4;  // waves per SIMD at two tiles per CU
using X = DevExec<TBLOCK>;
#define GNT_TBOUNDS __launch_bounds__(TBLOCK) __attribute__((amdgpu_waves_per_eu(TWAVES, TWAVES)))

// Workgroups are dispatched to the 8 XCDs round robin (workgroup w to XCD
// w mod 8), each XCD with its own L2.  A packing pass reads 2^(log_c - 2)
// limbs per tile row, half a 128-byte line when log_c = 5, and the other half
// belongs to the neighbouring tile: within every 16 workgroups this mapping
// gives tiles 2m and 2m+1 to workgroups w and w + 8, which run on the same XCD
// at about the same time, so the line is fetched from HBM once.
__device__ __forceinline__ unsigned xcd_paired_tile(unsigned w, unsigned tiles) {
    if (tiles & 15u) return w;
    return (w & ~15u) | ((w & 7u) << 1) | ((w >> 3) & 1u);
}

template <class G>
__global__ GNT_TBOUNDS void k_fwd_pass(u64 *__restrict__ x, const u64 *__restrict__ limbs, size_t n_limbs,
                                       PassDesc d, Tw tw, const u64 *__restrict__ T) {
    extern __shared__ u64 tile[];
    const unsigned t = limbs ? xcd_paired_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    body_fwd<X, G>(tile, make_geom<G>(d, t), x, limbs, n_limbs, tw, T);
}

// The same forward pass over two operands in one launch: blocks [0, tiles)
// work on A, [tiles, 2 tiles) on B.
template <class G>
__global__ GNT_TBOUNDS void k_fwd_pass_pair(u64 *__restrict__ xa, const u64 *__restrict__ la, size_t na,
                                            u64 *__restrict__ xb, const u64 *__restrict__ lb, size_t nb,
                                            unsigned tiles, PassDesc d, Tw tw, const u64 *__restrict__ T) {
    extern __shared__ u64 tile[];
    const bool b = blockIdx.x >= tiles;
    unsigned t = b ? blockIdx.x - tiles : blockIdx.x;
    if (la) t = xcd_paired_tile(t, tiles);  // tiles is a multiple of 16 or the mapping is the identity
    if (b) body_fwd<X, G>(tile, make_geom<G>(d, t), xb, lb, nb, tw, T);
    else body_fwd<X, G>(tile, make_geom<G>(d, t), xa, la, na, tw, T);
}

// Inverse pass; SUMS: the last one, writing limb sums.
template <class G, bool SUMS, int FORM>
__global__ GNT_TBOUNDS void k_inv_pass(u64 *__restrict__ x, PassDesc d, Tw tw, u64 *__restrict__ Ls,
                                       u64 *__restrict__ Hs, size_t n_coeffs, size_t n_limbs, int scale_e,
                                       StoreTw st) {
    extern __shared__ u64 tile[];
    body_inv<X, G, SUMS, FORM == INV_STEP, FORM == INV_TSTORE>(tile, make_geom<G>(d, blockIdx.x), x, tw, Ls, Hs,
                                                               n_coeffs, n_limbs, scale_e, st);
}

// Last forward pass of B (read from xb, or packed from limbs), pointwise with
// A (a_other; null squares), 1/N, first inverse pass; result to xa, or as
// limb sums when this is the only pass.
template <class G>
__global__ GNT_TBOUNDS void k_last_fused(const u64 *__restrict__ xb, const u64 *__restrict__ limbs,
                                         size_t n_limbs_in, u64 *xa, const u64 *a_other, PassDesc d, int log_n,
                                         Tw tw, u64 *__restrict__ Ls, u64 *__restrict__ Hs, size_t n_coeffs,
                                         size_t n_limbs, StoreTw st) {
    extern __shared__ u64 tile[];
    body_fused<X, G>(tile, make_geom<G>(d, blockIdx.x), xb, limbs, n_limbs_in, xa, a_other, log_n, tw, Ls, Hs,
                     n_coeffs, n_limbs, st);
}

// Last forward passes of A and B, pointwise, 1/N, first inverse pass; result
// to xa (multiplies with plans of two or more passes).
// G is the plan's last-pass shape; the kernel works on half-width tiles of it.
template <class G>
__global__ GNT_TBOUNDS void k_last_fused_ab(u64 *xa, const u64 *__restrict__ xb, PassDesc d, int log_n, Tw tw,
                                            StoreTw st) {
    extern __shared__ u64 tile[];
    using H = typename HalfGeom<G>::type;
    if constexpr (HalfGeom<G>::ok) body_fused_ab<X, H>(tile, make_geom<H>(d, blockIdx.x), xa, xb, log_n, tw, st);
}

// ---------------------------------------------------------------------------
// Carry: limb sums -> limbs (ntt.rs:669-791).  Limb j's value is
// y_j = L_j + H_{j-1} with a 1-bit overflow g_j; what is left is a 1-bit
// carry-lookahead scan over (generate g_j, propagate y_j == 2^64-1), done per
// tile (gnt_carry.hpp), then across tiles, then applied to the tiles that
// receive one.
// ---------------------------------------------------------------------------

// Tile flags: bit 0 generate, bit 1 propagate (gp_combine form), bit 2 the
// tile holds a nonzero limb.  The last tile stores its own trimmed length
// (1 + its highest nonzero limb, 0 if none); k_carry_scan replaces it when a
// lower tile is the highest flagged one, so no tile updates a shared word.
constexpr unsigned char TILE_NONZERO = 4;

__global__ __launch_bounds__(CARRY_BLOCK) void k_carry_tiles(const u64 *__restrict__ Ls, const u64 *__restrict__ Hs,
                                                             u64 *__restrict__ out, size_t n_limbs,
                                                             unsigned char *__restrict__ tile_gp,
                                                             u32 *__restrict__ out_len) {
    __shared__ u32 sh[2 * CARRY_WAVES];
    u32 gp, hi;
    carry_tile(SumsSrc{Ls, Hs}, n_limbs, blockIdx.x, out, sh, gp, hi);
    if (threadIdx.x == 0) {
        tile_gp[blockIdx.x] = (unsigned char)(gp | (hi ? TILE_NONZERO : 0));
        if (blockIdx.x == gridDim.x - 1) *out_len = hi;
    }
}

// 16 consecutive tile flags from p[c..c+16) (bytes past n read as 0) in one
// 16-byte load when they are all inside and p + c is 16-byte aligned.
__device__ __forceinline__ uint4 flags16(const unsigned char *__restrict__ p, size_t c, size_t n) {
    if (c + 16 <= n && ((uintptr_t)(p + c) & 15) == 0) return *(const uint4 *)(p + c);
    u32 w[4] = {0, 0, 0, 0};
    for (int k = 0; k < 16 && c + k < n; k++) w[k >> 2] |= (u32)p[c + k] << (8 * (k & 3));
    return make_uint4(w[0], w[1], w[2], w[3]);
}

__device__ __forceinline__ u32 flag_byte(const uint4 &v, int k) {
    const u32 w = k < 4 ? v.x : k < 8 ? v.y : k < 12 ? v.z : v.w;
    return (w >> (8 * (k & 3))) & 0xFF;
}

// One workgroup: exclusive carry-in of every tile, and the trimmed length
// before the ripple when the highest flagged tile is not the last (the
// product's top 2048 limbs are zero): 1 + the highest nonzero limb of that
// tile, read back from `out`.  A product's carry never leaves its top limb,
// so the ripple can only raise that length.  Each thread owns a run of
// consecutive tiles, a multiple of 16 long, and reads their flags 16 at a
// time (one load round trip per 16 tiles instead of one per tile).
constexpr int SCAN_BLOCK = 1024;
static_assert(CARRY_TILE == 2 * SCAN_BLOCK, "the length search reads two limbs per thread");
__global__ __launch_bounds__(SCAN_BLOCK) void k_carry_scan(const unsigned char *__restrict__ tile_gp,
                                                           unsigned char *__restrict__ tile_cin, size_t ntiles,
                                                           const u64 *__restrict__ out, size_t n_limbs,
                                                           u32 *__restrict__ out_len) {
    __shared__ u32 scan[SCAN_BLOCK];
    __shared__ u32 top_tile, top_limb;
    const size_t per = (ntiles + 16 * SCAN_BLOCK - 1) / (16 * SCAN_BLOCK) * 16;
    const size_t t0 = (size_t)threadIdx.x * per, t1 = min(t0 + per, ntiles);
    u32 agg = 2, top = 0;
    for (size_t c = t0; c < t1; c += 16) {
        const uint4 v = flags16(tile_gp, c, t1);
#pragma unroll
        for (int k = 0; k < 16; k++) {
            if (c + k >= t1) break;
            const u32 f = flag_byte(v, k);
            agg = gp_combine(agg, f);
            if (f & TILE_NONZERO) top = (u32)(c + k) + 1;
        }
    }
    scan[threadIdx.x] = agg;
    if (threadIdx.x == 0) top_tile = top_limb = 0;
    __syncthreads();
    // the highest thread with a flagged tile holds the top one
    const uint64_t has = __ballot(top != 0);
    if (has && (int)(threadIdx.x & 63) == 63 - __clzll((long long)has)) atomicMax(&top_tile, top);
    for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
        u32 v = scan[threadIdx.x];
        u32 w = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 2u;
        __syncthreads();
        scan[threadIdx.x] = gp_combine(w, v);
        __syncthreads();
    }
    u32 carry = threadIdx.x > 0 ? (scan[threadIdx.x - 1] & 1) : 0;
    for (size_t c = t0; c < t1; c += 16) {
        const uint4 v = flags16(tile_gp, c, t1);
#pragma unroll
        for (int k = 0; k < 16; k++) {
            if (c + k >= t1) break;
            tile_cin[c + k] = (unsigned char)carry;
            const u32 gp = flag_byte(v, k);
            carry = (gp & 1) | ((gp >> 1) & 1 & carry);
        }
    }
    const u32 tt = top_tile;  // final: written before the scan loop's barriers
    if (tt && tt < ntiles) {   // uniform
        const size_t j = (size_t)(tt - 1) * CARRY_TILE + 2 * threadIdx.x;
        const u64 v0 = j < n_limbs ? out[j] : 0, v1 = j + 1 < n_limbs ? out[j + 1] : 0;
        const u32 h = v1 ? (u32)(j + 2) : v0 ? (u32)(j + 1) : 0;
        if (h) atomicMax(&top_limb, h);
        __syncthreads();
        if (threadIdx.x == 0) *out_len = top_limb;
    }
}

// Tiles whose carry-in is 1 add it (one wave per tile).
__global__ __launch_bounds__(CARRY_RIPPLE_BLOCK) void k_carry_ripple(u64 *__restrict__ out, size_t n_limbs,
                                                                     const unsigned char *__restrict__ tile_cin,
                                                                     u32 *__restrict__ out_len) {
    if (tile_cin[blockIdx.x] == 0) return;  // uniform per workgroup
    const u32 hi = carry_ripple_tile(out, n_limbs, blockIdx.x);
    if (threadIdx.x == 0 && hi) atomicMax(out_len, hi);
}

__global__ void k_pass_table(u64 *__restrict__ T, int log_m, int log_r, int inv, int scale_e, Tw tw) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ((size_t)1 << log_m)) T[i] = pass_table_entry(tw, (u32)i, log_m, log_r, inv != 0, scale_e);
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------

static inline dim3 pass_grid(const PassDesc &d, int log_n) {
    return dim3((unsigned)(((size_t)1 << log_n) >> (d.log_r + d.log_c)));
}
static inline size_t pass_lds(const PassDesc &d) { return ((size_t)1 << (d.log_r + d.log_c)) * sizeof(u64); }

// The table the store of inverse pass p (or of the fused stage, p = npass-1)
// applies: that of inverse pass p-1, which runs next.
static StoreTw inv_store_tw(const Plan &pl, const PassTables &tabs, int p) {
    if (p < 1 || !tabs.inv[p - 1]) return StoreTw{nullptr, 0};
    return StoreTw{tabs.inv[p - 1], pl.pass[p - 1].log_m};
}

int launch_forward(const Plan &pl, int npasses, u64 *x, const u64 *limbs, size_t n_limbs, const Tw &tw,
                   const PassTables &tabs, hipStream_t st) {
    for (int p = 0; p < npasses; p++) {
        const PassDesc d = pl.pass[p];
        const u64 *src = p == 0 ? limbs : nullptr;
        switch (shape_key(d)) {
#define FWD_CASE(LR, LC, S)                                                                                     \
    case shape_key(LR, LC, S):                                                                                  \
        hipLaunchKernelGGL((k_fwd_pass<Geom<LR, LC, S>>), pass_grid(d, pl.log_n), dim3(TBLOCK), pass_lds(d), st, x, \
                           src, n_limbs, d, tw, tabs.fwd[p]);                                                                \
        break;
            GNT_FOR_EACH_SHAPE(FWD_CASE)
#undef FWD_CASE
            default: return -1;
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_forward_pair(const Plan &pl, int npasses, u64 *xa, const u64 *la, size_t na, u64 *xb, const u64 *lb,
                        size_t nb, const Tw &tw, const PassTables &tabs, hipStream_t st) {
    for (int p = 0; p < npasses; p++) {
        const PassDesc d = pl.pass[p];
        const unsigned tiles = pass_grid(d, pl.log_n).x;
        const u64 *sa = p == 0 ? la : nullptr, *sb = p == 0 ? lb : nullptr;
        switch (shape_key(d)) {
#define FWD2_CASE(LR, LC, S)                                                                                     \
    case shape_key(LR, LC, S):                                                                                   \
        hipLaunchKernelGGL((k_fwd_pass_pair<Geom<LR, LC, S>>), dim3(2 * tiles), dim3(TBLOCK), pass_lds(d), st, xa, \
                           sa, na, xb, sb, nb, tiles, d, tw, tabs.fwd[p]);                                                    \
        break;
            GNT_FOR_EACH_SHAPE(FWD2_CASE)
#undef FWD2_CASE
            default: return -1;
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_last_fused(const Plan &pl, const u64 *xb, const u64 *limbs_b, size_t nb, u64 *xa, bool square,
                      const Tw &tw, const PassTables &tabs, u64 *Ls, u64 *Hs, size_t n_coeffs, size_t n_limbs,
                      hipStream_t st) {
    const StoreTw stw = inv_store_tw(pl, tabs, pl.npass - 1);
    const PassDesc d = pl.pass[pl.npass - 1];
    const bool only = pl.npass == 1;
    const u64 *src = only ? limbs_b : nullptr;
    const u64 *other = square ? nullptr : xa;
    u64 *ls = only ? Ls : nullptr;
    u64 *hs = only ? Hs : nullptr;
    switch (shape_key(d)) {
#define FUSED_CASE(LR, LC, S)                                                                                   \
    case shape_key(LR, LC, S):                                                                                  \
        hipLaunchKernelGGL((k_last_fused<Geom<LR, LC, S>>), pass_grid(d, pl.log_n), dim3(TBLOCK), pass_lds(d), st, \
                           xb, src, nb, xa, other, d, pl.log_n, tw, ls, hs, n_coeffs, n_limbs, stw);                 \
        break;
        GNT_FOR_EACH_SHAPE(FUSED_CASE)
#undef FUSED_CASE
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_last_fused_ab(const Plan &pl, u64 *xa, const u64 *xb, const Tw &tw, const PassTables &tabs,
                         hipStream_t st) {
    if (pl.npass < 2) return -1;
    const StoreTw stw = inv_store_tw(pl, tabs, pl.npass - 1);
    const PassDesc d = pl.pass[pl.npass - 1];
    switch (shape_key(d)) {
#define FAB_CASE(LR, LC, S)                                                                                        \
    case shape_key(LR, LC, S):                                                                                     \
        if (!HalfGeom<Geom<LR, LC, S>>::ok) return -1;                                                            \
        hipLaunchKernelGGL((k_last_fused_ab<Geom<LR, LC, S>>), dim3(2 * pass_grid(d, pl.log_n).x), dim3(TBLOCK),  \
                           (fused_ab_lds_bytes<Geom<LR, LC, S>>()), st, xa, xb, d, pl.log_n, tw, stw);                   \
        break;
        GNT_FOR_EACH_SHAPE(FAB_CASE)
#undef FAB_CASE
        default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

// One inverse pass of a strided shape in the variant `form` (the last pass's
// inverse runs in the fused stage, so contiguous shapes never come here).
template <class G>
static int launch_inv_shape(const Plan &pl, const PassDesc &d, u64 *x, const Tw &tw, u64 *ls, u64 *hs,
                            size_t n_coeffs, size_t n_limbs, int form, const StoreTw &stw, hipStream_t st) {
    // pass 0 (block N) carries the transform's 1/N in its stepped twiddles
    const int scale_e = d.log_m == pl.log_n ? inv_scale_e(pl.log_n) : 0;
    if constexpr (!G::strided) {
        return -1;
    } else {
        const dim3 grid = pass_grid(d, pl.log_n);
        const size_t lds = pass_lds(d);
        if (ls && form == INV_STEP)
            hipLaunchKernelGGL((k_inv_pass<G, true, INV_STEP>), grid, dim3(TBLOCK), lds, st, x, d, tw, ls, hs,
                               n_coeffs, n_limbs, scale_e, stw);
        else if (ls)
            hipLaunchKernelGGL((k_inv_pass<G, true, INV_PLAIN>), grid, dim3(TBLOCK), lds, st, x, d, tw, ls, hs,
                               n_coeffs, n_limbs, scale_e, stw);
        else if (form == INV_STEP)
            hipLaunchKernelGGL((k_inv_pass<G, false, INV_STEP>), grid, dim3(TBLOCK), lds, st, x, d, tw, ls, hs,
                               n_coeffs, n_limbs, scale_e, stw);
        else if (form == INV_TSTORE)
            hipLaunchKernelGGL((k_inv_pass<G, false, INV_TSTORE>), grid, dim3(TBLOCK), lds, st, x, d, tw, ls, hs,
                               n_coeffs, n_limbs, scale_e, stw);
        else
            hipLaunchKernelGGL((k_inv_pass<G, false, INV_PLAIN>), grid, dim3(TBLOCK), lds, st, x, d, tw, ls, hs,
                               n_coeffs, n_limbs, scale_e, stw);
        return 0;
    }
}

int launch_inverse_rest(const Plan &pl, u64 *x, const Tw &tw, const PassTables &tabs, u64 *Ls, u64 *Hs,
                        size_t n_coeffs, size_t n_limbs, hipStream_t st) {
    for (int p = pl.npass - 2; p >= 0; p--) {
        const PassDesc d = pl.pass[p];
        u64 *ls = p == 0 ? Ls : nullptr;
        u64 *hs = p == 0 ? Hs : nullptr;
        const StoreTw stw = inv_store_tw(pl, tabs, p);
        // a pass with a table had its twiddles applied by the stage before
        const int form = !tabs.inv[p] ? INV_STEP : stw.T ? INV_TSTORE : INV_PLAIN;
        switch (shape_key(d)) {
#define INV_CASE(LR, LC, S)                                                                                  \
    case shape_key(LR, LC, S):                                                                               \
        if (launch_inv_shape<Geom<LR, LC, S>>(pl, d, x, tw, ls, hs, n_coeffs, n_limbs, form, stw, st)) return -1; \
        break;
            GNT_FOR_EACH_SHAPE(INV_CASE)
#undef INV_CASE
            default: return -1;
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_pass_table(u64 *T, int log_m, int log_r, bool inverse, int scale_e, const Tw &tw, hipStream_t st) {
    const size_t n = (size_t)1 << log_m;
    hipLaunchKernelGGL(k_pass_table, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, T, log_m, log_r,
                       inverse ? 1 : 0, scale_e, tw);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

size_t carry_tiles(size_t n_limbs) { return (n_limbs + CARRY_TILE - 1) / CARRY_TILE; }

int launch_carry(const u64 *Ls, const u64 *Hs, u64 *out, size_t n_limbs, unsigned char *tile_gp,
                 unsigned char *tile_cin, u32 *out_len, hipStream_t st) {
    size_t nt = carry_tiles(n_limbs);
    hipLaunchKernelGGL(k_carry_tiles, dim3((unsigned)nt), dim3(CARRY_BLOCK), 0, st, Ls, Hs, out, n_limbs, tile_gp,
                       out_len);
    hipLaunchKernelGGL(k_carry_scan, dim3(1), dim3(SCAN_BLOCK), 0, st, tile_gp, tile_cin, nt, out, n_limbs, out_len);
    hipLaunchKernelGGL(k_carry_ripple, dim3((unsigned)nt), dim3(CARRY_RIPPLE_BLOCK), 0, st, out, n_limbs, tile_cin,
                       out_len);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace gnt
