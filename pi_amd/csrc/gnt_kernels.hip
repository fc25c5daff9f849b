// HIP kernels of the Goldilocks-NTT multiply, written for gfx950 (wave64,
// 64 KiB LDS tiles).  Integer-only: no MFMA, no floating point.
//
// Pipeline for one product (reference: crates/bignum/src/ntt.rs:806-865):
//   fwd passes (pack fused into the first)  x2 operands (x1 for a square)
//   pointwise  c = a * b * N^-1               (N^-1 = 2^(192 - log N))
//   inv passes
//   carry: tile scan -> tile-carry scan -> sparse ripple  (exact limbs)
#include <hip/hip_runtime.h>

#include "gnt_kernels.hpp"
#include "gnt_pass.hpp"

namespace gnt {

constexpr int BLOCK = 256;

template <int LOGR, int I>
struct FwdSteps {
    __device__ __forceinline__ static void run(u64 *tile, int log_c, int tile_n, size_t cg0, int log_m, int log_s,
                                               const u64 *wlo, const u64 *whi) {
        const int ntask = tile_n >> Sub<LOGR>::log(I);
        for (int L = threadIdx.x; L < ntask; L += BLOCK)
            fwd_task<LOGR, I>(tile, L, log_c, cg0, log_m, log_s, wlo, whi);
        __syncthreads();
        FwdSteps<LOGR, I + 1>::run(tile, log_c, tile_n, cg0, log_m, log_s, wlo, whi);
    }
};
template <int LOGR>
struct FwdSteps<LOGR, (LOGR + 3) / 4> {
    __device__ __forceinline__ static void run(u64 *, int, int, size_t, int, int, const u64 *, const u64 *) {}
};

template <int LOGR, int I>
struct InvSteps {  // sub-steps I, I-1, ..., 0
    __device__ __forceinline__ static void run(u64 *tile, int log_c, int tile_n, size_t cg0, int log_m, int log_s,
                                               const u64 *wlo, const u64 *whi) {
        const int ntask = tile_n >> Sub<LOGR>::log(I);
        for (int L = threadIdx.x; L < ntask; L += BLOCK)
            inv_task<LOGR, I>(tile, L, log_c, cg0, log_m, log_s, wlo, whi);
        __syncthreads();
        InvSteps<LOGR, I - 1>::run(tile, log_c, tile_n, cg0, log_m, log_s, wlo, whi);
    }
};
template <int LOGR>
struct InvSteps<LOGR, -1> {
    __device__ __forceinline__ static void run(u64 *, int, int, size_t, int, int, const u64 *, const u64 *) {}
};

// Forward pass.  With `limbs` set this is the first pass: the 16-bit
// coefficients are unpacked from the operand's limbs as the tile is loaded.
template <int LOGR>
__global__ __launch_bounds__(BLOCK) void k_fwd_pass(u64 *__restrict__ x, const u64 *__restrict__ limbs,
                                                    size_t n_limbs, int log_m, int log_c,
                                                    const u64 *__restrict__ wlo, const u64 *__restrict__ whi) {
    extern __shared__ u64 tile[];
    const int log_s = log_m - LOGR;
    const int tile_n = 1 << (LOGR + log_c);
    const size_t cg0 = (size_t)blockIdx.x << log_c;
    for (int idx = threadIdx.x; idx < tile_n; idx += BLOCK) {
        int n, c;
        tile_rc(idx, LOGR, log_c, log_s, n, c);
        size_t a = tile_addr(cg0, n, c, log_m, log_s);
        tile[(n << log_c) + c] = limbs ? coeff_from_limbs(limbs, n_limbs, a) : x[a];
    }
    __syncthreads();
    FwdSteps<LOGR, 0>::run(tile, log_c, tile_n, cg0, log_m, log_s, wlo, whi);
    for (int idx = threadIdx.x; idx < tile_n; idx += BLOCK) {
        int n, c;
        tile_rc(idx, LOGR, log_c, log_s, n, c);
        x[tile_addr(cg0, n, c, log_m, log_s)] = tile[(n << log_c) + c];
    }
}

template <int LOGR>
__global__ __launch_bounds__(BLOCK) void k_inv_pass(u64 *__restrict__ x, int log_m, int log_c,
                                                    const u64 *__restrict__ wlo, const u64 *__restrict__ whi) {
    extern __shared__ u64 tile[];
    const int log_s = log_m - LOGR;
    const int tile_n = 1 << (LOGR + log_c);
    const size_t cg0 = (size_t)blockIdx.x << log_c;
    for (int idx = threadIdx.x; idx < tile_n; idx += BLOCK) {
        int n, c;
        tile_rc(idx, LOGR, log_c, log_s, n, c);
        tile[(n << log_c) + c] = x[tile_addr(cg0, n, c, log_m, log_s)];
    }
    __syncthreads();
    InvSteps<LOGR, (LOGR + 3) / 4 - 1>::run(tile, log_c, tile_n, cg0, log_m, log_s, wlo, whi);
    for (int idx = threadIdx.x; idx < tile_n; idx += BLOCK) {
        int n, c;
        tile_rc(idx, LOGR, log_c, log_s, n, c);
        x[tile_addr(cg0, n, c, log_m, log_s)] = tile[(n << log_c) + c];
    }
}

// c[i] = a[i] * b[i] * 2^(192 - log_n): the 1/N of the inverse (ntt.rs:376-391).
__global__ __launch_bounds__(BLOCK) void k_pointwise(u64 *__restrict__ a, const u64 *__restrict__ b, size_t n,
                                                     int log_n) {
    size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    size_t stride = (size_t)gridDim.x * BLOCK;
    const int e = 192 - log_n;
    for (; i < n; i += stride) a[i] = gl::mul_pow2(gl::mul(a[i], b[i]), e);
}

// ---------------------------------------------------------------------------
// Carry: coefficients (exact integers < 2^58) -> limbs (ntt.rs:669-791).
// Limb j collects S_j = L_j + 2^64 H_j; its value is y_j = L_j + H_{j-1} with a
// 1-bit overflow g_j; what is left is a 1-bit carry-lookahead scan over
// (generate g_j, propagate y_j == 2^64-1), done per tile, then across tiles,
// then applied to the (rare) tiles that receive a carry.
// ---------------------------------------------------------------------------

constexpr int CARRY_LIMBS_PER_THREAD = 4;
constexpr int CARRY_TILE = BLOCK * CARRY_LIMBS_PER_THREAD;  // limbs per workgroup

__global__ __launch_bounds__(BLOCK) void k_carry_tiles(const u64 *__restrict__ coeffs, size_t n_coeffs,
                                                       u64 *__restrict__ out, size_t n_limbs,
                                                       unsigned char *__restrict__ tile_gp, u32 *__restrict__ out_len) {
    __shared__ u32 scan[BLOCK];
    const size_t j0 = (size_t)blockIdx.x * CARRY_TILE + (size_t)threadIdx.x * CARRY_LIMBS_PER_THREAD;
    u64 y[CARRY_LIMBS_PER_THREAD];
    u32 gbits = 0, pbits = 0;
    u64 Hprev = 0, L, H;
    if (j0 > 0 && j0 - 1 < n_limbs) limb_sum(coeffs, n_coeffs, j0 - 1, L, Hprev);
#pragma unroll
    for (int k = 0; k < CARRY_LIMBS_PER_THREAD; k++) {
        size_t j = j0 + k;
        if (j < n_limbs) limb_sum(coeffs, n_coeffs, j, L, H);
        else { L = 0; H = 0; }
        u64 s = L + Hprev;
        gbits |= (u32)(s < L) << k;
        pbits |= (u32)(s == ~0ull) << k;
        y[k] = s;
        Hprev = H;
    }
    u32 agg = 2;
#pragma unroll
    for (int k = 0; k < CARRY_LIMBS_PER_THREAD; k++)
        agg = gp_combine(agg, ((gbits >> k) & 1) | (((pbits >> k) & 1) << 1));
    scan[threadIdx.x] = agg;
    __syncthreads();
    for (int off = 1; off < BLOCK; off <<= 1) {
        u32 v = scan[threadIdx.x];
        u32 w = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 2u;
        __syncthreads();
        scan[threadIdx.x] = gp_combine(w, v);
        __syncthreads();
    }
    u32 carry = threadIdx.x > 0 ? (scan[threadIdx.x - 1] & 1) : 0;
    u32 hi_nz = 0;
#pragma unroll
    for (int k = 0; k < CARRY_LIMBS_PER_THREAD; k++) {
        size_t j = j0 + k;
        u64 z = y[k] + carry;
        carry = ((gbits >> k) & 1) | (((pbits >> k) & 1) & carry);
        if (j < n_limbs) {
            out[j] = z;
            if (z != 0) hi_nz = (u32)(j + 1);
        }
    }
    if (hi_nz) atomicMax(out_len, hi_nz);
    if (threadIdx.x == BLOCK - 1) tile_gp[blockIdx.x] = (unsigned char)scan[BLOCK - 1];
}

// One workgroup: exclusive carry-in of every tile.
constexpr int SCAN_BLOCK = 1024;
__global__ __launch_bounds__(SCAN_BLOCK) void k_carry_scan(const unsigned char *__restrict__ tile_gp,
                                                           unsigned char *__restrict__ tile_cin, size_t ntiles) {
    __shared__ u32 scan[SCAN_BLOCK];
    const size_t per = (ntiles + SCAN_BLOCK - 1) / SCAN_BLOCK;
    const size_t t0 = (size_t)threadIdx.x * per;
    u32 agg = 2;
    for (size_t t = t0; t < t0 + per && t < ntiles; t++) agg = gp_combine(agg, tile_gp[t]);
    scan[threadIdx.x] = agg;
    __syncthreads();
    for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
        u32 v = scan[threadIdx.x];
        u32 w = threadIdx.x >= (unsigned)off ? scan[threadIdx.x - off] : 2u;
        __syncthreads();
        scan[threadIdx.x] = gp_combine(w, v);
        __syncthreads();
    }
    u32 carry = threadIdx.x > 0 ? (scan[threadIdx.x - 1] & 1) : 0;
    for (size_t t = t0; t < t0 + per && t < ntiles; t++) {
        tile_cin[t] = (unsigned char)carry;
        u32 gp = tile_gp[t];
        carry = (gp & 1) | ((gp >> 1) & carry);
    }
}

// Tiles whose carry-in is 1 add it: the +1 ripples through the run of
// all-ones limbs at the bottom of the tile.
__global__ __launch_bounds__(BLOCK) void k_carry_ripple(u64 *__restrict__ out, size_t n_limbs,
                                                        const unsigned char *__restrict__ tile_cin,
                                                        u32 *__restrict__ out_len) {
    if (tile_cin[blockIdx.x] == 0) return;  // uniform per workgroup
    __shared__ u32 allmax[BLOCK];
    const size_t j0 = (size_t)blockIdx.x * CARRY_TILE + (size_t)threadIdx.x * CARRY_LIMBS_PER_THREAD;
    u64 z[CARRY_LIMBS_PER_THREAD];
    u32 am = 1;
#pragma unroll
    for (int k = 0; k < CARRY_LIMBS_PER_THREAD; k++) {
        size_t j = j0 + k;
        z[k] = j < n_limbs ? out[j] : 0;
        am &= (j < n_limbs && z[k] == ~0ull) ? 1u : 0u;
    }
    allmax[threadIdx.x] = am;
    __syncthreads();
    for (int off = 1; off < BLOCK; off <<= 1) {
        u32 v = allmax[threadIdx.x];
        u32 w = threadIdx.x >= (unsigned)off ? allmax[threadIdx.x - off] : 1u;
        __syncthreads();
        allmax[threadIdx.x] = v & w;
        __syncthreads();
    }
    u32 carry = threadIdx.x > 0 ? allmax[threadIdx.x - 1] : 1u;
    u32 hi_nz = 0;
#pragma unroll
    for (int k = 0; k < CARRY_LIMBS_PER_THREAD; k++) {
        size_t j = j0 + k;
        if (carry && j < n_limbs) {
            z[k] += 1;
            carry = z[k] == 0;
            out[j] = z[k];
            if (z[k] != 0) hi_nz = (u32)(j + 1);
        }
    }
    if (hi_nz) atomicMax(out_len, hi_nz);
}

__global__ void k_zero_u32(u32 *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------

template <int LOGR>
static void launch_fwd(const PassDesc &d, u64 *x, const u64 *limbs, size_t n_limbs, int log_n, const u64 *wlo,
                       const u64 *whi, hipStream_t st) {
    size_t tile = (size_t)1 << (LOGR + d.log_c);
    size_t grid = ((size_t)1 << log_n) / tile;
    hipLaunchKernelGGL(k_fwd_pass<LOGR>, dim3((unsigned)grid), dim3(BLOCK), tile * sizeof(u64), st, x, limbs, n_limbs,
                       d.log_m, d.log_c, wlo, whi);
}

template <int LOGR>
static void launch_inv(const PassDesc &d, u64 *x, int log_n, const u64 *wlo, const u64 *whi, hipStream_t st) {
    size_t tile = (size_t)1 << (LOGR + d.log_c);
    size_t grid = ((size_t)1 << log_n) / tile;
    hipLaunchKernelGGL(k_inv_pass<LOGR>, dim3((unsigned)grid), dim3(BLOCK), tile * sizeof(u64), st, x, d.log_m,
                       d.log_c, wlo, whi);
}

#define GNT_DISPATCH_R(LR, CALL)                                                      \
    switch (LR) {                                                                     \
        case 1: CALL(1); break;  case 2: CALL(2); break;  case 3: CALL(3); break;     \
        case 4: CALL(4); break;  case 5: CALL(5); break;  case 6: CALL(6); break;     \
        case 7: CALL(7); break;  case 8: CALL(8); break;  case 9: CALL(9); break;     \
        case 10: CALL(10); break; case 11: CALL(11); break; case 12: CALL(12); break; \
        case 13: CALL(13); break;                                                     \
        default: return -1;                                                           \
    }

int launch_forward(const Plan &pl, u64 *x, const u64 *limbs, size_t n_limbs, const u64 *wlo, const u64 *whi,
                   hipStream_t st) {
    for (int p = 0; p < pl.npass; p++) {
        const PassDesc &d = pl.pass[p];
        const u64 *src = p == 0 ? limbs : nullptr;
#define FWD_CALL(R) launch_fwd<R>(d, x, src, n_limbs, pl.log_n, wlo, whi, st)
        GNT_DISPATCH_R(d.log_r, FWD_CALL)
#undef FWD_CALL
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_inverse(const Plan &pl, u64 *x, const u64 *wlo, const u64 *whi, hipStream_t st) {
    for (int p = pl.npass - 1; p >= 0; p--) {
        const PassDesc &d = pl.pass[p];
#define INV_CALL(R) launch_inv<R>(d, x, pl.log_n, wlo, whi, st)
        GNT_DISPATCH_R(d.log_r, INV_CALL)
#undef INV_CALL
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_pointwise(u64 *a, const u64 *b, int log_n, hipStream_t st) {
    size_t n = (size_t)1 << log_n;
    size_t grid = (n + BLOCK - 1) / BLOCK;
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(k_pointwise, dim3((unsigned)grid), dim3(BLOCK), 0, st, a, b, n, log_n);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

size_t carry_tiles(size_t n_limbs) { return (n_limbs + CARRY_TILE - 1) / CARRY_TILE; }

int launch_carry(const u64 *coeffs, size_t n_coeffs, u64 *out, size_t n_limbs, unsigned char *tile_gp,
                 unsigned char *tile_cin, u32 *out_len, hipStream_t st) {
    size_t nt = carry_tiles(n_limbs);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, out_len, 1);
    hipLaunchKernelGGL(k_carry_tiles, dim3((unsigned)nt), dim3(BLOCK), 0, st, coeffs, n_coeffs, out, n_limbs, tile_gp,
                       out_len);
    hipLaunchKernelGGL(k_carry_scan, dim3(1), dim3(SCAN_BLOCK), 0, st, tile_gp, tile_cin, nt);
    hipLaunchKernelGGL(k_carry_ripple, dim3((unsigned)nt), dim3(BLOCK), 0, st, out, n_limbs, tile_cin, out_len);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace gnt
