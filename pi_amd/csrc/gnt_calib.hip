// Calibration microbenchmarks for the roofline numbers (standalone binary).
//
//   gnt_calib copy  <bytes>   plain device copy: achievable HBM GB/s, and the
//                             reference kernel for scaling rocprofv3
//                             FETCH_SIZE / WRITE_SIZE counters
//   gnt_calib valu            Goldilocks mul / add+sub / shift-multiply
//                             throughput (G ops/s over the whole device)
//   gnt_calib isa             issue cost of single VALU instructions the
//                             field arithmetic is made of (SIMD cycles per
//                             wave64 instruction; 8 independent chains per
//                             lane, many waves per SIMD)
//   gnt_calib copy4 <bytes>   the copy with four 16-byte loads in flight per
//                             thread, plain and with non-temporal accesses
//   gnt_calib persistent      a 2 GiB read-modify-write in 64 KiB tiles: one
//                             workgroup per tile, against 512 persistent
//                             workgroups that claim tiles from a counter and
//                             signal each finished tile with a relaxed or a
//                             device-scope release atomic (what dependencies
//                             between tiles on different XCDs require)
//   gnt_calib field           field arithmetic as compiled for the device
//                             (add, sub, mul, the 128-bit reduction, every
//                             power-of-two multiply, radix-8/16 register DFTs)
//                             against the same functions compiled for the
//                             host, on edge values and random ones; exits 1
//                             on any difference
//   gnt_calib lds             one ds_read_b64 and one ds_write_b64 kernel per
//                             LDS access pattern (lane strides of 1..32
//                             words, and the down-column and last column-DFT
//                             walks of the 1024 x 4 tiles in the XOR-swizzled
//                             and the padded layout); run under rocprofv3
//                             --pmc SQ_LDS_BANK_CONFLICT SQ_INSTS_LDS to read
//                             the bank-conflict cycles of each pattern
// Prints one JSON line per measurement.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gl_field.hpp"
#include "gnt_pass.hpp"

using gl::u64;

#define CK(x)                                                                   \
    do {                                                                        \
        hipError_t e = (x);                                                     \
        if (e != hipSuccess) {                                                  \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));              \
            exit(1);                                                            \
        }                                                                       \
    } while (0)

__global__ __launch_bounds__(256) void k_copy(const ulonglong2 *__restrict__ src, ulonglong2 *__restrict__ dst,
                                              size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256;
    for (; i < n; i += stride) dst[i] = src[i];
}

// Four loads in flight per thread before the stores; NT: non-temporal.
typedef unsigned long long v2u64 __attribute__((ext_vector_type(2)));
template <bool NT>
__global__ __launch_bounds__(256) void k_copy4(const v2u64 *__restrict__ src, v2u64 *__restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += 4 * stride) {
        v2u64 v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const size_t j = i + k * stride;
            if (j < n) v[k] = NT ? __builtin_nontemporal_load(src + j) : src[j];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const size_t j = i + k * stride;
            if (j < n) {
                if (NT) __builtin_nontemporal_store(v[k], dst + j);
                else dst[j] = v[k];
            }
        }
    }
}

// 64 KiB tiles of x updated in place.  MODE 0: one workgroup per tile.
// MODE 1/2: persistent workgroups claim tiles from *next until all are taken
// (every workgroup leaves the loop once the counter passes the tile count)
// and count each finished tile in *done, relaxed (1) or with a device-scope
// release (2).
constexpr int PT_ELEMS = 8192;
template <int MODE>
__global__ __launch_bounds__(1024) void k_tiles(u64 *x, unsigned *done, unsigned long long *next,
                                                unsigned long long tiles) {
    __shared__ unsigned long long s_t;
    for (unsigned long long t = blockIdx.x;;) {
        if (MODE != 0) {
            if (threadIdx.x == 0) s_t = atomicAdd(next, 1ull);
            __syncthreads();
            t = s_t;
        }
        if (t >= tiles) break;
        u64 *p = x + t * PT_ELEMS;
        for (int i = threadIdx.x; i < PT_ELEMS; i += 1024) p[i] = p[i] * 3 + t;
        if (MODE == 0) break;
        __syncthreads();
        if (threadIdx.x == 0) {
            if (MODE == 2) __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            else __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();  // s_t is rewritten only after every thread has read it
    }
}

constexpr int CHAINS = 8;

__global__ __launch_bounds__(256) void k_mul(u64 *out, int iters, u64 seed) {
    u64 v[CHAINS];
    u64 w = seed + threadIdx.x;
#pragma unroll
    for (int c = 0; c < CHAINS; c++) v[c] = seed * (c + 1) + blockIdx.x;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < CHAINS; c++) v[c] = gl::mul(v[c], w);
    }
    u64 s = 0;
#pragma unroll
    for (int c = 0; c < CHAINS; c++) s ^= v[c];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_addsub(u64 *out, int iters, u64 seed) {
    u64 v[CHAINS];
    u64 w = seed + threadIdx.x;
#pragma unroll
    for (int c = 0; c < CHAINS; c++) v[c] = (seed * (c + 1) + blockIdx.x) % gl::P;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < CHAINS; c += 2) {
            u64 a = gl::add(v[c], v[c + 1]);
            u64 b = gl::sub(v[c], v[c + 1]);
            v[c] = a;
            v[c + 1] = gl::add(b, w);
        }
    }
    u64 s = 0;
#pragma unroll
    for (int c = 0; c < CHAINS; c++) s ^= v[c];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_shift(u64 *out, int iters, u64 seed) {
    u64 v[CHAINS];
#pragma unroll
    for (int c = 0; c < CHAINS; c++) v[c] = (seed * (c + 1) + blockIdx.x + threadIdx.x) % gl::P;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < CHAINS; c++) v[c] = gl::mul_pow2(v[c], 24);
    }
    u64 s = 0;
#pragma unroll
    for (int c = 0; c < CHAINS; c++) s ^= v[c];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// One instruction per chain step, written as inline assembly so that
// exactly that instruction is issued (ALU only: no memory access in asm).
constexpr int ISA_ITERS = 4096;
#define GNT_ISA_KERNEL(NAME, T, ASM, ...)                                         \
    __global__ __launch_bounds__(256) void NAME(u64 *out, unsigned seed) {       \
        T a[CHAINS];                                                              \
        const T b = (T)(seed ^ threadIdx.x);                                      \
        for (int c = 0; c < CHAINS; c++) a[c] = (T)(c + threadIdx.x);             \
        for (int it = 0; it < ISA_ITERS; it++) {                                  \
            _Pragma("unroll") for (int c = 0; c < CHAINS; c++) { ASM; }           \
        }                                                                         \
        u64 r = 0;                                                                \
        for (int c = 0; c < CHAINS; c++) r += (u64)a[c];                          \
        out[(size_t)blockIdx.x * 256 + threadIdx.x] = r;                          \
    }
GNT_ISA_KERNEL(k_isa_add_u32, unsigned, asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[c]) : "v"(b)))
GNT_ISA_KERNEL(k_isa_xor_b32, unsigned, asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a[c]) : "v"(b)))
GNT_ISA_KERNEL(k_isa_lshlrev_b32, unsigned, asm volatile("v_lshlrev_b32 %0, %1, %0" : "+v"(a[c]) : "v"(b & 31)))
GNT_ISA_KERNEL(k_isa_add_co_u32, unsigned,
               u64 sc; asm volatile("v_add_co_u32_e64 %0, %1, %0, %2" : "+v"(a[c]), "=s"(sc) : "v"(b)))
GNT_ISA_KERNEL(k_isa_cndmask_b32, unsigned,
               asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(a[c]) : "v"(b), "s"(~0ull)))
GNT_ISA_KERNEL(k_isa_mad_u64_u32, u64,
               u64 sc; asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a[c]), "=s"(sc) : "v"((unsigned)b), "v"((unsigned)b)))
GNT_ISA_KERNEL(k_isa_lshl_add_u64, u64, asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(a[c]) : "v"(b)))
GNT_ISA_KERNEL(k_isa_cmp_lt_u64, u64,
               u64 m; asm volatile("v_cmp_lt_u64_e64 %0, %1, %2" : "=s"(m) : "v"(a[c]), "v"(b)))
GNT_ISA_KERNEL(k_isa_lshlrev_b64, u64, asm volatile("v_lshlrev_b64 %0, %1, %0" : "+v"(a[c]) : "v"((unsigned)b & 31)))

static float time_ms(void (*launch)(void *), void *arg, int reps) {
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    launch(arg);
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(a));
    for (int i = 0; i < reps; i++) launch(arg);
    CK(hipEventRecord(b));
    CK(hipEventSynchronize(b));
    float ms;
    CK(hipEventElapsedTime(&ms, a, b));
    return ms / reps;
}

// ---- field check: device forms of the arithmetic against the host forms ----

__host__ __device__ inline u64 canon(u64 x) { return x >= gl::P ? x - gl::P : x; }

constexpr int FIELD_OUTS = 6;

// Per pair (a, b): add and sub of the canonical values, mul and the 128-bit
// reduction of the raw words, x * 2^e with a run-time e, and the shift
// reduction with a run-time shift.
__host__ __device__ inline void field_ops(u64 a, u64 b, u64 *o) {
    const u64 ca = canon(a), cb = canon(b);
    o[0] = gl::add(ca, cb);
    o[1] = gl::sub(ca, cb);
    o[2] = gl::mul(a, b);
    o[3] = gl::reduce128(a, b);
    o[4] = gl::mul_pow2(ca, (int)(b % 192));
    o[5] = gl::shl_reduce(a, 1 + (int)(b % 63));
}

// x * 2^e for every e < 192 with e a compile-time constant (the forms the
// register DFTs use), and a radix-16 forward DFT then inverse DFT of 16 values.
constexpr int POW2_OUTS = 192;
__host__ __device__ inline void pow2_ops(u64 x, u64 *o) {
#pragma unroll
    for (int e = 0; e < 192; e++) o[e] = gl::mul_pow2(canon(x), e);
}
__host__ __device__ inline void dft_ops(const u64 *x, u64 *o) {
    u64 v[16];
    for (int i = 0; i < 16; i++) v[i] = canon(x[i]);
    gl::dft_dif<16>(v);
    for (int i = 0; i < 16; i++) o[i] = v[i];
    gl::dft_dit_inv<16>(v);
    for (int i = 0; i < 16; i++) o[16 + i] = v[i];
    u64 w[8];
    for (int i = 0; i < 8; i++) w[i] = canon(x[i]);
    gl::dft_dif<8>(w);
    gl::dft_dit_inv<8>(w);
    for (int i = 0; i < 8; i++) o[32 + i] = w[i];
}

__global__ __launch_bounds__(256) void k_field(const u64 *a, const u64 *b, size_t n, u64 *out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) field_ops(a[i], b[i], out + i * FIELD_OUTS);
}
__global__ __launch_bounds__(256) void k_pow2(const u64 *x, size_t n, u64 *out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) pow2_ops(x[i], out + i * POW2_OUTS);
}
__global__ __launch_bounds__(256) void k_dft(const u64 *x, size_t n, u64 *out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dft_ops(x + i * 16, out + i * 40);
}

struct CopyArg { ulonglong2 *src, *dst; size_t n; };
static void launch_copy(void *p) {
    CopyArg *c = (CopyArg *)p;
    hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, 0, c->src, c->dst, c->n);
}

struct ValuArg { u64 *out; int iters; int kind; };
static const int VALU_BLOCKS = 256 * 32;
static void launch_valu(void *p) {
    ValuArg *v = (ValuArg *)p;
    if (v->kind == 0) hipLaunchKernelGGL(k_mul, dim3(VALU_BLOCKS), dim3(256), 0, 0, v->out, v->iters, 12345ull);
    else if (v->kind == 1) hipLaunchKernelGGL(k_addsub, dim3(VALU_BLOCKS), dim3(256), 0, 0, v->out, v->iters, 12345ull);
    else hipLaunchKernelGGL(k_shift, dim3(VALU_BLOCKS), dim3(256), 0, 0, v->out, v->iters, 12345ull);
}

struct IsaArg { u64 *out; void (*k)(u64 *, unsigned); };
static const int ISA_BLOCKS = 8192;
static void launch_isa(void *p) {
    IsaArg *a = (IsaArg *)p;
    hipLaunchKernelGGL(a->k, dim3(ISA_BLOCKS), dim3(256), 0, 0, a->out, 12345u);
}

// ---- LDS access patterns ----
// Word (8-byte) address of lane l's access in pattern P: strides of 2^P
// words for P < 6; then the 1024 x 4 tile's down-column walk (64 rows of one
// column) and last column-DFT sub-step walk (16 rows apart, 4 columns per
// row), XOR-swizzled (the contiguous tiles' layout) and padded (the fused
// multiply's half-width tiles).
constexpr int LDS_PATTERNS = 10, LDS_ITERS = 4096, LDS_BLOCKS = 1024;
using TileXor = gnt::Geom<10, 2, false, true, false>;
using TilePad = gnt::Geom<10, 2, false, true, true>;
static_assert(TilePad::padded && !TileXor::padded, "the two layouts");
template <int P>
__device__ __forceinline__ int lds_pattern(int l) {
    if constexpr (P < 6) return l << P;
    else if constexpr (P == 6) return gnt::tix(TileXor{}, l, 0);
    else if constexpr (P == 7) return gnt::tix(TilePad{}, l, 0);
    else if constexpr (P == 8) return gnt::tix(TileXor{}, 16 * (l >> 2) + 1, l & 3);
    else return gnt::tix(TilePad{}, 16 * (l >> 2) + 1, l & 3);
}
static const char *lds_pattern_name[LDS_PATTERNS] = {
    "stride_1", "stride_2", "stride_4", "stride_8", "stride_16", "stride_32",
    "xor_down_column", "padded_down_column", "xor_last_substep", "padded_last_substep"};

// Every lane reads (or writes) its pattern address LDS_ITERS times; the
// address is masked into the 8192-word array and moved by it * zero, a zero
// the compiler cannot see, so every iteration issues its own access.
template <int P, bool WRITE>
__global__ __launch_bounds__(256) void k_lds(u64 *out, int seed) {
    __shared__ u64 lds[8192];
    for (int i = threadIdx.x; i < 8192; i += 256) lds[i] = (u64)i * 0x9E3779B97F4A7C15ull + (u64)seed;
    __syncthreads();
    const int a0 = lds_pattern<P>(threadIdx.x & 63);
    const int zero = seed >> 20;
    u64 acc = 0;
    for (int it = 0; it < LDS_ITERS; it++) {
        const int a = (a0 + it * zero * 64) & 8191;
        if (WRITE) lds[a] = acc + (u64)it;
        else acc += lds[a];
    }
    __syncthreads();
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = acc + lds[threadIdx.x];
}

template <int P>
static void lds_run(u64 *out) {
    if constexpr (P < LDS_PATTERNS) {
        for (int w = 0; w < 2; w++) {
            hipEvent_t e0, e1;
            CK(hipEventCreate(&e0));
            CK(hipEventCreate(&e1));
            CK(hipEventRecord(e0, 0));
            if (w) hipLaunchKernelGGL((k_lds<P, true>), dim3(LDS_BLOCKS), dim3(256), 0, 0, out, 1);
            else hipLaunchKernelGGL((k_lds<P, false>), dim3(LDS_BLOCKS), dim3(256), 0, 0, out, 1);
            CK(hipGetLastError());
            CK(hipEventRecord(e1, 0));
            CK(hipEventSynchronize(e1));
            float ms = 0;
            CK(hipEventElapsedTime(&ms, e0, e1));
            printf("{\"calib\": \"lds\", \"pattern\": \"%s\", \"access\": \"%s\", "
                   "\"wave_accesses\": %llu, \"ms\": %.4f}\n",
                   lds_pattern_name[P], w ? "ds_write_b64" : "ds_read_b64",
                   (unsigned long long)LDS_BLOCKS * 4 * LDS_ITERS, ms);
            CK(hipEventDestroy(e0));
            CK(hipEventDestroy(e1));
        }
        lds_run<P + 1>(out);
    }
}

int main(int argc, char **argv) {
    const char *mode = argc > 1 ? argv[1] : "all";
    if (!strcmp(mode, "field")) {
        // edge values: around 0, 2^32 - 1 = EPS, 2^32, 2^63, p and 2^64, words
        // whose halves are 0 or all ones, and values whose high word equals or
        // exceeds the low word (the borrow cases of the reduction)
        std::vector<u64> edge = {0, 1, 2, 3, 0x7FFFFFFFull, 0x80000000ull, gl::EPS - 1, gl::EPS, gl::EPS + 1,
                                 gl::EPS + 2, 1ull << 33, (1ull << 33) - 1, 0x1FFFFFFFFull, 0x100000001ull,
                                 (1ull << 63) - 1, 1ull << 63, (1ull << 63) + 1, 0x80000000FFFFFFFFull,
                                 gl::P - 1, gl::P - 2, gl::P - 3, gl::P - gl::EPS, gl::P - gl::EPS - 1,
                                 gl::P - (1ull << 32), gl::P - (1ull << 33), 0xFFFFFFFE00000000ull,
                                 0xFFFFFFFEFFFFFFFFull, 0xFFFFFFFE00000001ull, gl::P, gl::P + 1, gl::P + 2,
                                 ~0ull, ~0ull - 1, ~0ull - gl::EPS, 0x0000000100000000ull * 5 + 4,
                                 0x0000000100000000ull * 5 + 5, 0x0000000100000000ull * 5 + 6,
                                 0xFFFFFFFF00000000ull + 0xFFFFFFFEull, 0x12345678ull << 32};
        std::vector<u64> A, B;
        for (u64 x : edge)
            for (u64 y : edge) { A.push_back(x); B.push_back(y); }
        u64 st = 0x9E3779B97F4A7C15ull;
        auto rnd = [&st]() {
            st += 0x9E3779B97F4A7C15ull;
            u64 z = st;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        // random words, and random words with forced halves (high word = low
        // word, all-ones or zero halves)
        for (int k = 0; k < (1 << 20); k++) {
            u64 x = rnd(), y = rnd();
            const int f = k & 7;
            if (f == 1) x = y >> 32;               // lo equal to hh
            if (f == 2) x = (y >> 32) - 1;         // lo just below hh
            if (f == 3) x |= 0xFFFFFFFF00000000ull;
            if (f == 4) y &= 0xFFFFFFFF00000000ull;  // hl = 0
            if (f == 5) y |= 0xFFFFFFFFull;          // hl = 2^32 - 1
            A.push_back(x);
            B.push_back(y);
        }
        const size_t n = A.size();
        u64 *dA, *dB, *dO;
        CK(hipMalloc(&dA, n * 8));
        CK(hipMalloc(&dB, n * 8));
        CK(hipMalloc(&dO, n * 8 * POW2_OUTS));
        CK(hipMemcpy(dA, A.data(), n * 8, hipMemcpyHostToDevice));
        CK(hipMemcpy(dB, B.data(), n * 8, hipMemcpyHostToDevice));
        std::vector<u64> got(n * POW2_OUTS), want(POW2_OUTS);
        size_t bad[3] = {0, 0, 0}, checked[3] = {0, 0, 0};
        hipLaunchKernelGGL(k_field, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dA, dB, n, dO);
        CK(hipDeviceSynchronize());
        CK(hipMemcpy(got.data(), dO, n * 8 * FIELD_OUTS, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) {
            field_ops(A[i], B[i], want.data());
            for (int k = 0; k < FIELD_OUTS; k++, checked[0]++)
                if (got[i * FIELD_OUTS + k] != want[k] && bad[0]++ < 8)
                    fprintf(stderr, "field op %d a=%016llx b=%016llx device %016llx host %016llx\n", k,
                            (unsigned long long)A[i], (unsigned long long)B[i],
                            (unsigned long long)got[i * FIELD_OUTS + k], (unsigned long long)want[k]);
        }
        const size_t np = edge.size() + 65536;
        hipLaunchKernelGGL(k_pow2, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, 0, dA + (n - np), np, dO);
        CK(hipDeviceSynchronize());
        CK(hipMemcpy(got.data(), dO, np * 8 * POW2_OUTS, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < np; i++) {
            pow2_ops(A[n - np + i], want.data());
            for (int e = 0; e < POW2_OUTS; e++, checked[1]++)
                if (got[i * POW2_OUTS + e] != want[e] && bad[1]++ < 8)
                    fprintf(stderr, "pow2 e=%d x=%016llx device %016llx host %016llx\n", e,
                            (unsigned long long)A[n - np + i], (unsigned long long)got[i * POW2_OUTS + e],
                            (unsigned long long)want[e]);
        }
        const size_t nd = n / 16;
        hipLaunchKernelGGL(k_dft, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, 0, dA, nd, dO);
        CK(hipDeviceSynchronize());
        CK(hipMemcpy(got.data(), dO, nd * 8 * 40, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nd; i++) {
            dft_ops(A.data() + i * 16, want.data());
            for (int k = 0; k < 40; k++, checked[2]++)
                if (got[i * 40 + k] != want[k] && bad[2]++ < 8)
                    fprintf(stderr, "dft vector %zu output %d device %016llx host %016llx\n", i, k,
                            (unsigned long long)got[i * 40 + k], (unsigned long long)want[k]);
        }
        printf("{\"calib\": \"field\", \"pairs\": %zu, \"ops_checked\": %zu, \"ops_differing\": %zu, "
               "\"pow2_checked\": %zu, \"pow2_differing\": %zu, \"dft_checked\": %zu, \"dft_differing\": %zu}\n",
               n, checked[0], bad[0], checked[1], bad[1], checked[2], bad[2]);
        CK(hipFree(dA));
        CK(hipFree(dB));
        CK(hipFree(dO));
        return (bad[0] | bad[1] | bad[2]) ? 1 : 0;
    }
    if (!strcmp(mode, "copy") || !strcmp(mode, "all")) {
        size_t bytes = argc > 2 ? strtoull(argv[2], 0, 10) : ((size_t)1 << 30);
        CopyArg c;
        c.n = bytes / sizeof(ulonglong2);
        CK(hipMalloc(&c.src, bytes));
        CK(hipMalloc(&c.dst, bytes));
        CK(hipMemset(c.src, 1, bytes));
        float ms = time_ms(launch_copy, &c, 20);
        printf("{\"calib\": \"copy\", \"bytes_each_way\": %zu, \"ms\": %.4f, \"GBps_read_plus_write\": %.1f}\n", bytes,
               ms, 2.0 * bytes / (ms * 1e-3) / 1e9);
        CK(hipFree(c.src));
        CK(hipFree(c.dst));
    }
    if (!strcmp(mode, "copy4") || !strcmp(mode, "all")) {
        size_t bytes = argc > 2 && !strcmp(mode, "copy4") ? strtoull(argv[2], 0, 10) : ((size_t)1 << 30);
        const size_t n = bytes / sizeof(v2u64);
        v2u64 *src, *dst;
        CK(hipMalloc(&src, bytes));
        CK(hipMalloc(&dst, bytes));
        CK(hipMemset(src, 1, bytes));
        hipEvent_t e0, e1;
        CK(hipEventCreate(&e0));
        CK(hipEventCreate(&e1));
        for (int nt = 0; nt < 2; nt++) {
            for (int grid : {1024, 4096, 16384}) {
                float best = 1e30f;
                for (int r = 0; r < 10; r++) {
                    CK(hipEventRecord(e0));
                    if (nt) hipLaunchKernelGGL(k_copy4<true>, dim3(grid), dim3(256), 0, 0, src, dst, n);
                    else hipLaunchKernelGGL(k_copy4<false>, dim3(grid), dim3(256), 0, 0, src, dst, n);
                    CK(hipEventRecord(e1));
                    CK(hipEventSynchronize(e1));
                    float ms;
                    CK(hipEventElapsedTime(&ms, e0, e1));
                    if (r > 0 && ms < best) best = ms;
                }
                printf("{\"calib\": \"copy4\", \"nontemporal\": %d, \"grid\": %d, \"bytes_each_way\": %zu, "
                       "\"best_ms\": %.4f, \"GBps_read_plus_write\": %.1f}\n",
                       nt, grid, bytes, best, 2.0 * bytes / (best * 1e-3) / 1e9);
            }
        }
        CK(hipFree(src));
        CK(hipFree(dst));
    }
    if (!strcmp(mode, "persistent") || !strcmp(mode, "all")) {
        const unsigned long long tiles = 32768;  // 2 GiB
        u64 *x;
        unsigned *done;
        unsigned long long *next;
        CK(hipMalloc(&x, tiles * PT_ELEMS * sizeof(u64)));
        CK(hipMemset(x, 0, tiles * PT_ELEMS * sizeof(u64)));
        CK(hipMalloc(&done, sizeof(unsigned)));
        CK(hipMalloc(&next, sizeof(unsigned long long)));
        hipEvent_t e0, e1;
        CK(hipEventCreate(&e0));
        CK(hipEventCreate(&e1));
        const char *names[3] = {"one_workgroup_per_tile", "persistent_relaxed", "persistent_release"};
        for (int mode_k = 0; mode_k < 3; mode_k++) {
            float best = 1e30f;
            for (int r = 0; r < 6; r++) {
                CK(hipMemset(next, 0, sizeof(unsigned long long)));
                CK(hipMemset(done, 0, sizeof(unsigned)));
                CK(hipDeviceSynchronize());
                CK(hipEventRecord(e0));
                if (mode_k == 0) hipLaunchKernelGGL(k_tiles<0>, dim3((unsigned)tiles), dim3(1024), 0, 0, x, done, next, tiles);
                else if (mode_k == 1) hipLaunchKernelGGL(k_tiles<1>, dim3(512), dim3(1024), 0, 0, x, done, next, tiles);
                else hipLaunchKernelGGL(k_tiles<2>, dim3(512), dim3(1024), 0, 0, x, done, next, tiles);
                CK(hipEventRecord(e1));
                CK(hipEventSynchronize(e1));
                float ms;
                CK(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0 && ms < best) best = ms;
            }
            unsigned d = 0;
            CK(hipMemcpy(&d, done, sizeof d, hipMemcpyDeviceToHost));
            printf("{\"calib\": \"persistent\", \"mode\": \"%s\", \"tiles\": %llu, \"tile_bytes\": %d, "
                   "\"best_ms\": %.4f, \"tiles_counted\": %u}\n",
                   names[mode_k], tiles, PT_ELEMS * 8, best, d);
        }
        CK(hipFree(x));
        CK(hipFree(done));
        CK(hipFree(next));
    }
    if (!strcmp(mode, "valu") || !strcmp(mode, "all")) {
        ValuArg v;
        v.iters = 2048;
        CK(hipMalloc(&v.out, (size_t)VALU_BLOCKS * 256 * sizeof(u64)));
        const char *names[3] = {"mul", "add_sub_add", "mul_pow2_24"};
        const double ops_per_iter[3] = {CHAINS, 1.5 * CHAINS, CHAINS};
        for (int k = 0; k < 3; k++) {
            v.kind = k;
            float ms = time_ms(launch_valu, &v, 5);
            double ops = (double)VALU_BLOCKS * 256 * v.iters * ops_per_iter[k];
            printf("{\"calib\": \"%s\", \"Gops_per_s\": %.1f, \"ms\": %.3f}\n", names[k], ops / (ms * 1e-3) / 1e9, ms);
        }
        CK(hipFree(v.out));
    }
    if (!strcmp(mode, "lds")) {
        u64 *out;
        CK(hipMalloc(&out, (size_t)LDS_BLOCKS * 256 * sizeof(u64)));
        lds_run<0>(out);
        CK(hipFree(out));
    }
    if (!strcmp(mode, "isa") || !strcmp(mode, "all")) {
        IsaArg a;
        CK(hipMalloc(&a.out, (size_t)ISA_BLOCKS * 256 * sizeof(u64)));
        int clk_khz = 0, cus = 0;
        CK(hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, 0));
        CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
        struct { const char *name; void (*k)(u64 *, unsigned); } ks[] = {
            {"v_add_u32", k_isa_add_u32},         {"v_xor_b32", k_isa_xor_b32},
            {"v_lshlrev_b32", k_isa_lshlrev_b32}, {"v_add_co_u32", k_isa_add_co_u32},
            {"v_cndmask_b32", k_isa_cndmask_b32}, {"v_mad_u64_u32", k_isa_mad_u64_u32},
            {"v_lshl_add_u64", k_isa_lshl_add_u64}, {"v_cmp_lt_u64", k_isa_cmp_lt_u64},
            {"v_lshlrev_b64", k_isa_lshlrev_b64}};
        for (auto &k : ks) {
            a.k = k.k;
            const float ms = time_ms(launch_isa, &a, 5);
            const double wave_insts = (double)ISA_BLOCKS * 256 / 64 * ISA_ITERS * CHAINS;
            const double simd_cycles = ms * 1e-3 * clk_khz * 1e3 * cus * 4;
            printf("{\"calib\": \"isa\", \"instruction\": \"%s\", \"simd_cycles_per_wave_instruction\": %.2f, "
                   "\"ms\": %.3f, \"clock_khz\": %d, \"cus\": %d}\n",
                   k.name, simd_cycles / wave_insts, ms, clk_khz, cus);
        }
        CK(hipFree(a.out));
    }
    return 0;
}
