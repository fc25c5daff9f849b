/*
 * CPU oracle for the Goldilocks-NTT big-integer multiply.
 *
 * TEST INFRASTRUCTURE ONLY (see gl_oracle.h).  Restates the reference's CPU
 * algorithm step for step so that its products, its knobs and its timing can
 * stand in for the reference, which is Rust and cannot be built here.
 * rayon's fork-join is restated with OpenMP: `par_chunks`/`par_iter` become
 * `omp parallel for` over the same chunking, `rayon::join` becomes two OpenMP
 * tasks.  Citations are to files of joelodom/pi.
 */
#define _GNU_SOURCE
#include "gl_oracle.h"

#include <omp.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ntt.rs:164-172 */
#define GL_P 0xFFFFFFFF00000001ULL
#define GL_EPS 0xFFFFFFFFULL
#define GL_GENERATOR 7ULL

/* config.rs:94-127 defaults */
static size_t cfg_target_task = (size_t)1 << 16;
static size_t cfg_parallel_pack = 1024;
static size_t cfg_parallel_pointwise = 1024;
static size_t cfg_four_step = (size_t)1 << 25;
static size_t cfg_ntt_threshold = 8192;
static size_t cfg_karatsuba = 32;
static size_t cfg_parallel_karatsuba = 512;
static int cfg_threads = 0; /* 0 = all cores */

static int nthreads(void) { return cfg_threads > 0 ? cfg_threads : omp_get_num_procs(); }

void orc_set_threads(int n) { cfg_threads = n; }
int orc_get_threads(void) { return nthreads(); }
void orc_set_four_step_threshold(size_t n) { cfg_four_step = n; }
size_t orc_get_four_step_threshold(void) { return cfg_four_step; }
void orc_set_target_task_size(size_t n) { cfg_target_task = n; }
void orc_set_ntt_threshold(size_t n) { cfg_ntt_threshold = n; }
void orc_set_karatsuba_threshold(size_t n) { cfg_karatsuba = n; }
void orc_set_parallel_karatsuba_threshold(size_t n) { cfg_parallel_karatsuba = n; }

/* crates/pi-cli/src/config_gen.rs:56-99: knob picks by core count. */
void orc_apply_host_knobs(int cores) {
    cfg_parallel_karatsuba = cores <= 12 ? 512 : cores <= 32 ? 256 : cores <= 64 ? 192 : 128;
    cfg_target_task = cores <= 12 ? 65536 : cores <= 32 ? 32768 : 16384;
    cfg_four_step = cores <= 12 ? ((size_t)1 << 26) : cores <= 64 ? ((size_t)1 << 25) : ((size_t)1 << 24);
}

/* ===================== field: ntt.rs:177-284 ===================== */

uint64_t orc_canonicalize(uint64_t x) { return x >= GL_P ? x - GL_P : x; }

uint64_t orc_add(uint64_t a, uint64_t b) {
    uint64_t s;
    int carry = __builtin_add_overflow(a, b, &s);
    uint64_t r = carry ? s + GL_EPS : s;
    return orc_canonicalize(r);
}

uint64_t orc_sub(uint64_t a, uint64_t b) {
    uint64_t d;
    int borrow = __builtin_sub_overflow(a, b, &d);
    return borrow ? d - GL_EPS : d;
}

static inline uint64_t reduce128(u128 x) {
    uint64_t x_lo = (uint64_t)x;
    uint64_t x_hi = (uint64_t)(x >> 64);
    uint64_t x_hi_hi = x_hi >> 32;
    uint64_t x_hi_lo = x_hi & 0xFFFFFFFFULL;
    uint64_t t0;
    int borrow = __builtin_sub_overflow(x_lo, x_hi_hi, &t0);
    if (borrow) t0 -= GL_EPS;
    t0 = orc_canonicalize(t0);
    uint64_t t1 = x_hi_lo * GL_EPS;
    uint64_t s;
    int carry = __builtin_add_overflow(t0, t1, &s);
    if (carry) s += GL_EPS;
    return orc_canonicalize(s);
}

uint64_t orc_reduce128(uint64_t lo, uint64_t hi) { return reduce128(((u128)hi << 64) | lo); }

uint64_t orc_mul(uint64_t a, uint64_t b) { return reduce128((u128)a * b); }

uint64_t orc_pow(uint64_t base, uint64_t exp) {
    uint64_t acc = 1;
    while (exp > 0) {
        if (exp & 1) acc = orc_mul(acc, base);
        exp >>= 1;
        if (exp > 0) base = orc_mul(base, base);
    }
    return acc;
}

uint64_t orc_inv(uint64_t a) { return orc_pow(a, GL_P - 2); }

uint64_t orc_primitive_root_of_unity(uint64_t n) { return orc_pow(GL_GENERATOR, (GL_P - 1) / n); }

#define F_ADD orc_add
#define F_SUB orc_sub
#define F_MUL orc_mul

/* ===================== bit reverse: ntt.rs:300-361 ===================== */

#define PARALLEL_BIT_REVERSE_THRESHOLD ((size_t)1 << 18)

static inline uint64_t rev_bits(uint64_t i, unsigned log_n) {
    uint64_t x = i;
    x = ((x >> 1) & 0x5555555555555555ULL) | ((x & 0x5555555555555555ULL) << 1);
    x = ((x >> 2) & 0x3333333333333333ULL) | ((x & 0x3333333333333333ULL) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((x & 0x0F0F0F0F0F0F0F0FULL) << 4);
    x = __builtin_bswap64(x);
    return x >> (64 - log_n);
}

void orc_bit_reverse(uint64_t *a, size_t n) {
    if (n <= 1) return;
    unsigned log_n = (unsigned)__builtin_ctzll(n);
    if (n >= PARALLEL_BIT_REVERSE_THRESHOLD) {
#pragma omp parallel for num_threads(nthreads()) schedule(static)
        for (size_t i = 0; i < n; i++) {
            size_t j = (size_t)rev_bits(i, log_n);
            if (i < j) { uint64_t t = a[i]; a[i] = a[j]; a[j] = t; }
        }
    } else {
        for (size_t i = 0; i < n; i++) {
            size_t j = (size_t)rev_bits(i, log_n);
            if (i < j) { uint64_t t = a[i]; a[i] = a[j]; a[j] = t; }
        }
    }
}

/* ===================== twiddles: ntt.rs:482-624 ===================== */

#define PARALLEL_TWIDDLE_THRESHOLD 8192
#define MAX_CACHED_TWIDDLE_COUNT ((size_t)1 << 28)

void orc_build_twiddles_serial(uint64_t omega, size_t count, uint64_t *out) {
    if (count == 0) return;
    uint64_t w = 1;
    out[0] = w;
    for (size_t i = 1; i < count; i++) { w = F_MUL(w, omega); out[i] = w; }
}

void orc_build_twiddles_parallel(uint64_t omega, size_t count, uint64_t *out) {
    const size_t MIN_CHUNK = 4096;
    size_t max_chunks = (size_t)nthreads();
    size_t num_chunks = count / MIN_CHUNK;
    if (num_chunks < 1) num_chunks = 1;
    if (num_chunks > max_chunks) num_chunks = max_chunks;
    size_t chunk = (count + num_chunks - 1) / num_chunks;
    uint64_t chunk_omega = orc_pow(omega, chunk);
    uint64_t *starts = (uint64_t *)malloc(num_chunks * sizeof(uint64_t));
    uint64_t w = 1;
    for (size_t c = 0; c < num_chunks; c++) { starts[c] = w; w = F_MUL(w, chunk_omega); }
#pragma omp parallel for num_threads(nthreads()) schedule(static, 1)
    for (size_t c = 0; c < num_chunks; c++) {
        size_t lo = c * chunk, hi = lo + chunk < count ? lo + chunk : count;
        uint64_t x = starts[c];
        for (size_t i = lo; i < hi; i++) { out[i] = x; x = F_MUL(x, omega); }
    }
    free(starts);
}

static void build_twiddles(uint64_t omega, size_t count, uint64_t *out) {
    if (count < PARALLEL_TWIDDLE_THRESHOLD) orc_build_twiddles_serial(omega, count, out);
    else orc_build_twiddles_parallel(omega, count, out);
}

/* Process-wide cache keyed by (count, inverse); count is a power of two, so
 * the key is (log2 count, inverse).  Tables are built outside the lock. */
static pthread_mutex_t tw_lock = PTHREAD_MUTEX_INITIALIZER;
static uint64_t *tw_cache[2][40];

static const uint64_t *twiddle_get(size_t count, int inverse, int *owned) {
    unsigned lg = (unsigned)__builtin_ctzll(count);
    *owned = 0;
    pthread_mutex_lock(&tw_lock);
    uint64_t *t = tw_cache[inverse][lg];
    pthread_mutex_unlock(&tw_lock);
    if (t) return t;
    size_t len = count * 2 < 2 ? 2 : count * 2;
    uint64_t base = orc_primitive_root_of_unity(len);
    uint64_t omega = inverse ? orc_inv(base) : base;
    uint64_t *built = (uint64_t *)malloc(count * sizeof(uint64_t));
    build_twiddles(omega, count, built);
    if (count <= MAX_CACHED_TWIDDLE_COUNT) {
        pthread_mutex_lock(&tw_lock);
        if (tw_cache[inverse][lg]) { free(built); built = tw_cache[inverse][lg]; }
        else tw_cache[inverse][lg] = built;
        pthread_mutex_unlock(&tw_lock);
        return built;
    }
    *owned = 1;
    return built;
}

void orc_twiddle_cache_clear(void) {
    pthread_mutex_lock(&tw_lock);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 40; j++) { free(tw_cache[i][j]); tw_cache[i][j] = NULL; }
    pthread_mutex_unlock(&tw_lock);
}

size_t orc_twiddle_cache_entries(void) {
    size_t c = 0;
    pthread_mutex_lock(&tw_lock);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 40; j++) c += tw_cache[i][j] != NULL;
    pthread_mutex_unlock(&tw_lock);
    return c;
}

/* ===================== butterflies: ntt.rs:406-480 ===================== */

static inline void butterfly_group(uint64_t *chunk, const uint64_t *tw, size_t half) {
    for (size_t j = 0; j < half; j++) {
        uint64_t u = chunk[j];
        uint64_t t = F_MUL(tw[j], chunk[j + half]);
        chunk[j] = F_ADD(u, t);
        chunk[j + half] = F_SUB(u, t);
    }
}

static void butterflies(uint64_t *a, size_t n, int inverse) {
    size_t target = cfg_target_task;
    for (size_t len = 2; len <= n; len *= 2) {
        size_t half = len / 2;
        int owned;
        const uint64_t *tw = twiddle_get(half, inverse, &owned);
        size_t groups = n / len;
        if (half >= target && groups < 4) {
            /* late pass: intra-group split of [0, half) into target-sized chunks */
            for (size_t g = 0; g < groups; g++) {
                uint64_t *lo = a + g * len, *hi = lo + half;
                size_t nchunks = (half + target - 1) / target;
#pragma omp parallel for num_threads(nthreads()) schedule(static, 1)
                for (size_t c = 0; c < nchunks; c++) {
                    size_t b = c * target, e = b + target < half ? b + target : half;
                    for (size_t j = b; j < e; j++) {
                        uint64_t u = lo[j];
                        uint64_t t = F_MUL(tw[j], hi[j]);
                        lo[j] = F_ADD(u, t);
                        hi[j] = F_SUB(u, t);
                    }
                }
            }
        } else {
            size_t task = len >= target ? len : (n >= target ? target : n);
            if (n > task) {
                size_t ntasks = n / task;
#pragma omp parallel for num_threads(nthreads()) schedule(static, 1)
                for (size_t c = 0; c < ntasks; c++) {
                    uint64_t *big = a + c * task;
                    for (size_t off = 0; off < task; off += len) butterfly_group(big + off, tw, half);
                }
            } else {
                for (size_t off = 0; off < n; off += len) butterfly_group(a + off, tw, half);
            }
        }
        if (owned) free((void *)tw);
    }
}

/* ntt.rs:367-391 */
void orc_ntt_forward(uint64_t *a, size_t n) {
    orc_bit_reverse(a, n);
    butterflies(a, n, 0);
}

void orc_ntt_inverse(uint64_t *a, size_t n) {
    orc_bit_reverse(a, n);
    butterflies(a, n, 1);
    uint64_t n_inv = orc_inv((uint64_t)n);
    if (n >= cfg_parallel_pointwise) {
#pragma omp parallel for num_threads(nthreads()) schedule(static)
        for (size_t i = 0; i < n; i++) a[i] = F_MUL(a[i], n_inv);
    } else {
        for (size_t i = 0; i < n; i++) a[i] = F_MUL(a[i], n_inv);
    }
}

/* ===================== four-step: ntt.rs:905-1105 ===================== */

static void pick_factorization(size_t n, size_t *n1, size_t *n2) {
    unsigned lg = (unsigned)__builtin_ctzll(n);
    unsigned l1 = lg / 2, l2 = lg - l1;
    *n1 = (size_t)1 << l1;
    *n2 = (size_t)1 << l2;
}

static void transpose(uint64_t *a, size_t rows, size_t cols) {
    if (rows == cols) {
        size_t m = rows;
#pragma omp parallel for num_threads(nthreads()) schedule(dynamic, 16)
        for (size_t i = 0; i < m; i++)
            for (size_t j = i + 1; j < m; j++) {
                uint64_t t = a[i * m + j]; a[i * m + j] = a[j * m + i]; a[j * m + i] = t;
            }
    } else {
        size_t n = rows * cols;
        uint64_t *scratch = (uint64_t *)malloc(n * sizeof(uint64_t));
#pragma omp parallel for num_threads(nthreads()) schedule(static)
        for (size_t r = 0; r < rows; r++)
            for (size_t c = 0; c < cols; c++) scratch[c * rows + r] = a[r * cols + c];
        memcpy(a, scratch, n * sizeof(uint64_t));
        free(scratch);
    }
}

static void apply_cross_twiddle(uint64_t *a, size_t n, size_t n1, size_t n2, int inverse) {
    uint64_t base = orc_primitive_root_of_unity(n);
    uint64_t omega = inverse ? orc_inv(base) : base;
    uint64_t *step = (uint64_t *)malloc(n2 * sizeof(uint64_t));
    uint64_t w = 1;
    for (size_t s = 0; s < n2; s++) { step[s] = w; w = F_MUL(w, omega); }
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (size_t s = 0; s < n2; s++) {
        uint64_t *row = a + s * n1;
        uint64_t tw = 1, rs = step[s];
        for (size_t i = 0; i < n1; i++) { row[i] = F_MUL(row[i], tw); tw = F_MUL(tw, rs); }
    }
    free(step);
}

/* Sub-FFTs of a four-step run serially inside each parallel row task, as
 * rayon's par_chunks_exact_mut does (each row call sees few elements). */
static void rows_forward(uint64_t *a, size_t rows, size_t len) {
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (size_t r = 0; r < rows; r++) orc_ntt_forward(a + r * len, len);
}

static void rows_inverse(uint64_t *a, size_t rows, size_t len) {
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (size_t r = 0; r < rows; r++) orc_ntt_inverse(a + r * len, len);
}

void orc_four_step_forward(uint64_t *a, size_t n) {
    if (n <= 1) return;
    size_t n1, n2;
    pick_factorization(n, &n1, &n2);
    transpose(a, n1, n2);
    rows_forward(a, n2, n1);
    apply_cross_twiddle(a, n, n1, n2, 0);
    transpose(a, n2, n1);
    rows_forward(a, n1, n2);
}

void orc_four_step_inverse(uint64_t *a, size_t n) {
    if (n <= 1) return;
    size_t n1, n2;
    pick_factorization(n, &n1, &n2);
    rows_inverse(a, n1, n2);
    transpose(a, n1, n2);
    apply_cross_twiddle(a, n, n1, n2, 1);
    rows_inverse(a, n2, n1);
    transpose(a, n2, n1);
}

/* ===================== pack / unpack: ntt.rs:633-791 ===================== */

#define COEFFS_PER_LIMB 4
#define PARALLEL_UNPACK_THRESHOLD ((size_t)1 << 20)

void orc_pack(const uint64_t *limbs, size_t nl, uint64_t *coeffs, size_t nc) {
    size_t needed = nl * COEFFS_PER_LIMB;
    if (nc < needed) abort(); /* the reference asserts here */
    if (nl >= cfg_parallel_pack) {
#pragma omp parallel for num_threads(nthreads()) schedule(static)
        for (size_t i = 0; i < nl; i++) {
            uint64_t l = limbs[i];
            coeffs[4 * i + 0] = l & 0xFFFF;
            coeffs[4 * i + 1] = (l >> 16) & 0xFFFF;
            coeffs[4 * i + 2] = (l >> 32) & 0xFFFF;
            coeffs[4 * i + 3] = (l >> 48) & 0xFFFF;
        }
#pragma omp parallel for num_threads(nthreads()) schedule(static)
        for (size_t i = needed; i < nc; i++) coeffs[i] = 0;
    } else {
        for (size_t i = 0; i < nl; i++) {
            uint64_t l = limbs[i];
            coeffs[4 * i + 0] = l & 0xFFFF;
            coeffs[4 * i + 1] = (l >> 16) & 0xFFFF;
            coeffs[4 * i + 2] = (l >> 32) & 0xFFFF;
            coeffs[4 * i + 3] = (l >> 48) & 0xFFFF;
        }
        for (size_t i = needed; i < nc; i++) coeffs[i] = 0;
    }
}

static size_t trim(const uint64_t *out, size_t len) {
    while (len > 0 && out[len - 1] == 0) len--;
    return len;
}

/* Output capacity needed: ceil(nc/4) + 2 limbs. */
size_t orc_unpack_serial(const uint64_t *coeffs, size_t nc, uint64_t *out, size_t cap) {
    if (nc == 0) return 0;
    size_t len = 0, i = 0;
    u128 carry = 0;
    while (i < nc) {
        for (int k = 0; k < COEFFS_PER_LIMB && i < nc; k++, i++) carry += (u128)coeffs[i] << (16 * k);
        if (len >= cap) abort();
        out[len++] = (uint64_t)carry;
        carry >>= 64;
    }
    while (carry != 0) {
        if (len >= cap) abort();
        out[len++] = (uint64_t)carry;
        carry >>= 64;
    }
    return trim(out, len);
}

size_t orc_unpack_parallel(const uint64_t *coeffs, size_t nc, uint64_t *out, size_t cap) {
    if (nc == 0) return 0;
    size_t total_blocks = nc / COEFFS_PER_LIMB;
    size_t max_chunks = (size_t)nthreads();
    size_t num_chunks = total_blocks < max_chunks ? total_blocks : max_chunks;
    if (num_chunks < 1) num_chunks = 1;
    size_t bpc = (total_blocks + num_chunks - 1) / num_chunks;
    if (bpc == 0) bpc = 1;
    size_t cpc = bpc * COEFFS_PER_LIMB;
    size_t nchunks = (nc + cpc - 1) / cpc;
    uint64_t **limbs = (uint64_t **)calloc(nchunks, sizeof(uint64_t *));
    size_t *counts = (size_t *)calloc(nchunks, sizeof(size_t));
    u128 *finals = (u128 *)calloc(nchunks, sizeof(u128));
#pragma omp parallel for num_threads(nthreads()) schedule(static, 1)
    for (size_t c = 0; c < nchunks; c++) {
        size_t lo = c * cpc, hi = lo + cpc < nc ? lo + cpc : nc;
        size_t cnt = (hi - lo + COEFFS_PER_LIMB - 1) / COEFFS_PER_LIMB;
        uint64_t *l = (uint64_t *)malloc((cnt ? cnt : 1) * sizeof(uint64_t));
        u128 carry = 0;
        size_t i = lo, k = 0;
        while (i < hi) {
            for (int t = 0; t < COEFFS_PER_LIMB && i < hi; t++, i++) carry += (u128)coeffs[i] << (16 * t);
            l[k++] = (uint64_t)carry;
            carry >>= 64;
        }
        limbs[c] = l;
        counts[c] = k;
        finals[c] = carry;
    }
    size_t len = 0;
    u128 spare = 0;
    for (size_t c = 0; c < nchunks; c++) {
        for (size_t k = 0; k < counts[c]; k++) {
            u128 s = (u128)limbs[c][k] + spare;
            if (len >= cap) abort();
            out[len++] = (uint64_t)s;
            spare = s >> 64;
        }
        spare += finals[c];
        free(limbs[c]);
    }
    while (spare != 0) {
        if (len >= cap) abort();
        out[len++] = (uint64_t)spare;
        spare >>= 64;
    }
    free(limbs); free(counts); free(finals);
    return trim(out, len);
}

size_t orc_unpack(const uint64_t *coeffs, size_t nc, uint64_t *out, size_t cap) {
    if (nc == 0) return 0;
    if (nc < PARALLEL_UNPACK_THRESHOLD) return orc_unpack_serial(coeffs, nc, out, cap);
    return orc_unpack_parallel(coeffs, nc, out, cap);
}

/* ===================== mul_mag_ntt: ntt.rs:806-865 ===================== */

size_t orc_mul_mag_ntt(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out) {
    if (na == 0 || nb == 0) return 0;
    size_t n_out = na * COEFFS_PER_LIMB + nb * COEFFS_PER_LIMB - 1;
    size_t n = 1;
    while (n < n_out) n <<= 1;
    uint64_t *pa = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *pb = (uint64_t *)malloc(n * sizeof(uint64_t));
    orc_pack(a, na, pa, n);
    orc_pack(b, nb, pb, n);
    int four = n >= cfg_four_step;
    if (four) { orc_four_step_forward(pa, n); orc_four_step_forward(pb, n); }
    else { orc_ntt_forward(pa, n); orc_ntt_forward(pb, n); }
    if (n >= cfg_parallel_pointwise) {
#pragma omp parallel for num_threads(nthreads()) schedule(static)
        for (size_t i = 0; i < n; i++) pa[i] = F_MUL(pa[i], pb[i]);
    } else {
        for (size_t i = 0; i < n; i++) pa[i] = F_MUL(pa[i], pb[i]);
    }
    free(pb);
    if (four) orc_four_step_inverse(pa, n);
    else orc_ntt_inverse(pa, n);
    /* the product has at most na+nb limbs; unpack's scratch may need +2 */
    size_t cap = na + nb + 2;
    uint64_t *tmp = (uint64_t *)malloc(cap * sizeof(uint64_t));
    size_t len = orc_unpack(pa, n_out, tmp, cap);
    memcpy(out, tmp, len * sizeof(uint64_t));
    free(tmp);
    free(pa);
    return len;
}

/* ===================== dispatch: integer.rs:351-635 ===================== */

size_t orc_mul_schoolbook(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out) {
    if (na == 0 || nb == 0) return 0;
    memset(out, 0, (na + nb) * sizeof(uint64_t));
    for (size_t i = 0; i < na; i++) {
        u128 carry = 0;
        u128 av = a[i];
        for (size_t j = 0; j < nb; j++) {
            u128 cur = (u128)out[i + j] + av * b[j] + carry;
            out[i + j] = (uint64_t)cur;
            carry = cur >> 64;
        }
        size_t k = i + nb;
        while (carry > 0) {
            u128 cur = (u128)out[k] + carry;
            out[k] = (uint64_t)cur;
            carry = cur >> 64;
            k++;
        }
    }
    return trim(out, na + nb);
}

/* add_mag (integer.rs:351-366): out holds max(na,nb)+1 limbs; length not
 * normalised, exactly as the reference (it only appends a final carry). */
static size_t add_mag(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out) {
    const uint64_t *lg = na >= nb ? a : b, *sh = na >= nb ? b : a;
    size_t nl = na >= nb ? na : nb, ns = na >= nb ? nb : na;
    u128 carry = 0;
    for (size_t i = 0; i < nl; i++) {
        u128 s = (u128)lg[i] + (i < ns ? sh[i] : 0) + carry;
        out[i] = (uint64_t)s;
        carry = s >> 64;
    }
    if (carry) { out[nl] = (uint64_t)carry; return nl + 1; }
    return nl;
}

/* sub_mag (integer.rs:369-389), a >= b; result normalised. */
static size_t sub_mag(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out) {
    uint64_t borrow = 0;
    for (size_t i = 0; i < na; i++) {
        uint64_t bi = i < nb ? b[i] : 0;
        uint64_t d1, d2;
        uint64_t b1 = __builtin_sub_overflow(a[i], bi, &d1);
        uint64_t b2 = __builtin_sub_overflow(d1, borrow, &d2);
        out[i] = d2;
        borrow = b1 | b2;
    }
    return trim(out, na);
}

/* add_at (integer.rs:612-635): dest has room for the carry by construction. */
static void add_at(uint64_t *dest, size_t dest_len, const uint64_t *src, size_t ns, size_t off) {
    u128 carry = 0;
    for (size_t i = 0; i < ns; i++) {
        u128 s = (u128)dest[off + i] + src[i] + carry;
        dest[off + i] = (uint64_t)s;
        carry = s >> 64;
    }
    size_t k = off + ns;
    while (carry > 0 && k < dest_len) {
        u128 s = (u128)dest[k] + carry;
        dest[k] = (uint64_t)s;
        carry = s >> 64;
        k++;
    }
}

static size_t g_ntt_max_coeffs = (size_t)1 << 32;

static size_t dispatch(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out,
                       size_t max_coeffs, size_t ntt_thr, size_t kara_thr);

/* mul_mag_karatsuba (integer.rs:552-608); out holds na+nb+1 limbs. */
static size_t karatsuba(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out,
                        size_t max_coeffs, size_t ntt_thr, size_t kara_thr) {
    size_t n = na > nb ? na : nb;
    size_t m = n / 2;
    size_t la = m < na ? m : na, lb = m < nb ? m : nb;
    const uint64_t *a_lo = a, *a_hi = a + la, *b_lo = b, *b_hi = b + lb;
    size_t na_hi = na - la, nb_hi = nb - lb;
    uint64_t *a_sum = (uint64_t *)malloc(((la > na_hi ? la : na_hi) + 1) * sizeof(uint64_t));
    uint64_t *b_sum = (uint64_t *)malloc(((lb > nb_hi ? lb : nb_hi) + 1) * sizeof(uint64_t));
    size_t nas = add_mag(a_lo, la, a_hi, na_hi, a_sum);
    size_t nbs = add_mag(b_lo, lb, b_hi, nb_hi, b_sum);
    uint64_t *z0 = (uint64_t *)malloc((la + lb + 1) * sizeof(uint64_t));
    uint64_t *z2 = (uint64_t *)malloc((na_hi + nb_hi + 1) * sizeof(uint64_t));
    uint64_t *z1 = (uint64_t *)malloc((nas + nbs + 1) * sizeof(uint64_t));
    size_t n0 = 0, n2 = 0, n1 = 0;
    int par = (na < nb ? na : nb) >= cfg_parallel_karatsuba;
    if (par) {
#pragma omp taskgroup
        {
#pragma omp task shared(n0) if (omp_in_parallel())
            n0 = dispatch(a_lo, la, b_lo, lb, z0, max_coeffs, ntt_thr, kara_thr);
#pragma omp task shared(n2) if (omp_in_parallel())
            n2 = dispatch(a_hi, na_hi, b_hi, nb_hi, z2, max_coeffs, ntt_thr, kara_thr);
            n1 = dispatch(a_sum, nas, b_sum, nbs, z1, max_coeffs, ntt_thr, kara_thr);
        }
    } else {
        n0 = dispatch(a_lo, la, b_lo, lb, z0, max_coeffs, ntt_thr, kara_thr);
        n2 = dispatch(a_hi, na_hi, b_hi, nb_hi, z2, max_coeffs, ntt_thr, kara_thr);
        n1 = dispatch(a_sum, nas, b_sum, nbs, z1, max_coeffs, ntt_thr, kara_thr);
    }
    size_t n02max = (n0 > n2 ? n0 : n2) + 1;
    uint64_t *z02 = (uint64_t *)malloc(n02max * sizeof(uint64_t));
    size_t n02 = add_mag(z0, n0, z2, n2, z02);
    uint64_t *z1d = (uint64_t *)malloc((n1 ? n1 : 1) * sizeof(uint64_t));
    size_t n1d = sub_mag(z1, n1, z02, n02, z1d);
    size_t olen = na + nb + 1;
    memset(out, 0, olen * sizeof(uint64_t));
    add_at(out, olen, z0, n0, 0);
    add_at(out, olen, z1d, n1d, m);
    add_at(out, olen, z2, n2, 2 * m);
    free(a_sum); free(b_sum); free(z0); free(z1); free(z2); free(z02); free(z1d);
    return trim(out, olen);
}

/* mul_mag_dispatch (integer.rs:460-486). */
static size_t dispatch(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out,
                       size_t max_coeffs, size_t ntt_thr, size_t kara_thr) {
    if (na == 0 || nb == 0) return 0;
    size_t smaller = na < nb ? na : nb;
    int fits = (na + nb) * COEFFS_PER_LIMB <= max_coeffs;
    if (fits && smaller >= ntt_thr) return orc_mul_mag_ntt(a, na, b, nb, out);
    if (smaller < kara_thr) return orc_mul_schoolbook(a, na, b, nb, out);
    return karatsuba(a, na, b, nb, out, max_coeffs, ntt_thr, kara_thr);
}

size_t orc_mul_karatsuba(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out) {
    if (na == 0 || nb == 0) return 0;
    return karatsuba(a, na, b, nb, out, g_ntt_max_coeffs, cfg_ntt_threshold, cfg_karatsuba);
}

size_t orc_mul_mag_dispatch(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out,
                            size_t ntt_max_coeffs, size_t ntt_threshold, size_t karatsuba_threshold) {
    if (na == 0 || nb == 0) return 0;
    size_t smaller = na < nb ? na : nb;
    /* An NTT or schoolbook leaf at the top runs its own parallel loops; only
     * a Karatsuba top opens a task region for its three-way fork. */
    if (((na + nb) * COEFFS_PER_LIMB <= ntt_max_coeffs && smaller >= ntt_threshold) ||
        smaller < karatsuba_threshold)
        return dispatch(a, na, b, nb, out, ntt_max_coeffs, ntt_threshold, karatsuba_threshold);
    size_t r = 0;
#pragma omp parallel num_threads(nthreads())
#pragma omp single
    r = dispatch(a, na, b, nb, out, ntt_max_coeffs, ntt_threshold, karatsuba_threshold);
    return r;
}

size_t orc_mul_mag(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out) {
    return orc_mul_mag_dispatch(a, na, b, nb, out, g_ntt_max_coeffs, cfg_ntt_threshold, cfg_karatsuba);
}

/* ===================== generators ===================== */

void orc_random_vec(size_t n, uint64_t seed, uint64_t *out) {
    uint64_t s = seed * 0x9E3779B97F4A7C15ULL + 1;
    for (size_t i = 0; i < n; i++) {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        out[i] = s % GL_P;
    }
}

void orc_make_int(uint64_t seed, size_t n, uint64_t *out) {
    uint64_t s = seed * 0x9E3779B97F4A7C15ULL;
    for (size_t i = 0; i < n; i++) {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        out[i] = s;
    }
    if (n > 0) out[n - 1] |= 1ULL << 63;
}

void orc_splitmix64_fill(uint64_t seed, size_t n, uint64_t *out) {
    uint64_t x = seed;
    for (size_t i = 0; i < n; i++) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        out[i] = z ^ (z >> 31);
    }
}

uint64_t orc_mod_small(const uint64_t *a, size_t n, uint64_t q) {
    /* Horner from the top limb: r = (r * 2^64 + limb) mod q */
    u128 r = 0;
    for (size_t i = n; i-- > 0;) r = ((r << 64) | a[i]) % q;
    return (uint64_t)r;
}
