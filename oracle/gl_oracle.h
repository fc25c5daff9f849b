/*
 * CPU oracle for the Goldilocks-NTT big-integer multiply.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a restatement in plain C of the
 * reference's CPU algorithm (joelodom/pi, crates/bignum/src/ntt.rs and the
 * multiply dispatch in crates/bignum/src/integer.rs).  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py may load it; the product library
 * (pi_amd/libgnt.so) never links, loads or calls anything here.
 *
 * All functions are exported with C linkage for ctypes.  Sizes are element
 * counts; limb arrays are little-endian uint64.
 */
#ifndef GL_ORACLE_H
#define GL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- field (ntt.rs:163-284) ---- */
uint64_t orc_canonicalize(uint64_t x);
uint64_t orc_add(uint64_t a, uint64_t b);
uint64_t orc_sub(uint64_t a, uint64_t b);
uint64_t orc_mul(uint64_t a, uint64_t b);
uint64_t orc_reduce128(uint64_t lo, uint64_t hi);
uint64_t orc_pow(uint64_t base, uint64_t exp);
uint64_t orc_inv(uint64_t a);
uint64_t orc_primitive_root_of_unity(uint64_t n);

/* ---- transforms (ntt.rs:300-391, 905-1000) ---- */
void orc_bit_reverse(uint64_t *a, size_t n);
void orc_ntt_forward(uint64_t *a, size_t n);
void orc_ntt_inverse(uint64_t *a, size_t n);
void orc_four_step_forward(uint64_t *a, size_t n);
void orc_four_step_inverse(uint64_t *a, size_t n);
void orc_build_twiddles_serial(uint64_t omega, size_t count, uint64_t *out);
void orc_build_twiddles_parallel(uint64_t omega, size_t count, uint64_t *out);
void orc_twiddle_cache_clear(void);
size_t orc_twiddle_cache_entries(void);

/* ---- pack / unpack (ntt.rs:633-791) ---- */
void orc_pack(const uint64_t *limbs, size_t nl, uint64_t *coeffs, size_t nc);
size_t orc_unpack(const uint64_t *coeffs, size_t nc, uint64_t *out, size_t cap);
size_t orc_unpack_serial(const uint64_t *coeffs, size_t nc, uint64_t *out, size_t cap);
size_t orc_unpack_parallel(const uint64_t *coeffs, size_t nc, uint64_t *out, size_t cap);

/* ---- multiply (ntt.rs:806-865; integer.rs:445-635) ----
 * `out` must hold na+nb limbs (na+nb+1 for karatsuba/dispatch).  Returns the
 * trimmed length (no most-significant zero limbs). */
size_t orc_mul_mag_ntt(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out);
size_t orc_mul_schoolbook(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out);
size_t orc_mul_karatsuba(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out);
size_t orc_mul_mag_dispatch(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out,
                            size_t ntt_max_coeffs, size_t ntt_threshold, size_t karatsuba_threshold);
/* orc_mul_mag_dispatch with the configured knobs and the 2^32 coefficient cap. */
size_t orc_mul_mag(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out);

/* ---- knobs (config.rs:94-127; config_gen.rs:56-99) ---- */
void orc_set_threads(int n);
int orc_get_threads(void);
void orc_set_four_step_threshold(size_t n);
size_t orc_get_four_step_threshold(void);
void orc_set_target_task_size(size_t n);
void orc_set_ntt_threshold(size_t n);
void orc_set_karatsuba_threshold(size_t n);
void orc_set_parallel_karatsuba_threshold(size_t n);
/* Pick knobs the way the reference's config generator does for `cores`. */
void orc_apply_host_knobs(int cores);

/* ---- deterministic generators ---- */
/* ntt.rs:1216-1226 test helper `random_vec` (LCG, values % P). */
void orc_random_vec(size_t n, uint64_t seed, uint64_t *out);
/* integer.rs:1644-1660 test helper `make_int` (LCG, top bit set). */
void orc_make_int(uint64_t seed, size_t n, uint64_t *out);
/* SplitMix64 stream (Steele, Lea, Flood 2014) used for the workload inputs. */
void orc_splitmix64_fill(uint64_t seed, size_t n, uint64_t *out);

/* ---- modular fingerprints (size-independent product checks) ---- */
uint64_t orc_mod_small(const uint64_t *a, size_t n, uint64_t q); /* value mod q, q < 2^63 */

#ifdef __cplusplus
}
#endif
#endif
