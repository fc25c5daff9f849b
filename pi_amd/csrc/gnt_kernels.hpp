// Launchers of the multiply kernels (device pointers, stream-ordered).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "gnt_plan.hpp"

namespace gnt {

// Forward transform in place on x (N = 2^pl.log_n u64).  With `limbs` set the
// first pass reads the operand's limbs and packs them (16 bits per
// coefficient, zero tail) instead of reading x.
int launch_forward(const Plan &pl, uint64_t *x, const uint64_t *limbs, size_t n_limbs, const uint64_t *wlo,
                   const uint64_t *whi, hipStream_t st);
// Inverse transform in place (unscaled; the pointwise step applies 1/N).
int launch_inverse(const Plan &pl, uint64_t *x, const uint64_t *wlo, const uint64_t *whi, hipStream_t st);
// a[i] = a[i] * b[i] / N  (b may alias a for a square).
int launch_pointwise(uint64_t *a, const uint64_t *b, int log_n, hipStream_t st);
// Exact carry of n_coeffs coefficients into n_limbs limbs; *out_len receives
// the trimmed length.  tile_gp / tile_cin hold carry_tiles(n_limbs) bytes.
size_t carry_tiles(size_t n_limbs);
int launch_carry(const uint64_t *coeffs, size_t n_coeffs, uint64_t *out, size_t n_limbs, unsigned char *tile_gp,
                 unsigned char *tile_cin, uint32_t *out_len, hipStream_t st);

}  // namespace gnt
