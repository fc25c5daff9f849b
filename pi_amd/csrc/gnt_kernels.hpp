// Launchers of the multiply kernels (device pointers, stream-ordered).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "gnt_pass.hpp"
#include "gnt_plan.hpp"

namespace gnt {

// Forward passes 0 .. npasses-1 in place on x (N = 2^pl.log_n u64).  With
// `limbs` set, pass 0 packs the operand's limbs (16 bits per coefficient,
// zero tail) instead of reading x.
int launch_forward(const Plan &pl, int npasses, uint64_t *x, const uint64_t *limbs, size_t n_limbs, const Tw &tw,
                   const PassTables &tabs, hipStream_t st);
// Forward passes 0 .. npasses-1 of two operands (same plan), one launch per pass.
int launch_forward_pair(const Plan &pl, int npasses, uint64_t *xa, const uint64_t *la, size_t na, uint64_t *xb,
                        const uint64_t *lb, size_t nb, const Tw &tw, const PassTables &tabs, hipStream_t st);
// Last forward pass of B + pointwise with A + 1/N + first inverse pass, result
// in xa.  square: B is A (xb == xa) and the pointwise product is A*A.  For a
// single-pass plan B is packed from limbs_b and the output is limb sums.
int launch_last_fused(const Plan &pl, const uint64_t *xb, const uint64_t *limbs_b, size_t nb, uint64_t *xa,
                      bool square, const Tw &tw, const PassTables &tabs, uint64_t *Ls, uint64_t *Hs,
                      size_t n_coeffs, size_t n_limbs, hipStream_t st);
// Multiply with a plan of two or more passes: last forward pass of A (xa) and
// of B (xb), pointwise, 1/N, first inverse pass; result in xa.
int launch_last_fused_ab(const Plan &pl, uint64_t *xa, const uint64_t *xb, const Tw &tw, const PassTables &tabs,
                         hipStream_t st);
// Inverse passes npass-2 .. 0; pass 0 writes limb sums Ls/Hs (n_limbs each).
// Twiddle tables (tabs): a forward pass with a table multiplies its twiddles
// from it on the way out; an inverse pass with a table has its inverse
// twiddles applied by the store of the stage before it (the fused stage or the
// previous inverse pass); passes without one step their twiddles.
int launch_inverse_rest(const Plan &pl, uint64_t *x, const Tw &tw, const PassTables &tabs, uint64_t *Ls,
                        uint64_t *Hs, size_t n_coeffs, size_t n_limbs, hipStream_t st);
// T[i] = pass_table_entry(i) for i < 2^log_m (a pass's twiddle table).
int launch_pass_table(uint64_t *T, int log_m, int log_r, bool inverse, int scale_e, const Tw &tw, hipStream_t st);
// Exact carry of the limb sums into n_limbs limbs; *out_len receives the
// trimmed length.  tile_gp / tile_cin hold carry_tiles(n_limbs) bytes.
size_t carry_tiles(size_t n_limbs);
int launch_carry(const uint64_t *Ls, const uint64_t *Hs, uint64_t *out, size_t n_limbs, unsigned char *tile_gp,
                 unsigned char *tile_cin, uint32_t *out_len, hipStream_t st);

}  // namespace gnt
