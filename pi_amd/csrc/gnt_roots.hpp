// Root of unity and twiddle tables of the library (host side).
#pragma once

#include <vector>

#include "gl_field.hpp"
#include "gnt_plan.hpp"

namespace gnt {

// r: a primitive 2^32-th root of unity with r^(2^26) = 8, found as a power
// of 7^((p-1)/2^32) (7 generates the multiplicative group, ntt.rs:172).
// With it the 64th root is 2^3 and every root of order <= 64 is a power of
// two.  Returns 0 if the search fails (it does not for this p).
inline gl::u64 library_root() {
    const gl::u64 r0 = gl::pow(7, (gl::P - 1) >> 32);
    const gl::u64 w64 = gl::pow(r0, (gl::u64)1 << 26);  // primitive 64th root = 8^m, m odd
    gl::u64 p8 = 1;
    int m = -1;
    for (int k = 0; k < 64; k++) {
        if (p8 == w64) { m = k; break; }
        p8 = gl::mul(p8, 8);
    }
    if (m < 0 || (m & 1) == 0) return 0;
    gl::u32 inv = (gl::u32)m;  // m^-1 mod 2^32 by Newton iteration
    for (int i = 0; i < 5; i++) inv *= 2u - (gl::u32)m * inv;
    return gl::pow(r0, inv);
}

// lo[i] = r^i, hi[i] = r^(i * 2^16), i < 2^16: r^e = lo[e & 0xFFFF] * hi[e >> 16].
// sm[2^LM + e] = omega_{2^LM}^e = r^(e * 2^(32-LM)) for LM <= TILE_LOG, e < 2^LM.
inline bool build_root_tables(std::vector<gl::u64> &lo, std::vector<gl::u64> &hi, std::vector<gl::u64> &sm) {
    const gl::u64 r = library_root();
    if (r == 0 || gl::pow(r, (gl::u64)1 << 26) != 8) return false;
    const size_t T = (size_t)1 << TW_SPLIT;
    lo.assign(T, 0);
    hi.assign(T, 0);
    gl::u64 w = 1;
    for (size_t i = 0; i < T; i++) { lo[i] = w; w = gl::mul(w, r); }
    const gl::u64 step = w;
    w = 1;
    for (size_t i = 0; i < T; i++) { hi[i] = w; w = gl::mul(w, step); }
    if (w != 1) return false;  // r^(2^32) == 1
    sm.assign((size_t)2 << TILE_LOG, 0);
    for (int lm = 0; lm <= TILE_LOG; lm++) {
        const gl::u64 om = gl::pow(r, (gl::u64)1 << (32 - lm));
        gl::u64 x = 1;
        for (size_t e = 0; e < ((size_t)1 << lm); e++) { sm[((size_t)1 << lm) + e] = x; x = gl::mul(x, om); }
    }
    return true;
}

}  // namespace gnt
