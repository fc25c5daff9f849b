// CPU emulation of libgnt's transform and carry kernels (test support).
//
// Runs the very same per-task code (pi_amd/csrc/gnt_pass.hpp) that the HIP
// kernels run, one task after another instead of one per thread, with the
// same plans and twiddle tables.  Tests use it to check the plan, index and
// twiddle logic of the GPU path on the CPU against the oracle and against a
// direct DFT; the GPU tests then check the device against the same oracle.
#include <cstring>
#include <vector>

#include "../../pi_amd/csrc/gnt_pass.hpp"
#include "../../pi_amd/csrc/gnt_roots.hpp"

using namespace gnt;

namespace {

std::vector<u64> g_lo, g_hi;

void tables() {
    if (g_lo.empty()) build_root_tables(g_lo, g_hi);
}

template <int LOGR, int I, bool FWD>
void steps(u64 *tile, int log_c, int tile_n, size_t cg0, int log_m, int log_s) {
    constexpr int count = (LOGR + 3) / 4;
    if constexpr (I < count) {
        constexpr int STEP = FWD ? I : count - 1 - I;
        const int ntask = tile_n >> Sub<LOGR>::log(STEP);
        for (int L = 0; L < ntask; L++) {
            if constexpr (FWD) fwd_task<LOGR, STEP>(tile, L, log_c, cg0, log_m, log_s, g_lo.data(), g_hi.data());
            else inv_task<LOGR, STEP>(tile, L, log_c, cg0, log_m, log_s, g_lo.data(), g_hi.data());
        }
        steps<LOGR, I + 1, FWD>(tile, log_c, tile_n, cg0, log_m, log_s);
    }
}

template <int LOGR, bool FWD>
void pass(const PassDesc &d, int log_n, u64 *x, const u64 *limbs, size_t n_limbs) {
    const int log_s = d.log_m - LOGR;
    const int tile_n = 1 << (LOGR + d.log_c);
    std::vector<u64> tile(tile_n);
    const size_t grid = ((size_t)1 << log_n) / tile_n;
    for (size_t blk = 0; blk < grid; blk++) {
        const size_t cg0 = blk << d.log_c;
        for (int idx = 0; idx < tile_n; idx++) {
            int n, c;
            tile_rc(idx, LOGR, d.log_c, log_s, n, c);
            size_t a = tile_addr(cg0, n, c, d.log_m, log_s);
            tile[(n << d.log_c) + c] = limbs ? coeff_from_limbs(limbs, n_limbs, a) : x[a];
        }
        steps<LOGR, 0, FWD>(tile.data(), d.log_c, tile_n, cg0, d.log_m, log_s);
        for (int idx = 0; idx < tile_n; idx++) {
            int n, c;
            tile_rc(idx, LOGR, d.log_c, log_s, n, c);
            x[tile_addr(cg0, n, c, d.log_m, log_s)] = tile[(n << d.log_c) + c];
        }
    }
}

#define DISPATCH(LR, CALL)                                                            \
    switch (LR) {                                                                     \
        case 1: CALL(1); break;  case 2: CALL(2); break;  case 3: CALL(3); break;     \
        case 4: CALL(4); break;  case 5: CALL(5); break;  case 6: CALL(6); break;     \
        case 7: CALL(7); break;  case 8: CALL(8); break;  case 9: CALL(9); break;     \
        case 10: CALL(10); break; case 11: CALL(11); break; case 12: CALL(12); break; \
        case 13: CALL(13); break;                                                     \
        default: break;                                                               \
    }

void forward(const Plan &pl, u64 *x, const u64 *limbs, size_t n_limbs) {
    for (int p = 0; p < pl.npass; p++) {
        const PassDesc &d = pl.pass[p];
        const u64 *src = p == 0 ? limbs : nullptr;
#define F(R) pass<R, true>(d, pl.log_n, x, src, n_limbs)
        DISPATCH(d.log_r, F)
#undef F
    }
}

void inverse(const Plan &pl, u64 *x) {
    for (int p = pl.npass - 1; p >= 0; p--) {
        const PassDesc &d = pl.pass[p];
#define F(R) pass<R, false>(d, pl.log_n, x, nullptr, 0)
        DISPATCH(d.log_r, F)
#undef F
    }
}

int ceil_log2(size_t x) {
    int l = 0;
    while (((size_t)1 << l) < x) l++;
    return l;
}

}  // namespace

extern "C" {

// Forward transform of x (length 2^log_n) in place, as the GPU computes it.
void emu_forward(u64 *x, int log_n) {
    tables();
    forward(make_plan(log_n), x, nullptr, 0);
}

void emu_inverse(u64 *x, int log_n) {
    tables();
    inverse(make_plan(log_n), x);
}

// omega_N of the library (primitive N-th root used by the transforms).
u64 emu_root(int log_n) {
    tables();
    u64 e = (u64)1 << (ROOT_LOG - log_n);
    return root_pow(g_lo.data(), g_hi.data(), (u32)e);
}

void emu_plan(int log_n, int *npass, int *log_r, int *log_c) {
    Plan pl = make_plan(log_n);
    *npass = pl.npass;
    for (int p = 0; p < 4; p++) {
        log_r[p] = p < pl.npass ? pl.pass[p].log_r : 0;
        log_c[p] = p < pl.npass ? pl.pass[p].log_c : 0;
    }
}

// Whole product as the GPU pipeline does it; returns the trimmed length.
// out holds na+nb limbs.  b == nullptr squares a.
size_t emu_mul(const u64 *a, size_t na, const u64 *b, size_t nb, u64 *out) {
    tables();
    const bool sqr = b == nullptr;
    if (sqr) nb = na;
    if (na == 0 || nb == 0) return 0;
    const size_t n_out = 4 * (na + nb) - 1;
    const int log_n = ceil_log2(n_out);
    const Plan pl = make_plan(log_n);
    const size_t N = (size_t)1 << log_n;
    std::vector<u64> xa(N), xb(sqr ? 0 : N);
    forward(pl, xa.data(), a, na);
    if (!sqr) forward(pl, xb.data(), b, nb);
    const u64 *pb = sqr ? xa.data() : xb.data();
    for (size_t i = 0; i < N; i++) xa[i] = gl::mul_pow2(gl::mul(xa[i], pb[i]), 192 - log_n);
    inverse(pl, xa.data());
    // carry, the same limb sums and carry rules as the kernels, run serially
    const size_t nl = na + nb;
    u64 Hprev = 0;
    u32 carry = 0;
    size_t len = 0;
    for (size_t j = 0; j < nl; j++) {
        u64 L, H;
        limb_sum(xa.data(), n_out, j, L, H);
        u64 s = L + Hprev;
        u32 g = s < L;
        u32 p = s == ~0ull;
        out[j] = s + carry;
        carry = g | (p & carry);
        Hprev = H;
        if (out[j]) len = j + 1;
    }
    return len;
}

// Field helpers for direct checks of the device arithmetic.
u64 emu_add(u64 a, u64 b) { return gl::add(a, b); }
u64 emu_sub(u64 a, u64 b) { return gl::sub(a, b); }
u64 emu_mul_mod(u64 a, u64 b) { return gl::mul(a, b); }
u64 emu_reduce128(u64 lo, u64 hi) { return gl::reduce128(lo, hi); }
u64 emu_mul_pow2(u64 x, int e) { return gl::mul_pow2(x, e); }

}  // extern "C"
