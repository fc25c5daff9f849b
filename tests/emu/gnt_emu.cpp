// CPU emulation of libgnt's transform and carry kernels (test support).
//
// Runs the very same per-task code (pi_amd/csrc/gnt_pass.hpp) that the HIP
// kernels run, one task after another instead of one per thread, with the
// same plans and twiddle tables.  Tests use it to check the plan, index and
// twiddle logic of the GPU path on the CPU against the oracle and against a
// direct DFT; the GPU tests then check the device against the same oracle.
#include <cstring>
#include <map>
#include <vector>

#include "../../pi_amd/csrc/gnt_pass.hpp"
#include "../../pi_amd/csrc/gnt_roots.hpp"

using namespace gnt;

namespace {

std::vector<u64> g_lo, g_hi, g_sm;

void tables() {
    if (g_lo.empty()) build_root_tables(g_lo, g_hi, g_sm);
}

Tw tw() { return Tw{g_lo.data(), g_hi.data(), g_sm.data()}; }

// Pass twiddle tables, as the runtime builds them on the device.
std::map<int, std::vector<u64>> g_ptab;

PassTables tables_for(const Plan &pl) {
    PassTables t;
    for (int p = 0; p < pl.npass; p++) {
        const PassDesc &d = pl.pass[p];
        if (!pass_has_table(d)) continue;
        for (int inv = 0; inv < 2; inv++) {
            const int scale_e = (inv && p == 0) ? inv_scale_e(d.log_m) : 0;
            std::vector<u64> &T = g_ptab[(d.log_m << 8) | ((scale_e ? 1 : 0) << 6) | (d.log_r << 1) | inv];
            if (T.empty()) {
                T.resize((size_t)1 << d.log_m);
                for (size_t i = 0; i < T.size(); i++)
                    T[i] = pass_table_entry(tw(), (u32)i, d.log_m, d.log_r, inv != 0, scale_e);
            }
            (inv ? t.inv : t.fwd)[p] = T.data();
        }
    }
    return t;
}

StoreTw inv_store_tw(const Plan &pl, const PassTables &tabs, int p) {
    if (p < 1 || !tabs.inv[p - 1]) return StoreTw{nullptr, 0};
    return StoreTw{tabs.inv[p - 1], pl.pass[p - 1].log_m};
}

using HX = HostExec;

// The three kernel bodies of gnt_kernels.hip, run tile after tile.
template <class G>
void k_fwd_pass(const PassDesc &d, int log_n, u64 *x, const u64 *limbs, size_t n_limbs, const u64 *T) {
    std::vector<u64> tile(G::tile_n);
    const size_t grid = ((size_t)1 << log_n) / G::tile_n;
    for (size_t t = 0; t < grid; t++) body_fwd<HX, G>(tile.data(), make_geom<G>(d, t), x, limbs, n_limbs, tw(), T);
}

template <class G, bool SUMS, bool STEP, bool TSTORE>
void k_inv_variant(const PassDesc &d, int log_n, u64 *x, u64 *Ls, u64 *Hs, size_t n_coeffs, size_t n_limbs,
                   const StoreTw &st) {
    std::vector<u64> tile(G::tile_n);
    const size_t grid = ((size_t)1 << log_n) / G::tile_n;
    const int scale_e = d.log_m == log_n ? inv_scale_e(log_n) : 0;  // pass 0 carries 1/N
    for (size_t t = 0; t < grid; t++)
        body_inv<HX, G, SUMS, STEP, TSTORE>(tile.data(), make_geom<G>(d, t), x, tw(), Ls, Hs, n_coeffs, n_limbs,
                                            scale_e, st);
}

// The variant the library launches for this pass (launch_inverse_rest).
template <class G>
void k_inv_pass(const PassDesc &d, int log_n, u64 *x, u64 *Ls, u64 *Hs, size_t n_coeffs, size_t n_limbs,
                bool step_tw, const StoreTw &st) {
    step_tw = step_tw && G::strided;  // a contiguous pass has no twiddles to step
    if (Ls && step_tw) k_inv_variant<G, true, true, false>(d, log_n, x, Ls, Hs, n_coeffs, n_limbs, st);
    else if (Ls) k_inv_variant<G, true, false, false>(d, log_n, x, Ls, Hs, n_coeffs, n_limbs, st);
    else if (step_tw) k_inv_variant<G, false, true, false>(d, log_n, x, Ls, Hs, n_coeffs, n_limbs, st);
    else if (st.T) k_inv_variant<G, false, false, true>(d, log_n, x, Ls, Hs, n_coeffs, n_limbs, st);
    else k_inv_variant<G, false, false, false>(d, log_n, x, Ls, Hs, n_coeffs, n_limbs, st);
}

template <class G>
void k_last_fused(const PassDesc &d, int log_n, const u64 *xb, const u64 *limbs, size_t n_in, u64 *xa,
                  const u64 *other, u64 *Ls, u64 *Hs, size_t n_coeffs, size_t n_limbs, const StoreTw &st) {
    std::vector<u64> tile(G::tile_n);
    const size_t grid = ((size_t)1 << log_n) / G::tile_n;
    for (size_t t = 0; t < grid; t++)
        body_fused<HX, G>(tile.data(), make_geom<G>(d, t), xb, limbs, n_in, xa, other, log_n, tw(), Ls, Hs,
                          n_coeffs, n_limbs, st);
}

template <class G>
void k_last_fused_ab(const PassDesc &d, int log_n, u64 *xa, const u64 *xb, const StoreTw &st) {
    using H = typename HalfGeom<G>::type;
    if constexpr (HalfGeom<G>::ok) {
        std::vector<u64> tile(2 * H::tile_words);
        const size_t grid = ((size_t)1 << log_n) / H::tile_n;
        for (size_t t = 0; t < grid; t++)
            body_fused_ab<HX, H>(tile.data(), make_geom<H>(d, t), xa, xb, log_n, tw(), st);
    }
}

#define DISPATCH(D, CALL)                                            \
    switch (shape_key(D)) {                                          \
        GNT_FOR_EACH_SHAPE(CALL)                                     \
        default: break;                                              \
    }

void forward(const Plan &pl, int npasses, u64 *x, const u64 *limbs, size_t n_limbs, const PassTables &tabs) {
    for (int p = 0; p < npasses; p++) {
        const PassDesc &d = pl.pass[p];
        const u64 *src = p == 0 ? limbs : nullptr;
#define F(LR, LC, S) \
    case shape_key(LR, LC, S): k_fwd_pass<Geom<LR, LC, S>>(d, pl.log_n, x, src, n_limbs, tabs.fwd[p]); break;
        DISPATCH(d, F)
#undef F
    }
}

void inverse(const Plan &pl, int from, u64 *x, u64 *Ls, u64 *Hs, size_t n_coeffs, size_t n_limbs,
             const PassTables &tabs) {
    for (int p = from; p >= 0; p--) {
        const PassDesc &d = pl.pass[p];
        u64 *ls = p == 0 ? Ls : nullptr;
        u64 *hs = p == 0 ? Hs : nullptr;
        const bool step_tw = tabs.inv[p] == nullptr;
        const StoreTw st = inv_store_tw(pl, tabs, p);
#define F(LR, LC, S)                                                                                          \
    case shape_key(LR, LC, S):                                                                                \
        k_inv_pass<Geom<LR, LC, S>>(d, pl.log_n, x, ls, hs, n_coeffs, n_limbs, step_tw, st);                 \
        break;
        DISPATCH(d, F)
#undef F
    }
}

int ceil_log2(size_t x) {
    int l = 0;
    while (((size_t)1 << l) < x) l++;
    return l;
}

}  // namespace

// Layout checks of the fused multiply kernel's half-width tiles (see
// emu_fused_tile_layout_check).
template <class G, int I = 0>
long tile_tasks_check(const G &g) {
    if constexpr (I >= G::S::count) {
        return 0;
    } else {
        constexpr int A = G::S::log(I), LM = G::S::log_m(I), LT = LM - A;
        long bad = 0;
        for (int L = 0; L < (G::tile_n >> A); L++) {
            int c, jj, blk;
            task_coords<G, I>(L, c, jj, blk);
            const int base = (blk << LM) + jj;
            for (int u = 0; u < (1 << A); u++)
                bad += tix(g, base + (u << LT), c) != tix(g, base, c) + prow_off<G>(u << LT);
        }
        return bad + tile_tasks_check<G, I + 1>(g);
    }
}

template <class FullG>
long half_tile_check() {
    if constexpr (!HalfGeom<FullG>::ok) {
        return 0;
    } else {
        using H = typename HalfGeom<FullG>::type;
        const H g{};
        std::vector<int> seen(H::tile_words, 0);
        long bad = 0;
        for (int n = 0; n < (1 << H::LOGR); n++)
            for (int c = 0; c < (1 << H::log_c); c++) {
                const int t = tix(g, n, c);
                if (t < 0 || t >= H::tile_words || seen[t]++) bad++;
            }
        if (H::padded) bad += tile_tasks_check(g);
        bad += 2 * 2 * H::tile_words * 8 > 160 * 1024;
        return bad;
    }
}

extern "C" {

// Forward transform of x (length 2^log_n) in place, as the GPU computes it.
void emu_forward(u64 *x, int log_n) {
    tables();
    Plan pl = make_plan(log_n);
    forward(pl, pl.npass, x, nullptr, 0, tables_for(pl));
}

// Inverse transform as the library runs it: a plan of two or more passes
// includes the 1/N (in the inverse twiddles of pass 0), a single-pass plan
// leaves it to the pointwise product (so its inverse is unscaled).
void emu_inverse(u64 *x, int log_n) {
    tables();
    Plan pl = make_plan(log_n);
    inverse(pl, pl.npass - 1, x, nullptr, nullptr, 0, 0, tables_for(pl));
}

// omega_N of the library (primitive N-th root used by the transforms).
u64 emu_root(int log_n) {
    tables();
    u64 e = (u64)1 << (ROOT_LOG - log_n);
    return root_pow(tw(), (u32)e);
}

// Install a pass split for one length in this emulation (npass 0: built-in rule).
int emu_set_plan(int log_n, int npass, const int *log_r) {
    if (npass == 0) {
        set_plan_choice(log_n, 0, nullptr);
        return 0;
    }
    if (!split_valid(log_n, npass, log_r)) return -1;
    set_plan_choice(log_n, npass, log_r);
    return 0;
}

// Entry i of the twiddle table of a pass (block 2^log_m, radix 2^log_r).
u64 emu_pass_table_entry(u32 i, int log_m, int log_r, int inv) {
    tables();
    return pass_table_entry(tw(), i, log_m, log_r, inv != 0);
}

// Largest pass block size (log) that takes its twiddles from a table; -1: none.
void emu_set_table_max_log(int max_log) { g_table_max_log.store(max_log); }

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
    const size_t nl = na + nb;
    std::vector<u64> Ls(nl), Hs(nl);
    const int P = pl.npass;
    const PassTables tabs = tables_for(pl);
    const StoreTw fst = inv_store_tw(pl, tabs, P - 1);
    // the same launch sequence as Ctx::enqueue_product
    const bool ab = !sqr && P > 1;
    if (sqr) {
        if (P > 1) forward(pl, P - 1, xa.data(), a, na, tabs);
    } else {
        forward(pl, ab ? P - 1 : P, xa.data(), a, na, tabs);
        if (P > 1) forward(pl, P - 1, xb.data(), b, nb, tabs);
    }
    if (ab) {
        const PassDesc &d = pl.pass[P - 1];
#define F(LR, LC, S) \
    case shape_key(LR, LC, S): k_last_fused_ab<Geom<LR, LC, S>>(d, log_n, xa.data(), xb.data(), fst); break;
        DISPATCH(d, F)
#undef F
    } else {
        const PassDesc &d = pl.pass[P - 1];
        const bool only = P == 1;
        const u64 *bsrc = sqr ? xa.data() : xb.data();
        const u64 *blimbs = only ? (sqr ? a : b) : nullptr;
        const u64 *other = sqr ? nullptr : xa.data();
#define F(LR, LC, S)                                                                                       \
    case shape_key(LR, LC, S):                                                                             \
        k_last_fused<Geom<LR, LC, S>>(d, log_n, bsrc, blimbs, sqr ? na : nb, xa.data(), other,             \
                                      only ? Ls.data() : nullptr, only ? Hs.data() : nullptr, n_out, nl, fst); \
        break;
        DISPATCH(d, F)
#undef F
    }
    inverse(pl, P - 2, xa.data(), Ls.data(), Hs.data(), n_out, nl, tabs);
    // carry: the same limb rule as the kernels, run serially
    u64 Hprev = 0;
    u32 carry = 0;
    size_t len = 0;
    for (size_t j = 0; j < nl; j++) {
        u64 s = Ls[j] + Hprev;
        u32 g = s < Ls[j];
        u32 p = s == ~0ull;
        out[j] = s + carry;
        carry = g | (p & carry);
        Hprev = Hs[j];
        if (out[j]) len = j + 1;
    }
    return len;
}

// Layout checks of every tile shape the fused multiply kernel halves (its
// half-width tiles): tix maps the tile one-to-one into [0, tile_words); in a
// padded tile every column-DFT task's rows sit at one base plus the
// compile-time offsets prow_off; and the fused kernel's two tiles fit two per
// CU in 160 KiB of LDS.  Returns the number of violations.
long emu_fused_tile_layout_check() {
    long bad = 0;
#define F(LR, LC, S) bad += half_tile_check<Geom<LR, LC, S>>();
    GNT_FOR_EACH_SHAPE(F)
#undef F
    return bad;
}

// Field helpers for direct checks of the device arithmetic.
u64 emu_add(u64 a, u64 b) { return gl::add(a, b); }
u64 emu_sub(u64 a, u64 b) { return gl::sub(a, b); }
u64 emu_mul_mod(u64 a, u64 b) { return gl::mul(a, b); }
u64 emu_reduce128(u64 lo, u64 hi) { return gl::reduce128(lo, hi); }
u64 emu_mul_pow2(u64 x, int e) { return gl::mul_pow2(x, e); }

}  // extern "C"
