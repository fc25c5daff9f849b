// Transform plans shared by the HIP kernels, the host runtime and the CPU
// emulation used in tests.
//
// A length-N transform (N = 2^log_n) is a chain of passes.  Pass p applies a
// radix-R_p decimation-in-frequency step in the Bailey form over blocks of
// M_p elements: for every column j < s_p = M_p / R_p of every block, an
// R_p-point DFT over the elements j + u*s_p (u < R_p), then the twiddle
// omega_{M_p}^(j*k).  The next pass works on blocks of s_p elements.  All
// radices are powers of two, so the forward output is in bit-reversed order;
// the inverse passes run in reverse order and return natural order.
//
// A pass keeps a tile of R_p x C_p elements (C_p adjacent columns) in LDS;
// the R_p-point DFT is done as register sub-steps of radix 2..8.
#pragma once

#include <stdint.h>

#include <atomic>

namespace gnt {

constexpr int TILE_LOG = 13;         // 8192 u64 = 64 KiB of LDS per workgroup
constexpr int MAX_PASSES = 4;
constexpr int MAX_LOG_N = 32;        // transform length cap: 4(na+nb) <= 2^32, as ntt_fits (integer.rs:476-478)
constexpr int ROOT_LOG = 32;         // twiddle tables hold powers of a 2^32-th root
constexpr int TW_SPLIT = 16;         // two-level table: 2^16 low x 2^16 high

struct PassDesc {
    int log_r;   // radix of the pass
    int log_m;   // block size M of the pass
    int log_s;   // column stride s = M / R
    int log_c;   // columns per workgroup tile
};

struct Plan {
    int log_n;
    int npass;
    PassDesc pass[MAX_PASSES];
};

// Sub-step radices of an R-point column DFT: split log_r into parts of at
// most 3 bits, as evenly as possible, larger parts first.
static inline int substep_count(int log_r) { return (log_r + 2) / 3; }
static inline int substep_log(int log_r, int i) {
    int k = substep_count(log_r);
    int base = log_r / k, extra = log_r % k;
    return base + (i < extra ? 1 : 0);
}

// Fill pass geometry for the radix split lr[0..np) of a length-2^log_n
// transform (earlier passes are strided, the last is contiguous).
static inline Plan make_plan_split(int log_n, int np, const int *lr) {
    Plan pl;
    pl.log_n = log_n;
    pl.npass = np;
    int rem = log_n;
    for (int p = 0; p < np; p++) {
        pl.pass[p].log_r = lr[p];
        pl.pass[p].log_m = rem;
        pl.pass[p].log_s = rem - lr[p];
        rem -= lr[p];
    }
    for (int p = 0; p < np; p++) {
        PassDesc &d = pl.pass[p];
        int cols = TILE_LOG - d.log_r;
        if (cols < 0) cols = 0;
        if (d.log_s > 0 && cols > d.log_s) cols = d.log_s;
        // a tile never spans more than the whole array
        if (d.log_r + cols > log_n) cols = log_n - d.log_r;
        d.log_c = cols;
    }
    return pl;
}

// Built-in rule: one pass up to a tile, else 2, 3 or (above 2^30) 4 passes
// with the smaller radices first (strided passes then load longer contiguous
// column runs).
static inline void default_split(int log_n, int *np_out, int *lr) {
    int np = log_n <= TILE_LOG ? 1 : (log_n <= 20 ? 2 : (log_n <= 30 ? 3 : 4));
    int rem = log_n;
    for (int p = 0; p < np; p++) {
        lr[p] = rem / (np - p);
        rem -= lr[p];
    }
    *np_out = np;
}

// A split the compiled kernels cover: one pass of any radix up to a tile, or
// 2..4 passes of radix 7..10 (the 2^TILE_LOG-element tile shapes).
static inline bool split_valid(int log_n, int np, const int *lr) {
    if (np < 1 || np > MAX_PASSES || log_n < 1 || log_n > MAX_LOG_N) return false;
    int sum = 0;
    for (int p = 0; p < np; p++) sum += lr[p];
    if (sum != log_n) return false;
    if (np == 1) return log_n <= TILE_LOG;
    for (int p = 0; p < np; p++)
        if (lr[p] < 7 || lr[p] > 10) return false;
    return true;
}

// Per-length choices installed by the autotuner or a plan cache (host side;
// 0 keeps the built-in rule).  Each choice is one word (pass count in bits
// 0-3, radix p in bits 4+4p..7+4p), stored and loaded atomically, so a thread
// planning a product while another installs a plan never pairs the new pass
// count with the old radices.
inline std::atomic<uint32_t> g_plan_word[MAX_LOG_N + 1] = {};

static inline void set_plan_choice(int log_n, int np, const int *lr) {
    uint32_t w = (uint32_t)np;
    for (int p = 0; p < np; p++) w |= (uint32_t)lr[p] << (4 + 4 * p);
    g_plan_word[log_n].store(w, std::memory_order_release);
}
// The installed split of log_n into lr[], or 0 passes (built-in rule).
static inline int get_plan_choice(int log_n, int *lr) {
    if (log_n < 0 || log_n > MAX_LOG_N) return 0;
    const uint32_t w = g_plan_word[log_n].load(std::memory_order_acquire);
    const int np = (int)(w & 15u);
    for (int p = 0; p < np; p++) lr[p] = (int)((w >> (4 + 4 * p)) & 15u);
    return np;
}

// A strided pass whose block size M is at most 2^g_table_max_log takes its
// twiddles from a precomputed table of M entries per direction (one multiply
// per element); larger blocks step their twiddles (two multiplies), since
// their table would be as large as the data.  Host side; -1 disables tables.
inline std::atomic<int> g_table_max_log{22};
static inline bool pass_has_table(const PassDesc &d) {
    return d.log_m > d.log_r && d.log_m <= g_table_max_log.load(std::memory_order_relaxed);
}

// Twiddle tables of a plan's passes (device pointers; null where a pass steps
// its twiddles), filled by the runtime.
struct PassTables {
    const uint64_t *fwd[MAX_PASSES] = {};
    const uint64_t *inv[MAX_PASSES] = {};
};

static inline Plan make_plan(int log_n) {
    int lr[MAX_PASSES] = {0, 0, 0, 0};
    int np = get_plan_choice(log_n, lr);
    if (!np) default_split(log_n, &np, lr);
    return make_plan_split(log_n, np, lr);
}

}  // namespace gnt

// Every tile shape the planner can produce, as (log_r, log_c, strided):
// single-tile transforms (log_n <= TILE_LOG) and the 7..10-bit passes of the
// 2- and 3-pass plans, whose tiles are always 2^TILE_LOG elements.  Kernels are
// instantiated for exactly these; shape_key() selects one at run time.
#define GNT_FOR_EACH_SHAPE(M)                                                                      \
    M(1, 0, 0) M(2, 0, 0) M(3, 0, 0) M(4, 0, 0) M(5, 0, 0) M(6, 0, 0) M(7, 0, 0) M(8, 0, 0)        \
    M(9, 0, 0) M(10, 0, 0) M(11, 0, 0) M(12, 0, 0) M(13, 0, 0)                                     \
    M(7, 6, 0) M(7, 6, 1) M(8, 5, 0) M(8, 5, 1) M(9, 4, 0) M(9, 4, 1) M(10, 3, 0) M(10, 3, 1)

namespace gnt {
constexpr int shape_key(int log_r, int log_c, int strided) { return (log_r << 8) | (log_c << 1) | strided; }
static inline int shape_key(const PassDesc &d) { return shape_key(d.log_r, d.log_c, d.log_s > 0 ? 1 : 0); }
}  // namespace gnt
