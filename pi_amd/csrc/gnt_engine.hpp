// Stage engine: many products of uneven size advanced together.
//
// Every product is a fixed sequence of stages (forward passes, the fused
// last pass, inverse passes, three carry steps).  One engine iteration
// advances every product in flight by one stage; all tile-passes of that
// iteration that share a kernel (kind, radix) go out as ONE launch over a
// device table of items, so a level of N independent products costs a handful
// of launches instead of ~10 per product.  The same iterations serve the
// batched API (all products admitted at once, optionally captured as one
// hipGraph and replayed) and the serving loop (products admitted between
// iterations as memory frees up).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "gnt_pass.hpp"
#include "gnt_plan.hpp"

namespace gnt {

// Kinds of tile-pass items; each is its own kernel.  Inverse passes: K_INV
// steps its twiddles after the load (a pass without a table), K_INV_PLAIN
// found them applied by the stage before, K_INV_TSTORE in addition applies
// the next inverse pass's table on the way out; the _SUMS kinds are the last
// inverse pass, writing limb sums.
enum ItemKind : int {
    K_FWD = 0,
    K_INV = 1,
    K_FUSED = 2,
    K_FUSED_AB = 3,
    K_INV_SUMS = 4,
    K_INV_PLAIN = 5,
    K_INV_TSTORE = 6,
    K_INV_SUMS_PLAIN = 7
};
// Group kinds of the carry steps: K_CARRY + step (0: tiles, 1: scan, 2: ripple).
constexpr int K_CARRY = 16;
static_assert(K_INV_SUMS_PLAIN < K_CARRY, "item kinds and carry-step kinds must not overlap");

// One tile-pass over one array (device descriptor).
struct TItem {
    uint64_t *x;            // FWD/INV: array in place; FUSED: B's array (source); FUSED_AB: A's array
    const uint64_t *limbs;  // pack source for a first pass, else null
    uint64_t *xa;           // FUSED: A's transform and output
    const uint64_t *other;  // FUSED: A (null squares); FUSED_AB: B's array
    uint64_t *Ls, *Hs;      // limb sums (last inverse pass), else null
    uint64_t n_limbs_in;
    uint64_t n_coeffs;
    uint64_t n_limbs;
    uint64_t tile0;         // first launch-wide tile index of this item
    const uint64_t *T;      // FWD: the pass's twiddle table; INV/FUSED*: the next inverse pass's (null: none)
    PassDesc d;
    int log_n;
    int log_t;              // log2 of T's length
    int pad;
};

// Carry modes: limb sums of a product, or the sum / difference of two numbers.
enum CarryMode : int { C_SUMS = 0, C_ADD = 1, C_SUB = 2 };

// One product's carry (device descriptor).  C_SUMS: limb j of the result is
// Ls[j] + Hs[j-1] plus carries.  C_ADD / C_SUB: Ls and Hs are the operands
// (na and nb limbs); C_SUB computes Ls - Hs over n_limbs = na limbs as
// Ls + ~Hs + 1, which needs Ls >= Hs.
struct CItem {
    const uint64_t *Ls, *Hs;
    uint64_t *out;
    unsigned char *gp, *cin;  // per carry tile
    uint32_t *hi;             // per carry tile: highest nonzero limb + 1
    uint32_t *out_len;
    uint32_t *len_host;       // last carry step: copy *out_len here too (pinned host word), or null
    uint64_t n_limbs;
    uint64_t ntiles;
    uint64_t tile0;
    uint64_t na, nb;  // C_ADD / C_SUB operand lengths
    int mode;
    int pad;
};

enum ProdOp : int { OP_MUL = 0, OP_SQR = 1, OP_ADD = 2, OP_SUB = 3 };

// A product (or a sum / difference) in flight (host side).
struct Prod {
    const uint64_t *da = nullptr, *db = nullptr;  // device limbs (db null: square)
    size_t na = 0, nb = 0;
    uint64_t *dout = nullptr;
    uint32_t *dlen = nullptr;
    Plan pl;
    PassTables tabs;  // twiddle tables of pl's passes (filled by the runtime)
    size_t N = 0, n_out = 0, nl = 0, nt = 0;
    uint64_t *xa = nullptr, *xb = nullptr, *lh = nullptr;
    unsigned char *cs = nullptr;  // carry scratch: gp[nt] cin[nt] pad, hi[nt] u32
    int stage = 0;
    int nstages = 0;
    int op = OP_MUL;  // OP_ADD / OP_SUB: carry stages only, no transform
    uint32_t *len_host = nullptr;  // pinned word that receives the trimmed length (or null)
    // A multiply whose B reaches the device after A (serving): A's forward
    // passes run ahead, B's follow once b_ready (b_next of them done), and the
    // fused stage waits for both.
    bool b_late = false;
    bool b_ready = true;
    int b_next = 0;
    bool sqr() const { return db == nullptr; }
    bool linear() const { return op == OP_ADD || op == OP_SUB; }
    size_t scratch_bytes() const;            // xa + xb + lh + cs
    size_t cs_bytes() const;
};

// Fill the derived fields of p (plan, sizes, stage count).  False if too large.
bool prod_init(Prod &p);

// Launch groups of one iteration.
struct Group {
    int kind;   // an ItemKind, or K_CARRY + k for carry step k
    int shape;  // shape_key of the group's tiles
    size_t first, count;  // range in the iteration's item table
    uint64_t tiles;       // total tiles of the launch
    int lds_log;          // log2 of the largest tile (u64 elements) in the group
};

struct Iteration {
    std::vector<TItem> titems;
    std::vector<CItem> citems;
    std::vector<Group> groups;
    bool want_lens = false;  // some last carry step has a len_host: copy after the ripple
};

// Append the next stage of product p to `it` and advance p.stage: one stage,
// or the last transform stage with the three carry steps after it.
void add_stage(Iteration &it, Prod &p);
// Stages the next add_stage of p takes (its last call finishes the product).
int next_stage_span(const Prod &p);
// True when p (b_late) has nothing to launch until B is on the device.
bool waits_for_b(const Prod &p);
// Sort the items into launch groups (kind, radix) and fill tile offsets.
void finalize_iteration(Iteration &it);
// Bounds the work a stream has queued: tile launches are cut into pieces of
// at most max_tiles workgroups, and before each launch the stream waits (on
// the host) until fewer than depth earlier launches are unfinished.  Streams
// may share a hardware queue, whose packets run in order, so a stream that
// queues a whole large product makes another stream's small launches wait for
// all of it; paced, they wait for at most depth pieces.
struct LaunchPacer {
    uint64_t max_tiles = 4096;
    int depth = 2;
    hipStream_t st = nullptr;
    std::vector<hipEvent_t> inflight;  // oldest first
    std::vector<hipEvent_t> spare;
    hipError_t before();  // wait until fewer than depth launches are unfinished
    hipError_t after();   // mark the launch just issued
    void release();       // destroy the events (the stream drained)
};

// Launch one iteration whose items were copied to dev_t / dev_c, paced by
// `pace` when given (one launch per group otherwise).
int launch_iteration(const Iteration &it, const TItem *dev_t, const CItem *dev_c, const Tw &tw, hipStream_t st,
                     LaunchPacer *pace = nullptr);

}  // namespace gnt
