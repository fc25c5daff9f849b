// Internal runtime types of libgnt.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/gnt.h"
#include "gnt_pass.hpp"
#include "gnt_plan.hpp"
#include "gnt_stage.hpp"

namespace gnt {

int set_error(int code, const char *msg);

// Idle device memory a pool keeps for reuse (buffers above SLAB_PIECE_MAX).
constexpr size_t POOL_IDLE_MAX = (size_t)32 << 30;
// Buffers up to SLAB_PIECE_MAX are carved from SLAB_BYTES slabs and always
// kept for reuse: a level of binary splitting holds tens of thousands of small
// buffers (lengths, carry flags, small products), which would otherwise cost
// one hipMalloc and one hipFree each.
constexpr size_t SLAB_PIECE_MAX = (size_t)4 << 20;
constexpr size_t SLAB_BYTES = (size_t)64 << 20;

class Pool {
public:
    static size_t size_class(size_t bytes);
    void *acquire(size_t bytes);
    void release(void *p, size_t bytes);
    void trim();         // give idle large buffers back to the device
    void release_all();  // everything, slabs included (nothing may be in use)
    void absorb(Pool &from);  // take over from's buffers and slabs (nothing of from in use)
    size_t live_bytes = 0;    // allocated from the device (pooled or in use)
    size_t pooled_bytes = 0;  // of which idle in the free lists
private:
    std::mutex mu;
    std::map<size_t, std::vector<void *>> free_;
    std::vector<void *> slabs_;
    char *slab_cur_ = nullptr;
    size_t slab_left_ = 0;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;  // downloads of finished buffers, off the compute stream
    std::mutex mu;  // serialises host entry points on this context
    uint64_t *wlo = nullptr, *whi = nullptr, *wsm = nullptr;
    Tw tw() const { return Tw{wlo, whi, wsm}; }
    Pool pool;         // buffers used in compute-stream order
    Pool upload_pool;  // serving operands written on the copy stream; freed only after completion
    Stager up{true}, down{false};  // host <-> device copies of caller memory
    size_t hbm_cap = 0;
    int serve_split_log_n = 22;  // serving: largest transform (log2) on the small-request lane
    size_t serve_large_inflight = 1;  // serving: products the large lane advances at once
    size_t serve_device_limbs = 2048; // serving: largest result (limbs) of the device-resident loop; 0 = off
    bool timing = false;
    double phase_ms[5] = {0, 0, 0, 0, 0};
    uint64_t products = 0;
    void *server = nullptr;  // serving loop (gnt_serve.hip) while running
    uint64_t *noalloc_scratch = nullptr;  // this device's gnt_mul_noalloc scratch array (device address)

    int init_tables();
    // Twiddle tables of pl's strided passes that have one (pass_has_table),
    // built on first use on the compute stream and kept until shutdown; a
    // pass whose table cannot be allocated steps its twiddles instead.
    void pass_tables(const Plan &pl, PassTables &out);
    void free_pass_tables();
    std::mutex tab_mu;
    std::map<int, uint64_t *> ptab;  // key (log_m, log_r, inverse)
    int enqueue_product(const uint64_t *da, size_t na, const uint64_t *db, size_t nb, uint64_t *dout,
                        uint32_t *dlen);
};

bool plan_for(size_t na, size_t nb, Plan &pl);

// Device scratch of one product (sizes in elements): coefficient arrays xa,
// xb (N u64 each; xb unused for a square), limb sums (2 n_limbs u64), carry
// flags (2 nt bytes).
struct ProductScratch {
    size_t N = 0, n_limbs = 0, nt = 0;
    bool sqr = false;
};
ProductScratch product_scratch_layout(size_t na, size_t nb, bool sqr, int log_n);
int launch_product(const Plan &pl, const PassTables &tabs, const Tw &tw, const uint64_t *da, size_t na,
                   const uint64_t *db, size_t nb, uint64_t *dout, uint32_t *dlen, const ProductScratch &s,
                   uint64_t *xa, uint64_t *xb, uint64_t *lh, unsigned char *gp, hipStream_t st, hipEvent_t *ev);
int run_batch(Ctx *c, gnt_job *jobs, size_t n);
// Looks up this device's gnt_mul_noalloc scratch address (device current).
int noalloc_setup(Ctx *c);

}  // namespace gnt
