// Device-resident request loop for small serving requests.
//
// A persistent kernel of a few 1024-thread workgroups polls a ring of request
// descriptors in pinned host memory.  A workgroup claims the next published
// request (a compare-and-swap on a device counter, only below the host's
// published count), computes it whole in its LDS -- a product's two transforms,
// pointwise product, inverse transform, limb sums and carry, or a sum or
// difference's carry -- writes the result limbs to device memory and then the
// completion word (trimmed length and sequence number) to pinned host memory,
// and polls again.  No host thread launches anything per request: the host
// writes a descriptor and one counter, and reads one word.
//
// Every wave of the kernel exits: a workgroup leaves when the host sets the
// stop word, after DL_IDLE_US without a request to claim, or once it has run
// DL_LIFE_US (a persistent kernel holds its hardware queue, which other
// streams of the process may share; the lifetime bounds how long their work
// waits).  The host relaunches the kernel whenever published requests are
// unclaimed and the previous launch has ended, so a request published while
// the workgroups were leaving is never stranded.
//
// Memory ordering: the host fills the descriptor and the operand limbs, then
// stores the published count (release); the claiming thread loads it with a
// system-scope acquire.  Result limbs are made visible with a system-scope
// fence by every thread before one thread stores the completion word with a
// system-scope release.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "gnt_carry.hpp"
#include "gnt_pass.hpp"

namespace gnt {

constexpr int DL_MAX_LOG_N = 13;                                       // products: N <= 2^13
constexpr size_t DL_MAX_LIMBS = size_t(1) << (DL_MAX_LOG_N - 2);        // result limbs (na + nb, or a sum's)
constexpr int DL_WAVES = 16;                                            // 1024 threads per workgroup
constexpr int DL_BLOCK = 64 * DL_WAVES;
constexpr uint32_t DL_SLOTS = 128;                                      // ring slots
constexpr size_t DL_SLOT_LIMBS = 2 * DL_MAX_LIMBS;                      // host operand limbs per slot
constexpr int DL_GROUPS = 16;                                           // persistent workgroups
constexpr uint64_t DL_IDLE_US = 100;
constexpr uint64_t DL_LIFE_US = 1000;
static_assert(DL_MAX_LIMBS <= (size_t)64 * DL_WAVES * CARRY_ROWS, "one carry tile per request");

// One request (written by the host before publication).  a, b: operand limbs
// in device memory or pinned host memory (b null squares); out: result limbs
// in device memory, nl of them; len: where the trimmed length goes (device
// memory; null in the serving ring, whose completion word carries it); log_n:
// the product's transform length.
struct DevReq {
    const uint64_t *a, *b;
    uint64_t *out;
    uint32_t *len;
    uint32_t na, nb;
    uint32_t op;  // GNT_OP_*
    uint32_t log_n;
    uint32_t nl;
    uint32_t pad[3];
};
static_assert(sizeof(DevReq) == 64, "descriptor of one cache line");

// Kernel argument: the ring and its counters.
struct DevRing {
    const DevReq *req;        // DL_SLOTS descriptors (pinned host)
    const uint64_t *tail;     // requests published (pinned host)
    uint64_t *done;           // per slot: (len << 32) | (seq + 1) once request seq is finished (pinned host)
    const uint32_t *stop;     // nonzero: leave (pinned host)
    uint64_t *claimed;        // requests claimed by workgroups (device)
    uint64_t idle_ticks;      // s_memrealtime ticks (100 MHz)
    uint64_t life_ticks;
};

// Whether a request runs in the device loop (result limbs and op).
inline bool dl_fits(int op, size_t na, size_t nb) {
    if (op == 0 || op == 1) return na + nb <= DL_MAX_LIMBS;            // GNT_OP_MUL, GNT_OP_SQR (nb = na)
    if (op == 2) return (na > nb ? na : nb) + 1 <= DL_MAX_LIMBS;        // GNT_OP_ADD
    return na <= DL_MAX_LIMBS;                                          // GNT_OP_SUB
}

// The batched form of the same routine: one persistent launch whose
// workgroups claim the n requests of reqs (device memory) through the counter
// *next, each computed whole in LDS with its trimmed length written to
// req.len; workgroups leave when no request is left.  *next is zeroed by a
// kernel launched first on the same stream (no memset node: the pair may be
// captured in a graph and replayed).
int launch_devbatch(const DevReq *reqs, uint32_t n, uint32_t *next, const Tw &tw, hipStream_t st);
// Workgroups launch_devbatch uses for n requests.
inline uint32_t devbatch_groups(uint32_t n) { return n < 256 ? n : 256; }

// Host side: the ring in pinned memory and the kernel's stream.  Submitting
// threads publish their own requests (in sequence order) and launch the
// kernel when no launch is running; a lane thread publishes requests that had
// to wait for context-stream work, relaunches while requests are unfinished,
// and completes requests as their words appear.
struct DevLane {
    struct Item {
        void *owner = nullptr;       // the serving request
        bool staged = false;         // descriptor and operands written
        hipEvent_t ready = nullptr;  // context-stream work the request must follow, or null
    };
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ended = nullptr;      // recorded after each launch
    bool launched = false;
    DevReq *req = nullptr;           // pinned host
    uint64_t *tail = nullptr;        // pinned host
    uint64_t *done = nullptr;        // pinned host
    uint32_t *stop = nullptr;        // pinned host
    uint64_t *arena = nullptr;       // pinned host: DL_SLOT_LIMBS per slot
    uint64_t *claimed = nullptr;     // device
    Tw tw{};
    // Sequence numbers: [0, published) published, [finished, next) unfinished.
    uint64_t next = 0, published = 0, finished = 0;
    Item items[DL_SLOTS];
    uint64_t launches = 0, requests = 0;
    std::mutex mu;
    std::condition_variable cv;
    bool quit = false;
    std::thread th;
    // Called with a finished request (owner, trimmed length) by the lane thread.
    void (*complete)(void *server, void *owner, uint32_t len) = nullptr;
    void *server = nullptr;

    int init(int dev, const Tw &t, void (*cb)(void *, void *, uint32_t), void *srv);
    void destroy();  // requests all finished
    // Reserve the next sequence number for owner (false when the ring is full).
    bool reserve(void *owner, uint64_t &seq);
    uint64_t *operands(uint64_t seq) { return arena + (seq % DL_SLOTS) * DL_SLOT_LIMBS; }
    // Descriptor written: the request may be published once `ready` (if any) completes.
    void stage(uint64_t seq, const DevReq &d, hipEvent_t ready);
    size_t unfinished();
    void loop();
    void publish_locked();
    int launch();  // mu held
};

}  // namespace gnt
