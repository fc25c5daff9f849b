// Device-resident request loop: the persistent kernel and its host lane.
#include "gnt_devloop.hpp"

#include <chrono>
#include <cstring>

#include "../../include/gnt.h"

namespace gnt {
namespace {

using X = DevExec<DL_BLOCK>;

// Operand limbs into the tile, 16 bits per coefficient (ntt.rs:640-665): one
// task per limb writes its four coefficients; coefficients from 4n on are 0.
__device__ __forceinline__ void dl_pack(uint64_t *tile, int N, const uint64_t *limbs, uint32_t n) {
    GNT_FOR(X, l, N >> 2) {
        const uint64_t v = (uint32_t)l < n ? limbs[l] : 0;
#pragma unroll
        for (int t = 0; t < 4; t++) tile[4 * l + t] = (v >> (16 * t)) & 0xFFFFull;
    }
    X::sync();
}

// A whole product (or square) of transform length 2^LOGN in one workgroup:
// both forward transforms in LDS (one pass of radix 2^LOGN), the pointwise
// product with the inverse transform's 1/N, the inverse transform, limb sums
// in B's tile, and one carry tile into r.out.  Returns the trimmed length.
template <int LOGN>
__device__ uint32_t dl_product(uint64_t *ta, uint64_t *tb, const DevReq &r, const Tw &tw, uint32_t *sh) {
    using G = Geom<LOGN, 0, false>;
    constexpr int N = 1 << LOGN;
    G g;
    g.log_m = LOGN;
    g.log_s = 0;
    g.cg0 = 0;
    const bool sq = r.b == nullptr;
    dl_pack(ta, N, r.a, r.na);
    tile_fwd<X, G, 0>(ta, g, tw);
    if (!sq) {
        dl_pack(tb, N, r.b, r.nb);
        tile_fwd<X, G, 0>(tb, g, tw);
    }
    const int e = inv_scale_e(LOGN);  // the single pass carries 1/N
    GNT_FOR(X, i, N) {
        const uint64_t v = ta[i];
        ta[i] = gl::mul_pow2(gl::mul(v, sq ? v : tb[i]), e);
    }
    X::sync();
    tile_inv<X, G, G::S::count - 1>(ta, g, tw);
    uint64_t *Ls = tb, *Hs = tb + r.nl;  // B's tile is free now (2 nl <= N)
    tile_limb_sums<X, G>(ta, g, 4 * (size_t)r.nl - 1, r.nl, Ls, Hs);
    X::sync();
    uint32_t gp, hi;
    carry_tile<SumsSrc, DL_WAVES>(SumsSrc{Ls, Hs}, r.nl, 0, r.out, sh, gp, hi);
    return hi;
}

__device__ uint32_t dl_run(uint64_t *ta, uint64_t *tb, const DevReq &r, const Tw &tw, uint32_t *sh) {
    if (r.op == GNT_OP_ADD || r.op == GNT_OP_SUB) {
        uint32_t gp, hi;
        carry_tile<LinearSrc, DL_WAVES>(LinearSrc{r.a, r.b, r.na, r.nb, r.op == GNT_OP_SUB}, r.nl, 0, r.out, sh,
                                        gp, hi);
        return hi;
    }
    switch (r.log_n) {
        case 3: return dl_product<3>(ta, tb, r, tw, sh);
        case 4: return dl_product<4>(ta, tb, r, tw, sh);
        case 5: return dl_product<5>(ta, tb, r, tw, sh);
        case 6: return dl_product<6>(ta, tb, r, tw, sh);
        case 7: return dl_product<7>(ta, tb, r, tw, sh);
        case 8: return dl_product<8>(ta, tb, r, tw, sh);
        case 9: return dl_product<9>(ta, tb, r, tw, sh);
        case 10: return dl_product<10>(ta, tb, r, tw, sh);
        case 11: return dl_product<11>(ta, tb, r, tw, sh);
        case 12: return dl_product<12>(ta, tb, r, tw, sh);
        default: return dl_product<13>(ta, tb, r, tw, sh);  // the host admits 3..13 only
    }
}

// The next published request for this workgroup, or ~0 to leave: the stop
// word is set, the workgroup has run life_ticks, or nothing was published for
// idle_ticks since t_idle.  One thread; every path ends.
__device__ uint64_t dl_claim(const DevRing &g, uint64_t t_start, uint64_t t_idle) {
    for (;;) {
        const uint64_t now = __builtin_amdgcn_s_memrealtime();
        if (now - t_start > g.life_ticks || now - t_idle > g.idle_ticks) return ~0ull;
        if (__hip_atomic_load(g.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) return ~0ull;
        const uint64_t tail = __hip_atomic_load(g.tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
        uint64_t k = __hip_atomic_load(g.claimed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k < tail) {
            if (__hip_atomic_compare_exchange_strong(g.claimed, &k, k + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT))
                return k;
            continue;  // another workgroup took it
        }
        __builtin_amdgcn_s_sleep(4);
    }
}

__global__ __launch_bounds__(DL_BLOCK) void k_devloop(DevRing ring, Tw tw) {
    __shared__ uint64_t lds[2 << DL_MAX_LOG_N];  // A's and B's tiles: 128 KiB
    __shared__ uint32_t sh[2 * DL_WAVES];
    __shared__ DevReq cur;
    __shared__ uint64_t cur_seq;
    const uint64_t t_start = __builtin_amdgcn_s_memrealtime();
    uint64_t t_idle = t_start;
    for (;;) {
        if (threadIdx.x == 0) {
            const uint64_t k = dl_claim(ring, t_start, t_idle);
            cur_seq = k;
            if (k != ~0ull) cur = ring.req[k % DL_SLOTS];  // after the acquire of the published count
        }
        __syncthreads();
        const uint64_t seq = cur_seq;
        if (seq == ~0ull) return;  // uniform: every thread leaves
        const DevReq r = cur;
        const uint32_t len = dl_run(lds, lds + (1 << DL_MAX_LOG_N), r, tw, sh);
        // the result limbs reach memory before the completion word says so
        __threadfence_system();
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_store(&ring.done[seq % DL_SLOTS], ((uint64_t)len << 32) | (seq + 1), __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        t_idle = __builtin_amdgcn_s_memrealtime();
    }
}

// Batched: claim requests through *next until none is left.
__global__ __launch_bounds__(DL_BLOCK) void k_devbatch(const DevReq *__restrict__ reqs, uint32_t n, uint32_t *next,
                                                       Tw tw) {
    __shared__ uint64_t lds[2 << DL_MAX_LOG_N];
    __shared__ uint32_t sh[2 * DL_WAVES];
    __shared__ uint32_t cur;
    for (;;) {
        if (threadIdx.x == 0) cur = atomicAdd(next, 1u);
        __syncthreads();
        const uint32_t i = cur;
        if (i >= n) return;  // uniform
        const DevReq r = reqs[i];
        const uint32_t len = dl_run(lds, lds + (1 << DL_MAX_LOG_N), r, tw, sh);
        if (threadIdx.x == 0) *r.len = len;
        __syncthreads();  // cur and the tiles are rewritten by the next claim
    }
}

__global__ void k_devbatch_reset(uint32_t *next) { *next = 0; }

template <class T>
int pinned(T **p, size_t n) {
    void *h = nullptr;
    if (hipHostMalloc(&h, n * sizeof(T), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) return -1;
    std::memset(h, 0, n * sizeof(T));
    *p = (T *)h;
    return 0;
}

}  // namespace

int launch_devbatch(const DevReq *reqs, uint32_t n, uint32_t *next, const Tw &tw, hipStream_t st) {
    if (n == 0) return GNT_OK;
    hipLaunchKernelGGL(k_devbatch_reset, dim3(1), dim3(1), 0, st, next);
    hipLaunchKernelGGL(k_devbatch, dim3(devbatch_groups(n)), dim3(DL_BLOCK), 0, st, reqs, n, next, tw);
    return hipGetLastError() == hipSuccess ? GNT_OK : GNT_E_DEVICE;
}

int DevLane::init(int dev, const Tw &t, void (*cb)(void *, void *, uint32_t), void *srv) {
    device = dev;
    tw = t;
    complete = cb;
    server = srv;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return GNT_E_DEVICE;
    if (hipEventCreateWithFlags(&ended, hipEventDisableTiming) != hipSuccess) return GNT_E_DEVICE;
    if (pinned(&req, DL_SLOTS) || pinned(&tail, 1) || pinned(&done, DL_SLOTS) || pinned(&stop, 1) ||
        pinned(&arena, DL_SLOTS * DL_SLOT_LIMBS))
        return GNT_E_NOMEM;
    if (hipMalloc(&claimed, sizeof(uint64_t)) != hipSuccess) return GNT_E_NOMEM;
    if (hipMemsetAsync(claimed, 0, sizeof(uint64_t), stream) != hipSuccess) return GNT_E_DEVICE;
    if (hipStreamSynchronize(stream) != hipSuccess) return GNT_E_DEVICE;
    th = std::thread([this] { loop(); });
    return GNT_OK;
}

void DevLane::destroy() {
    if (th.joinable()) {
        {
            std::lock_guard<std::mutex> g(mu);
            quit = true;
        }
        cv.notify_all();
        th.join();
    }
    if (stop) __atomic_store_n(stop, 1u, __ATOMIC_RELEASE);
    if (stream) (void)hipStreamSynchronize(stream);  // the last launch has ended
    if (claimed) (void)hipFree(claimed);
    for (void *p : {(void *)req, (void *)tail, (void *)done, (void *)stop, (void *)arena})
        if (p) (void)hipHostFree(p);
    if (ended) (void)hipEventDestroy(ended);
    if (stream) (void)hipStreamDestroy(stream);
    for (Item &it : items)
        if (it.ready) (void)hipEventDestroy(it.ready);
    req = nullptr;
    tail = done = nullptr;
    stop = nullptr;
    arena = claimed = nullptr;
    ended = nullptr;
    stream = nullptr;
}

bool DevLane::reserve(void *owner, uint64_t &seq) {
    std::lock_guard<std::mutex> g(mu);
    if (quit || next - finished >= DL_SLOTS) return false;
    seq = next++;
    Item &it = items[seq % DL_SLOTS];
    it.owner = owner;
    it.staged = false;
    it.ready = nullptr;
    return true;
}

void DevLane::stage(uint64_t seq, const DevReq &d, hipEvent_t ready) {
    req[seq % DL_SLOTS] = d;
    {
        std::lock_guard<std::mutex> g(mu);
        Item &it = items[seq % DL_SLOTS];
        it.ready = ready;
        it.staged = true;
        requests++;
        publish_locked();  // from the submitting thread: no hand-off before the device sees it
    }
    cv.notify_one();
}

size_t DevLane::unfinished() {
    std::lock_guard<std::mutex> g(mu);
    return next - finished;
}

int DevLane::launch() {
    DevRing r;
    r.req = req;
    r.tail = tail;
    r.done = done;
    r.stop = stop;
    r.claimed = claimed;
    r.idle_ticks = DL_IDLE_US * 100;  // s_memrealtime runs at 100 MHz
    r.life_ticks = DL_LIFE_US * 100;
    hipLaunchKernelGGL(k_devloop, dim3(DL_GROUPS), dim3(DL_BLOCK), 0, stream, r, tw);
    if (hipGetLastError() != hipSuccess) return GNT_E_DEVICE;
    if (hipEventRecord(ended, stream) != hipSuccess) return GNT_E_DEVICE;
    launched = true;
    launches++;
    return GNT_OK;
}

// Publish staged requests in sequence order (each after its context-stream
// event), and launch the kernel when published requests are unfinished and
// the previous launch has ended (requests published as its workgroups left
// would otherwise wait for the next one).  mu held.
void DevLane::publish_locked() {
    uint64_t p = published;
    while (p < next) {
        Item &it = items[p % DL_SLOTS];
        if (!it.staged) break;
        if (it.ready) {
            if (hipEventQuery(it.ready) != hipSuccess) break;
            (void)hipEventDestroy(it.ready);
            it.ready = nullptr;
        }
        p++;
    }
    if (p != published) {
        published = p;
        __atomic_store_n(tail, p, __ATOMIC_RELEASE);
    }
    if (published > finished && (!launched || hipEventQuery(ended) == hipSuccess)) (void)launch();
}

// Publishes what the submitting threads could not (requests behind a
// context-stream event), keeps a launch alive, and completes requests whose
// completion word has appeared, oldest first.  While requests are unfinished
// the thread polls, yielding between rounds; otherwise it sleeps on the
// condition variable.
void DevLane::loop() {
    (void)hipSetDevice(device);
    std::vector<std::pair<void *, uint32_t>> fin;
    for (;;) {
        fin.clear();
        {
            std::unique_lock<std::mutex> g(mu);
            cv.wait(g, [&] {
                return quit || published > finished || (published < next && items[published % DL_SLOTS].staged);
            });
            if (quit && next == finished) return;
            publish_locked();
            while (finished < published) {
                const uint64_t w = __atomic_load_n(&done[finished % DL_SLOTS], __ATOMIC_ACQUIRE);
                if ((w & 0xFFFFFFFFull) != ((finished + 1) & 0xFFFFFFFFull)) break;
                fin.emplace_back(items[finished % DL_SLOTS].owner, (uint32_t)(w >> 32));
                items[finished % DL_SLOTS].owner = nullptr;
                finished++;
            }
        }
        for (auto &f : fin) complete(server, f.first, f.second);
        if (fin.empty()) std::this_thread::yield();
    }
}

}  // namespace gnt
