// Serving loop: continuous batching of multiply/square requests.
//
// Two lanes, each a scheduler thread with its own stream: products up to a
// transform length go to the small lane, larger ones to the large lane, so
// a small request never waits behind a large product's stages while both
// streams share the device (the large lane keeps little queued on its stream,
// since the runtime may put both streams on one in-order hardware queue).  Each loop iteration of a lane (1) admits its
// waiting requests in arrival order while the bytes in flight (both lanes)
// stay under the HBM cap, (2) advances every admitted product by one stage
// through the stage engine (one launch per kernel shape for the whole
// iteration), and (3) retires products whose last stage has completed on the
// device.  Requests that arrive while others run join at the next iteration,
// so their work overlaps on the device.
// Results stay in HBM as gnt_buf handles (handed out once complete, and not
// counted against the cap afterwards); a request may take such a result as an
// operand without a host round trip.  Host operands are uploaded by an
// uploader thread on a separate copy stream, so PCIe transfers overlap
// compute: gnt_submit returns at once, and the uploader moves one slice at a
// time of the request with the fewest bytes left, so a small request never
// waits behind the whole upload of a large one.
// Requests small enough to compute in one workgroup's LDS (products of at most
// 2048 result limbs, sums and differences of up to 2047-limb operands) bypass
// the lanes: they go to the device-resident loop (gnt_devloop.hpp), a
// persistent kernel that claims them from a ring in pinned host memory, with
// host operands copied into the ring's pinned slots by gnt_submit itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "../../include/gnt.h"
#include "gnt_devloop.hpp"
#include "gnt_engine.hpp"
#include "gnt_runtime.hpp"

using u64 = uint64_t;
using gnt::Ctx;

// gnt_buf is defined by the runtime; the serving loop only builds new ones
// through these helpers (gnt_api.hip).
namespace gnt {
gnt_buf *buf_wrap(Ctx *c, u64 *d, size_t cap, uint32_t *dlen, int64_t len);
bool buf_info(gnt_buf *b, u64 **d, size_t *cap);
}

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
    uint64_t ticket = 0;
    gnt_req r{};
    Clock::time_point t_submit, t_done;
    gnt::Prod p;
    u64 *own_a = nullptr, *own_b = nullptr;  // uploaded host operands
    size_t bytes = 0;                        // footprint counted against the cap
    bool done = false;
    int status = GNT_OK;
    int len_slot = -1;                       // pinned word the trimmed length is copied to
    int64_t len = -1;                        // trimmed result length once known
    size_t up_total = 0, up_done = 0;        // host operand bytes to upload, uploaded so far
    hipEvent_t uploaded = nullptr;           // recorded on the copy stream after the last slice
    // A multiply whose B is uploaded after the request is admitted: its lane
    // runs A's forward passes meanwhile (Prod::b_late).
    bool b_split = false;
    bool b_arrived = false;                  // B's last slice issued (Server::mu)
    hipEvent_t uploaded_b = nullptr;         // recorded after B's last slice
    void *lane = nullptr;                    // the lane that runs it
    bool dev = false;                        // runs in the device-resident loop
};

struct Slot {  // descriptor staging for one iteration in flight
    void *host = nullptr, *dev = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
    bool used = false;
};

}  // namespace

namespace gnt {

// Host B operands from this size on are uploaded after their request is
// admitted, while A's forward passes run.
constexpr size_t B_SPLIT_MIN_BYTES = size_t(1) << 20;

struct Server;

// One scheduler thread over one stream.  It admits its own requests in the
// order their operands reach the device (up to max_inflight products at once)
// while the server's bytes in flight stay under the cap, advances every
// admitted product one stage per iteration, and retires the finished ones an
// iteration at a time.  Scratch comes from the lane's own pool, whose reuse
// follows the lane's stream order.
struct Lane {
    Server *s = nullptr;
    Ctx *c = nullptr;
    hipStream_t stream = nullptr;
    Pool scratch;
    std::thread th;
    std::condition_variable cv_work;
    std::deque<Req *> waiting;   // guarded by Server::mu
    std::vector<Req *> active;   // admitted, stages still to enqueue
    // Products whose last stage went out in one iteration, with the event
    // recorded after it; retired in stream order.
    struct Done {
        hipEvent_t ev;
        std::vector<Req *> reqs;
    };
    std::deque<Done> draining;
    std::vector<hipEvent_t> free_events;
    static constexpr int NSLOT = 2;
    Slot slots[NSLOT];
    int slot_next = 0;
    // Pinned words that receive each finished product's trimmed length on the
    // lane's stream, so results are handed out with their length known.
    static constexpr int NLEN = 1 << 15;
    uint32_t *lens = nullptr;
    std::vector<int> lens_free;  // guarded by Server::mu
    hipEvent_t ctx_ev = nullptr; // orders admissions after the context stream's work
    size_t max_inflight = 0;     // products advanced at once (0: any number)
    size_t pending_admit = 0;    // admitted in this pass, not yet in `active`
    // The large-request lane cuts its tile launches into pieces of 4096 tiles
    // (0.2-0.5 ms at 2^27 elements) and keeps at most two unfinished: the
    // runtime may give both lanes' streams one hardware queue, whose packets
    // run in order, and a small request's launches would then wait behind
    // every launch of the large products queued before them.  Measured at 400
    // requests/s: p99 of requests below 2^16 limbs 4.7-5.4 ms unpaced,
    // 1.1-1.4 ms paced (0.7-1.0 ms with pieces of 2048 tiles, which cost 9%
    // of the throughput at saturation against 4% for 4096).
    bool paced = false;
    LaunchPacer pace;

    int init(Server *srv, Ctx *ctx);
    void destroy();
    bool admissible() const;     // Server::mu held
    bool can_launch() const;     // Server::mu held: some admitted product has work to launch
    int admit(Req *q);
    void free_scratch(Req *q);
    void fail_request(Req *q, int code);
    Slot &slot_for(size_t bytes);
    void loop();
    void run_iteration();
    void retire();
};

struct Server {
    Ctx *c;
    std::mutex mu;         // queues, tickets, counters
    std::condition_variable cv_done;
    bool stop = false;
    std::map<uint64_t, Req *> tickets;
    uint64_t next_ticket = 1;
    size_t bytes_in_flight = 0, peak_bytes = 0;
    uint64_t iterations = 0, launches = 0, completed = 0, peak_active = 0;
    uint64_t n_admitted = 0;  // admitted by either lane and not yet finished
    uint64_t lane_iterations[2] = {0, 0};
    std::condition_variable cv_up;
    std::deque<Req *> uploads;   // host operands still to upload (guarded by mu)
    bool stop_up = false;
    std::thread up_th;
    static constexpr size_t UPLOAD_SLICE = size_t(8) << 20;
    void upload_loop();
    void fail_upload(Req *q, int code);
    // Products whose transform length is at most 2^small_max_log_n run on
    // lane 0, larger ones on lane 1: a small request then waits for small
    // stages only, while the two streams share the device.
    int small_max_log_n;
    Lane lanes[2];
    // The device-resident loop (null when off) and the largest result (limbs)
    // it takes.
    DevLane *dl = nullptr;
    size_t dl_max_limbs = 0;

    explicit Server(Ctx *ctx, int split) : c(ctx), small_max_log_n(split) {}

    Lane &lane_for(const gnt::Prod &p) {
        const size_t N = p.linear() ? 4 * p.nl : p.N;
        return N <= (size_t(1) << small_max_log_n) ? lanes[0] : lanes[1];
    }

    size_t footprint(Req *q) {
        gnt::Prod &p = q->p;
        return p.scratch_bytes() + p.nl * 8 + (q->own_a ? p.na * 8 : 0) + (q->own_b ? p.nb * 8 : 0);
    }

    // Operands uploaded on the copy stream: released only once the product
    // has finished, so no later upload can overwrite them under a kernel.
    void free_operands(Req *q) {
        c->upload_pool.release(q->own_a, q->p.na * 8);
        c->upload_pool.release(q->own_b, q->p.nb * 8);
        q->own_a = q->own_b = nullptr;
    }

    // Bytes came free: every lane may now admit its head request.
    void wake_lanes() {
        for (Lane &l : lanes) l.cv_work.notify_one();
    }
};

void Server::fail_upload(Req *q, int code) {
    (void)hipStreamSynchronize(c->copy_stream);
    for (hipEvent_t *e : {&q->uploaded, &q->uploaded_b}) {
        if (*e) (void)hipEventDestroy(*e);
        *e = nullptr;
    }
    free_operands(q);
    {
        std::lock_guard<std::mutex> g(mu);
        q->status = code;
        q->done = true;
        q->t_done = Clock::now();
    }
    cv_done.notify_all();
}

// One slice at a time of the pending request with the fewest bytes left
// (operand a, then b), through the context's staged copies on the copy
// stream.  A request whose last slice is issued moves to its lane's queue
// with an event its lane waits on before the first stage.
void Server::upload_loop() {
    (void)hipSetDevice(c->device);
    for (;;) {
        Req *q;
        size_t done;
        {
            std::unique_lock<std::mutex> g(mu);
            cv_up.wait(g, [&] { return stop_up || !uploads.empty(); });
            if (uploads.empty()) return;  // stopping, and nothing left
            q = *std::min_element(uploads.begin(), uploads.end(), [](const Req *x, const Req *y) {
                return x->up_total - x->up_done < y->up_total - y->up_done;
            });
            done = q->up_done;
        }
        const size_t a_bytes = q->own_a ? q->p.na * 8 : 0;
        size_t len;
        hipError_t e;
        if (done < a_bytes) {
            len = std::min(UPLOAD_SLICE, a_bytes - done);
            e = c->up.copy((char *)q->own_a + done, (const char *)q->r.a + done, len, c->copy_stream);
        } else {
            const size_t o = done - a_bytes;
            len = std::min(UPLOAD_SLICE, q->p.nb * 8 - o);
            e = c->up.copy((char *)q->own_b + o, (const char *)q->r.b + o, len, c->copy_stream);
        }
        bool finished = done + len == q->up_total;
        // a split request goes to its lane once A is on the device; from then
        // on its lane owns the operands, so an error only ends B's upload and
        // is reported when the product completes
        const bool a_now = q->b_split && a_bytes && done < a_bytes && done + len == a_bytes;
        const bool handed = q->b_split && done >= a_bytes;
        hipEvent_t *ev = handed ? &q->uploaded_b : &q->uploaded;
        if (e == hipSuccess && (finished || a_now)) {
            e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventRecord(*ev, c->copy_stream);
        }
        if (e != hipSuccess && handed) {
            if (*ev) (void)hipEventDestroy(*ev);
            *ev = nullptr;
            q->status = GNT_E_DEVICE;
            e = hipSuccess;
            finished = true;
        }
        Lane *lane = (Lane *)q->lane;
        {
            std::lock_guard<std::mutex> g(mu);
            q->up_done = done + len;
            if (e != hipSuccess || finished) uploads.erase(std::find(uploads.begin(), uploads.end(), q));
            if (e == hipSuccess && (a_now || (finished && !q->b_split))) lane->waiting.push_back(q);
            if (e == hipSuccess && finished && q->b_split) q->b_arrived = true;
        }
        if (e != hipSuccess) fail_upload(q, GNT_E_DEVICE);
        else if (finished || a_now) lane->cv_work.notify_one();
    }
}

int Lane::init(Server *srv, Ctx *ctx) {
    s = srv;
    c = ctx;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return GNT_E_DEVICE;
    pace.st = stream;
    if (hipEventCreateWithFlags(&ctx_ev, hipEventDisableTiming) != hipSuccess) return GNT_E_DEVICE;
    if (hipHostMalloc((void **)&lens, NLEN * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
        lens = nullptr;
    if (lens)
        for (int i = NLEN - 1; i >= 0; i--) lens_free.push_back(i);
    return GNT_OK;
}

void Lane::destroy() {
    if (stream) (void)hipStreamSynchronize(stream);
    pace.release();
    for (auto &slot : slots) {
        if (slot.ev) (void)hipEventDestroy(slot.ev);
        if (slot.host) (void)hipHostFree(slot.host);
        if (slot.dev) (void)hipFree(slot.dev);
        slot = Slot{};
    }
    if (lens) (void)hipHostFree(lens);
    lens = nullptr;
    for (hipEvent_t e : free_events) (void)hipEventDestroy(e);
    free_events.clear();
    if (ctx_ev) (void)hipEventDestroy(ctx_ev);
    ctx_ev = nullptr;
    // the idle scratch serves the context's later products (the stream is
    // drained, so no work of this lane still uses it)
    c->pool.absorb(scratch);
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
}

// The head request fits beside the bytes in flight (an empty device admits
// anything).
bool Lane::admissible() const {
    if (waiting.empty()) return false;
    if (max_inflight && active.size() + pending_admit >= max_inflight) return false;
    return s->bytes_in_flight == 0 || s->bytes_in_flight + waiting.front()->bytes <= c->hbm_cap;
}

bool Lane::can_launch() const {
    for (const Req *q : active)
        if (!waits_for_b(q->p) || q->b_arrived) return true;
    return false;
}

// Buffers of an admitted request (lane thread only).  Host operands have
// device buffers from submission (filled by the uploader on the copy stream),
// so only the result and scratch are allocated: the result from the context's
// pool (it outlives the request), scratch from the lane's.
int Lane::admit(Req *q) {
    gnt::Prod &p = q->p;
    const bool sq = q->r.op == GNT_OP_SQR;
    u64 *d;
    size_t cap;
    if (q->r.a_buf) { buf_info(q->r.a_buf, &d, &cap); p.da = d; }
    else p.da = q->own_a;
    if (sq) p.db = nullptr;
    else if (q->r.b_buf) { buf_info(q->r.b_buf, &d, &cap); p.db = d; }
    else p.db = q->own_b;
    p.dout = (u64 *)c->pool.acquire(p.nl * 8);
    p.dlen = (uint32_t *)c->pool.acquire(4);
    if (p.linear()) {
        p.cs = (unsigned char *)scratch.acquire(p.cs_bytes());
        return (p.dout && p.dlen && p.cs) ? GNT_OK : GNT_E_NOMEM;
    }
    p.xa = (u64 *)scratch.acquire(p.N * 8);
    p.xb = sq ? nullptr : (u64 *)scratch.acquire(p.N * 8);
    p.lh = (u64 *)scratch.acquire(p.nl * 16);
    p.cs = (unsigned char *)scratch.acquire(p.cs_bytes());
    if (!p.dout || !p.dlen || !p.xa || (!sq && !p.xb) || !p.lh || !p.cs) return GNT_E_NOMEM;
    return GNT_OK;
}

void Lane::free_scratch(Req *q) {
    gnt::Prod &p = q->p;
    scratch.release(p.xa, p.N * 8);
    scratch.release(p.xb, p.N * 8);
    scratch.release(p.lh, p.nl * 16);
    scratch.release(p.cs, p.cs_bytes());
    p.xa = p.xb = p.lh = nullptr;
    p.cs = nullptr;
}

Slot &Lane::slot_for(size_t bytes) {
    Slot &sl = slots[slot_next];
    slot_next = (slot_next + 1) % NSLOT;
    if (sl.used) (void)hipEventSynchronize(sl.ev);  // its copy and kernels are done
    if (sl.bytes < bytes) {
        if (sl.host) (void)hipHostFree(sl.host);
        if (sl.dev) (void)hipFree(sl.dev);
        sl.bytes = std::max(bytes, (size_t)1 << 16);
        if (hipHostMalloc(&sl.host, sl.bytes, hipHostMallocDefault) != hipSuccess) sl.host = nullptr;
        if (hipMalloc(&sl.dev, sl.bytes) != hipSuccess) sl.dev = nullptr;
    }
    if (!sl.ev) (void)hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
    sl.used = true;
    return sl;
}

void Lane::fail_request(Req *q, int code) {
    (void)hipStreamSynchronize(stream);
    free_scratch(q);
    // a late B may still be uploading into the operand buffers
    for (;;) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            if (!q->b_split || q->b_arrived) break;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    for (hipEvent_t *e : {&q->uploaded, &q->uploaded_b}) {
        if (*e) {
            (void)hipEventSynchronize(*e);
            (void)hipEventDestroy(*e);
        }
        *e = nullptr;
    }
    s->free_operands(q);
    c->pool.release(q->p.dout, q->p.nl * 8);
    c->pool.release(q->p.dlen, 4);
    q->p.dout = nullptr;
    q->p.dlen = nullptr;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->bytes_in_flight -= q->bytes;
        s->n_admitted--;
        q->status = code;
        q->done = true;
        q->t_done = Clock::now();
    }
    s->cv_done.notify_all();
    s->wake_lanes();
}

void Lane::loop() {
    (void)hipSetDevice(c->device);
    for (;;) {
        // pace with the device: at most two iterations queued, so requests
        // that arrive meanwhile join the next iteration
        Slot &oldest = slots[slot_next];
        if (oldest.used) (void)hipEventSynchronize(oldest.ev);
        std::vector<Req *> admitted;
        {
            std::unique_lock<std::mutex> g(s->mu);
            cv_work.wait(g, [&] { return s->stop || admissible() || can_launch() || !draining.empty(); });
            if (s->stop && waiting.empty() && active.empty() && draining.empty()) return;
            // admission in arrival order while the bytes in flight fit the cap
            // (bytes_in_flight counts everything admitted by either lane and
            // not retired, including requests admitted earlier in this pass)
            pending_admit = 0;
            while (admissible()) {
                Req *q = waiting.front();
                waiting.pop_front();
                pending_admit++;
                s->bytes_in_flight += q->bytes;
                s->peak_bytes = std::max(s->peak_bytes, s->bytes_in_flight);
                s->peak_active = std::max(s->peak_active, ++s->n_admitted);
                admitted.push_back(q);
            }
        }
        {
            std::lock_guard<std::mutex> cg(c->mu);
            if (!admitted.empty()) {
                // result buffers come from the context's pool, whose earlier
                // users ran on the context stream
                (void)hipEventRecord(ctx_ev, c->stream);
                (void)hipStreamWaitEvent(stream, ctx_ev, 0);
            }
            for (Req *q : admitted) {
                // the operands' last slices may still be on the copy stream
                if (q->uploaded) (void)hipStreamWaitEvent(stream, q->uploaded, 0);
                int rc = admit(q);
                if (rc) fail_request(q, rc);
                else active.push_back(q);
            }
            pending_admit = 0;  // now counted in `active`
        }
        // the iteration touches only this lane's stream, scratch and slots, so
        // it runs without the context lock (a paced lane waits in it)
        if (!active.empty()) run_iteration();
        retire();
    }
}

void Lane::run_iteration() {
    // products that finish in this iteration write their trimmed length to a
    // pinned word from the device (one small kernel after the carry)
    std::vector<Req *> finishing;
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (Req *q : active)
            if (q->p.b_late && !q->p.b_ready && q->b_arrived) {
                // B's last slices may still be on the copy stream
                if (q->uploaded_b) (void)hipStreamWaitEvent(stream, q->uploaded_b, 0);
                q->p.b_ready = true;
            }
        for (Req *q : active)
            if (q->p.stage + next_stage_span(q->p) >= q->p.nstages) {
                finishing.push_back(q);
                if (lens_free.empty()) continue;  // length found later by a device read
                q->len_slot = lens_free.back();
                lens_free.pop_back();
                q->p.len_host = lens + q->len_slot;
            }
    }
    Iteration it;
    for (Req *q : active) add_stage(it, q->p);
    if (it.titems.empty() && it.citems.empty()) return;  // every product waits for its B
    finalize_iteration(it);
    const size_t tb = (it.titems.size() * sizeof(TItem) + 255) & ~(size_t)255;
    const size_t bytes = tb + it.citems.size() * sizeof(CItem) + 256;
    Slot &sl = slot_for(bytes);
    std::memcpy(sl.host, it.titems.data(), it.titems.size() * sizeof(TItem));
    std::memcpy((unsigned char *)sl.host + tb, it.citems.data(), it.citems.size() * sizeof(CItem));
    (void)hipMemcpyAsync(sl.dev, sl.host, bytes, hipMemcpyHostToDevice, stream);
    launch_iteration(it, (const TItem *)sl.dev, (const CItem *)((unsigned char *)sl.dev + tb), c->tw(), stream,
                     paced ? &pace : nullptr);
    (void)hipEventRecord(sl.ev, stream);
    if (!finishing.empty()) {
        // one event marks every product that finishes in this iteration
        Done d;
        if (!free_events.empty()) {
            d.ev = free_events.back();
            free_events.pop_back();
        } else {
            (void)hipEventCreateWithFlags(&d.ev, hipEventDisableTiming);
        }
        (void)hipEventRecord(d.ev, stream);
        for (Req *q : finishing) free_scratch(q);  // later work on this lane is stream-ordered after it
        d.reqs = std::move(finishing);
        draining.push_back(std::move(d));
    }
    std::vector<Req *> still;
    for (Req *q : active)
        if (q->p.stage < q->p.nstages) still.push_back(q);
    active.swap(still);
    std::lock_guard<std::mutex> g(s->mu);
    s->iterations++;
    s->lane_iterations[this == &s->lanes[0] ? 0 : 1]++;
    s->launches += it.groups.size();
}

// Retire the iterations whose products have finished, oldest first (their
// events complete in stream order).  With nothing left to launch, wait for
// the oldest one rather than spin.
void Lane::retire() {
    bool idle;  // nothing to launch before an iteration finishes or a B arrives
    {
        std::lock_guard<std::mutex> g(s->mu);
        idle = !can_launch();
    }
    if (idle && !draining.empty()) (void)hipEventSynchronize(draining.front().ev);
    bool freed = false;
    while (!draining.empty() && hipEventQuery(draining.front().ev) == hipSuccess) {
        Done d = std::move(draining.front());
        draining.pop_front();
        free_events.push_back(d.ev);
        for (Req *q : d.reqs) {
            for (hipEvent_t *e : {&q->uploaded, &q->uploaded_b}) {
                if (*e) (void)hipEventDestroy(*e);
                *e = nullptr;
            }
            s->free_operands(q);
        }
        {
            std::lock_guard<std::mutex> g(s->mu);
            const auto now = Clock::now();
            for (Req *q : d.reqs) {
                if (q->len_slot >= 0) {
                    q->len = lens[q->len_slot];
                    lens_free.push_back(q->len_slot);
                    q->len_slot = -1;
                }
                s->bytes_in_flight -= q->bytes;
                s->n_admitted--;
                q->t_done = now;
                q->done = true;
                s->completed++;
            }
        }
        freed = true;
    }
    if (freed) {
        s->cv_done.notify_all();
        s->wake_lanes();
    }
}

Server *server_of(Ctx *c) { return reinterpret_cast<Server *>(c->server); }

// A device-loop request has finished (its result limbs are in device memory).
static void dl_complete(void *srv, void *owner, uint32_t len) {
    Server *s = static_cast<Server *>(srv);
    Req *q = static_cast<Req *>(owner);
    {
        std::lock_guard<std::mutex> g(s->mu);
        q->len = len;
        q->t_done = Clock::now();
        q->done = true;
        s->completed++;
    }
    s->cv_done.notify_all();
}

// Route q to the device-resident loop when it fits and the ring has a free
// slot: result buffers from the context's pool (after any context-stream work
// queued before, through an event the loop waits on), host operands copied
// into the slot's pinned limbs, device operands by address.  Returns false to
// leave q to the lanes (nothing allocated then).
static bool dl_submit(Ctx *c, Server *s, Req *q) {
    gnt::Prod &p = q->p;
    const gnt_req &r = q->r;
    if (!s->dl || p.nl > s->dl_max_limbs || !dl_fits(r.op, p.na, p.nb)) return false;
    hipEvent_t ready = nullptr;
    {
        std::lock_guard<std::mutex> cg(c->mu);
        p.dout = (u64 *)c->pool.acquire(p.nl * 8);
        p.dlen = (uint32_t *)c->pool.acquire(4);
        if (p.dout && p.dlen && hipStreamQuery(c->stream) != hipSuccess &&
            (hipEventCreateWithFlags(&ready, hipEventDisableTiming) != hipSuccess ||
             hipEventRecord(ready, c->stream) != hipSuccess)) {
            if (ready) (void)hipEventDestroy(ready);
            ready = nullptr;
            (void)hipStreamSynchronize(c->stream);  // no event: wait here instead
        }
    }
    uint64_t seq = 0;
    if (!p.dout || !p.dlen || !s->dl->reserve(q, seq)) {
        std::lock_guard<std::mutex> cg(c->mu);
        c->pool.release(p.dout, p.nl * 8);
        c->pool.release(p.dlen, 4);
        p.dout = nullptr;
        p.dlen = nullptr;
        if (ready) (void)hipEventDestroy(ready);
        return false;
    }
    const bool sq = r.op == GNT_OP_SQR;
    u64 *ops = s->dl->operands(seq);
    DevReq d{};
    u64 *dp;
    size_t cap;
    if (r.a_buf) {
        buf_info(r.a_buf, &dp, &cap);
        d.a = dp;
    } else {
        std::memcpy(ops, r.a, p.na * 8);
        d.a = ops;
    }
    if (sq) {
        d.b = nullptr;
    } else if (r.b_buf) {
        buf_info(r.b_buf, &dp, &cap);
        d.b = dp;
    } else {
        std::memcpy(ops + p.na, r.b, p.nb * 8);
        d.b = ops + p.na;
    }
    d.out = p.dout;
    d.na = (uint32_t)p.na;
    d.nb = (uint32_t)p.nb;
    d.op = (uint32_t)r.op;
    d.log_n = p.linear() ? 0 : (uint32_t)p.pl.log_n;
    d.nl = (uint32_t)p.nl;
    q->dev = true;
    {
        std::lock_guard<std::mutex> g(s->mu);
        q->ticket = s->next_ticket++;
        s->tickets[q->ticket] = q;
    }
    s->dl->stage(seq, d, ready);
    return true;
}

}  // namespace gnt

extern "C" {

int gnt_serve_start(gnt_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return gnt::set_error(GNT_E_ARG, "null context");
    if (c->server) return GNT_OK;
    gnt::Server *s = new gnt::Server(c, c->serve_split_log_n);
    s->lanes[1].max_inflight = c->serve_large_inflight;
    s->lanes[1].paced = true;
    for (gnt::Lane &l : s->lanes)
        if (l.init(s, c) != GNT_OK) {
            for (gnt::Lane &m : s->lanes) m.destroy();
            delete s;
            return gnt::set_error(GNT_E_DEVICE, "serving lane setup failed");
        }
    if (c->serve_device_limbs) {
        s->dl = new gnt::DevLane();
        s->dl_max_limbs = std::min(c->serve_device_limbs, gnt::DL_MAX_LIMBS);
        if (s->dl->init(c->device, c->tw(), gnt::dl_complete, s) != GNT_OK) {
            s->dl->destroy();
            delete s->dl;
            for (gnt::Lane &m : s->lanes) m.destroy();
            delete s;
            return gnt::set_error(GNT_E_DEVICE, "device-resident loop setup failed");
        }
    }
    c->server = s;
    for (gnt::Lane &l : s->lanes) l.th = std::thread([&l] { l.loop(); });
    s->up_th = std::thread([s] { s->upload_loop(); });
    return GNT_OK;
}

int gnt_serve_stop(gnt_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !c->server) return GNT_OK;
    gnt::Server *s = gnt::server_of(c);
    // the device-resident loop finishes its requests and its last launch ends
    if (s->dl) {
        s->dl->destroy();
        delete s->dl;
        s->dl = nullptr;
    }
    // pending uploads finish and reach their lanes before the lanes drain
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop_up = true;
    }
    s->cv_up.notify_all();
    s->up_th.join();
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->wake_lanes();
    for (gnt::Lane &l : s->lanes) l.th.join();
    for (gnt::Lane &l : s->lanes) l.destroy();
    // results never claimed by gnt_wait have no gnt_buf: their buffers go back
    // to the pool here (the lanes have drained, so every request is done)
    for (auto &kv : s->tickets) {
        Req *q = kv.second;
        if (q->p.dout) c->pool.release(q->p.dout, q->p.nl * 8);
        if (q->p.dlen) c->pool.release(q->p.dlen, 4);
        delete q;
    }
    delete s;
    c->server = nullptr;
    return GNT_OK;
}

int gnt_submit(gnt_ctx *ctx, const gnt_req *req, uint64_t *ticket) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !req || !ticket) return gnt::set_error(GNT_E_ARG, "null argument");
    if (!c->server) return gnt::set_error(GNT_E_STATE, "serving loop not running");
    gnt::Server *s = gnt::server_of(c);
    Req *q = new Req();
    q->r = *req;
    gnt::Prod &p = q->p;
    if (req->op < GNT_OP_MUL || req->op > GNT_OP_SUB) {
        delete q;
        return gnt::set_error(GNT_E_ARG, "unknown request op");
    }
    const bool sq = req->op == GNT_OP_SQR;
    p.op = req->op;
    size_t cap;
    u64 *d;
    p.na = req->a_buf ? (gnt::buf_info(req->a_buf, &d, &cap), cap) : req->na;
    p.nb = sq ? p.na : (req->b_buf ? (gnt::buf_info(req->b_buf, &d, &cap), cap) : req->nb);
    if (p.na == 0 || p.nb == 0 || (!req->a_buf && !req->a) || (!sq && !req->b_buf && !req->b)) {
        delete q;
        return gnt::set_error(GNT_E_ARG, "empty or null operand");
    }
    if (req->op == GNT_OP_SUB && p.nb > p.na) {
        delete q;
        return gnt::set_error(GNT_E_ARG, "difference needs a >= b (b longer than a)");
    }
    p.db = sq ? nullptr : (const u64 *)1;  // shape only; real pointers at admission
    if (!gnt::prod_init(p)) {
        delete q;
        return gnt::set_error(GNT_E_TOO_LARGE, "product exceeds the transform cap");
    }
    if (!p.linear()) c->pass_tables(p.pl, p.tabs);
    q->own_a = req->a_buf ? nullptr : (u64 *)1;
    q->own_b = (sq || req->b_buf) ? nullptr : (u64 *)1;
    q->bytes = s->footprint(q);
    q->own_a = q->own_b = nullptr;
    if (q->bytes > c->hbm_cap) {
        delete q;
        return gnt::set_error(GNT_E_NOMEM, "request exceeds the HBM admission cap");
    }
    q->t_submit = Clock::now();
    if (gnt::dl_submit(c, s, q)) {  // small enough for the device-resident loop
        *ticket = q->ticket;
        return GNT_OK;
    }
    // Device buffers for host operands now (an allocation failure is this
    // call's error); the uploader thread fills them.  They count against the
    // cap once admitted.
    if (!req->a_buf) q->own_a = (u64 *)c->upload_pool.acquire(p.na * 8);
    if (!sq && !req->b_buf) q->own_b = (u64 *)c->upload_pool.acquire(p.nb * 8);
    if ((!req->a_buf && !q->own_a) || (!sq && !req->b_buf && !q->own_b)) {
        s->free_operands(q);
        delete q;
        return gnt::set_error(GNT_E_NOMEM, "no device memory for the operands");
    }
    q->up_total = (q->own_a ? p.na * 8 : 0) + (q->own_b ? p.nb * 8 : 0);
    // a multiply of two or more passes whose host B is large starts on A while
    // B is still uploading (A first: the uploader moves a, then b)
    q->b_split = req->op == GNT_OP_MUL && q->own_b && p.pl.npass >= 2 && p.nb * 8 >= gnt::B_SPLIT_MIN_BYTES;
    p.b_late = q->b_split;
    p.b_ready = !q->b_split;
    gnt::Lane &lane = s->lane_for(p);
    q->lane = &lane;
    const bool to_lane = !q->up_total || (q->b_split && !q->own_a);  // A needs no upload
    {
        std::lock_guard<std::mutex> g(s->mu);
        q->ticket = s->next_ticket++;
        s->tickets[q->ticket] = q;
        if (q->up_total) s->uploads.push_back(q);
        if (to_lane) lane.waiting.push_back(q);
    }
    if (q->up_total) s->cv_up.notify_one();
    if (to_lane) lane.cv_work.notify_one();
    *ticket = q->ticket;
    return GNT_OK;
}

int gnt_poll(gnt_ctx *ctx, uint64_t ticket, int *done) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !done) return gnt::set_error(GNT_E_ARG, "null argument");
    if (!c->server) return gnt::set_error(GNT_E_STATE, "serving loop not running");
    gnt::Server *s = gnt::server_of(c);
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return gnt::set_error(GNT_E_ARG, "unknown ticket");
    *done = it->second->done ? 1 : 0;
    return GNT_OK;
}

int gnt_wait(gnt_ctx *ctx, uint64_t ticket, int64_t timeout_us, gnt_buf **result, double *latency_us) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !result) return gnt::set_error(GNT_E_ARG, "null argument");
    if (!c->server) return gnt::set_error(GNT_E_STATE, "serving loop not running");
    gnt::Server *s = gnt::server_of(c);
    std::unique_lock<std::mutex> g(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return gnt::set_error(GNT_E_ARG, "unknown ticket");
    Req *q = it->second;
    auto pred = [&] { return q->done; };
    if (timeout_us < 0) s->cv_done.wait(g, pred);
    else if (!s->cv_done.wait_for(g, std::chrono::microseconds(timeout_us), pred))
        return gnt::set_error(GNT_E_TIMEOUT, "request still in flight");
    s->tickets.erase(it);
    const int st = q->status;
    if (latency_us) *latency_us = std::chrono::duration<double, std::micro>(q->t_done - q->t_submit).count();
    if (st == GNT_OK) {
        *result = gnt::buf_wrap(c, q->p.dout, q->p.nl, q->p.dlen, q->len);
    } else {  // a request that ran to the end on a failed upload returns its result buffer
        c->pool.release(q->p.dout, q->p.nl * 8);
        c->pool.release(q->p.dlen, 4);
    }
    delete q;
    if (st != GNT_OK) return gnt::set_error(st, "request failed");
    return GNT_OK;
}

int gnt_serve_stats(gnt_ctx *ctx, uint64_t *stats, int n) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !stats) return gnt::set_error(GNT_E_ARG, "null argument");
    if (!c->server) return gnt::set_error(GNT_E_STATE, "serving loop not running");
    gnt::Server *s = gnt::server_of(c);
    std::lock_guard<std::mutex> g(s->mu);
    uint64_t dl_req = 0, dl_launch = 0;
    if (s->dl) {
        std::lock_guard<std::mutex> dg(s->dl->mu);
        dl_req = s->dl->requests;
        dl_launch = s->dl->launches;
    }
    uint64_t v[10] = {s->iterations, s->launches,       s->completed,         s->peak_active,
                      s->peak_bytes, s->bytes_in_flight, s->lane_iterations[0], s->lane_iterations[1],
                      dl_req,        dl_launch};
    for (int i = 0; i < n && i < 10; i++) stats[i] = v[i];
    return GNT_OK;
}

}  // extern "C"
