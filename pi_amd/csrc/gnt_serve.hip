// Serving loop: continuous batching of multiply/square requests.
//
// A scheduler thread owns the context's stream while serving.  Each loop
// iteration (1) admits waiting requests in arrival order while the bytes in
// flight stay under the HBM cap, uploading their host operands, (2) advances
// every admitted product by one stage through the stage engine (one launch
// per kernel shape for the whole iteration), and (3) retires products whose
// last stage has completed on the device.  Requests that arrive while others
// run join at the next iteration, so their work overlaps on the device.
// Results stay in HBM as gnt_buf handles (handed out once complete, and not
// counted against the cap afterwards); a request may take such a result as an
// operand without a host round trip.  Host operands are uploaded by
// gnt_submit on a separate copy stream, so PCIe transfers overlap compute.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "../../include/gnt.h"
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
    hipEvent_t last = nullptr;               // recorded after the product's last stage
    int len_slot = -1;                       // pinned word the trimmed length is copied to
    int64_t len = -1;                        // trimmed result length once known
};

struct Slot {  // descriptor staging for one iteration in flight
    void *host = nullptr, *dev = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
    bool used = false;
};

}  // namespace

namespace gnt {

struct Server {
    Ctx *c;
    std::thread th;
    std::mutex mu;         // queue, tickets
    std::mutex upload_mu;  // submissions share the copy stream
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    std::deque<Req *> waiting;
    std::vector<Req *> active;                  // admitted, stages still to enqueue
    std::vector<Req *> draining;                // all stages enqueued, waiting for completion
    std::map<uint64_t, Req *> tickets;
    uint64_t next_ticket = 1;
    size_t bytes_in_flight = 0, peak_bytes = 0;
    uint64_t iterations = 0, launches = 0, completed = 0, peak_active = 0;
    static constexpr int NSLOT = 2;
    Slot slots[NSLOT];
    int slot_next = 0;
    // Pinned words that receive each finished product's trimmed length on the
    // compute stream, so results are handed out with their length known.
    static constexpr int NLEN = 1 << 16;
    uint32_t *lens = nullptr;
    std::vector<int> lens_free;

    explicit Server(Ctx *ctx) : c(ctx) {
        if (hipHostMalloc((void **)&lens, NLEN * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
            lens = nullptr;
        if (lens)
            for (int i = NLEN - 1; i >= 0; i--) lens_free.push_back(i);
    }

    size_t footprint(Req *q) {
        gnt::Prod &p = q->p;
        return p.scratch_bytes() + p.nl * 8 + (q->own_a ? p.na * 8 : 0) + (q->own_b ? p.nb * 8 : 0);
    }

    // Buffers of an admitted request (scheduler thread only).  Host operands
    // were uploaded at submission (copy stream), so only scratch is allocated.
    int admit(Req *q) {
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
            p.cs = (unsigned char *)c->pool.acquire(p.cs_bytes());
            return (p.dout && p.dlen && p.cs) ? GNT_OK : GNT_E_NOMEM;
        }
        p.xa = (u64 *)c->pool.acquire(p.N * 8);
        p.xb = sq ? nullptr : (u64 *)c->pool.acquire(p.N * 8);
        p.lh = (u64 *)c->pool.acquire(p.nl * 16);
        p.cs = (unsigned char *)c->pool.acquire(p.cs_bytes());
        if (!p.dout || !p.dlen || !p.xa || (!sq && !p.xb) || !p.lh || !p.cs) return GNT_E_NOMEM;
        return GNT_OK;
    }

    void free_scratch(Req *q) {
        gnt::Prod &p = q->p;
        c->pool.release(p.xa, p.N * 8);
        c->pool.release(p.xb, p.N * 8);
        c->pool.release(p.lh, p.nl * 16);
        c->pool.release(p.cs, p.cs_bytes());
        p.xa = p.xb = p.lh = nullptr;
        p.cs = nullptr;
    }

    // Operands uploaded on the copy stream: released only once the product
    // has finished, so no later upload can overwrite them under a kernel.
    void free_operands(Req *q) {
        c->upload_pool.release(q->own_a, q->p.na * 8);
        c->upload_pool.release(q->own_b, q->p.nb * 8);
        q->own_a = q->own_b = nullptr;
    }

    Slot &slot_for(size_t bytes) {
        Slot &s = slots[slot_next];
        slot_next = (slot_next + 1) % NSLOT;
        if (s.used) (void)hipEventSynchronize(s.ev);  // its copy and kernels are done
        if (s.bytes < bytes) {
            if (s.host) (void)hipHostFree(s.host);
            if (s.dev) (void)hipFree(s.dev);
            s.bytes = std::max(bytes, (size_t)1 << 16);
            if (hipHostMalloc(&s.host, s.bytes, hipHostMallocDefault) != hipSuccess) s.host = nullptr;
            if (hipMalloc(&s.dev, s.bytes) != hipSuccess) s.dev = nullptr;
        }
        if (!s.ev) (void)hipEventCreateWithFlags(&s.ev, hipEventDisableTiming);
        s.used = true;
        return s;
    }

    void fail_request(Req *q, int code) {
        (void)hipStreamSynchronize(c->stream);
        free_scratch(q);
        free_operands(q);
        c->pool.release(q->p.dout, q->p.nl * 8);
        c->pool.release(q->p.dlen, 4);
        std::lock_guard<std::mutex> g(mu);
        bytes_in_flight -= q->bytes;
        q->status = code;
        q->done = true;
        q->t_done = Clock::now();
        cv_done.notify_all();
    }

    void loop() {
        (void)hipSetDevice(c->device);
        for (;;) {
            // pace with the device: at most two iterations queued, so requests
            // that arrive meanwhile join the next iteration
            Slot &pace = slots[slot_next];
            if (pace.used) (void)hipEventSynchronize(pace.ev);
            std::vector<Req *> admitted;
            {
                std::unique_lock<std::mutex> g(mu);
                cv_work.wait(g, [&] { return stop || !waiting.empty() || !active.empty() || !draining.empty(); });
                if (stop && waiting.empty() && active.empty() && draining.empty()) return;
                // admission in arrival order while the bytes in flight fit the cap
                while (!waiting.empty()) {
                    Req *q = waiting.front();
                    // bytes_in_flight counts everything admitted and not retired, including
                    // requests admitted earlier in this pass; an empty device admits anything
                    if (bytes_in_flight > 0 && bytes_in_flight + q->bytes > c->hbm_cap) break;
                    waiting.pop_front();
                    bytes_in_flight += q->bytes;
                    peak_bytes = std::max(peak_bytes, bytes_in_flight);
                    admitted.push_back(q);
                }
            }
            {
                std::lock_guard<std::mutex> cg(c->mu);
                for (Req *q : admitted) {
                    int rc = admit(q);
                    if (rc) fail_request(q, rc);
                    else active.push_back(q);
                }
                peak_active = std::max<uint64_t>(peak_active, active.size());
                if (!active.empty()) run_iteration();
            }
            retire();
        }
    }

    void run_iteration() {
        Iteration it;
        for (Req *q : active) add_stage(it, q->p);
        finalize_iteration(it);
        const size_t tb = (it.titems.size() * sizeof(TItem) + 255) & ~(size_t)255;
        const size_t bytes = tb + it.citems.size() * sizeof(CItem) + 256;
        Slot &s = slot_for(bytes);
        std::memcpy(s.host, it.titems.data(), it.titems.size() * sizeof(TItem));
        std::memcpy((unsigned char *)s.host + tb, it.citems.data(), it.citems.size() * sizeof(CItem));
        (void)hipMemcpyAsync(s.dev, s.host, bytes, hipMemcpyHostToDevice, c->stream);
        launch_iteration(it, (const TItem *)s.dev, (const CItem *)((unsigned char *)s.dev + tb), c->tw(), c->stream);
        (void)hipEventRecord(s.ev, c->stream);
        iterations++;
        launches += it.groups.size();
        std::vector<Req *> still;
        for (Req *q : active) {
            if (q->p.stage >= q->p.nstages) {
                if (!lens_free.empty()) {
                    q->len_slot = lens_free.back();
                    lens_free.pop_back();
                    (void)hipMemcpyAsync(lens + q->len_slot, q->p.dlen, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                         c->stream);
                }
                (void)hipEventCreateWithFlags(&q->last, hipEventDisableTiming);
                (void)hipEventRecord(q->last, c->stream);
                free_scratch(q);  // later work is stream-ordered after this product
                draining.push_back(q);
            } else {
                still.push_back(q);
            }
        }
        active.swap(still);
    }

    // Retire products whose last stage has finished.  With nothing left to
    // launch, wait for the oldest one rather than spin.
    void retire() {
        if (active.empty() && !draining.empty()) (void)hipEventSynchronize(draining.front()->last);
        std::vector<Req *> keep;
        for (Req *q : draining) {
            if (hipEventQuery(q->last) != hipSuccess) { keep.push_back(q); continue; }
            (void)hipEventDestroy(q->last);
            q->last = nullptr;
            free_operands(q);
            {
                std::lock_guard<std::mutex> g(mu);
                if (q->len_slot >= 0) {
                    q->len = lens[q->len_slot];
                    lens_free.push_back(q->len_slot);
                    q->len_slot = -1;
                }
                bytes_in_flight -= q->bytes;
                q->t_done = Clock::now();
                q->done = true;
                completed++;
            }
            cv_done.notify_all();
        }
        draining.swap(keep);
    }
};

Server *server_of(Ctx *c) { return reinterpret_cast<Server *>(c->server); }

}  // namespace gnt

extern "C" {

int gnt_serve_start(gnt_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return gnt::set_error(GNT_E_ARG, "null context");
    if (c->server) return GNT_OK;
    gnt::Server *s = new gnt::Server(c);
    c->server = s;
    s->th = std::thread([s] { s->loop(); });
    return GNT_OK;
}

int gnt_serve_stop(gnt_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !c->server) return GNT_OK;
    gnt::Server *s = gnt::server_of(c);
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->th.join();
    for (auto &slot : s->slots) {
        if (slot.ev) (void)hipEventDestroy(slot.ev);
        if (slot.host) (void)hipHostFree(slot.host);
        if (slot.dev) (void)hipFree(slot.dev);
    }
    for (auto &kv : s->tickets) delete kv.second;  // unclaimed results stay owned by their gnt_buf
    if (s->lens) (void)hipHostFree(s->lens);
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
    // Upload host operands now, on the copy stream, so PCIe transfers overlap
    // the scheduler's kernels.  They count against the cap once admitted.
    {
        std::lock_guard<std::mutex> ug(s->upload_mu);
        int rc = GNT_OK;
        if (!req->a_buf) {
            q->own_a = (u64 *)c->upload_pool.acquire(p.na * 8);
            if (!q->own_a || c->up.copy(q->own_a, req->a, p.na * 8, c->copy_stream))
                rc = GNT_E_DEVICE;
        }
        if (!rc && !sq && !req->b_buf) {
            q->own_b = (u64 *)c->upload_pool.acquire(p.nb * 8);
            if (!q->own_b || c->up.copy(q->own_b, req->b, p.nb * 8, c->copy_stream))
                rc = GNT_E_DEVICE;
        }
        if (!rc && hipStreamSynchronize(c->copy_stream) != hipSuccess) rc = GNT_E_DEVICE;
        if (rc) {
            c->upload_pool.release(q->own_a, p.na * 8);
            c->upload_pool.release(q->own_b, p.nb * 8);
            delete q;
            return gnt::set_error(rc, "operand upload failed");
        }
    }
    {
        std::lock_guard<std::mutex> g(s->mu);
        q->ticket = s->next_ticket++;
        s->tickets[q->ticket] = q;
        s->waiting.push_back(q);
    }
    s->cv_work.notify_one();
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
    if (st == GNT_OK) *result = gnt::buf_wrap(c, q->p.dout, q->p.nl, q->p.dlen, q->len);
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
    uint64_t v[6] = {s->iterations, s->launches, s->completed, s->peak_active, s->peak_bytes, s->bytes_in_flight};
    for (int i = 0; i < n && i < 6; i++) stats[i] = v[i];
    return GNT_OK;
}

}  // extern "C"
