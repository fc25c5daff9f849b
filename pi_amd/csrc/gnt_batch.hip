// Batched products: a plan for a fixed set of product shapes.  Products small
// enough for one workgroup's LDS (at most 2048 result limbs) run in one
// persistent launch whose workgroups claim them one after another
// (gnt_devloop.hpp); the others run through the stage engine (gnt_engine.hpp)
// as level-grouped launches over device descriptor tables.  The whole run is
// optionally captured once as a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/gnt.h"
#include "gnt_devloop.hpp"
#include "gnt_engine.hpp"
#include "gnt_kernels.hpp"
#include "gnt_runtime.hpp"

using u64 = uint64_t;
using gnt::Ctx;

struct gnt_batch {
    Ctx *c = nullptr;
    std::vector<gnt::Prod> prods;
    std::vector<u64 *> in_a, in_b;             // device operands: slices of in_blob
    // Operands, results and trimmed lengths of all products live in three
    // device ranges, so a batch crosses the link in a few large transfers.
    u64 *in_blob = nullptr, *out_blob = nullptr;
    uint32_t *len_blob = nullptr;
    size_t in_bytes = 0, out_bytes = 0;
    std::vector<size_t> out_off;               // byte offset of each result in out_blob
    std::vector<gnt::Iteration> iters;
    std::vector<size_t> t_off, c_off;          // item offsets of each iteration in the blob
    void *blob = nullptr;                      // device descriptor tables
    size_t blob_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t scratch_bytes = 0;
    // products of the persistent launch: descriptors and claim counter in the blob
    uint32_t n_small = 0;
    size_t small_off = 0, next_off = 0;  // byte offsets in blob
};

namespace gnt {

static void release_batch(gnt_batch *b) {
    Ctx *c = b->c;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    c->pool.release(b->in_blob, b->in_bytes);
    c->pool.release(b->out_blob, b->out_bytes);
    c->pool.release(b->len_blob, b->prods.size() * 4);
    for (size_t i = 0; i < b->prods.size(); i++) {
        Prod &p = b->prods[i];
        c->pool.release(p.xa, p.N * 8);
        c->pool.release(p.xb, p.sqr() ? 0 : p.N * 8);
        c->pool.release(p.lh, p.nl * 16);
        c->pool.release(p.cs, p.cs_bytes());
    }
    c->pool.release(b->blob, b->blob_bytes);
    delete b;
}

// The persistent launch over the plan's small products (none: nothing).
static int launch_small(gnt_batch *b) {
    unsigned char *blob = (unsigned char *)b->blob;
    return launch_devbatch((const DevReq *)(blob + b->small_off), b->n_small, (uint32_t *)(blob + b->next_off),
                           b->c->tw(), b->c->stream);
}

// Allocate buffers, build the schedule and upload the descriptor tables.
static int create_batch(Ctx *c, const size_t *na, const size_t *nb, const int *is_sqr, size_t n, bool graph,
                        gnt_batch **out) {
    gnt_batch *b = new gnt_batch();
    b->c = c;
    b->prods.resize(n);
    b->in_a.assign(n, nullptr);
    b->in_b.assign(n, nullptr);
    b->out_off.assign(n, 0);
    // shapes first, then one range each for operands, results and lengths
    // (pieces on 256-byte boundaries)
    auto rup = [](size_t bytes) { return (bytes + 255) & ~size_t(255); };
    std::vector<size_t> a_off(n), b_off(n);
    for (size_t i = 0; i < n; i++) {
        Prod &p = b->prods[i];
        const bool sq = is_sqr && is_sqr[i];
        p.na = na[i];
        p.nb = sq ? na[i] : nb[i];
        if (p.na == 0 || p.nb == 0) { release_batch(b); return set_error(GNT_E_ARG, "empty operand in batch plan"); }
        p.db = sq ? nullptr : (const u64 *)1;  // shape only
        if (!prod_init(p)) { release_batch(b); return set_error(GNT_E_TOO_LARGE, "product exceeds the transform cap"); }
        a_off[i] = b->in_bytes;
        b->in_bytes += rup(p.na * 8);
        b_off[i] = b->in_bytes;
        if (!sq) b->in_bytes += rup(p.nb * 8);
        b->out_off[i] = b->out_bytes;
        b->out_bytes += rup(p.nl * 8);
    }
    if (n) {
        b->in_blob = (u64 *)c->pool.acquire(b->in_bytes);
        b->out_blob = (u64 *)c->pool.acquire(b->out_bytes);
        b->len_blob = (uint32_t *)c->pool.acquire(n * 4);
        if (!b->in_blob || !b->out_blob || !b->len_blob) {
            release_batch(b);
            return set_error(GNT_E_NOMEM, "device memory exhausted");
        }
    }
    std::vector<DevReq> small;
    for (size_t i = 0; i < n; i++) {
        Prod &p = b->prods[i];
        const bool sq = p.sqr();
        b->in_a[i] = (u64 *)((char *)b->in_blob + a_off[i]);
        if (!sq) b->in_b[i] = (u64 *)((char *)b->in_blob + b_off[i]);
        p.da = b->in_a[i];
        p.db = sq ? nullptr : b->in_b[i];
        p.dout = (u64 *)((char *)b->out_blob + b->out_off[i]);
        p.dlen = b->len_blob + i;
        if (dl_fits(sq ? GNT_OP_SQR : GNT_OP_MUL, p.na, p.nb)) {
            // whole in one workgroup of the persistent launch: no scratch, no stages
            DevReq d{};
            d.a = p.da;
            d.b = p.db;
            d.out = p.dout;
            d.len = p.dlen;
            d.na = (uint32_t)p.na;
            d.nb = (uint32_t)p.nb;
            d.op = sq ? GNT_OP_SQR : GNT_OP_MUL;
            d.log_n = (uint32_t)p.pl.log_n;
            d.nl = (uint32_t)p.nl;
            small.push_back(d);
            p.stage = p.nstages;
            b->scratch_bytes += p.na * 8 + p.nb * 8 + p.nl * 8;
            continue;
        }
        c->pass_tables(p.pl, p.tabs);
        p.xa = (u64 *)c->pool.acquire(p.N * 8);
        if (!sq) p.xb = (u64 *)c->pool.acquire(p.N * 8);
        p.lh = (u64 *)c->pool.acquire(p.nl * 16);
        p.cs = (unsigned char *)c->pool.acquire(p.cs_bytes());
        b->scratch_bytes += p.scratch_bytes() + p.na * 8 + p.nb * 8 + p.nl * 8;
        if (!p.xa || (!sq && !p.xb) || !p.lh || !p.cs) {
            release_batch(b);
            return set_error(GNT_E_NOMEM, "device memory exhausted");
        }
    }
    // schedule: every product advances one stage per iteration
    size_t nt = 0, nc = 0;
    for (;;) {
        Iteration it;
        for (Prod &p : b->prods)
            if (p.stage < p.nstages) add_stage(it, p);
        if (it.titems.empty() && it.citems.empty()) break;
        finalize_iteration(it);
        b->t_off.push_back(nt);
        b->c_off.push_back(nc);
        nt += it.titems.size();
        nc += it.citems.size();
        b->iters.push_back(std::move(it));
    }
    const size_t tbytes = (nt * sizeof(TItem) + 255) & ~(size_t)255;
    b->n_small = (uint32_t)small.size();
    b->small_off = (tbytes + nc * sizeof(CItem) + 255) & ~(size_t)255;
    b->next_off = b->small_off + small.size() * sizeof(DevReq);
    b->blob_bytes = b->next_off + 256;
    b->blob = c->pool.acquire(b->blob_bytes);
    if (!b->blob) { release_batch(b); return set_error(GNT_E_NOMEM, "device memory exhausted"); }
    std::vector<unsigned char> host(b->blob_bytes, 0);
    for (size_t k = 0; k < b->iters.size(); k++) {
        const Iteration &it = b->iters[k];
        if (!it.titems.empty())
            std::memcpy(host.data() + b->t_off[k] * sizeof(TItem), it.titems.data(), it.titems.size() * sizeof(TItem));
        if (!it.citems.empty())
            std::memcpy(host.data() + tbytes + b->c_off[k] * sizeof(CItem), it.citems.data(),
                        it.citems.size() * sizeof(CItem));
    }
    if (!small.empty()) std::memcpy(host.data() + b->small_off, small.data(), small.size() * sizeof(DevReq));
    // on the context stream, which orders the copy after earlier users of the
    // pooled buffer; the host vector lives until the copy is done
    if (hipMemcpyAsync(b->blob, host.data(), b->blob_bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        release_batch(b);
        return set_error(GNT_E_DEVICE, "descriptor upload failed");
    }
    if (graph) {
        const TItem *dt = (const TItem *)b->blob;
        const CItem *dc = (const CItem *)((unsigned char *)b->blob + tbytes);
        int rc = 0;
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = -1;
        if (!rc) rc = launch_small(b);
        for (size_t k = 0; !rc && k < b->iters.size(); k++)
            rc = launch_iteration(b->iters[k], dt + b->t_off[k], dc + b->c_off[k], c->tw(), c->stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc || e != hipSuccess || !g || hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            b->exec = nullptr;
            release_batch(b);
            return set_error(GNT_E_DEVICE, "graph capture failed");
        }
        b->graph = g;
    }
    *out = b;
    return GNT_OK;
}

static int run_batch_plan(gnt_batch *b) {
    Ctx *c = b->c;
    if (b->exec) return hipGraphLaunch(b->exec, c->stream) == hipSuccess ? GNT_OK
                                                                         : set_error(GNT_E_DEVICE, "graph launch failed");
    if (launch_small(b)) return set_error(GNT_E_DEVICE, "batch launch failed");
    size_t nt = b->t_off.empty() ? 0 : b->t_off.back() + b->iters.back().titems.size();
    const size_t tbytes = (nt * sizeof(TItem) + 255) & ~(size_t)255;
    const TItem *dt = (const TItem *)b->blob;
    const CItem *dc = (const CItem *)((unsigned char *)b->blob + tbytes);
    for (size_t k = 0; k < b->iters.size(); k++) {
        int rc = launch_iteration(b->iters[k], dt + b->t_off[k], dc + b->c_off[k], c->tw(), c->stream);
        if (rc) return set_error(GNT_E_DEVICE, "batch launch failed");
    }
    return GNT_OK;
}

// The unbatched comparison: the per-product launch sequence, product by product.
static int run_batch_serial(gnt_batch *b) {
    Ctx *c = b->c;
    for (Prod &p : b->prods) {
        int rc = c->enqueue_product(p.da, p.na, p.db, p.nb, p.dout, p.dlen);
        if (rc) return rc;
    }
    return GNT_OK;
}

// Every operand in one packed transfer of in_blob.
static int upload_inputs(gnt_batch *b, const gnt_job *jobs) {
    Ctx *c = b->c;
    std::vector<Seg> segs;
    segs.reserve(2 * b->prods.size());
    for (size_t i = 0; i < b->prods.size(); i++) {
        const Prod &p = b->prods[i];
        if (!jobs[i].a || (!p.sqr() && !jobs[i].b)) return set_error(GNT_E_ARG, "null operand");
        segs.push_back(Seg{size_t((char *)b->in_a[i] - (char *)b->in_blob), (void *)jobs[i].a, p.na * 8});
        if (!p.sqr()) segs.push_back(Seg{size_t((char *)b->in_b[i] - (char *)b->in_blob), (void *)jobs[i].b, p.nb * 8});
    }
    if (c->up.gather(b->in_blob, b->in_bytes, segs, c->stream) != hipSuccess)
        return set_error(GNT_E_DEVICE, "upload failed");
    return GNT_OK;
}

// The lengths in one transfer, then every trimmed result in one packed
// transfer of out_blob.
static int download_outputs(gnt_batch *b, gnt_job *jobs) {
    Ctx *c = b->c;
    const size_t n = b->prods.size();
    if (n == 0) return GNT_OK;
    std::vector<uint32_t> lens(n);
    if (hipMemcpyAsync(lens.data(), b->len_blob, n * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return set_error(GNT_E_DEVICE, "download failed");
    std::vector<Seg> segs;
    segs.reserve(n);
    for (size_t i = 0; i < n; i++) {
        if (jobs[i].out_cap < lens[i]) { jobs[i].status = GNT_E_ARG; continue; }
        if (lens[i]) segs.push_back(Seg{b->out_off[i], jobs[i].out, (size_t)lens[i] * 8});
        jobs[i].out_len = lens[i];
    }
    if (c->down.scatter(b->out_blob, b->out_bytes, segs, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return set_error(GNT_E_DEVICE, "download failed");
    return GNT_OK;
}

// gnt_mul_batch: jobs grouped into plans that fit the HBM admission cap.
int run_batch(Ctx *c, gnt_job *jobs, size_t n) {
    if (!c) return set_error(GNT_E_ARG, "null context");
    if (n && !jobs) return set_error(GNT_E_ARG, "null jobs");
    if (hipSetDevice(c->device) != hipSuccess) return set_error(GNT_E_DEVICE, "hipSetDevice failed");
    std::lock_guard<std::mutex> g(c->mu);
    std::vector<size_t> idx;
    for (size_t i = 0; i < n; i++) {
        gnt_job &j = jobs[i];
        j.status = GNT_OK;
        j.out_len = 0;
        size_t nb = j.is_sqr ? j.na : j.nb;
        if (j.na == 0 || nb == 0) continue;  // empty product: out_len 0
        if (!j.out || j.out_cap < j.na + nb || !j.a || (!j.is_sqr && !j.b)) { j.status = GNT_E_ARG; continue; }
        Plan pl;
        if (!plan_for(j.na, nb, pl)) { j.status = GNT_E_TOO_LARGE; continue; }
        idx.push_back(i);
    }
    int first = GNT_OK;
    size_t k = 0;
    while (k < idx.size()) {
        // admit jobs in order while their footprint fits the cap
        std::vector<size_t> na, nbv;
        std::vector<int> sq;
        std::vector<gnt_job> sub;
        size_t bytes = 0;
        while (k < idx.size()) {
            const gnt_job &j = jobs[idx[k]];
            Prod p;
            p.na = j.na;
            p.nb = j.is_sqr ? j.na : j.nb;
            p.db = j.is_sqr ? nullptr : (const u64 *)1;
            prod_init(p);
            size_t need = p.scratch_bytes() + 8 * (p.na + p.nb + p.nl);
            if (!sub.empty() && bytes + need > c->hbm_cap) break;
            bytes += need;
            na.push_back(j.na);
            nbv.push_back(p.nb);
            sq.push_back(j.is_sqr ? 1 : 0);
            sub.push_back(j);
            k++;
        }
        gnt_batch *b = nullptr;
        int rc = create_batch(c, na.data(), nbv.data(), sq.data(), sub.size(), false, &b);
        if (!rc) rc = upload_inputs(b, sub.data());
        if (!rc) rc = run_batch_plan(b);
        if (!rc) rc = download_outputs(b, sub.data());
        if (b) release_batch(b);
        for (size_t q = 0; q < sub.size(); q++) {
            gnt_job &dst = jobs[idx[k - sub.size() + q]];
            dst.status = rc ? rc : sub[q].status;
            dst.out_len = rc ? 0 : sub[q].out_len;
        }
        if (rc && !first) first = rc;
    }
    for (size_t i = 0; i < n; i++)
        if (jobs[i].status && !first) first = jobs[i].status;
    if (first) set_error(first, "one or more batch jobs failed");
    return first;
}

}  // namespace gnt

extern "C" {

int gnt_batch_create(gnt_ctx *ctx, const size_t *na, const size_t *nb, const int *is_sqr, size_t n, int use_graph,
                     gnt_batch **out) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !out || (n && (!na || !nb))) return gnt::set_error(GNT_E_ARG, "null argument");
    if (hipSetDevice(c->device) != hipSuccess) return gnt::set_error(GNT_E_DEVICE, "hipSetDevice failed");
    std::lock_guard<std::mutex> g(c->mu);
    return gnt::create_batch(c, na, nb, is_sqr, n, use_graph != 0, out);
}

int gnt_batch_upload(gnt_batch *b, const gnt_job *jobs) {
    if (!b || !jobs) return gnt::set_error(GNT_E_ARG, "null argument");
    std::lock_guard<std::mutex> g(b->c->mu);
    int rc = gnt::upload_inputs(b, jobs);
    if (!rc && hipStreamSynchronize(b->c->stream) != hipSuccess) rc = gnt::set_error(GNT_E_DEVICE, "stream failure");
    return rc;
}

int gnt_batch_run(gnt_batch *b, int mode) {
    if (!b) return gnt::set_error(GNT_E_ARG, "null batch");
    std::lock_guard<std::mutex> g(b->c->mu);
    if (mode == 1) return gnt::run_batch_serial(b);
    return gnt::run_batch_plan(b);
}

int gnt_batch_download(gnt_batch *b, gnt_job *jobs) {
    if (!b || !jobs) return gnt::set_error(GNT_E_ARG, "null argument");
    std::lock_guard<std::mutex> g(b->c->mu);
    return gnt::download_outputs(b, jobs);
}

int gnt_batch_info(gnt_batch *b, size_t *launches_per_run, size_t *iterations, size_t *device_bytes) {
    if (!b) return gnt::set_error(GNT_E_ARG, "null batch");
    size_t L = b->n_small ? 2 : 0;  // the counter reset and the persistent launch
    for (const auto &it : b->iters) L += it.groups.size();
    if (launches_per_run) *launches_per_run = L;
    if (iterations) *iterations = b->iters.size();
    if (device_bytes) *device_bytes = b->scratch_bytes;
    return GNT_OK;
}

int gnt_batch_destroy(gnt_batch *b) {
    if (!b) return GNT_OK;
    Ctx *c = b->c;
    std::lock_guard<std::mutex> g(c->mu);
    (void)hipStreamSynchronize(c->stream);
    gnt::release_batch(b);
    return GNT_OK;
}

}  // extern "C"
