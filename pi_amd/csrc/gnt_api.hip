// C ABI of libgnt.so: contexts, device memory pool, single / device-resident /
// batched products and the host-side multiply dispatch.  See include/gnt.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gnt.h"
#include "gl_field.hpp"
#include "gnt_kernels.hpp"
#include "gnt_plan.hpp"
#include "gnt_roots.hpp"
#include "gnt_runtime.hpp"

using gl::u32;
using gl::u64;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(GNT_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

int ceil_log2(size_t x) {
    int l = 0;
    while (((size_t)1 << l) < x) l++;
    return l;
}

}  // namespace

namespace gnt {

int set_error(int code, const char *msg) {
    g_last_error = msg;
    return code;
}

// ---------------------------------------------------------------------------
// Device memory pool (the reference keeps a host scratch pool, ntt.rs:51-155).
// Buffers are handed back in stream order, so reuse by later work on the same
// stream is safe without a sync.  Size classes: 256-byte granules up to
// 64 KiB, then eight classes per octave (at most 12.5% slack), so products of
// uneven sizes reuse each other's buffers instead of going back to
// hipMalloc/hipFree (hipFree waits for the device).
// ---------------------------------------------------------------------------
size_t Pool::size_class(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes <= ((size_t)1 << 16)) return bytes;
    int top = 63 - __builtin_clzll(bytes);
    size_t step = (size_t)1 << (top - 3);
    return (bytes + step - 1) & ~(step - 1);
}

void *Pool::acquire(size_t bytes) {
    bytes = size_class(bytes);
    for (int attempt = 0; attempt < 2; attempt++) {
        std::unique_lock<std::mutex> g(mu);
        auto it = free_.find(bytes);
        if (it != free_.end() && !it->second.empty()) {
            void *p = it->second.back();
            it->second.pop_back();
            pooled_bytes -= bytes;
            return p;
        }
        if (bytes <= SLAB_PIECE_MAX) {
            if (slab_left_ < bytes) {
                void *s = nullptr;
                if (hipMalloc(&s, SLAB_BYTES) != hipSuccess) {
                    g.unlock();
                    if (attempt == 0) trim();  // idle large buffers back to the device, then retry
                    continue;
                }
                slabs_.push_back(s);
                live_bytes += SLAB_BYTES;
                slab_cur_ = (char *)s;
                slab_left_ = SLAB_BYTES;
            }
            void *p = slab_cur_;  // classes are multiples of 256 bytes: pieces stay aligned
            slab_cur_ += bytes;
            slab_left_ -= bytes;
            return p;
        }
        break;
    }
    if (bytes <= SLAB_PIECE_MAX) return nullptr;
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        // give cached buffers back to the device and try once more
        trim();
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    }
    std::lock_guard<std::mutex> g(mu);
    live_bytes += bytes;
    return p;
}

void Pool::release(void *p, size_t bytes) {
    if (!p) return;
    bytes = size_class(bytes);
    std::lock_guard<std::mutex> g(mu);
    auto &v = free_[bytes];
    // slab pieces are always kept; large buffers while the idle total stays
    // under POOL_IDLE_MAX
    if (bytes <= SLAB_PIECE_MAX || (pooled_bytes + bytes <= POOL_IDLE_MAX && v.size() < 4096)) {
        v.push_back(p);
        pooled_bytes += bytes;
    } else {
        (void)hipFree(p);
        live_bytes -= bytes;
    }
}

void Pool::trim() {
    std::lock_guard<std::mutex> g(mu);
    for (auto it = free_.begin(); it != free_.end();) {
        if (it->first <= SLAB_PIECE_MAX) {
            ++it;
            continue;
        }
        for (void *p : it->second) {
            (void)hipFree(p);
            live_bytes -= it->first;
            pooled_bytes -= it->first;
        }
        it = free_.erase(it);
    }
}

void Pool::release_all() {
    trim();
    std::lock_guard<std::mutex> g(mu);
    for (void *s : slabs_) (void)hipFree(s);
    live_bytes -= slabs_.size() * SLAB_BYTES;
    slabs_.clear();
    free_.clear();
    pooled_bytes = 0;
    slab_cur_ = nullptr;
    slab_left_ = 0;
}

// Take over everything `from` holds (free lists and slabs); nothing of
// `from` may be in use or have work pending on it.
void Pool::absorb(Pool &from) {
    if (&from == this) return;
    std::scoped_lock g(mu, from.mu);
    for (auto &kv : from.free_) {
        auto &v = free_[kv.first];
        v.insert(v.end(), kv.second.begin(), kv.second.end());
    }
    slabs_.insert(slabs_.end(), from.slabs_.begin(), from.slabs_.end());
    live_bytes += from.live_bytes;
    pooled_bytes += from.pooled_bytes;
    from.free_.clear();
    from.slabs_.clear();
    from.live_bytes = from.pooled_bytes = 0;
    from.slab_cur_ = nullptr;
    from.slab_left_ = 0;
}

int Ctx::init_tables() {
    std::vector<u64> lo, hi, sm;
    if (!build_root_tables(lo, hi, sm)) return set_error(GNT_E_DEVICE, "root table construction failed");
    const size_t T = lo.size();
    HIP_TRY(hipMalloc(&wlo, T * sizeof(u64)));
    HIP_TRY(hipMalloc(&whi, T * sizeof(u64)));
    HIP_TRY(hipMalloc(&wsm, sm.size() * sizeof(u64)));
    HIP_TRY(hipMemcpy(wlo, lo.data(), T * sizeof(u64), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(whi, hi.data(), T * sizeof(u64), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(wsm, sm.data(), sm.size() * sizeof(u64), hipMemcpyHostToDevice));
    return GNT_OK;
}

void Ctx::pass_tables(const Plan &pl, PassTables &out) {
    out = PassTables{};
    std::lock_guard<std::mutex> lk(tab_mu);
    for (int p = 0; p < pl.npass; p++) {
        const PassDesc &d = pl.pass[p];
        if (!pass_has_table(d)) continue;
        for (int inv = 0; inv < 2; inv++) {
            // the inverse of pass 0 carries the transform's 1/N (inv_scale_e)
            const int scale_e = (inv && p == 0) ? inv_scale_e(d.log_m) : 0;
            const int key = (d.log_m << 8) | ((scale_e ? 1 : 0) << 6) | (d.log_r << 1) | inv;
            auto it = ptab.find(key);
            uint64_t *T = nullptr;
            if (it != ptab.end()) {
                T = it->second;
            } else {
                if (hipMalloc(&T, sizeof(uint64_t) << d.log_m) != hipSuccess) {
                    (void)hipGetLastError();
                    T = nullptr;
                } else if (launch_pass_table(T, d.log_m, d.log_r, inv != 0, scale_e, tw(), stream) != 0 ||
                           hipStreamSynchronize(stream) != hipSuccess) {
                    // built before first use: the serving lanes read it on their own streams
                    (void)hipFree(T);
                    T = nullptr;
                }
                if (!T) continue;
                ptab[key] = T;
            }
            (inv ? out.inv : out.fwd)[p] = T;
        }
    }
}

void Ctx::free_pass_tables() {
    std::lock_guard<std::mutex> lk(tab_mu);
    for (auto &kv : ptab) (void)hipFree(kv.second);
    ptab.clear();
}

bool plan_for(size_t na, size_t nb, Plan &pl) {
    size_t n_out = 4 * (na + nb) - 1;
    int log_n = ceil_log2(n_out);
    if (log_n > MAX_LOG_N) return false;
    pl = make_plan(log_n);
    return true;
}

// Device scratch of one product: two coefficient arrays (one for a square),
// the limb sums and the carry flags.
ProductScratch product_scratch_layout(size_t na, size_t nb, bool sqr, int log_n) {
    ProductScratch s;
    s.N = (size_t)1 << log_n;
    s.n_limbs = na + nb;
    s.nt = carry_tiles(s.n_limbs);
    s.sqr = sqr;
    return s;
}

// The launch sequence of one product on `st`, all buffers given: forward
// passes, fused last pass + pointwise + first inverse pass, remaining inverse
// passes, carry.  ev (if not null) receives 5 phase events.
int launch_product(const Plan &pl, const PassTables &tabs, const Tw &tw, const u64 *da, size_t na, const u64 *db,
                   size_t nb, u64 *dout, u32 *dlen, const ProductScratch &s, u64 *xa, u64 *xb, u64 *lh,
                   unsigned char *gp, hipStream_t st, hipEvent_t *ev) {
    const bool sqr = s.sqr;
    const size_t n_out = 4 * (na + nb) - 1;
    u64 *Ls = lh, *Hs = lh + s.n_limbs;
    if (ev) (void)hipEventRecord(ev[0], st);
    const int P = pl.npass;
    const bool ab = !sqr && P > 1;  // both last forward passes in the fused kernel
    int rc = 0;
    if (sqr) {
        if (P > 1) rc = launch_forward(pl, P - 1, xa, da, na, tw, tabs, st);
    } else if (ab) {
        rc = launch_forward_pair(pl, P - 1, xa, da, na, xb, db, nb, tw, tabs, st);
    } else {
        rc = launch_forward(pl, P, xa, da, na, tw, tabs, st);
    }
    if (ev) (void)hipEventRecord(ev[1], st);
    if (!rc && ab) {
        rc = launch_last_fused_ab(pl, xa, xb, tw, tabs, st);
    } else if (!rc) {
        const u64 *bsrc = sqr ? xa : xb;
        const u64 *blimbs = sqr ? da : db;
        rc = launch_last_fused(pl, bsrc, blimbs, sqr ? na : nb, xa, sqr, tw, tabs, Ls, Hs, n_out, s.n_limbs, st);
    }
    if (ev) (void)hipEventRecord(ev[2], st);
    if (!rc) rc = launch_inverse_rest(pl, xa, tw, tabs, Ls, Hs, n_out, s.n_limbs, st);
    if (ev) (void)hipEventRecord(ev[3], st);
    if (!rc) rc = launch_carry(Ls, Hs, dout, s.n_limbs, gp, gp + s.nt, dlen, st);
    if (ev) (void)hipEventRecord(ev[4], st);
    return rc;
}

// Enqueue one product of device limb arrays on ctx->stream.  b == nullptr
// squares a.  out holds na+nb limbs; *len receives the trimmed length.
int Ctx::enqueue_product(const u64 *da, size_t na, const u64 *db, size_t nb, u64 *dout, u32 *dlen) {
    const bool sqr = db == nullptr;
    if (sqr) nb = na;
    Plan pl;
    if (!plan_for(na, nb, pl)) return set_error(GNT_E_TOO_LARGE, "product exceeds the transform cap");
    PassTables tabs;
    pass_tables(pl, tabs);
    const ProductScratch s = product_scratch_layout(na, nb, sqr, pl.log_n);
    u64 *xa = (u64 *)pool.acquire(s.N * sizeof(u64));
    u64 *xb = sqr ? nullptr : (u64 *)pool.acquire(s.N * sizeof(u64));
    u64 *lh = (u64 *)pool.acquire(2 * s.n_limbs * sizeof(u64));
    unsigned char *gp = (unsigned char *)pool.acquire(2 * s.nt);
    auto give_back = [&]() {
        pool.release(xa, s.N * sizeof(u64));
        pool.release(xb, s.N * sizeof(u64));
        pool.release(lh, 2 * s.n_limbs * sizeof(u64));
        pool.release(gp, 2 * s.nt);
    };
    if (!xa || (!sqr && !xb) || !lh || !gp) {
        give_back();
        return set_error(GNT_E_NOMEM, "device memory exhausted");
    }
    hipEvent_t ev[5];
    if (timing)
        for (auto &e : ev) (void)hipEventCreate(&e);
    const int rc = launch_product(pl, tabs, tw(), da, na, db, nb, dout, dlen, s, xa, xb, lh, gp, stream,
                                  timing ? ev : nullptr);
    give_back();
    if (timing) {
        (void)hipEventSynchronize(ev[4]);
        float t;
        (void)hipEventElapsedTime(&t, ev[0], ev[1]); phase_ms[0] += t;
        (void)hipEventElapsedTime(&t, ev[1], ev[2]); phase_ms[1] += t;
        (void)hipEventElapsedTime(&t, ev[2], ev[3]); phase_ms[2] += t;
        (void)hipEventElapsedTime(&t, ev[3], ev[4]); phase_ms[3] += t;
        (void)hipEventElapsedTime(&t, ev[0], ev[4]); phase_ms[4] += t;
        products++;
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
    if (rc) return set_error(GNT_E_DEVICE, "kernel launch failed");
    return GNT_OK;
}

}  // namespace gnt

using gnt::Ctx;

struct gnt_buf {
    u64 *d = nullptr;
    size_t cap = 0;         // limbs allocated (untrimmed length of the value)
    u32 *dlen = nullptr;    // device-side trimmed length
    size_t len = 0;
    bool len_known = false;
    bool complete = false;  // contents final (no pending work writes them)
};

namespace gnt {
gnt_buf *buf_wrap(Ctx *c, u64 *d, size_t cap, u32 *dlen, int64_t len) {
    (void)c;
    gnt_buf *b = new gnt_buf();
    b->d = d;
    b->cap = cap;
    b->dlen = dlen;
    b->complete = true;  // handed out only after its product finished
    if (len >= 0) {
        b->len = (size_t)len;
        b->len_known = true;
    }
    return b;
}
bool buf_info(gnt_buf *b, u64 **d, size_t *cap) {
    if (!b) return false;
    *d = b->d;
    *cap = b->len_known ? b->len : b->cap;
    return true;
}
}  // namespace gnt

extern "C" {

const char *gnt_version(void) { return "gnt 0.1 (gfx950)"; }
const char *gnt_last_error(void) { return g_last_error.c_str(); }

int gnt_init(gnt_ctx **out, const gnt_config *cfg) {
    if (!out) return fail(GNT_E_ARG, "null ctx pointer");
    int dev = cfg ? cfg->device : 0;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(GNT_E_DEVICE, "no HIP device");
    if (dev < 0 || dev >= count) return fail(GNT_E_ARG, "device %d out of range", dev);
    HIP_TRY(hipSetDevice(dev));
    Ctx *c = new Ctx();
    c->device = dev;
    c->hbm_cap = (cfg && cfg->hbm_cap_bytes) ? cfg->hbm_cap_bytes : (size_t)200 * 1000 * 1000 * 1000;
    if (cfg && cfg->serve_split_log_n > (size_t)gnt::MAX_LOG_N) {
        delete c;
        return fail(GNT_E_ARG, "serve_split_log_n above %d", gnt::MAX_LOG_N);
    }
    c->serve_split_log_n = (cfg && cfg->serve_split_log_n) ? (int)cfg->serve_split_log_n : 22;
    // one large product at a time: overlapping large products share the
    // device and all finish late (measured: p99 at 400 req/s 18 ms against 23)
    c->serve_large_inflight = (cfg && cfg->serve_large_inflight) ? cfg->serve_large_inflight : 1;
    c->serve_device_limbs = !cfg || !cfg->serve_device_limbs            ? 2048
                            : cfg->serve_device_limbs == GNT_SERVE_DEVICE_OFF ? 0
                                                                            : cfg->serve_device_limbs;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(GNT_E_DEVICE, "stream creation failed");
    }
    int rc = c->init_tables();
    // the no-allocation entry point's scratch address on this device; without
    // it only gnt_mul_noalloc fails (with GNT_E_DEVICE)
    if (!rc) (void)gnt::noalloc_setup(c);
    if (!rc && cfg && cfg->plan_cache_path && *cfg->plan_cache_path) rc = gnt_load_plans(cfg->plan_cache_path);
    if (rc) {
        (void)hipStreamDestroy(c->stream);
        (void)hipStreamDestroy(c->copy_stream);
        delete c;
        return rc;
    }
    *out = reinterpret_cast<gnt_ctx *>(c);
    return GNT_OK;
}

int gnt_shutdown(gnt_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return GNT_OK;
    gnt_serve_stop(ctx);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->up.release();
    c->down.release();
    c->pool.release_all();
    c->upload_pool.release_all();
    c->free_pass_tables();
    (void)hipFree(c->wlo);
    (void)hipFree(c->whi);
    (void)hipFree(c->wsm);
    (void)hipStreamDestroy(c->stream);
    (void)hipStreamDestroy(c->copy_stream);
    delete c;
    return GNT_OK;
}

static int check_ctx(Ctx *c) {
    if (!c) return fail(GNT_E_ARG, "null context");
    if (hipSetDevice(c->device) != hipSuccess) return fail(GNT_E_DEVICE, "hipSetDevice failed");
    return GNT_OK;
}

// One product of host operands (b == nullptr squares a).
// While the serving loop runs, host-buffer products from any thread go through
// it, so concurrent callers share launches (continuous batching) instead of
// taking turns on the context.
static int served_product(Ctx *c, const u64 *a, size_t na, const u64 *b, size_t nb, u64 *out, size_t out_cap,
                          size_t *out_len) {
    gnt_ctx *ctx = reinterpret_cast<gnt_ctx *>(c);
    gnt_req r{};
    r.op = b ? GNT_OP_MUL : GNT_OP_SQR;
    r.a = a;
    r.na = na;
    r.b = b;
    r.nb = nb;
    uint64_t ticket = 0;
    if (int rc = gnt_submit(ctx, &r, &ticket)) return rc;
    gnt_buf *res = nullptr;
    if (int rc = gnt_wait(ctx, ticket, -1, &res, nullptr)) return rc;
    size_t len = 0;
    int rc = gnt_buf_download(ctx, res, out, out_cap, &len);
    gnt_buf_release(ctx, res);
    if (rc) return rc;
    *out_len = len;
    return GNT_OK;
}

static int host_product(Ctx *c, const u64 *a, size_t na, const u64 *b, size_t nb, u64 *out, size_t out_cap,
                        size_t *out_len) {
    const bool sqr = b == nullptr;
    if (sqr) nb = na;
    if (!out_len) return fail(GNT_E_ARG, "null out_len");
    if (na == 0 || nb == 0) {
        *out_len = 0;
        return GNT_OK;
    }
    if (!a || (!sqr && !b) || !out) return fail(GNT_E_ARG, "null operand");
    if (out_cap < na + nb) return fail(GNT_E_ARG, "out_cap %zu < %zu", out_cap, na + nb);
    gnt::Plan pl;
    if (!gnt::plan_for(na, nb, pl)) return fail(GNT_E_TOO_LARGE, "4*(na+nb) exceeds 2^%d", gnt::MAX_LOG_N);
    if (c->server) return served_product(c, a, na, sqr ? nullptr : b, nb, out, out_cap, out_len);
    std::lock_guard<std::mutex> g(c->mu);
    const size_t nl = na + nb;
    u64 *da = (u64 *)c->pool.acquire(na * 8);
    u64 *db = sqr ? nullptr : (u64 *)c->pool.acquire(nb * 8);
    u64 *dout = (u64 *)c->pool.acquire(nl * 8);
    u32 *dlen = (u32 *)c->pool.acquire(sizeof(u32));
    int rc = GNT_OK;
    if (!da || (!sqr && !db) || !dout || !dlen) rc = fail(GNT_E_NOMEM, "device memory exhausted");
    u32 len = 0;
    if (!rc && c->up.copy(da, a, na * 8, c->stream) != hipSuccess) rc = fail(GNT_E_DEVICE, "upload failed");
    if (!rc && !sqr && c->up.copy(db, b, nb * 8, c->stream) != hipSuccess) rc = fail(GNT_E_DEVICE, "upload failed");
    if (!rc) rc = c->enqueue_product(da, na, db, nb, dout, dlen);
    if (!rc && hipMemcpyAsync(&len, dlen, sizeof(u32), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = fail(GNT_E_DEVICE, "download failed");
    if (!rc && c->down.copy(out, dout, nl * 8, c->stream) != hipSuccess) rc = fail(GNT_E_DEVICE, "download failed");
    hipError_t se = hipStreamSynchronize(c->stream);
    if (!rc && se != hipSuccess) rc = fail(GNT_E_DEVICE, "stream: %s", hipGetErrorString(se));
    c->pool.release(da, na * 8);
    c->pool.release(db, nb * 8);
    c->pool.release(dout, nl * 8);
    c->pool.release(dlen, sizeof(u32));
    if (rc) return rc;
    *out_len = len;
    return GNT_OK;
}

int gnt_mul(gnt_ctx *ctx, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out,
            size_t out_cap, size_t *out_len) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_ctx(c)) return rc;
    if (na && nb && !b) return fail(GNT_E_ARG, "null operand");
    return host_product(c, a, na, b, nb, out, out_cap, out_len);
}

int gnt_sqr(gnt_ctx *ctx, const uint64_t *a, size_t na, uint64_t *out, size_t out_cap, size_t *out_len) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_ctx(c)) return rc;
    return host_product(c, a, na, nullptr, na, out, out_cap, out_len);
}

// ---- device-resident operands ----

static gnt_buf *new_buf(Ctx *c, size_t n) {
    gnt_buf *b = new gnt_buf();
    b->cap = n;
    b->d = n ? (u64 *)c->pool.acquire(n * 8) : nullptr;
    b->dlen = (u32 *)c->pool.acquire(sizeof(u32));
    if ((n && !b->d) || !b->dlen) {
        c->pool.release(b->d, n * 8);
        c->pool.release(b->dlen, sizeof(u32));
        delete b;
        return nullptr;
    }
    return b;
}

int gnt_buf_upload(gnt_ctx *ctx, const uint64_t *limbs, size_t n, gnt_buf **out) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_ctx(c)) return rc;
    if (!out || (n && !limbs)) return fail(GNT_E_ARG, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    size_t len = n;
    while (len > 0 && limbs[len - 1] == 0) len--;
    gnt_buf *b = new_buf(c, len);
    if (!b) return fail(GNT_E_NOMEM, "device memory exhausted");
    if (len) HIP_TRY(c->up.copy(b->d, limbs, len * 8, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    b->len = len;
    b->len_known = true;
    b->complete = true;
    *out = b;
    return GNT_OK;
}

// Length of a buffer.  A finished buffer is read over the copy stream without
// waiting for unrelated work on the compute stream.
static int buf_len_locked(Ctx *c, gnt_buf *b, size_t *len) {
    if (!b->len_known) {
        u32 l = 0;
        hipStream_t st = b->complete ? c->copy_stream : c->stream;
        HIP_TRY(hipMemcpyAsync(&l, b->dlen, sizeof(u32), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        b->len = l;
        b->len_known = true;
    }
    *len = b->len;
    return GNT_OK;
}

int gnt_buf_len(gnt_ctx *ctx, gnt_buf *b, size_t *len) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_ctx(c)) return rc;
    if (!b || !len) return fail(GNT_E_ARG, "null argument");
    if (b->complete) return buf_len_locked(c, b, len);
    std::lock_guard<std::mutex> g(c->mu);
    return buf_len_locked(c, b, len);
}

int gnt_buf_download(gnt_ctx *ctx, gnt_buf *b, uint64_t *out, size_t cap, size_t *out_len) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_ctx(c)) return rc;
    if (!b || !out_len) return fail(GNT_E_ARG, "null argument");
    std::unique_lock<std::mutex> g(c->mu, std::defer_lock);
    if (!b->complete) g.lock();
    size_t len;
    if (int rc = buf_len_locked(c, b, &len)) return rc;
    if (cap < len || (len && !out)) return fail(GNT_E_ARG, "capacity %zu < %zu", cap, len);
    if (len) {
        hipStream_t st = b->complete ? c->copy_stream : c->stream;
        HIP_TRY(c->down.copy(out, b->d, len * 8, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    *out_len = len;
    return GNT_OK;
}

int gnt_buf_release(gnt_ctx *ctx, gnt_buf *b) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_ctx(c)) return rc;
    if (!b) return GNT_OK;
    std::lock_guard<std::mutex> g(c->mu);
    c->pool.release(b->d, b->cap * 8);
    c->pool.release(b->dlen, sizeof(u32));
    delete b;
    return GNT_OK;
}

static int buf_product(Ctx *c, gnt_buf *a, gnt_buf *b, gnt_buf **out) {
    if (!a || !out) return fail(GNT_E_ARG, "null argument");
    const bool sqr = (b == nullptr || b == a);
    // Operand lengths: the trimmed length when known, else the capacity (the
    // extra top limbs are zeros and leave the product unchanged).
    size_t na = a->len_known ? a->len : a->cap;
    size_t nb = sqr ? na : (b->len_known ? b->len : b->cap);
    if (na == 0 || nb == 0) {
        gnt_buf *r = new_buf(c, 0);
        if (!r) return fail(GNT_E_NOMEM, "device memory exhausted");
        r->len = 0;
        r->len_known = true;
        *out = r;
        return GNT_OK;
    }
    gnt::Plan pl;
    if (!gnt::plan_for(na, nb, pl)) return fail(GNT_E_TOO_LARGE, "4*(na+nb) exceeds 2^%d", gnt::MAX_LOG_N);
    gnt_buf *r = new_buf(c, na + nb);
    if (!r) return fail(GNT_E_NOMEM, "device memory exhausted");
    int rc = c->enqueue_product(a->d, na, sqr ? nullptr : b->d, nb, r->d, r->dlen);
    if (rc) {
        c->pool.release(r->d, r->cap * 8);
        c->pool.release(r->dlen, sizeof(u32));
        delete r;
        return rc;
    }
    *out = r;
    return GNT_OK;
}

int gnt_mul_buf(gnt_ctx *ctx, gnt_buf *a, gnt_buf *b, gnt_buf **out) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_ctx(c)) return rc;
    if (!b) return fail(GNT_E_ARG, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    return buf_product(c, a, b, out);
}

int gnt_sqr_buf(gnt_ctx *ctx, gnt_buf *a, gnt_buf **out) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_ctx(c)) return rc;
    std::lock_guard<std::mutex> g(c->mu);
    return buf_product(c, a, nullptr, out);
}

int gnt_sync(gnt_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_ctx(c)) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GNT_OK;
}

// ---- batches ----

int gnt_mul_batch_serial(gnt_ctx *ctx, gnt_job *jobs, size_t n) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = check_ctx(c)) return rc;
    if (n && !jobs) return fail(GNT_E_ARG, "null jobs");
    int first = GNT_OK;
    for (size_t i = 0; i < n; i++) {
        gnt_job &j = jobs[i];
        j.status = host_product(c, j.a, j.na, j.is_sqr ? nullptr : j.b, j.is_sqr ? j.na : j.nb, j.out, j.out_cap,
                                &j.out_len);
        if (j.status && !first) first = j.status;
    }
    return first;
}

int gnt_mul_batch(gnt_ctx *ctx, gnt_job *jobs, size_t n) {
    return gnt::run_batch(reinterpret_cast<Ctx *>(ctx), jobs, n);
}

// ---- introspection ----

int gnt_plan_info(size_t na, size_t nb, int *log_n, int *npass, int *log_r) {
    if (na == 0 || nb == 0) return fail(GNT_E_ARG, "empty operand");
    gnt::Plan pl;
    if (!gnt::plan_for(na, nb, pl)) return fail(GNT_E_TOO_LARGE, "too large");
    if (log_n) *log_n = pl.log_n;
    if (npass) *npass = pl.npass;
    if (log_r)
        for (int p = 0; p < 4; p++) log_r[p] = p < pl.npass ? pl.pass[p].log_r : 0;
    return GNT_OK;
}

int gnt_set_timing(gnt_ctx *ctx, int on) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return fail(GNT_E_ARG, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    c->timing = on != 0;
    for (double &t : c->phase_ms) t = 0;
    c->products = 0;
    return GNT_OK;
}

int gnt_get_timing(gnt_ctx *ctx, double *ms, int n, uint64_t *products) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return fail(GNT_E_ARG, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    for (int i = 0; i < n && i < 5; i++) ms[i] = c->phase_ms[i];
    if (products) *products = c->products;
    return GNT_OK;
}

}  // extern "C"
