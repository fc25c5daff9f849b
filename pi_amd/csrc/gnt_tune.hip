// Autotuner: pass radices per transform length, timed on the device.
//
// For every length 2^log_n with more than one valid split (every composition
// of log_n into 2..4 passes of radix 2^7..2^10, the tile shapes the kernels
// are compiled for), a product whose transform has exactly that length is
// timed with each split, and the fastest split is installed for that length.
// The choices persist as a small JSON plan cache:
//   {"device": "...", "plans": {"14": [7, 7], ...}, "us": {"14": {"7,7": 12.3, ...}}}
// which gnt_init loads from gnt_config.plan_cache_path.  No reference
// counterpart: the reference's only transform knob is the four-step
// threshold (crates/bignum/src/config.rs:81-89), a CPU cache-size choice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gnt.h"
#include "gnt_plan.hpp"
#include "gnt_stage.hpp"
#include "gnt_runtime.hpp"
#include "gnt_stage.hpp"

using u64 = uint64_t;

namespace gnt {
namespace {

std::mutex g_plan_mu;  // serialises writers of the plan words and the table limit

void compositions(int log_n, int np, std::vector<int> &cur, std::vector<std::vector<int>> &out) {
    if ((int)cur.size() == np) {
        int s = 0;
        for (int x : cur) s += x;
        if (s == log_n) out.push_back(cur);
        return;
    }
    for (int r = 7; r <= 10; r++) {
        cur.push_back(r);
        compositions(log_n, np, cur, out);
        cur.pop_back();
    }
}

std::vector<std::vector<int>> candidates(int log_n) {
    std::vector<std::vector<int>> out;
    if (log_n <= TILE_LOG) {
        out.push_back({log_n});
        return out;
    }
    for (int np = 2; np <= MAX_PASSES; np++) {
        std::vector<int> cur;
        compositions(log_n, np, cur, out);
    }
    return out;
}

void install(int log_n, int np, const int *lr) {
    std::lock_guard<std::mutex> g(g_plan_mu);
    set_plan_choice(log_n, np, lr);
}

// Fingerprints of the integer with little-endian limbs x: x mod (2^64 - 1)
// (2^64 = 1 there: the limb sum with end-around carry) and x mod (2^61 - 1)
// (2^64 = 8 there; Horner from the top limb).
struct Fp {
    u64 m64, m61;
    bool operator==(const Fp &o) const { return m64 == o.m64 && m61 == o.m61; }
};
constexpr u64 M61 = (1ull << 61) - 1;
u64 mod61(unsigned __int128 v) {
    u64 r = (u64)(v & M61) + (u64)((v >> 61) & M61) + (u64)(v >> 122);
    while (r >= M61) r -= M61;
    return r;
}
Fp fingerprint(const u64 *x, size_t n) {
    unsigned __int128 s = 0;
    u64 h = 0;
    for (size_t i = 0; i < n; i++) s += x[i];
    for (size_t i = n; i-- > 0;) h = mod61((unsigned __int128)h * 8 + x[i]);
    u64 m = (u64)s + (u64)(s >> 64);
    if (m < (u64)s) m++;             // end-around carry
    if (m == ~0ull) m = 0;           // 2^64 - 1 = 0
    return Fp{m, h};
}
// c == a * b by the two fingerprints and the low limb.
bool product_consistent(const u64 *a, size_t na, const u64 *b, size_t nb, const u64 *c, size_t nc) {
    const Fp fa = fingerprint(a, na), fb = fingerprint(b, nb), fc = fingerprint(c, nc);
    unsigned __int128 p = (unsigned __int128)fa.m64 * fb.m64;
    u64 m = (u64)p + (u64)(p >> 64);
    if (m < (u64)p) m++;
    if (m == ~0ull) m = 0;
    const u64 h = mod61((unsigned __int128)fa.m61 * fb.m61);
    const u64 low = nc ? c[0] : 0;
    return m == fc.m64 && h == fc.m61 && low == a[0] * b[0];
}

std::string key(const std::vector<int> &v) {
    std::string s;
    for (size_t i = 0; i < v.size(); i++) s += (i ? "," : "") + std::to_string(v[i]);
    return s;
}

}  // namespace
}  // namespace gnt

using gnt::Ctx;

extern "C" {

int gnt_set_pass_tables(int max_log_m, int *previous) {
    if (max_log_m < -1 || max_log_m > gnt::MAX_LOG_N) return gnt::set_error(GNT_E_ARG, "max_log_m out of range");
    std::lock_guard<std::mutex> g(gnt::g_plan_mu);
    if (previous) *previous = gnt::g_table_max_log.load();
    gnt::g_table_max_log.store(max_log_m);
    return GNT_OK;
}

int gnt_set_staging(int threads, size_t chunk_bytes, size_t min_bytes) {
    if (threads < 0 || threads > 64) return gnt::set_error(GNT_E_ARG, "threads out of range (0..64)");
    if (threads > 0 && (chunk_bytes < (size_t(1) << 16) || chunk_bytes > (size_t(1) << 30)))
        return gnt::set_error(GNT_E_ARG, "chunk_bytes out of range (64 KiB..1 GiB)");
    gnt::StageConfig s;
    s.threads = threads;
    s.chunk = chunk_bytes;
    s.min_bytes = min_bytes;
    gnt::set_stage_config(s);
    return GNT_OK;
}

int gnt_set_plan(int log_n, int npass, const int *log_r) {
    if (log_n < 1 || log_n > gnt::MAX_LOG_N) return gnt::set_error(GNT_E_ARG, "log_n out of range");
    if (npass == 0) {  // back to the built-in rule
        std::lock_guard<std::mutex> g(gnt::g_plan_mu);
        gnt::set_plan_choice(log_n, 0, nullptr);
        return GNT_OK;
    }
    if (!log_r || !gnt::split_valid(log_n, npass, log_r))
        return gnt::set_error(GNT_E_ARG, "split not covered by the compiled tile shapes");
    gnt::install(log_n, npass, log_r);
    return GNT_OK;
}

int gnt_load_plans(const char *path) {
    if (!path) return gnt::set_error(GNT_E_ARG, "null path");
    FILE *f = std::fopen(path, "rb");
    if (!f) return gnt::set_error(GNT_E_ARG, "cannot open plan cache");
    std::string s;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
    std::fclose(f);
    size_t at = s.find("\"plans\"");
    if (at == std::string::npos) return gnt::set_error(GNT_E_ARG, "plan cache has no \"plans\"");
    at = s.find('{', at);
    const size_t end = at == std::string::npos ? at : s.find('}', at);
    if (end == std::string::npos) return gnt::set_error(GNT_E_ARG, "malformed plan cache");
    // entries "<log_n>": [r0, r1, ...]
    size_t i = at + 1;
    int loaded = 0;
    while (i < end) {
        size_t q = s.find('"', i);
        if (q == std::string::npos || q >= end) break;
        const int log_n = std::atoi(s.c_str() + q + 1);
        size_t lb = s.find('[', q), rb = s.find(']', q);
        if (lb == std::string::npos || rb == std::string::npos || rb > end) return gnt::set_error(GNT_E_ARG, "malformed plan cache");
        int lr[gnt::MAX_PASSES] = {0, 0, 0, 0};
        int np = 0;
        for (size_t k = lb + 1; k < rb && np <= gnt::MAX_PASSES;) {
            while (k < rb && (s[k] == ' ' || s[k] == ',')) k++;
            if (k >= rb) break;
            if (np == gnt::MAX_PASSES) return gnt::set_error(GNT_E_ARG, "too many passes in plan cache");
            lr[np++] = std::atoi(s.c_str() + k);
            while (k < rb && s[k] != ',') k++;
        }
        if (int rc = gnt_set_plan(log_n, np, lr)) return rc;
        loaded++;
        i = rb + 1;
    }
    return loaded > 0 ? GNT_OK : gnt::set_error(GNT_E_ARG, "empty plan cache");
}

int gnt_tune(gnt_ctx *ctx, int min_log_n, int max_log_n, int reps, const char *cache_path) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return gnt::set_error(GNT_E_ARG, "null context");
    if (min_log_n < 4 || max_log_n > 28 || min_log_n > max_log_n) return gnt::set_error(GNT_E_ARG, "range must lie in [4, 28]");
    if (reps < 1) reps = 3;
    if (hipSetDevice(c->device) != hipSuccess) return gnt::set_error(GNT_E_DEVICE, "hipSetDevice failed");
    std::lock_guard<std::mutex> g(c->mu);
    // operands for the largest length: na = nb = 2^(log_n-3) fills the transform
    // exactly; two different SplitMix64 streams, so the multiply is no square
    const size_t nmax = (size_t)1 << (max_log_n - 3);
    std::vector<u64> ha(nmax), hb(nmax), hout(2 * nmax);
    u64 st = 0x243F6A8885A308D3ull;
    for (size_t i = 0; i < 2 * nmax; i++) {
        st += 0x9E3779B97F4A7C15ull;
        u64 z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        (i < nmax ? ha[i] : hb[i - nmax]) = z ^ (z >> 31);
    }
    u64 *da = nullptr, *db = nullptr, *dout = nullptr;
    uint32_t *dlen = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = GNT_OK;
    std::string plans, times, times_sq;
    if (hipMalloc(&da, nmax * 8) || hipMalloc(&db, nmax * 8) || hipMalloc(&dout, 2 * nmax * 8) ||
        hipMalloc(&dlen, 4) || hipMemcpy(da, ha.data(), nmax * 8, hipMemcpyHostToDevice) ||
        hipMemcpy(db, hb.data(), nmax * 8, hipMemcpyHostToDevice) || hipEventCreate(&e0) || hipEventCreate(&e1))
        rc = gnt::set_error(GNT_E_DEVICE, "tuner allocation failed");
    // median device time (us) of reps products after one warm-up, and the
    // product itself checked (fingerprints mod 2^64-1 and 2^61-1, low limb)
    auto time_product = [&](size_t n, bool sqr, float &med) -> int {
        std::vector<float> us;
        for (int r = 0; r <= reps; r++) {
            if (hipEventRecord(e0, c->stream) || c->enqueue_product(da, n, sqr ? nullptr : db, n, dout, dlen) ||
                hipEventRecord(e1, c->stream) || hipEventSynchronize(e1))
                return gnt::set_error(GNT_E_DEVICE, "tuner product failed");
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (r > 0) us.push_back(ms * 1000.0f);
        }
        std::sort(us.begin(), us.end());
        med = us[us.size() / 2];
        if (hipMemcpy(hout.data(), dout, 2 * n * 8, hipMemcpyDeviceToHost))
            return gnt::set_error(GNT_E_DEVICE, "tuner download failed");
        const u64 *b = sqr ? ha.data() : hb.data();
        if (!gnt::product_consistent(ha.data(), n, b, n, hout.data(), 2 * n))
            return gnt::set_error(GNT_E_DEVICE, "tuner: a candidate split gave a wrong product");
        return GNT_OK;
    };
    for (int log_n = min_log_n; !rc && log_n <= max_log_n; log_n++) {
        const size_t n = (size_t)1 << (log_n - 3);
        const auto cands = gnt::candidates(log_n);
        std::vector<int> best;
        float best_us = 0;
        std::string row, row_sq;
        for (const auto &cand : cands) {
            gnt::install(log_n, (int)cand.size(), cand.data());
            float mul_us = 0, sqr_us = 0;
            if ((rc = time_product(n, false, mul_us)) || (rc = time_product(n, true, sqr_us))) break;
            char b[64];
            std::snprintf(b, sizeof b, "%s\"%s\": %.2f", row.empty() ? "" : ", ", gnt::key(cand).c_str(), mul_us);
            row += b;
            std::snprintf(b, sizeof b, "%s\"%s\": %.2f", row_sq.empty() ? "" : ", ", gnt::key(cand).c_str(), sqr_us);
            row_sq += b;
            // one split per length serves multiplies and squares: the
            // fastest for a multiply and a square together
            if (best.empty() || mul_us + sqr_us < best_us) {
                best = cand;
                best_us = mul_us + sqr_us;
            }
        }
        if (rc) break;
        gnt::install(log_n, (int)best.size(), best.data());
        std::string l = std::to_string(log_n);
        plans += (plans.empty() ? "" : ",\n  ") + ("\"" + l + "\": [" + gnt::key(best) + "]");
        times += (times.empty() ? "" : ",\n  ") + ("\"" + l + "\": {" + row + "}");
        times_sq += (times_sq.empty() ? "" : ",\n  ") + ("\"" + l + "\": {" + row_sq + "}");
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dout);
    (void)hipFree(dlen);
    if (rc) return rc;
    if (cache_path && *cache_path) {
        hipDeviceProp_t prop;
        std::string dev = hipGetDeviceProperties(&prop, c->device) == hipSuccess ? prop.gcnArchName : "unknown";
        FILE *f = std::fopen(cache_path, "wb");
        if (!f) return gnt::set_error(GNT_E_ARG, "cannot write plan cache");
        std::fprintf(f, "{\"device\": \"%s\",\n \"plans\": {\n  %s\n },\n \"us\": {\n  %s\n },\n \"us_sqr\": {\n  %s\n }\n}\n",
                     dev.c_str(), plans.c_str(), times.c_str(), times_sq.c_str());
        std::fclose(f);
    }
    return GNT_OK;
}

}  // extern "C"
