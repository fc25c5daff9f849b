// Pinned staging of host <-> device copies (gnt_stage.hpp).
#include "gnt_stage.hpp"

#include <algorithm>
#include <cstring>

namespace gnt {

static std::mutex g_stage_mu;
static StageConfig g_stage;

StageConfig stage_config() {
    std::lock_guard<std::mutex> g(g_stage_mu);
    return g_stage;
}

void set_stage_config(const StageConfig &s) {
    std::lock_guard<std::mutex> g(g_stage_mu);
    g_stage = s;
}

// ---- CopyTeam ----

void CopyTeam::worker(int idx, uint64_t seen) {
    for (;;) {
        char *d;
        const char *s;
        size_t n;
        {
            std::unique_lock<std::mutex> g(mu_);
            go_.wait(g, [&] { return quit_ || gen_ != seen; });
            if (quit_) return;
            seen = gen_;
            const size_t off = slice_ * size_t(idx + 1);
            d = dst_ + off;
            s = src_ + off;
            n = off < bytes_ ? std::min(slice_, bytes_ - off) : 0;
        }
        if (n) std::memcpy(d, s, n);
        std::lock_guard<std::mutex> g(mu_);
        if (--left_ == 0) done_.notify_one();
    }
}

void CopyTeam::copy(void *dst, const void *src, size_t bytes, int threads) {
    const int helpers = std::max(0, threads - 1);
    if (helpers == 0 || bytes < (size_t(256) << 10)) {
        std::memcpy(dst, src, bytes);
        return;
    }
    // a new helper starts at the current generation, so it takes only jobs
    // posted after it exists (copy() is called under the stager's lock)
    while ((int)th_.size() < helpers) th_.emplace_back(&CopyTeam::worker, this, (int)th_.size(), gen_);
    // slices of whole cache lines; helper i copies slice i+1
    size_t slice = (bytes + helpers) / (helpers + 1);
    slice = (slice + 63) & ~size_t(63);
    {
        std::lock_guard<std::mutex> g(mu_);
        dst_ = (char *)dst;
        src_ = (const char *)src;
        slice_ = slice;
        bytes_ = bytes;
        left_ = (int)th_.size();
        gen_++;
    }
    go_.notify_all();
    std::memcpy(dst, src, std::min(slice, bytes));
    std::unique_lock<std::mutex> g(mu_);
    done_.wait(g, [&] { return left_ == 0; });
}

void CopyTeam::stop() {
    {
        std::lock_guard<std::mutex> g(mu_);
        quit_ = true;
    }
    go_.notify_all();
    for (auto &t : th_) t.join();
    th_.clear();
    std::lock_guard<std::mutex> g(mu_);
    quit_ = false;
}

// ---- Stager ----

hipError_t Stager::ensure(size_t chunk) {
    if (chunk_ == chunk && slot_[0]) return hipSuccess;
    for (int s = 0; s < SLOTS; s++) {
        if (busy_[s]) (void)hipEventSynchronize(ev_[s]);
        busy_[s] = false;
        if (slot_[s]) (void)hipHostFree(slot_[s]);
        slot_[s] = nullptr;
    }
    chunk_ = 0;
    for (int s = 0; s < SLOTS; s++) {
        if (!ev_[s]) {
            hipError_t e = hipEventCreateWithFlags(&ev_[s], hipEventDisableTiming);
            if (e != hipSuccess) return e;
        }
        hipError_t e = hipHostMalloc(&slot_[s], chunk, hipHostMallocDefault);
        if (e != hipSuccess) {
            slot_[s] = nullptr;
            return e;
        }
    }
    chunk_ = chunk;
    return hipSuccess;
}

hipError_t Stager::copy(void *dst, const void *src, size_t bytes, hipStream_t st) {
    const hipMemcpyKind kind = up_ ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    if (bytes == 0) return hipSuccess;
    const StageConfig cfg = stage_config();
    std::lock_guard<std::mutex> g(mu_);
    if (cfg.threads <= 0 || bytes < cfg.min_bytes || cfg.chunk == 0 || ensure(cfg.chunk) != hipSuccess) {
        return hipMemcpyAsync(dst, src, bytes, kind, st);
    }
    const size_t ch = chunk_;
    const size_t n = (bytes + ch - 1) / ch;
    auto len = [&](size_t k) { return std::min(ch, bytes - k * ch); };
    if (up_) {
        for (size_t k = 0; k < n; k++) {
            const int s = int(k % SLOTS);
            if (busy_[s]) {
                hipError_t e = hipEventSynchronize(ev_[s]);
                if (e != hipSuccess) return e;
                busy_[s] = false;
            }
            team_.copy(slot_[s], (const char *)src + k * ch, len(k), cfg.threads);
            hipError_t e = hipMemcpyAsync((char *)dst + k * ch, slot_[s], len(k), kind, st);
            if (e == hipSuccess) e = hipEventRecord(ev_[s], st);
            if (e != hipSuccess) return e;
            busy_[s] = true;
        }
        return hipSuccess;
    }
    auto issue = [&](size_t k) {
        const int s = int(k % SLOTS);
        hipError_t e = hipMemcpyAsync(slot_[s], (const char *)src + k * ch, len(k), kind, st);
        if (e == hipSuccess) e = hipEventRecord(ev_[s], st);
        busy_[s] = e == hipSuccess;
        return e;
    };
    hipError_t err = hipSuccess;
    for (size_t k = 0; k < std::min(n, size_t(SLOTS)) && err == hipSuccess; k++) err = issue(k);
    for (size_t k = 0; k < n && err == hipSuccess; k++) {
        const int s = int(k % SLOTS);
        err = hipEventSynchronize(ev_[s]);
        busy_[s] = false;
        if (err != hipSuccess) break;
        team_.copy((char *)dst + k * ch, slot_[s], len(k), cfg.threads);
        if (k + SLOTS < n) err = issue(k + SLOTS);
    }
    // leave no transfer in flight into a slot after an error
    for (int s = 0; s < SLOTS; s++)
        if (busy_[s]) {
            (void)hipEventSynchronize(ev_[s]);
            busy_[s] = false;
        }
    return err;
}

void Stager::release() {
    std::lock_guard<std::mutex> g(mu_);
    for (int s = 0; s < SLOTS; s++) {
        if (busy_[s]) (void)hipEventSynchronize(ev_[s]);
        busy_[s] = false;
        if (slot_[s]) (void)hipHostFree(slot_[s]);
        slot_[s] = nullptr;
        if (ev_[s]) (void)hipEventDestroy(ev_[s]);
        ev_[s] = nullptr;
    }
    chunk_ = 0;
    team_.stop();
}

}  // namespace gnt
