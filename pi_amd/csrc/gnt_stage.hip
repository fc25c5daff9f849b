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

// Below this many bytes a chunk is copied by the calling thread alone.
static constexpr size_t TEAM_MIN_BYTES = size_t(256) << 10;

// ---- CopyTeam ----

void CopyTeam::worker(int idx, uint64_t seen) {
    for (;;) {
        const std::function<void(int, int)> *fn;
        int parts;
        {
            std::unique_lock<std::mutex> g(mu_);
            go_.wait(g, [&] { return quit_ || gen_ != seen; });
            if (quit_) return;
            seen = gen_;
            fn = fn_;
            parts = parts_;
        }
        if (idx + 1 < parts) (*fn)(idx + 1, parts);
        std::lock_guard<std::mutex> g(mu_);
        if (--left_ == 0) done_.notify_one();
    }
}

void CopyTeam::run(int threads, const std::function<void(int, int)> &fn) {
    const int parts = std::max(1, threads);
    if (parts == 1) {
        fn(0, 1);
        return;
    }
    // a new helper starts at the current generation, so it takes only jobs
    // posted after it exists (run() is called under the stager's lock)
    while ((int)th_.size() < parts - 1) th_.emplace_back(&CopyTeam::worker, this, (int)th_.size(), gen_);
    {
        std::lock_guard<std::mutex> g(mu_);
        fn_ = &fn;
        parts_ = parts;
        left_ = (int)th_.size();
        gen_++;
    }
    go_.notify_all();
    fn(0, parts);
    std::unique_lock<std::mutex> g(mu_);
    done_.wait(g, [&] { return left_ == 0; });
}

// [lo, hi) of [0, n) for part p of parts, on 64-byte boundaries
static void part_range(size_t n, int p, int parts, size_t &lo, size_t &hi) {
    size_t slice = (n + parts - 1) / parts;
    slice = (slice + 63) & ~size_t(63);
    lo = std::min(n, slice * size_t(p));
    hi = std::min(n, lo + slice);
}

void CopyTeam::copy(void *dst, const void *src, size_t bytes, int threads) {
    if (threads <= 1 || bytes < TEAM_MIN_BYTES) {
        std::memcpy(dst, src, bytes);
        return;
    }
    run(threads, [&](int p, int parts) {
        size_t lo, hi;
        part_range(bytes, p, parts, lo, hi);
        if (hi > lo) std::memcpy((char *)dst + lo, (const char *)src + lo, hi - lo);
    });
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

hipError_t Stager::direct(void *dst, const void *src, size_t bytes, hipStream_t st) {
    return hipMemcpyAsync(dst, src, bytes, up_ ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost, st);
}

hipError_t Stager::up_chunks(void *dev, size_t total, int threads,
                             const std::function<void(char *, size_t, size_t)> &fill, hipStream_t st) {
    const size_t ch = chunk_;
    for (size_t lo = 0, k = 0; lo < total; lo += ch, k++) {
        const size_t hi = std::min(total, lo + ch);
        const int s = int(k % SLOTS);
        if (busy_[s]) {
            hipError_t e = hipEventSynchronize(ev_[s]);
            if (e != hipSuccess) return e;
            busy_[s] = false;
        }
        char *slot = (char *)slot_[s];
        if (hi - lo < TEAM_MIN_BYTES) fill(slot, lo, hi);
        else
            team_.run(threads, [&](int p, int parts) {
                size_t a, b;
                part_range(hi - lo, p, parts, a, b);
                if (b > a) fill(slot + a, lo + a, lo + b);
            });
        hipError_t e = hipMemcpyAsync((char *)dev + lo, slot, hi - lo, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipEventRecord(ev_[s], st);
        if (e != hipSuccess) return e;
        busy_[s] = true;
    }
    return hipSuccess;
}

hipError_t Stager::down_chunks(const void *dev, size_t total, int threads,
                               const std::function<void(const char *, size_t, size_t)> &drain, hipStream_t st) {
    const size_t ch = chunk_;
    const size_t n = (total + ch - 1) / ch;
    auto issue = [&](size_t k) {
        const int s = int(k % SLOTS);
        const size_t lo = k * ch, hi = std::min(total, lo + ch);
        hipError_t e = hipMemcpyAsync(slot_[s], (const char *)dev + lo, hi - lo, hipMemcpyDeviceToHost, st);
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
        const size_t lo = k * ch, hi = std::min(total, lo + ch);
        const char *slot = (const char *)slot_[s];
        if (hi - lo < TEAM_MIN_BYTES) drain(slot, lo, hi);
        else
            team_.run(threads, [&](int p, int parts) {
                size_t a, b;
                part_range(hi - lo, p, parts, a, b);
                if (b > a) drain(slot + a, lo + a, lo + b);
            });
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

hipError_t Stager::copy(void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return hipSuccess;
    const StageConfig cfg = stage_config();
    std::lock_guard<std::mutex> g(mu_);
    if (cfg.threads <= 0 || bytes < cfg.min_bytes || cfg.chunk == 0 || ensure(cfg.chunk) != hipSuccess)
        return direct(dst, src, bytes, st);
    if (up_)
        return up_chunks(dst, bytes, cfg.threads,
                         [&](char *slot, size_t lo, size_t hi) { std::memcpy(slot, (const char *)src + lo, hi - lo); },
                         st);
    return down_chunks(src, bytes, cfg.threads,
                       [&](const char *slot, size_t lo, size_t hi) { std::memcpy((char *)dst + lo, slot, hi - lo); },
                       st);
}

// The pieces of segs inside [lo, hi) of the device range, between the range
// and buf (which holds [lo, hi)).
static void seg_copy(const std::vector<Seg> &segs, char *buf, size_t lo, size_t hi, bool to_buf) {
    auto it = std::partition_point(segs.begin(), segs.end(), [&](const Seg &x) { return x.off + x.bytes <= lo; });
    for (; it != segs.end() && it->off < hi; ++it) {
        const size_t a = std::max(lo, it->off), b = std::min(hi, it->off + it->bytes);
        if (a >= b) continue;
        if (to_buf) std::memcpy(buf + (a - lo), (const char *)it->host + (a - it->off), b - a);
        else std::memcpy((char *)it->host + (a - it->off), buf + (a - lo), b - a);
    }
}

hipError_t Stager::gather(void *dev, size_t total, const std::vector<Seg> &segs, hipStream_t st) {
    if (!up_) return hipErrorInvalidValue;
    if (segs.empty() || total == 0) return hipSuccess;
    const StageConfig cfg = stage_config();
    std::lock_guard<std::mutex> g(mu_);
    if (cfg.threads <= 0 || cfg.chunk == 0 || ensure(cfg.chunk) != hipSuccess) {
        for (const Seg &x : segs)
            if (x.bytes)
                if (hipError_t e = direct((char *)dev + x.off, x.host, x.bytes, st)) return e;
        return hipSuccess;
    }
    return up_chunks(dev, total, cfg.threads,
                     [&](char *slot, size_t lo, size_t hi) { seg_copy(segs, slot, lo, hi, true); }, st);
}

hipError_t Stager::scatter(const void *dev, size_t total, const std::vector<Seg> &segs, hipStream_t st) {
    if (up_) return hipErrorInvalidValue;
    if (segs.empty() || total == 0) return hipSuccess;
    const StageConfig cfg = stage_config();
    std::lock_guard<std::mutex> g(mu_);
    if (cfg.threads <= 0 || cfg.chunk == 0 || ensure(cfg.chunk) != hipSuccess) {
        for (const Seg &x : segs)
            if (x.bytes)
                if (hipError_t e = direct(x.host, (const char *)dev + x.off, x.bytes, st)) return e;
        return hipSuccess;
    }
    return down_chunks(dev, total, cfg.threads,
                       [&](const char *slot, size_t lo, size_t hi) { seg_copy(segs, (char *)slot, lo, hi, false); },
                       st);
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
