// Host <-> device copies of pageable host memory through pinned staging.
//
// The runtime moves pageable memory through its own staging on one thread,
// at about half the link rate (host to device) or less (device to host).  A
// Stager splits a copy into chunks: a team of host threads copies chunk k
// between the caller's memory and a pinned slot while the DMA engine moves
// chunk k-1 (or k+1) over the link, so each side runs at its own rate and
// the link only sees pinned transfers.  Copies below a size threshold, and
// every copy when staging is off, go straight to hipMemcpyAsync.
#pragma once
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace gnt {

// Process-wide staging settings (gnt_set_staging).  threads = host threads
// that copy a chunk (the calling thread included); 0 turns staging off.
struct StageConfig {
    int threads = 4;
    size_t chunk = size_t(8) << 20;
    size_t min_bytes = size_t(4) << 20;
};
StageConfig stage_config();
void set_stage_config(const StageConfig &s);

// A few host threads that share one job; the calling thread takes part 0.
class CopyTeam {
  public:
    ~CopyTeam() { stop(); }
    // fn(part, parts) for every part in [0, parts), parts = max(threads, 1)
    void run(int threads, const std::function<void(int, int)> &fn);
    // memcpy in slices of whole cache lines
    void copy(void *dst, const void *src, size_t bytes, int threads);
    void stop();

  private:
    void worker(int idx, uint64_t seen);
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable go_, done_;
    const std::function<void(int, int)> *fn_ = nullptr;
    int parts_ = 0;
    uint64_t gen_ = 0;
    int left_ = 0;
    bool quit_ = false;
};

// One piece of a packed transfer: bytes at offset off of a device range, and
// the host memory they come from (to the device) or go to (to the host).
struct Seg {
    size_t off;
    void *host;
    size_t bytes;
};

class Stager {
  public:
    explicit Stager(bool to_device) : up_(to_device) {}
    ~Stager() { release(); }
    // Ordered on st like hipMemcpyAsync, and like it complete once st is
    // synchronised.  A staged copy to the device returns once the caller's
    // memory has been read (its last transfers may still be in flight); a
    // staged copy to the host returns with dst already written.
    hipError_t copy(void *dst, const void *src, size_t bytes, hipStream_t st);
    // Many pieces of one device range [dev, dev + total) in one pass: the
    // pieces (sorted by offset, not overlapping) are packed into the pinned
    // chunks, and each chunk crosses the link as one transfer; bytes between
    // pieces are left undefined on the device and ignored on the host.  The
    // same ordering and completion rules as copy().
    hipError_t gather(void *dev, size_t total, const std::vector<Seg> &segs, hipStream_t st);
    hipError_t scatter(const void *dev, size_t total, const std::vector<Seg> &segs, hipStream_t st);
    void release();  // frees the pinned slots and stops the team

  private:
    static constexpr int SLOTS = 4;
    hipError_t ensure(size_t chunk);
    hipError_t direct(void *dst, const void *src, size_t bytes, hipStream_t st);
    // pipelined chunks of [0, total): fill(slot, lo, hi) packs a chunk to
    // send, drain(slot, lo, hi) unpacks a received one
    hipError_t up_chunks(void *dev, size_t total, int threads,
                         const std::function<void(char *, size_t, size_t)> &fill, hipStream_t st);
    hipError_t down_chunks(const void *dev, size_t total, int threads,
                           const std::function<void(const char *, size_t, size_t)> &drain, hipStream_t st);
    const bool up_;
    std::mutex mu_;
    CopyTeam team_;
    size_t chunk_ = 0;
    void *slot_[SLOTS] = {};
    hipEvent_t ev_[SLOTS] = {};
    bool busy_[SLOTS] = {};
};

}  // namespace gnt
