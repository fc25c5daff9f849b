// Host-side multiply dispatch with the GPU product as its transform leaf.
//
// gnt_mul_mag restates bignum's mul_mag / mul_mag_dispatch
// (crates/bignum/src/integer.rs:445-486): schoolbook below the Karatsuba
// threshold, Karatsuba in between, the transform product when it fits and the
// smaller operand reaches the NTT threshold, and Karatsuba above the transform
// cap so that oversized products become several transform leaves
// (integer.rs:433-444, 552-608).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/gnt.h"
#include "gnt_kernels.hpp"
#include "gnt_runtime.hpp"

typedef unsigned __int128 u128;
using u64 = uint64_t;

namespace gnt {
namespace {

// integer.rs:492-518
std::vector<u64> schoolbook(const u64 *a, size_t na, const u64 *b, size_t nb) {
    std::vector<u64> out(na + nb, 0);
    for (size_t i = 0; i < na; i++) {
        u128 carry = 0;
        u128 av = a[i];
        for (size_t j = 0; j < nb; j++) {
            u128 cur = (u128)out[i + j] + av * b[j] + carry;
            out[i + j] = (u64)cur;
            carry = cur >> 64;
        }
        size_t k = i + nb;
        while (carry) {
            u128 cur = (u128)out[k] + carry;
            out[k] = (u64)cur;
            carry = cur >> 64;
            k++;
        }
    }
    while (!out.empty() && out.back() == 0) out.pop_back();
    return out;
}

// integer.rs:351-366 (length not normalised, as in the reference)
std::vector<u64> add_mag(const u64 *a, size_t na, const u64 *b, size_t nb) {
    const u64 *lg = na >= nb ? a : b, *sh = na >= nb ? b : a;
    size_t nl = na >= nb ? na : nb, ns = na >= nb ? nb : na;
    std::vector<u64> out;
    out.reserve(nl + 1);
    u128 carry = 0;
    for (size_t i = 0; i < nl; i++) {
        u128 s = (u128)lg[i] + (i < ns ? sh[i] : 0) + carry;
        out.push_back((u64)s);
        carry = s >> 64;
    }
    if (carry) out.push_back((u64)carry);
    return out;
}

// integer.rs:369-389, a >= b
std::vector<u64> sub_mag(const std::vector<u64> &a, const std::vector<u64> &b) {
    std::vector<u64> out(a.size());
    u64 borrow = 0;
    for (size_t i = 0; i < a.size(); i++) {
        u64 bi = i < b.size() ? b[i] : 0;
        u64 d1 = a[i] - bi;
        u64 b1 = a[i] < bi;
        u64 d2 = d1 - borrow;
        u64 b2 = d1 < borrow;
        out[i] = d2;
        borrow = b1 | b2;
    }
    while (!out.empty() && out.back() == 0) out.pop_back();
    return out;
}

// integer.rs:612-635
void add_at(std::vector<u64> &dest, const std::vector<u64> &src, size_t off) {
    if (src.empty()) return;
    if (dest.size() < off + src.size()) dest.resize(off + src.size(), 0);
    u128 carry = 0;
    for (size_t i = 0; i < src.size(); i++) {
        u128 s = (u128)dest[off + i] + src[i] + carry;
        dest[off + i] = (u64)s;
        carry = s >> 64;
    }
    size_t k = off + src.size();
    while (carry) {
        if (dest.size() <= k) dest.push_back(0);
        u128 s = (u128)dest[k] + carry;
        dest[k] = (u64)s;
        carry = s >> 64;
        k++;
    }
}

// Karatsuba's three sub-products run concurrently (z0 on the calling thread,
// z2 and z1 on two more) while the smaller operand has at least
// PAR_KARATSUBA limbs, as the reference forks them with rayon::join above
// parallel_karatsuba_threshold (integer.rs:565-593, config.rs:100), and at
// most PAR_DEPTH levels deep (3^PAR_DEPTH threads at most).  Concurrent GPU
// leaves are batched together by the serving loop when it runs; otherwise they
// take turns on the device while their host work (uploads, sums, schoolbook
// leaves) overlaps.
constexpr size_t PAR_KARATSUBA = 512;
constexpr int PAR_DEPTH = 3;

struct Dispatch {
    gnt_ctx *ctx;
    size_t max_coeffs;
    size_t ntt_thr;
    size_t kara_thr;
    std::atomic<int> err{GNT_OK};

    void fail(int rc) {
        int ok = GNT_OK;
        err.compare_exchange_strong(ok, rc);
    }

    std::vector<u64> gpu(const u64 *a, size_t na, const u64 *b, size_t nb) {
        std::vector<u64> out(na + nb);
        size_t len = 0;
        int rc = (a == b && na == nb) ? gnt_sqr(ctx, a, na, out.data(), out.size(), &len)
                                      : gnt_mul(ctx, a, na, b, nb, out.data(), out.size(), &len);
        if (rc) {
            fail(rc);
            return {};
        }
        out.resize(len);
        return out;
    }

    // integer.rs:552-608
    std::vector<u64> karatsuba(const u64 *a, size_t na, const u64 *b, size_t nb, int depth) {
        const size_t n = na > nb ? na : nb;
        const size_t m = n / 2;
        const size_t la = m < na ? m : na, lb = m < nb ? m : nb;
        const std::vector<u64> as = add_mag(a, la, a + la, na - la);
        const std::vector<u64> bs = add_mag(b, lb, b + lb, nb - lb);
        std::vector<u64> z0, z2, z1p;
        const size_t smaller = na < nb ? na : nb;
        if (smaller >= PAR_KARATSUBA && depth < PAR_DEPTH) {
            std::thread t2([&] { z2 = mul(a + la, na - la, b + lb, nb - lb, depth + 1); });
            std::thread t1([&] { z1p = mul(as.data(), as.size(), bs.data(), bs.size(), depth + 1); });
            z0 = mul(a, la, b, lb, depth + 1);
            t2.join();
            t1.join();
        } else {
            z0 = mul(a, la, b, lb, depth + 1);
            z2 = mul(a + la, na - la, b + lb, nb - lb, depth + 1);
            z1p = mul(as.data(), as.size(), bs.data(), bs.size(), depth + 1);
        }
        if (err.load()) return {};
        std::vector<u64> z02 = add_mag(z0.data(), z0.size(), z2.data(), z2.size());
        std::vector<u64> z1 = sub_mag(z1p, z02);
        std::vector<u64> out(na + nb + 1, 0);
        add_at(out, z0, 0);
        add_at(out, z1, m);
        add_at(out, z2, 2 * m);
        while (!out.empty() && out.back() == 0) out.pop_back();
        return out;
    }

    // integer.rs:460-486
    std::vector<u64> mul(const u64 *a, size_t na, const u64 *b, size_t nb, int depth = 0) {
        if (na == 0 || nb == 0 || err.load()) return {};
        size_t smaller = na < nb ? na : nb;
        bool fits = (na + nb) <= max_coeffs / 4;  // 4(na+nb) <= max_coeffs, without overflow
        if (fits && smaller >= ntt_thr) return gpu(a, na, b, nb);
        if (smaller < kara_thr) return schoolbook(a, na, b, nb);
        return karatsuba(a, na, b, nb, depth);
    }
};

}  // namespace

}  // namespace gnt

extern "C" int gnt_mul_mag_dispatch(gnt_ctx *ctx, const uint64_t *a, size_t na, const uint64_t *b, size_t nb,
                                    uint64_t *out, size_t out_cap, size_t *out_len, size_t ntt_max_coeffs,
                                    size_t ntt_threshold, size_t karatsuba_threshold) {
    if (!ctx || !out_len) return gnt::set_error(GNT_E_ARG, "null argument");
    if ((na && !a) || (nb && !b)) return gnt::set_error(GNT_E_ARG, "null operand");
    const size_t cap = (size_t)1 << gnt::MAX_LOG_N;
    if (karatsuba_threshold == 1) return gnt::set_error(GNT_E_ARG, "karatsuba_threshold must be 0 or >= 2");
    gnt::Dispatch d;
    d.ctx = ctx;
    d.max_coeffs = ntt_max_coeffs && ntt_max_coeffs < cap ? ntt_max_coeffs : cap;
    d.ntt_thr = ntt_threshold ? ntt_threshold : 8192;
    d.kara_thr = karatsuba_threshold ? karatsuba_threshold : 32;
    std::vector<u64> r = d.mul(a, na, b, nb);
    if (int rc = d.err.load()) return rc;
    if (r.size() > out_cap) return gnt::set_error(GNT_E_ARG, "out_cap too small");
    if (!r.empty()) std::memcpy(out, r.data(), r.size() * 8);
    *out_len = r.size();
    return GNT_OK;
}

extern "C" int gnt_mul_mag(gnt_ctx *ctx, const uint64_t *a, size_t na, const uint64_t *b, size_t nb,
                           uint64_t *out, size_t out_cap, size_t *out_len, size_t ntt_threshold,
                           size_t karatsuba_threshold) {
    return gnt_mul_mag_dispatch(ctx, a, na, b, nb, out, out_cap, out_len, 0, ntt_threshold, karatsuba_threshold);
}
