// Multi-threaded forms of the whole-number operations the final assembly and
// the decimal conversion spend their host time in, on numbers of millions of
// limbs: shifts, truncation to low bits, signed sums and differences, a
// multiply by one word, and powers of two.  GMP runs each of these on one
// thread at memory speed; here the limb range is cut into one contiguous
// chunk per thread and the carries or borrows that cross chunk boundaries are
// settled afterwards, chunk by chunk.  Values below cfg.min_limbs go to GMP.
// Results are built in a fresh number and swapped in, so the result may be
// one of the operands.
#pragma once

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "gmp_abi.hpp"

namespace par {

struct Config {
    int threads = 1;
    size_t min_limbs = size_t(1) << 17;  // below this, GMP on the calling thread
};
inline Config cfg;
// Threads available to the calling thread (0: cfg.threads); parallel callers
// give each of their workers a share.
inline thread_local int tl_threads = 0;

inline size_t nl(const gmp_mpz *x) { return (size_t)(x->size < 0 ? -x->size : x->size); }

inline int threads_here() { return tl_threads > 0 ? tl_threads : cfg.threads; }

inline bool small(size_t n) { return n < cfg.min_limbs || threads_here() <= 1; }

// f(t, lo, hi) for the chunks of [0, n), one per thread (the caller runs
// chunk 0); returns the chunk boundaries.
template <class F>
std::vector<size_t> chunks(size_t n, F f) {
    const int T = std::max(1, std::min<int>(threads_here(), (int)((n + 4095) / 4096)));
    std::vector<size_t> b(T + 1);
    for (int t = 0; t < T; t++) b[t] = std::min(n, ((n * (size_t)t / (size_t)T) + 7) & ~size_t(7));
    b[T] = n;
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back([&, t] { f(t, b[t], b[t + 1]); });
    f(0, b[0], b[1]);
    for (auto &x : th) x.join();
    return b;
}

inline void reserve(gmp_mpz *r, size_t limbs) {
    if ((size_t)r->alloc < limbs) __gmpz_realloc2(r, (unsigned long)limbs * 64);
}

inline void set_size(gmp_mpz *r, size_t n, bool neg) {
    while (n > 0 && r->d[n - 1] == 0) n--;
    r->size = neg ? -(int)n : (int)n;
}

// r = x * 2^s; s < 0 truncates toward zero (GMP's mul_2exp / tdiv_q_2exp).
inline void shift(gmp_mpz *r, const gmp_mpz *x, long s) {
    const size_t n = nl(x);
    if (small(n)) {
        if (s >= 0) __gmpz_mul_2exp(r, x, (unsigned long)s);
        else __gmpz_tdiv_q_2exp(r, x, (unsigned long)(-s));
        return;
    }
    const bool neg = x->size < 0;
    const uint64_t *xp = x->d;
    Mpz t;
    if (s >= 0) {
        const size_t w = (size_t)s >> 6;
        const unsigned b = (unsigned)(s & 63);
        const size_t rn = n + w + 1;
        reserve(&t.z, rn);
        uint64_t *rp = t.z.d;
        chunks(rn, [&](int, size_t lo, size_t hi) {
            for (size_t i = lo; i < hi; i++) {
                if (i < w) {
                    rp[i] = 0;
                    continue;
                }
                const size_t j = i - w;  // limb of x landing here
                uint64_t v = j < n ? xp[j] << b : 0;
                if (b && j >= 1 && j - 1 < n) v |= xp[j - 1] >> (64 - b);
                rp[i] = v;
            }
        });
        set_size(&t.z, rn, neg);
    } else {
        const size_t sh = (size_t)(-s), w = sh >> 6;
        const unsigned b = (unsigned)(sh & 63);
        if (w >= n) {
            __gmpz_set_ui(r, 0);
            return;
        }
        const size_t rn = n - w;
        reserve(&t.z, rn);
        uint64_t *rp = t.z.d;
        chunks(rn, [&](int, size_t lo, size_t hi) {
            for (size_t i = lo; i < hi; i++) {
                uint64_t v = xp[w + i] >> b;
                if (b && w + i + 1 < n) v |= xp[w + i + 1] << (64 - b);
                rp[i] = v;
            }
        });
        set_size(&t.z, rn, neg);
    }
    __gmpz_swap(r, &t.z);
}

// r = x mod 2^k with the sign of x (GMP's tdiv_r_2exp).
inline void trunc(gmp_mpz *r, const gmp_mpz *x, unsigned long k) {
    const size_t n = nl(x);
    if (small(n)) {
        __gmpz_tdiv_r_2exp(r, x, k);
        return;
    }
    const size_t w = k >> 6;
    const unsigned b = (unsigned)(k & 63);
    const size_t rn = std::min(n, w + (b ? 1 : 0));
    Mpz t;
    reserve(&t.z, rn ? rn : 1);
    uint64_t *rp = t.z.d;
    const uint64_t *xp = x->d;
    chunks(rn, [&](int, size_t lo, size_t hi) { std::memcpy(rp + lo, xp + lo, (hi - lo) * 8); });
    if (b && rn == w + 1) rp[w] &= (uint64_t(1) << b) - 1;
    set_size(&t.z, rn, x->size < 0);
    __gmpz_swap(r, &t.z);
}

// r = 2^k.
inline void pow2(gmp_mpz *r, unsigned long k) {
    const size_t rn = (k >> 6) + 1;
    if (small(rn)) {
        __gmpz_set_ui(r, 1);
        __gmpz_mul_2exp(r, r, k);
        return;
    }
    Mpz t;
    reserve(&t.z, rn);
    uint64_t *rp = t.z.d;
    chunks(rn, [&](int, size_t lo, size_t hi) { std::memset(rp + lo, 0, (hi - lo) * 8); });
    rp[rn - 1] = uint64_t(1) << (k & 63);
    t.z.size = (int)rn;
    __gmpz_swap(r, &t.z);
}

// |a| > |b| (both limb arrays, an >= bn): rp[0..an) = |a| - |b|.
inline void mag_sub(uint64_t *rp, const uint64_t *ap, size_t an, const uint64_t *bp, size_t bn) {
    std::vector<unsigned char> bw(threads_here() + 1, 0);
    const std::vector<size_t> bnd = chunks(an, [&](int t, size_t lo, size_t hi) {
        unsigned char borrow = 0;
        for (size_t i = lo; i < hi; i++) {
            const uint64_t x = ap[i], y = i < bn ? bp[i] : 0;
            const uint64_t d = x - y;
            const unsigned char b1 = x < y;
            rp[i] = d - borrow;
            borrow = b1 | (unsigned char)(d < borrow);
        }
        bw[t] = borrow;
    });
    // a chunk's borrow-out and a borrow rippling through it never come together
    unsigned char borrow = 0;
    for (size_t t = 0; t + 1 < bnd.size(); t++) {
        unsigned char through = 0;
        if (borrow) {  // take the borrow into this chunk
            size_t i = bnd[t];
            for (; i < bnd[t + 1]; i++)
                if (rp[i]-- != 0) break;
            through = i == bnd[t + 1];
        }
        borrow = bw[t] | through;
    }
}

// |a| + |b| (an >= bn) into rp[0..an]; returns the limb count an or an + 1.
inline size_t mag_add(uint64_t *rp, const uint64_t *ap, size_t an, const uint64_t *bp, size_t bn) {
    std::vector<unsigned char> cy(threads_here() + 1, 0);
    const std::vector<size_t> bnd = chunks(an, [&](int t, size_t lo, size_t hi) {
        unsigned char carry = 0;
        for (size_t i = lo; i < hi; i++) {
            const uint64_t x = ap[i], y = i < bn ? bp[i] : 0;
            const uint64_t s = x + y;
            const unsigned char c1 = s < x;
            rp[i] = s + carry;
            carry = c1 | (unsigned char)(rp[i] < s);
        }
        cy[t] = carry;
    });
    unsigned char carry = 0;
    for (size_t t = 0; t + 1 < bnd.size(); t++) {
        unsigned char through = 0;
        if (carry) {  // add the carry into this chunk
            size_t i = bnd[t];
            for (; i < bnd[t + 1]; i++)
                if (++rp[i] != 0) break;
            through = i == bnd[t + 1];
        }
        carry = cy[t] | through;
    }
    if (carry) rp[an] = 1;
    return carry ? an + 1 : an;
}

// r = a + b, or a - b with sub (GMP's add / sub).
inline void addsub(gmp_mpz *r, const gmp_mpz *a, const gmp_mpz *b, bool sub) {
    size_t an = nl(a), bn = nl(b);
    if (small(std::max(an, bn))) {
        if (sub) __gmpz_sub(r, a, b);
        else __gmpz_add(r, a, b);
        return;
    }
    bool aneg = a->size < 0, bneg = (b->size < 0) != sub;
    const uint64_t *ap = a->d, *bp = b->d;
    Mpz t;
    if (aneg == bneg) {
        if (an < bn) {
            std::swap(an, bn);
            std::swap(ap, bp);
        }
        reserve(&t.z, an + 1);
        const size_t rn = mag_add(t.z.d, ap, an, bp, bn);
        set_size(&t.z, rn, aneg);
    } else {
        int cmp = an != bn ? (an > bn ? 1 : -1) : 0;
        for (size_t i = an; cmp == 0 && i-- > 0;)
            if (ap[i] != bp[i]) cmp = ap[i] > bp[i] ? 1 : -1;
        if (cmp == 0) {
            __gmpz_set_ui(r, 0);
            return;
        }
        if (cmp < 0) {
            std::swap(an, bn);
            std::swap(ap, bp);
            aneg = bneg;
        }
        reserve(&t.z, an);
        mag_sub(t.z.d, ap, an, bp, bn);
        set_size(&t.z, an, aneg);
    }
    __gmpz_swap(r, &t.z);
}

inline void add(gmp_mpz *r, const gmp_mpz *a, const gmp_mpz *b) { addsub(r, a, b, false); }
inline void sub(gmp_mpz *r, const gmp_mpz *a, const gmp_mpz *b) { addsub(r, a, b, true); }

// r = x * m (GMP's mul_ui).
inline void mul_ui(gmp_mpz *r, const gmp_mpz *x, unsigned long m) {
    const size_t n = nl(x);
    if (small(n) || m == 0) {
        __gmpz_mul_ui(r, x, m);
        return;
    }
    const uint64_t *xp = x->d;
    Mpz t;
    reserve(&t.z, n + 1);
    uint64_t *rp = t.z.d;
    std::vector<uint64_t> hi_limb(threads_here() + 1, 0);
    const std::vector<size_t> bnd = chunks(n, [&](int c, size_t lo, size_t hi) {
        uint64_t carry = 0;
        for (size_t i = lo; i < hi; i++) {
            const unsigned __int128 p = (unsigned __int128)xp[i] * m + carry;
            rp[i] = (uint64_t)p;
            carry = (uint64_t)(p >> 64);
        }
        hi_limb[c] = carry;
    });
    uint64_t carry = 0;
    for (size_t c = 0; c + 1 < bnd.size(); c++) {
        uint64_t through = 0;
        if (carry) {  // add the previous chunk's top limb into this chunk
            size_t i = bnd[c];
            for (; i < bnd[c + 1] && carry; i++) {
                rp[i] += carry;
                carry = rp[i] < carry;
            }
            through = carry;
        }
        carry = hi_limb[c] + through;  // < m + 1: no overflow
    }
    rp[n] = carry;
    set_size(&t.z, n + 1, x->size < 0);
    __gmpz_swap(r, &t.z);
}

}  // namespace par
