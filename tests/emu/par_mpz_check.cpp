// Brute-force check of the multi-threaded whole-number helpers
// (pi_amd/csrc/par_mpz.hpp) against GMP on the same inputs: shifts both ways,
// truncation, powers of two, signed sums and differences (with aliasing,
// cancellation and all-ones / zero limbs), multiply by a word.  Seven threads
// and no size floor, so numbers above 4096 limbs take the chunked paths.
#include "../../pi_amd/csrc/par_mpz.hpp"
#include <cstdio>
#include <random>
static std::mt19937_64 rng(7);  // fixed inputs
static void rnd(Mpz &z, size_t n, bool neg, int kind) {
    __gmpz_realloc2(&z, n * 64 + 64);
    for (size_t i = 0; i < n; i++) z.z.d[i] = kind == 1 ? ~0ull : (kind == 2 ? 0 : rng());
    if (n) z.z.d[n - 1] |= 1;
    z.z.size = neg ? -(int)n : (int)n;
}
static int fails = 0;
static void eq(const Mpz &a, const Mpz &b, const char *what, size_t n) {
    if (__gmpz_cmp(&a, &b) != 0) { fails++; if (fails < 20) printf("FAIL %s n=%zu\n", what, n); }
}
int main() {
    par::cfg.threads = 7;
    par::cfg.min_limbs = 1;
    size_t sizes[] = {1, 2, 7, 8, 9, 63, 64, 65, 1000, 4095, 4096, 4097, 20000, 100003};
    for (size_t n : sizes) for (int kind = 0; kind < 3; kind++) for (int sg = 0; sg < 4; sg++) {
        Mpz a, b, r1, r2;
        rnd(a, n, sg & 1, kind);
        size_t m = n > 3 ? n - rng() % (n / 2 + 1) : n;
        rnd(b, m, sg & 2, kind == 2 ? 1 : kind);
        for (long s : {0L, 1L, 63L, 64L, 65L, 1000L, -1L, -63L, -64L, -65L, -1000L, -(long)(64 * n), -(long)(64 * n + 5)}) {
            par::shift(&r1, &a, s);
            if (s >= 0) __gmpz_mul_2exp(&r2, &a, s); else __gmpz_tdiv_q_2exp(&r2, &a, -s);
            eq(r1, r2, "shift", n);
        }
        for (unsigned long k : {0UL, 1UL, 64UL, 65UL, 6400UL, 64UL * n + 3}) {
            par::trunc(&r1, &a, k); __gmpz_tdiv_r_2exp(&r2, &a, k); eq(r1, r2, "trunc", n);
            par::pow2(&r1, k); __gmpz_set_ui(&r2, 1); __gmpz_mul_2exp(&r2, &r2, k); eq(r1, r2, "pow2", n);
        }
        par::add(&r1, &a, &b); __gmpz_add(&r2, &a, &b); eq(r1, r2, "add", n);
        par::add(&r1, &b, &a); __gmpz_add(&r2, &b, &a); eq(r1, r2, "add2", n);
        par::sub(&r1, &a, &b); __gmpz_sub(&r2, &a, &b); eq(r1, r2, "sub", n);
        par::sub(&r1, &b, &a); __gmpz_sub(&r2, &b, &a); eq(r1, r2, "sub2", n);
        par::sub(&r1, &a, &a); __gmpz_sub(&r2, &a, &a); eq(r1, r2, "subself", n);
        // aliasing
        Mpz c; __gmpz_set(&c, &a); par::add(&c, &c, &b); __gmpz_add(&r2, &a, &b); eq(c, r2, "alias add", n);
        __gmpz_set(&c, &a); par::sub(&c, &c, &b); __gmpz_sub(&r2, &a, &b); eq(c, r2, "alias sub", n);
        __gmpz_set(&c, &a); par::shift(&c, &c, 77); __gmpz_mul_2exp(&r2, &a, 77); eq(c, r2, "alias shl", n);
        __gmpz_set(&c, &a); par::shift(&c, &c, -77); __gmpz_tdiv_q_2exp(&r2, &a, 77); eq(c, r2, "alias shr", n);
        for (unsigned long mm : {1UL, 2UL, 10005UL, 426880UL, ~0UL}) {
            par::mul_ui(&r1, &a, mm); __gmpz_mul_ui(&r2, &a, mm); eq(r1, r2, "mul_ui", n);
        }
        // a - b where they share long equal top parts (cancellation), and 2^k - t
        Mpz p2, t; __gmpz_set_ui(&p2, 1); __gmpz_mul_2exp(&p2, &p2, 64 * n); __gmpz_sub(&t, &p2, &b);
        par::sub(&r1, &p2, &t); __gmpz_sub(&r2, &p2, &t); eq(r1, r2, "pow-sub", n);
        __gmpz_add(&t, &p2, &b);
        par::sub(&r1, &p2, &t); __gmpz_sub(&r2, &p2, &t); eq(r1, r2, "pow-sub2", n);
    }
    printf("fails %d\n", fails);
    return fails != 0;
}
