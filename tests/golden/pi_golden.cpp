// Independent pi digits for the end-to-end fixtures (test infrastructure).
//
// Two algorithms on GMP integers, neither sharing code with the driver under
// test:
//   chudnovsky      the same series, summed by a plain recursive split and
//                   assembled in decimal fixed point:
//                   floor(426880 sqrt(10005 * 10^(2E)) Q / (A Q + T)), E = D-1+G
//   gauss-legendre  Brent-Salamin AGM in binary fixed point (the reference's
//                   cross-check algorithm, gauss_legendre.rs:61-166)
// Output: "3." + (D-1) digits + "\n" (output.rs:33-61), digits truncated
// (util.rs:48-56).  G guard digits are dropped after checking that they do
// not sit next to a carry into the kept digits.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../pi_amd/csrc/gmp_abi.hpp"

namespace {

void split(unsigned long a, unsigned long b, gmp_mpz *P, gmp_mpz *Q, gmp_mpz *T) {
    if (b - a == 1) {
        // p = -(6a-5)(2a-1)(6a-1); q = a^3 640320^3 / 24; t = p (13591409 + 545140134 a)
        __gmpz_set_si(P, -(long)(6 * a - 5));
        __gmpz_mul_ui(P, P, 2 * a - 1);
        __gmpz_mul_ui(P, P, 6 * a - 1);
        __gmpz_set_ui(Q, 640320ul * 640320ul);
        __gmpz_mul_ui(Q, Q, 640320ul / 24ul);
        __gmpz_mul_ui(Q, Q, a);
        __gmpz_mul_ui(Q, Q, a);
        __gmpz_mul_ui(Q, Q, a);
        __gmpz_mul_ui(T, P, 13591409ul + 545140134ul * a);
        return;
    }
    unsigned long m = a + (b - a) / 2;
    Mpz P2, Q2, T2;
    split(a, m, P, Q, T);
    split(m, b, &P2, &Q2, &T2);
    __gmpz_mul(T, T, &Q2);
    __gmpz_mul(&T2, &T2, P);
    __gmpz_add(T, T, &T2);
    __gmpz_mul(P, P, &P2);
    __gmpz_mul(Q, Q, &Q2);
}

// floor(pi * 10^E) by the Chudnovsky series in decimal fixed point.
void pi_chudnovsky(unsigned long E, gmp_mpz *out) {
    const unsigned long terms = (unsigned long)((double)E / 14.18) + 4;
    Mpz P, Q, T;
    split(1, terms + 1, &P, &Q, &T);
    // denominator A Q + T
    Mpz den;
    __gmpz_mul_ui(&den, &Q, 13591409ul);
    __gmpz_add(&den, &den, &T);
    // 426880 sqrt(10005 * 10^(2E)) Q
    Mpz r;
    __gmpz_ui_pow_ui(&r, 10, 2 * E);
    __gmpz_mul_ui(&r, &r, 10005);
    __gmpz_sqrt(&r, &r);
    __gmpz_mul(&r, &r, &Q);
    __gmpz_mul_ui(&r, &r, 426880);
    __gmpz_tdiv_q(out, &r, &den);
}

// floor(pi * 2^F) by the Gauss-Legendre AGM in binary fixed point.
void pi_gauss_legendre(unsigned long F, gmp_mpz *out) {
    Mpz a, b, t, x, y;
    __gmpz_set_ui(&a, 1);
    __gmpz_mul_2exp(&a, &a, F);  // a = 1
    __gmpz_set_ui(&b, 1);
    __gmpz_mul_2exp(&b, &b, 2 * F - 1);
    __gmpz_sqrt(&b, &b);  // b = 1/sqrt(2)
    __gmpz_set_ui(&t, 1);
    __gmpz_mul_2exp(&t, &t, F - 2);  // t = 1/4
    unsigned long p = 0;             // p = 2^p
    const int iters = (int)std::ceil(std::log2((double)F)) + 2;
    for (int i = 0; i < iters; i++) {
        __gmpz_add(&x, &a, &b);
        __gmpz_tdiv_q_2exp(&x, &x, 1);  // a' = (a + b) / 2
        __gmpz_mul(&y, &a, &b);
        __gmpz_sqrt(&b, &y);            // b' = sqrt(a b)
        __gmpz_sub(&y, &a, &x);
        __gmpz_mul(&y, &y, &y);
        __gmpz_tdiv_q_2exp(&y, &y, F);
        __gmpz_mul_2exp(&y, &y, p);     // p (a - a')^2
        __gmpz_sub(&t, &t, &y);
        __gmpz_set(&a, &x);
        p++;
    }
    __gmpz_add(&x, &a, &b);
    __gmpz_mul(&x, &x, &x);             // (a + b)^2, scale 2^(2F)
    __gmpz_mul_2exp(&t, &t, 2);         // 4 t, scale 2^F
    __gmpz_tdiv_q(out, &x, &t);         // pi, scale 2^F
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: pi_golden chudnovsky|gauss-legendre DIGITS OUT\n");
        return 2;
    }
    const std::string algo = argv[1];
    const unsigned long D = std::strtoul(argv[2], nullptr, 10);
    const unsigned long G = 40;  // guard digits
    if (D < 1) return 2;
    Mpz v;
    if (algo == "chudnovsky") {
        pi_chudnovsky(D - 1 + G, &v);
    } else if (algo == "gauss-legendre") {
        const unsigned long F = (unsigned long)((double)(D - 1 + G) * 3.3219280948873623) + 64;
        Mpz p, s;
        pi_gauss_legendre(F, &p);
        __gmpz_ui_pow_ui(&s, 10, D - 1 + G);
        __gmpz_mul(&v, &p, &s);
        __gmpz_tdiv_q_2exp(&v, &v, F);
    } else {
        return 2;
    }
    char *s = __gmpz_get_str(nullptr, 10, &v);
    const size_t n = std::strlen(s);
    if (n != D + G) {
        std::fprintf(stderr, "unexpected length %zu\n", n);
        return 1;
    }
    // the kept digits are settled unless the guard digits are all 9s or all 0s
    bool all9 = true, all0 = true;
    for (size_t i = D; i < n; i++) {
        all9 &= s[i] == '9';
        all0 &= s[i] == '0';
    }
    if (all9 || all0) {
        std::fprintf(stderr, "guard digits undecided\n");
        return 1;
    }
    FILE *f = std::fopen(argv[3], "wb");
    if (!f) return 1;
    std::fputc(s[0], f);
    if (D > 1) {
        std::fputc('.', f);
        std::fwrite(s + 1, 1, D - 1, f);
    }
    std::fputc('\n', f);
    std::fclose(f);
    std::free(s);
    return 0;
}
