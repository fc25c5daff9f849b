// Declarations of the subset of GMP's documented integer API used by the
// host driver (the image ships libgmp.so.10 without its development header).
// Names are the library's exported symbols behind the documented mpz_* /
// mpn_* macros; the struct is the documented mpz_t layout on LP64 with 64-bit
// limbs.
#pragma once

#include <stddef.h>
#include <stdint.h>

extern "C" {
typedef struct {
    int alloc;
    int size;  // |size| limbs in use; the sign of the value is the sign of size
    uint64_t *d;
} gmp_mpz;

void __gmpz_init(gmp_mpz *);
void __gmpz_init2(gmp_mpz *, unsigned long bits);
void __gmpz_clear(gmp_mpz *);
void *__gmpz_realloc2(gmp_mpz *, unsigned long bits);
void __gmpz_set(gmp_mpz *, const gmp_mpz *);
void __gmpz_set_ui(gmp_mpz *, unsigned long);
void __gmpz_set_si(gmp_mpz *, long);
void __gmpz_swap(gmp_mpz *, gmp_mpz *);
void __gmpz_add(gmp_mpz *, const gmp_mpz *, const gmp_mpz *);
void __gmpz_sub(gmp_mpz *, const gmp_mpz *, const gmp_mpz *);
void __gmpz_mul(gmp_mpz *, const gmp_mpz *, const gmp_mpz *);
void __gmpz_mul_ui(gmp_mpz *, const gmp_mpz *, unsigned long);
void __gmpz_mul_si(gmp_mpz *, const gmp_mpz *, long);
void __gmpz_neg(gmp_mpz *, const gmp_mpz *);
void __gmpz_mul_2exp(gmp_mpz *, const gmp_mpz *, unsigned long);
void __gmpz_tdiv_q(gmp_mpz *, const gmp_mpz *, const gmp_mpz *);
void __gmpz_tdiv_q_2exp(gmp_mpz *, const gmp_mpz *, unsigned long);
void __gmpz_tdiv_r_2exp(gmp_mpz *, const gmp_mpz *, unsigned long);
void __gmpz_sqrt(gmp_mpz *, const gmp_mpz *);
void __gmpz_ui_pow_ui(gmp_mpz *, unsigned long, unsigned long);
int __gmpz_cmp(const gmp_mpz *, const gmp_mpz *);
size_t __gmpz_sizeinbase(const gmp_mpz *, int);
char *__gmpz_get_str(char *, int, const gmp_mpz *);
uint64_t __gmpn_mul(uint64_t *rp, const uint64_t *ap, long an, const uint64_t *bp, long bn);
}

// RAII holder.
struct Mpz {
    gmp_mpz z;
    Mpz() { __gmpz_init(&z); }
    ~Mpz() { __gmpz_clear(&z); }
    Mpz(const Mpz &) = delete;
    Mpz &operator=(const Mpz &) = delete;
    Mpz(Mpz &&o) noexcept {
        __gmpz_init(&z);
        __gmpz_swap(&z, &o.z);
    }
    Mpz &operator=(Mpz &&o) noexcept {
        __gmpz_swap(&z, &o.z);
        return *this;
    }
    gmp_mpz *operator&() { return &z; }
    const gmp_mpz *operator&() const { return &z; }
    size_t limbs() const { return (size_t)(z.size < 0 ? -z.size : z.size); }
    int sign() const { return z.size < 0 ? -1 : (z.size > 0 ? 1 : 0); }
    const uint64_t *data() const { return z.d; }
    // Set to the magnitude held in n little-endian limbs (trimmed by the caller or not).
    void set_limbs(const uint64_t *v, size_t n) {
        while (n > 0 && v[n - 1] == 0) n--;
        __gmpz_realloc2(&z, (unsigned long)(n ? n : 1) * 64);
        for (size_t i = 0; i < n; i++) z.d[i] = v[i];
        z.size = (int)n;
    }
};
