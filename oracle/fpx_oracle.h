/* fpx_oracle — CPU restatement of fastpoly's exact NTT-convolution path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU build is
 * compared against: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in fastpoly_amd/ links or calls it.
 *
 * Every routine cites the reference file:line it restates (paths relative to
 * the fastpoly repository root).  Arithmetic is plain scalar C on Montgomery
 * limbs (x * 2^32 mod p, one uint32_t per coefficient), exactly the
 * representation of include/fastpoly/modint.hpp:54-98.
 *
 * Pinning (how each function is checked, see tests/test_oracle.py):
 *   field ops ........ closed forms against unsigned __int128 / Python ints
 *   ntt forward ...... published ntt-forward FNV checksums (998244353 2^10..2^20,
 *                      aux primes 2^10..2^17) + brute-force bit-reversed DFT
 *   conv / crt ....... published conv checksums (998244353, 1e9+7) + O(n^2)
 *                      integer products
 *   inv/log/exp/sqrt . published checksums + coefficient recurrences + identities
 *   pow / divmod ..... published pow / div / mod / divmod checksums + products
 *                      (a^j a^k = a^(j+k), b q + r = a with deg r < deg b)
 */
#ifndef FPX_ORACLE_H
#define FPX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FPO_OK = 0,
  FPO_E_DOMAIN = 1, /* poly::domain_error (detail/poly/base.hpp:26-30) */
  FPO_E_SIZE = 2,   /* ntt_size_error / std::length_error */
  FPO_E_NOMEM = 3,  /* std::bad_alloc */
  FPO_E_ARG = 4     /* invalid argument to the oracle itself */
};

typedef struct {
  uint32_t p;     /* prime, 2 <= p < 2^31 */
  uint32_t ninv;  /* -p^-1 mod 2^32 (modint.hpp:66-67) */
  uint32_t one;   /* 2^32 mod p, Montgomery one (modint.hpp:69-70) */
  uint32_t r2;    /* one^2 mod p (modint.hpp:72-73) */
  uint32_t root;  /* primitive root used for NTT plans (modint.hpp:65) */
  int maxlog;     /* v2(p-1): largest NTT is 2^maxlog (ntt.hpp:48-56) */
} fpo_field;

/* Field set-up; root == 0 discovers the smallest primitive root
 * (modint.hpp:28-52).  Returns FPO_E_ARG for p < 2 or p >= 2^31. */
int fpo_field_init(fpo_field* f, uint32_t p, uint32_t root);
/* Root of the reference's named aliases (modint.hpp:195-205), else 0. */
uint32_t fpo_alias_root(uint32_t p);

uint32_t fpo_reduce(const fpo_field* f, uint64_t t);
uint32_t fpo_from_int(const fpo_field* f, uint64_t x);
uint32_t fpo_val(const fpo_field* f, uint32_t limb);
uint32_t fpo_mul(const fpo_field* f, uint32_t a, uint32_t b);
uint32_t fpo_add(const fpo_field* f, uint32_t a, uint32_t b);
uint32_t fpo_sub(const fpo_field* f, uint32_t a, uint32_t b);
uint32_t fpo_pow(const fpo_field* f, uint32_t a, uint64_t e);
uint32_t fpo_minv(const fpo_field* f, uint32_t a);
/* Tonelli-Shanks root choice of modint.hpp:139-166; returns 0 for a
 * non-residue, 1 with *out set otherwise. */
int fpo_sqrt(const fpo_field* f, uint32_t a, uint32_t* out);
uint32_t fpo_primitive_root(uint32_t p);

/* std::mt19937_64 (the generator of bench/bench.cpp:216-220). */
typedef struct {
  uint64_t mt[312];
  int idx;
} fpo_mt64;
void fpo_mt64_seed(fpo_mt64* g, uint64_t seed);
uint64_t fpo_mt64_next(fpo_mt64* g);
/* a[i] = Mont::from_int(rng() % p), bench/bench.cpp:216-220. */
void fpo_rnd(const fpo_field* f, fpo_mt64* g, uint32_t* out, size_t n);
/* FNV-1a over raw limbs, then ^= size (bench/bench.cpp:152-159). */
uint64_t fpo_checksum(const uint32_t* a, size_t n);

/* NttPlan<M>::forward / inverse (ntt.hpp:142-143, 179) restated with the
 * reference's radix-4 DIF / DIT stage structure; canonical outputs. */
int fpo_ntt_forward(const fpo_field* f, uint32_t* a, uint32_t n);
int fpo_ntt_inverse(const fpo_field* f, uint32_t* a, uint32_t n);

/* poly::conv (detail/poly/convolution.hpp:17-147).  Output length
 * min(min(la,n)+min(lb,n)-1, n) (0 if an input is empty or n == 0) is stored
 * in *out_len; out must hold it. */
int fpo_conv(const fpo_field* f, const uint32_t* a, size_t la, const uint32_t* b,
             size_t lb, size_t n, uint32_t* out, size_t* out_len);
/* The same product forced through the three-prime CRT path
 * (detail/poly/crt_convolution.hpp:187-205) with a chosen block size. */
int fpo_conv_crt(const fpo_field* f, const uint32_t* a, size_t la, const uint32_t* b,
                 size_t lb, size_t out_count, size_t block, uint32_t* out);
/* crt_prime_count (crt_convolution.hpp:78-83). */
int fpo_crt_prime_count(uint32_t p, size_t terms);

/* Series (n output limbs each). */
int fpo_inv(const fpo_field* f, const uint32_t* a, size_t la, size_t n, uint32_t* out);
int fpo_log(const fpo_field* f, const uint32_t* a, size_t la, size_t n, uint32_t* out);
int fpo_exp(const fpo_field* f, const uint32_t* a, size_t la, size_t n, uint32_t* out);
int fpo_sqrt_series(const fpo_field* f, const uint32_t* a, size_t la, size_t n, uint32_t* out);

/* poly::pow (detail/poly/power.hpp:17-87): a^k mod x^n, n limbs. */
int fpo_poly_pow(const fpo_field* f, const uint32_t* a, size_t la, uint64_t k, size_t n, uint32_t* out);
/* poly::divmod (detail/poly/division.hpp:12-104): a = b q + r, deg r < deg b,
 * q and r trimmed; q and r must each hold la limbs.  FPO_E_DOMAIN for b == 0. */
int fpo_divmod(const fpo_field* f, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
               uint32_t* q, size_t* qlen, uint32_t* r, size_t* rlen);

/* O(la*lb) product with exact integer accumulation (the independent pin). */
int fpo_conv_naive(const fpo_field* f, const uint32_t* a, size_t la, const uint32_t* b,
                   size_t lb, size_t out_count, uint32_t* out);

#ifdef __cplusplus
}
#endif

#endif /* FPX_ORACLE_H */
