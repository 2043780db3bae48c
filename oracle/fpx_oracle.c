/* fpx_oracle — CPU restatement of fastpoly's exact NTT-convolution path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the GPU build is compared against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it; the product library (fastpoly_amd/) never links or calls it.
 *
 * Parity status: every exported function is pinned by tests/test_oracle.py
 * (published FNV checksums of the reference's bench CSVs, brute-force DFT and
 * O(n^2) integer products, coefficient recurrences).  See fpx_oracle.h.
 *
 * Plain scalar C11 with unsigned __int128 only for the CRT recombination.
 */
#include "fpx_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Field arithmetic: include/fastpoly/modint.hpp:20-180                      */
/* ------------------------------------------------------------------------ */

/* inv_mod_2_32, modint.hpp:20-24 (Newton iteration for a^-1 mod 2^32). */
static uint32_t inv_mod_2_32(uint32_t a) {
  uint32_t x = 1;
  for (int i = 0; i < 5; ++i) x *= 2u - a * x;
  return x;
}

static uint64_t powmod_plain(uint64_t a, uint64_t e, uint64_t p) {
  uint64_t r = 1 % p;
  a %= p;
  for (; e; e >>= 1, a = a * a % p)
    if (e & 1) r = r * a % p;
  return r;
}

/* primitive_root, modint.hpp:28-52: smallest g with g^((p-1)/q) != 1 for
 * every prime factor q of p-1. */
uint32_t fpo_primitive_root(uint32_t p) {
  if (p == 2) return 1;
  uint32_t factors[16];
  int count = 0;
  uint32_t rem = p - 1;
  for (uint32_t d = 2; (uint64_t)d * d <= rem; ++d) {
    if (rem % d) continue;
    factors[count++] = d;
    do rem /= d; while (rem % d == 0);
  }
  if (rem > 1) factors[count++] = rem;
  for (uint32_t g = 2; g < p; ++g) {
    int ok = 1;
    for (int i = 0; i < count; ++i)
      if (powmod_plain(g, (p - 1) / factors[i], p) == 1) { ok = 0; break; }
    if (ok) return g;
  }
  return 0;
}

/* Named aliases with explicit roots, modint.hpp:195-205. */
uint32_t fpo_alias_root(uint32_t p) {
  switch (p) {
    case 998244353u: case 1004535809u: case 469762049u:
    case 167772161u: case 1224736769u: return 3;
    case 754974721u: return 11;
    default: return 0;
  }
}

int fpo_field_init(fpo_field* f, uint32_t p, uint32_t root) {
  if (!f || p < 2 || p >= (1u << 31)) return FPO_E_ARG;
  f->p = p;
  f->ninv = (uint32_t)(0u - inv_mod_2_32(p));
  f->one = p == 2 ? 1u : (uint32_t)(((uint64_t)1 << 32) % p);
  f->r2 = (uint32_t)((uint64_t)f->one * f->one % p);
  f->root = root ? root : fpo_primitive_root(p);
  uint32_t t = p - 1;
  int k = 0;
  while (t && (t & 1) == 0) { t >>= 1; ++k; }
  f->maxlog = k;
  return FPO_OK;
}

/* Mont::reduce, modint.hpp:88-93. */
uint32_t fpo_reduce(const fpo_field* f, uint64_t t) {
  if (f->p == 2) return (uint32_t)(t & 1u);
  uint32_t m = (uint32_t)t * f->ninv;
  uint64_t u = (t + (uint64_t)m * f->p) >> 32;
  return (uint32_t)(u >= f->p ? u - f->p : u);
}

/* Mont::from_int, modint.hpp:171-173. */
uint32_t fpo_from_int(const fpo_field* f, uint64_t x) {
  return (uint32_t)((uint64_t)(uint32_t)(x % f->p) * f->one % f->p);
}
uint32_t fpo_val(const fpo_field* f, uint32_t limb) { return fpo_reduce(f, limb); }
uint32_t fpo_mul(const fpo_field* f, uint32_t a, uint32_t b) {
  return fpo_reduce(f, (uint64_t)a * b);
}
/* operator+ / operator-, modint.hpp:100-106. */
uint32_t fpo_add(const fpo_field* f, uint32_t a, uint32_t b) {
  uint32_t s = a + b;
  return s >= f->p ? s - f->p : s;
}
uint32_t fpo_sub(const fpo_field* f, uint32_t a, uint32_t b) {
  return a >= b ? a - b : a - b + f->p;
}
static uint32_t fpo_neg(const fpo_field* f, uint32_t a) { return a == 0 ? 0 : f->p - a; }

/* Mont::pow, modint.hpp:123-130. */
uint32_t fpo_pow(const fpo_field* f, uint32_t a, uint64_t e) {
  uint32_t r = fpo_from_int(f, 1), b = a;
  for (; e; e >>= 1) {
    if (e & 1) r = fpo_mul(f, r, b);
    b = fpo_mul(f, b, b);
  }
  return r;
}
/* Mont::inv, modint.hpp:133-136 (0 maps to 0). */
uint32_t fpo_minv(const fpo_field* f, uint32_t a) {
  if (f->p == 2) return a;
  return fpo_pow(f, a, f->p - 2);
}

/* Mont::sqrt, modint.hpp:139-166: Euler criterion, the p = 3 mod 4 shortcut,
 * otherwise Tonelli-Shanks with the non-residue search starting at z = 2. */
int fpo_sqrt(const fpo_field* f, uint32_t a, uint32_t* out) {
  const uint32_t p = f->p;
  if (p == 2) { *out = a; return 1; }
  const uint32_t one = fpo_from_int(f, 1);
  if (a == 0) { *out = 0; return 1; }
  if (fpo_pow(f, a, (p - 1) / 2) != one) return 0;
  if (p % 4 == 3) { *out = fpo_pow(f, a, ((uint64_t)p + 1) / 4); return 1; }
  uint64_t q = p - 1;
  int s = 0;
  while (q % 2 == 0) { q /= 2; ++s; }
  uint32_t z = fpo_from_int(f, 2);
  while (fpo_pow(f, z, (p - 1) / 2) != fpo_neg(f, one)) z = fpo_add(f, z, one);
  uint32_t c = fpo_pow(f, z, q), x = fpo_pow(f, a, (q + 1) / 2), t = fpo_pow(f, a, q);
  int m = s;
  while (t != one) {
    int i = 0;
    for (uint32_t t2 = t; t2 != one; t2 = fpo_mul(f, t2, t2)) ++i;
    uint32_t b = fpo_pow(f, c, (uint64_t)1 << (m - i - 1));
    x = fpo_mul(f, x, b);
    c = fpo_mul(f, b, b);
    t = fpo_mul(f, t, c);
    m = i;
  }
  *out = x;
  return 1;
}

/* ------------------------------------------------------------------------ */
/* std::mt19937_64 and the bench input/checksum (bench/bench.cpp:152-220)    */
/* ------------------------------------------------------------------------ */

void fpo_mt64_seed(fpo_mt64* g, uint64_t seed) {
  g->mt[0] = seed;
  for (int i = 1; i < 312; ++i)
    g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
  g->idx = 312;
}

uint64_t fpo_mt64_next(fpo_mt64* g) {
  static const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
  static const uint64_t MATRIX_A = 0xB5026F5AA96619E9ULL;
  if (g->idx >= 312) {
    for (int i = 0; i < 312; ++i) {
      uint64_t x = (g->mt[i] & UM) | (g->mt[(i + 1) % 312] & LM);
      uint64_t xa = x >> 1;
      if (x & 1) xa ^= MATRIX_A;
      g->mt[i] = g->mt[(i + 156) % 312] ^ xa;
    }
    g->idx = 0;
  }
  uint64_t x = g->mt[g->idx++];
  x ^= (x >> 29) & 0x5555555555555555ULL;
  x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
  x ^= (x << 37) & 0xFFF7EEE000000000ULL;
  x ^= x >> 43;
  return x;
}

void fpo_rnd(const fpo_field* f, fpo_mt64* g, uint32_t* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = fpo_from_int(f, fpo_mt64_next(g) % f->p);
}

uint64_t fpo_checksum(const uint32_t* a, size_t n) {
  uint64_t h = 14695981039346656037ULL;
  for (size_t i = 0; i < n; ++i) { h ^= a[i]; h *= 1099511628211ULL; }
  return h ^ (uint64_t)n;
}

/* ------------------------------------------------------------------------ */
/* NTT: include/fastpoly/ntt.hpp                                             */
/* ------------------------------------------------------------------------ */

static int is_pow2(uint64_t n) { return n >= 1 && (n & (n - 1)) == 0; }
static int ctz64(uint64_t n) { int k = 0; while (!(n & 1)) { n >>= 1; ++k; } return k; }
static uint64_t next_pow2(uint64_t n) { uint64_t r = 1; while (r < n) r <<= 1; return r; }

/* cached_plan's validation, ntt.hpp:250-255. */
static int plan_ok(const fpo_field* f, uint32_t n) {
  if (n < 2 || !is_pow2(n)) return 0;
  if (f->maxlog >= 31 || ((uint64_t)1 << f->maxlog) < n) return 0;
  return 1;
}

/* Twiddle tables of one radix-4 stage (ntt.hpp:398-446): A[k]=w^k, B[k]=w^2k,
 * C[k]=w^3k for k < m, each chain seeded with `start` (Montgomery one, or
 * inv_n for the normalised inverse stage 0, ntt.hpp:418). */
static void stage_tables(const fpo_field* f, uint32_t w, uint32_t start, uint32_t m,
                         uint32_t* A, uint32_t* B, uint32_t* C) {
  const uint32_t w2 = fpo_mul(f, w, w), w3 = fpo_mul(f, w2, w);
  uint32_t c1 = start, c2 = start, c3 = start;
  for (uint32_t k = 0; k < m; ++k) {
    A[k] = c1; B[k] = c2; C[k] = c3;
    c1 = fpo_mul(f, c1, w); c2 = fpo_mul(f, c2, w2); c3 = fpo_mul(f, c3, w3);
  }
}

/* Forward DIF, ntt.hpp:283-308 with the butterflies of radix4_fwd
 * (ntt.hpp:858-902) and the trailing radix2_fwd (ntt.hpp:955-984):
 * stage lens n, n/4, ..., then a radix-2 on consecutive pairs when log2 n is
 * odd.  Output: canonical limbs in base-4 digit-reversed (= bit-reversed)
 * order. */
int fpo_ntt_forward(const fpo_field* f, uint32_t* a, uint32_t n) {
  if (!plan_ok(f, n)) return FPO_E_SIZE;
  const int logn = ctz64(n);
  if (n >= 4) {
    const uint32_t wn = fpo_pow(f, fpo_from_int(f, f->root), (f->p - 1) / n);
    const uint32_t w4 = fpo_pow(f, wn, n / 4);
    uint32_t* tab = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (n / 4));
    if (!tab) return FPO_E_NOMEM;
    uint32_t w = wn;
    for (uint32_t len = n; len >= 4; len >>= 2) {
      const uint32_t m = len / 4;
      uint32_t *A = tab, *B = tab + m, *C = tab + 2 * m;
      stage_tables(f, w, fpo_from_int(f, 1), m, A, B, C);
      for (uint32_t base = 0; base < n; base += len) {
        uint32_t* p = a + base;
        for (uint32_t k = 0; k < m; ++k) {
          uint32_t x0 = p[k], x1 = p[k + m], x2 = p[k + 2 * m], x3 = p[k + 3 * m];
          uint32_t t0 = fpo_add(f, x0, x2), t1 = fpo_sub(f, x0, x2);
          uint32_t t2 = fpo_add(f, x1, x3), t3 = fpo_mul(f, fpo_sub(f, x1, x3), w4);
          p[k] = fpo_add(f, t0, t2);
          p[k + m] = fpo_mul(f, fpo_sub(f, t0, t2), B[k]);
          p[k + 2 * m] = fpo_mul(f, fpo_add(f, t1, t3), A[k]);
          p[k + 3 * m] = fpo_mul(f, fpo_sub(f, t1, t3), C[k]);
        }
      }
      const uint32_t w2 = fpo_mul(f, w, w);
      w = fpo_mul(f, w2, w2);
    }
    free(tab);
  }
  if (logn & 1) {
    for (uint32_t k = 0; k + 1 < n; k += 2) {
      uint32_t x = a[k], y = a[k + 1];
      a[k] = fpo_add(f, x, y);
      a[k + 1] = fpo_sub(f, x, y);
    }
  }
  return FPO_OK;
}

/* Inverse DIT, ntt.hpp:195-229 with radix4_inv (ntt.hpp:905-946):
 * radix-2 first when log2 n is odd, then stages from the smallest len up;
 * stage 0 multiplies b0 by 1/n and uses twiddles prescaled by 1/n. */
int fpo_ntt_inverse(const fpo_field* f, uint32_t* a, uint32_t n) {
  if (!plan_ok(f, n)) return FPO_E_SIZE;
  const int logn = ctz64(n);
  const uint32_t inv_n = fpo_pow(f, fpo_from_int(f, (f->p + 1) / 2), (uint64_t)logn);
  if (n == 2) {
    const uint32_t x = a[0], y = a[1];
    a[0] = fpo_mul(f, fpo_add(f, x, y), inv_n);
    a[1] = fpo_mul(f, fpo_sub(f, x, y), inv_n);
    return FPO_OK;
  }
  const uint32_t wn = fpo_pow(f, fpo_from_int(f, f->root), (f->p - 1) / n);
  const uint32_t w4 = fpo_pow(f, wn, n / 4);
  const uint32_t w4_inv = f->p - w4;
  int stages = 0;
  uint32_t lens[16], ws[16];
  {
    uint32_t w = wn;
    for (uint32_t len = n; len >= 4; len >>= 2) {
      lens[stages] = len;
      ws[stages++] = w;
      const uint32_t w2 = fpo_mul(f, w, w);
      w = fpo_mul(f, w2, w2);
    }
  }
  uint32_t* tab = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (n / 4));
  if (!tab) return FPO_E_NOMEM;
  if (logn & 1) {
    for (uint32_t k = 0; k + 1 < n; k += 2) {
      uint32_t x = a[k], y = a[k + 1];
      a[k] = fpo_add(f, x, y);
      a[k + 1] = fpo_sub(f, x, y);
    }
  }
  for (int s = stages - 1; s >= 0; --s) {
    const uint32_t len = lens[s], m = len / 4;
    uint32_t *IA = tab, *IB = tab + m, *IC = tab + 2 * m;
    stage_tables(f, fpo_minv(f, ws[s]), s == 0 ? inv_n : fpo_from_int(f, 1), m, IA, IB, IC);
    for (uint32_t base = 0; base < n; base += len) {
      uint32_t* p = a + base;
      for (uint32_t k = 0; k < m; ++k) {
        uint32_t b0 = p[k], b1 = p[k + m], b2 = p[k + 2 * m], b3 = p[k + 3 * m];
        if (s == 0) b0 = fpo_mul(f, b0, inv_n);
        uint32_t c1 = fpo_mul(f, b1, IB[k]), c2 = fpo_mul(f, b2, IA[k]), c3 = fpo_mul(f, b3, IC[k]);
        uint32_t d0 = fpo_add(f, b0, c1), d1 = fpo_sub(f, b0, c1);
        uint32_t d2 = fpo_add(f, c2, c3), e = fpo_mul(f, fpo_sub(f, c2, c3), w4_inv);
        p[k] = fpo_add(f, d0, d2);
        p[k + m] = fpo_add(f, d1, e);
        p[k + 2 * m] = fpo_sub(f, d0, d2);
        p[k + 3 * m] = fpo_sub(f, d1, e);
      }
    }
  }
  free(tab);
  return FPO_OK;
}

/* ------------------------------------------------------------------------ */
/* Convolution: detail/poly/convolution.hpp, detail/poly/crt_convolution.hpp */
/* ------------------------------------------------------------------------ */

#define NAIVE_THRESHOLD 40 /* detail/poly/base.hpp:89 */

/* Exact O(la*lb) product: integer dot products accumulated in 128 bits,
 * reduced once.  Montgomery limbs multiply to limb*limb = x*y*R^2, one REDC
 * yields the product's limb. */
int fpo_conv_naive(const fpo_field* f, const uint32_t* a, size_t la, const uint32_t* b,
                   size_t lb, size_t out_count, uint32_t* out) {
  for (size_t k = 0; k < out_count; ++k) {
    unsigned __int128 acc = 0;
    size_t lo = k >= lb ? k - lb + 1 : 0;
    size_t hi = k < la ? k : la - 1;
    for (size_t i = lo; i <= hi && la && lb; ++i) acc += (uint64_t)a[i] * b[k - i];
    out[k] = fpo_reduce(f, (uint64_t)(acc % f->p));
  }
  return FPO_OK;
}

/* Cyclic-free NTT product mod a prime with 2^maxlog >= N (the native branch of
 * conv_limbs, convolution.hpp:105-139): forward both zero-padded operands,
 * multiply pointwise (detail/poly/base.hpp:106-133), inverse. */
static int ntt_product(const fpo_field* f, const uint32_t* a, size_t la, const uint32_t* b,
                       size_t lb, size_t out_count, uint32_t* out) {
  const uint32_t N = (uint32_t)next_pow2(la + lb - 1);
  const uint32_t n = N < 2 ? 2 : N;
  uint32_t* fa = (uint32_t*)calloc(n, 4);
  uint32_t* fb = (uint32_t*)calloc(n, 4);
  if (!fa || !fb) { free(fa); free(fb); return FPO_E_NOMEM; }
  memcpy(fa, a, la * 4);
  memcpy(fb, b, lb * 4);
  int rc = fpo_ntt_forward(f, fa, n);
  if (!rc) rc = fpo_ntt_forward(f, fb, n);
  if (!rc) {
    for (uint32_t i = 0; i < n; ++i) fa[i] = fpo_mul(f, fa[i], fb[i]);
    rc = fpo_ntt_inverse(f, fa, n);
  }
  if (!rc) memcpy(out, fa, out_count * 4);
  free(fa);
  free(fb);
  return rc;
}

static const uint32_t CRT_P[3] = {167772161u, 469762049u, 754974721u}; /* crt_convolution.hpp:14-19 */
#define CRT_BLOCK ((size_t)1 << 23)                                    /* crt_convolution.hpp:19 */

int fpo_crt_prime_count(uint32_t p, size_t terms) {
  const uint64_t square = (uint64_t)(p - 1) * (p - 1);
  if (terms == 0) terms = 1;
  if (square <= (CRT_P[0] - 1) / terms) return 1;
  if (square <= ((uint64_t)CRT_P[0] * CRT_P[1] - 1) / terms) return 2;
  return 3;
}

/* auxiliary_product (crt_convolution.hpp:37-74): lift raw target limbs as
 * integers into the auxiliary field (lift_limbs, :25-33), multiply exactly,
 * return the plain residues of the integer product window [0, count). */
static int aux_product(const fpo_field* P, const uint32_t* a, size_t la, const uint32_t* b,
                       size_t lb, size_t count, uint32_t* plain) {
  uint32_t* la_l = (uint32_t*)malloc((la ? la : 1) * 4);
  uint32_t* lb_l = (uint32_t*)malloc((lb ? lb : 1) * 4);
  uint32_t* prod = (uint32_t*)malloc((count ? count : 1) * 4);
  if (!la_l || !lb_l || !prod) { free(la_l); free(lb_l); free(prod); return FPO_E_NOMEM; }
  for (size_t i = 0; i < la; ++i) la_l[i] = fpo_reduce(P, (uint64_t)a[i] * P->r2);
  for (size_t i = 0; i < lb; ++i) lb_l[i] = fpo_reduce(P, (uint64_t)b[i] * P->r2);
  int rc;
  if ((la < lb ? la : lb) <= NAIVE_THRESHOLD)
    rc = fpo_conv_naive(P, la_l, la, lb_l, lb, count, prod);
  else
    rc = ntt_product(P, la_l, la, lb_l, lb, count, prod);
  if (!rc)
    for (size_t i = 0; i < count; ++i) plain[i] = fpo_reduce(P, prod[i]);
  free(la_l); free(lb_l); free(prod);
  return rc;
}

/* crt_product_window (crt_convolution.hpp:131-182) with restore_crt's Garner
 * digits (:88-127): x1 = c1, x2 = (c2-x1)/p1 mod p2, x3 = (c3-x1-p1 x2)/(p1 p2)
 * mod p3; the integer x1 + p1 x2 + p1 p2 x3 carries two target Montgomery
 * factors, one REDC leaves the product's limb, accumulated into dst. */
static int crt_window(const fpo_field* f, const uint32_t* a, size_t na, const uint32_t* b,
                      size_t nb, size_t count, uint32_t* dst) {
  const int primes = fpo_crt_prime_count(f->p, na < nb ? na : nb);
  uint32_t* c[3] = {0, 0, 0};
  fpo_field P[3];
  int rc = FPO_OK;
  for (int k = 0; k < primes && !rc; ++k) {
    fpo_field_init(&P[k], CRT_P[k], fpo_alias_root(CRT_P[k]));
    c[k] = (uint32_t*)malloc((count ? count : 1) * 4);
    if (!c[k]) { rc = FPO_E_NOMEM; break; }
    rc = aux_product(&P[k], a, na, b, nb, count, c[k]);
  }
  if (!rc) {
    const uint64_t p1 = CRT_P[0], p2 = CRT_P[1], p3 = CRT_P[2];
    const uint64_t inv12 = powmod_plain(p1 % p2, p2 - 2, p2);
    const uint64_t inv123 = powmod_plain((p1 * p2) % p3, p3 - 2, p3);
    for (size_t i = 0; i < count; ++i) {
      unsigned __int128 value = c[0][i];
      if (primes >= 2) {
        const uint64_t x1 = c[0][i];
        const uint64_t x2 = (c[1][i] + p2 - x1 % p2) % p2 * inv12 % p2;
        const uint64_t x12 = x1 + p1 * x2;
        value = x12;
        if (primes == 3) {
          const uint64_t x3 = (c[2][i] + p3 - x12 % p3) % p3 * inv123 % p3;
          value = (unsigned __int128)x12 + (unsigned __int128)(p1 * p2) * x3;
        }
      }
      const uint32_t limb = fpo_reduce(f, (uint64_t)(value % f->p));
      dst[i] = f->p == 2 ? (dst[i] ^ limb) & 1u : fpo_add(f, dst[i], limb);
    }
  }
  for (int k = 0; k < 3; ++k) free(c[k]);
  return rc;
}

/* crt_convolution (crt_convolution.hpp:187-205): block pairs of at most
 * `block` coefficients, each window accumulated at offset ai+bi. */
int fpo_conv_crt(const fpo_field* f, const uint32_t* a, size_t la, const uint32_t* b,
                 size_t lb, size_t out, size_t block, uint32_t* r) {
  if (block == 0 || block > CRT_BLOCK) return FPO_E_ARG;
  memset(r, 0, out * 4);
  if (la > out) la = out;
  if (lb > out) lb = out;
  for (size_t ai = 0; ai < la;) {
    const size_t na = block < la - ai ? block : la - ai;
    for (size_t bi = 0; bi < lb && bi < out - ai;) {
      const size_t nb = block < lb - bi ? block : lb - bi;
      size_t count = na + nb - 1;
      if (count > out - ai - bi) count = out - ai - bi;
      int rc = crt_window(f, a + ai, na, b + bi, nb, count, r + ai + bi);
      if (rc) return rc;
      bi += nb;
    }
    ai += na;
  }
  return FPO_OK;
}

/* conv_limbs (convolution.hpp:17-141).  Dispatch: direct product when the
 * short side is <= 40 (:33-95); native NTT when next_pow2(full) fits the root
 * bound (:105-139); otherwise CRT (:99-100, :108).  The reference's tail
 * repair (:102-137) is a cost optimisation with the identical exact result
 * and is not restated. */
int fpo_conv(const fpo_field* f, const uint32_t* a, size_t la, const uint32_t* b,
             size_t lb, size_t n, uint32_t* out, size_t* out_len) {
  *out_len = 0;
  if (la == 0 || lb == 0 || n == 0) return FPO_OK;
  if (la > n) la = n;
  if (lb > n) lb = n;
  if (la > SIZE_MAX - (lb - 1)) return FPO_E_SIZE;
  const size_t full = la + lb - 1;
  const size_t outc = full < n ? full : n;
  *out_len = outc;
  const size_t shorter = la < lb ? la : lb;
  if (shorter <= NAIVE_THRESHOLD) return fpo_conv_naive(f, a, la, b, lb, outc, out);
  if (f->p != 2 && f->maxlog < 31 && next_pow2(full) <= ((uint64_t)1 << f->maxlog) &&
      ((uint64_t)1 << f->maxlog) > NAIVE_THRESHOLD)
    return ntt_product(f, a, la, b, lb, outc, out);
  return fpo_conv_crt(f, a, la, b, lb, outc, CRT_BLOCK, out);
}

/* ------------------------------------------------------------------------ */
/* Series: detail/poly/{inverse,exponential,square_root,calculus}.hpp        */
/* ------------------------------------------------------------------------ */

/* conv of two prefixes truncated to lim, zero-filled to exactly lim limbs. */
static int conv_fill(const fpo_field* f, const uint32_t* a, size_t la, const uint32_t* b,
                     size_t lb, size_t lim, uint32_t* out) {
  size_t got = 0;
  memset(out, 0, lim * 4);
  if (la == 0 || lb == 0 || lim == 0) return FPO_OK;
  uint32_t* tmp = (uint32_t*)malloc((la + lb) * 4);
  if (!tmp) return FPO_E_NOMEM;
  int rc = fpo_conv(f, a, la, b, lb, lim, tmp, &got);
  if (!rc) memcpy(out, tmp, got * 4);
  free(tmp);
  return rc;
}

/* inverse_seed (inverse.hpp:68-77): c[i] = -(sum_{j>=1} a[j] c[i-j]) / a0. */
static void inverse_seed(const fpo_field* f, const uint32_t* a, size_t la, uint32_t* c,
                         size_t n) {
  c[0] = fpo_minv(f, a[0]);
  for (size_t i = 1; i < n; ++i) {
    uint32_t s = 0;
    for (size_t j = 1; j <= i && j < la; ++j) s = fpo_add(f, s, fpo_mul(f, a[j], c[i - j]));
    c[i] = fpo_mul(f, fpo_neg(f, s), c[0]);
  }
}

/* extend_inverse (inverse.hpp:18-65): extend c from k to wanted <= 2k limbs.
 * A tail of <= 8 limbs uses the recurrence; otherwise the middle product
 * e = (a c)[k, wanted) and c[k, wanted) = -(e c) mod x^hlen. */
static int extend_inverse(const fpo_field* f, const uint32_t* a, size_t la, uint32_t* c,
                          size_t k, size_t wanted) {
  const size_t hlen = wanted - k;
  if (hlen <= 8) {
    const uint32_t inv0 = c[0];
    for (size_t i = k; i < wanted; ++i) {
      uint32_t s = 0;
      for (size_t j = 1; j <= i && j < la; ++j) s = fpo_add(f, s, fpo_mul(f, a[j], c[i - j]));
      c[i] = fpo_mul(f, fpo_neg(f, s), inv0);
    }
    return FPO_OK;
  }
  uint32_t* prod = (uint32_t*)malloc(wanted * 4);
  uint32_t* err = (uint32_t*)malloc(hlen * 4);
  if (!prod || !err) { free(prod); free(err); return FPO_E_NOMEM; }
  int rc = conv_fill(f, a, la < wanted ? la : wanted, c, k, wanted, prod);
  if (!rc) {
    for (size_t i = 0; i < hlen; ++i) err[i] = fpo_neg(f, prod[k + i]);
    rc = conv_fill(f, err, hlen, c, k, hlen, c + k);
  }
  free(prod); free(err);
  return rc;
}

/* Newton doubling from a 32-limb seed (inverse.hpp:178-187). */
static int inverse_newton(const fpo_field* f, const uint32_t* a, size_t la, size_t n,
                          uint32_t* out) {
  const size_t seed = n < 32 ? n : 32;
  inverse_seed(f, a, la, out, seed);
  for (size_t m = seed; m < n; m <<= 1) {
    const size_t want = 2 * m < n ? 2 * m : n;
    int rc = extend_inverse(f, a, la, out, m, want);
    if (rc) return rc;
  }
  return FPO_OK;
}

/* poly::inv (inverse.hpp:163-188).  The constant, linear and short special
 * cases (:169-177) have the same unique result as Newton and are folded in. */
int fpo_inv(const fpo_field* f, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  if (n == 0) return FPO_OK;
  if (la == 0 || a[0] == 0) return FPO_E_DOMAIN;
  size_t cnt = la < n ? la : n;
  while (cnt > 1 && a[cnt - 1] == 0) --cnt;
  if (cnt == 1) {
    memset(out, 0, n * 4);
    out[0] = fpo_minv(f, a[0]);
    return FPO_OK;
  }
  return inverse_newton(f, a, cnt, n, out);
}

/* inv_series (calculus.hpp:53-94): Montgomery limbs of 1/1..1/n by
 * 1/i = -(p/i) * (1/(p mod i)). */
static int inv_series(const fpo_field* f, size_t n, uint32_t** res) {
  if (n >= f->p) return FPO_E_DOMAIN;
  uint32_t* r = (uint32_t*)calloc(n + 1, 4);
  if (!r) return FPO_E_NOMEM;
  if (n >= 1) r[1] = fpo_from_int(f, 1);
  for (size_t i = 2; i <= n; ++i) {
    const uint32_t q = f->p / (uint32_t)i, rem = f->p % (uint32_t)i;
    r[i] = f->p - (uint32_t)((uint64_t)q * r[rem] % f->p);
  }
  *res = r;
  return FPO_OK;
}

/* derivative_prefix (calculus.hpp:17-39): r[i-1] = a[i] * i, i < cnt. */
static size_t derivative_prefix(const fpo_field* f, const uint32_t* a, size_t la, size_t cnt,
                                uint32_t* r) {
  if (cnt > la) cnt = la;
  if (cnt <= 1) return 0;
  for (size_t i = 1; i < cnt; ++i) r[i - 1] = fpo_mul(f, a[i], fpo_from_int(f, i));
  return cnt - 1;
}

/* poly::log (exponential.hpp:85-98): integral(a' * inv(a, n-1) mod x^(n-1)). */
int fpo_log(const fpo_field* f, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  if (n == 0) return FPO_OK;
  if (la == 0 || a[0] != fpo_from_int(f, 1)) return FPO_E_DOMAIN;
  if (n > f->p) return FPO_E_DOMAIN;
  memset(out, 0, n * 4);
  if (n == 1) return FPO_OK;
  uint32_t* d = (uint32_t*)malloc(n * 4);
  uint32_t* ia = (uint32_t*)malloc(n * 4);
  uint32_t* pr = (uint32_t*)malloc(n * 4);
  uint32_t* den = 0;
  int rc = (!d || !ia || !pr) ? FPO_E_NOMEM : FPO_OK;
  size_t ld = 0;
  if (!rc) ld = derivative_prefix(f, a, la, n, d);
  if (!rc) rc = fpo_inv(f, a, la, n - 1, ia);
  size_t lp = 0;
  if (!rc && ld) {
    rc = fpo_conv(f, d, ld, ia, n - 1, n - 1, pr, &lp);
  }
  if (!rc && lp) rc = inv_series(f, lp, &den);
  if (!rc)
    for (size_t i = 0; i < lp; ++i) out[i + 1] = fpo_mul(f, pr[i], den[i + 1]);
  free(d); free(ia); free(pr); free(den);
  return rc;
}

/* poly::exp (exponential.hpp:107-207): recurrence seed, then Newton steps
 * m -> m2 with the high residual r = (a' b)[m-1, m-1+hlen), the inverse
 * prefix c = 1/b mod x^hlen, t = (r c)/(m+i), b[m, m2) = (t b) mod x^hlen.
 * NEON's exp_linear (:27-79) yields the same unique coefficients. */
int fpo_exp(const fpo_field* f, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  if (n == 0) return FPO_OK;
  if (la != 0 && a[0] != 0) return FPO_E_DOMAIN;
  if (n - 1 >= f->p) return FPO_E_DOMAIN;
  memset(out, 0, n * 4);
  uint32_t* ap = (uint32_t*)malloc((n ? n : 1) * 4);
  if (!ap) return FPO_E_NOMEM;
  size_t lap = derivative_prefix(f, a, la, n, ap);
  while (lap && ap[lap - 1] == 0) --lap; /* trim (basic.hpp:15-17) */
  if (lap == 0) {
    out[0] = fpo_from_int(f, 1);
    free(ap);
    return FPO_OK;
  }
  uint32_t* den = 0;
  int rc = inv_series(f, n - 1, &den);
  if (rc) { free(ap); return rc; }
  uint32_t* b = out;
  b[0] = fpo_from_int(f, 1);
  const size_t seed = lap <= NAIVE_THRESHOLD ? n : (n < 32 ? n : 32);
  for (size_t i = 1; i < seed; ++i) {
    uint32_t s = 0;
    for (size_t j = 1; j <= i && j <= lap; ++j) s = fpo_add(f, s, fpo_mul(f, ap[j - 1], b[i - j]));
    b[i] = fpo_mul(f, s, den[i]);
  }
  uint32_t *c = 0, *r = 0, *t = 0, *prod = 0;
  if (seed < n) {
    c = (uint32_t*)malloc(n * 4);
    r = (uint32_t*)malloc(n * 4);
    t = (uint32_t*)malloc(n * 4);
    prod = (uint32_t*)malloc(2 * n * 4);
    if (!c || !r || !t || !prod) rc = FPO_E_NOMEM;
  }
  for (size_t m = seed; m < n && !rc; m *= 2) {
    const size_t m2 = 2 * m < n ? 2 * m : n, hlen = m2 - m;
    if (hlen <= 8) {
      for (size_t i = m; i < m2; ++i) {
        uint32_t s = 0;
        for (size_t j = 1; j <= i && j <= lap; ++j) s = fpo_add(f, s, fpo_mul(f, ap[j - 1], b[i - j]));
        b[i] = fpo_mul(f, s, den[i]);
      }
      continue;
    }
    rc = fpo_inv(f, b, m, hlen, c);
    if (!rc) {
      const size_t use = lap < m2 - 1 ? lap : m2 - 1;
      rc = conv_fill(f, ap, use, b, m, m - 1 + hlen, prod);
    }
    if (!rc) {
      memcpy(r, prod + (m - 1), hlen * 4);
      rc = conv_fill(f, r, hlen, c, hlen, hlen, t);
    }
    if (!rc) {
      for (size_t i = 0; i < hlen; ++i) t[i] = fpo_mul(f, t[i], den[m + i]);
      rc = conv_fill(f, t, hlen, b, m, hlen, b + m);
    }
  }
  free(ap); free(den); free(c); free(r); free(t); free(prod);
  return rc;
}

/* poly::sqrt (square_root.hpp:12-112): valuation, Tonelli-Shanks root0,
 * 32-limb recurrence seed, then Newton b[m, m2) = ((a - b^2)[m, m2) / b) / 2. */
int fpo_sqrt_series(const fpo_field* f, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  if (n == 0) return FPO_OK;
  memset(out, 0, n * 4);
  if (f->p == 2) { /* Frobenius, square_root.hpp:15-25 */
    for (size_t i = 0; i < la && i < n; ++i) {
      if (i & 1) {
        if (a[i] != 0) return FPO_E_DOMAIN;
      } else {
        out[i / 2] = a[i];
      }
    }
    return FPO_OK;
  }
  size_t v = 0;
  while (v < la && a[v] == 0) ++v;
  if (v == la) return FPO_OK;
  if (v & 1) return FPO_E_DOMAIN;
  uint32_t root0;
  if (!fpo_sqrt(f, a[v], &root0)) return FPO_E_DOMAIN;
  const size_t half = v / 2;
  if (half >= n) return FPO_OK;
  const size_t inner = n - half;
  const size_t ucnt = la - v < inner ? la - v : inner;
  int all_zero = 1;
  for (size_t i = 1; i < ucnt; ++i) if (a[v + i]) { all_zero = 0; break; }
  uint32_t* b = out + half;
  if (all_zero) { b[0] = root0; return FPO_OK; }
  const uint32_t half_inv = fpo_from_int(f, (f->p + 1) / 2);
  const uint32_t tri = fpo_mul(f, fpo_minv(f, root0), half_inv);
  const size_t seed = inner < 32 ? inner : 32;
  const size_t au = la - v; /* coefficients of the unit part available */
  b[0] = root0;
  for (size_t i = 1; i < seed; ++i) {
    uint32_t res = i < au ? a[v + i] : 0;
    for (size_t j = 1; j < i; ++j) res = fpo_sub(f, res, fpo_mul(f, b[j], b[i - j]));
    b[i] = fpo_mul(f, res, tri);
  }
  int rc = FPO_OK;
  uint32_t *c = 0, *sq = 0, *res = 0;
  if (seed < inner) {
    c = (uint32_t*)malloc(inner * 4);
    sq = (uint32_t*)malloc(2 * inner * 4);
    res = (uint32_t*)malloc(inner * 4);
    if (!c || !sq || !res) rc = FPO_E_NOMEM;
  }
  for (size_t m = seed; m < inner && !rc; m *= 2) {
    const size_t m2 = 2 * m < inner ? 2 * m : inner, hlen = m2 - m;
    if (hlen <= 8) {
      for (size_t i = m; i < m2; ++i) {
        uint32_t r0 = i < au ? a[v + i] : 0;
        for (size_t j = 1; j < i; ++j) r0 = fpo_sub(f, r0, fpo_mul(f, b[j], b[i - j]));
        b[i] = fpo_mul(f, r0, tri);
      }
      continue;
    }
    rc = fpo_inv(f, b, m, hlen, c);
    if (!rc) rc = conv_fill(f, b, m, b, m, m + hlen, sq);
    if (!rc) {
      for (size_t i = 0; i < hlen; ++i)
        res[i] = fpo_sub(f, m + i < au ? a[v + m + i] : 0, sq[m + i]);
      rc = conv_fill(f, res, hlen, c, hlen, hlen, b + m);
    }
    if (!rc)
      for (size_t i = 0; i < hlen; ++i) b[m + i] = fpo_mul(f, b[m + i], half_inv);
  }
  free(c); free(sq); free(res);
  return rc;
}

/* ------------------------------------------------------------------------
 * poly::pow (detail/poly/power.hpp:17-87): a^k mod x^n.
 * ---------------------------------------------------------------------- */
static int conv_trunc(const fpo_field* f, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
                      size_t n, uint32_t** res, size_t* len) {
  const size_t cap = (la && lb) ? (la + lb - 1 < n ? la + lb - 1 : n) : 0;
  uint32_t* out = (uint32_t*)malloc((cap ? cap : 1) * 4);
  if (!out) return FPO_E_NOMEM;
  size_t got = 0;
  const int rc = fpo_conv(f, a, la, b, lb, n, out, &got);
  if (rc) { free(out); return rc; }
  *res = out;
  *len = got;
  return FPO_OK;
}

int fpo_poly_pow(const fpo_field* f, const uint32_t* a, size_t la, uint64_t k, size_t n, uint32_t* out) {
  if (n == 0) return FPO_OK;
  memset(out, 0, n * 4);
  if (k == 0) { out[0] = fpo_from_int(f, 1); return FPO_OK; }                       /* :20-24 */
  if (k == 1) { memcpy(out, a, (la < n ? la : n) * 4); return FPO_OK; }             /* :25-29 */
  size_t v = 0;
  while (v < la && a[v] == 0) ++v;                                                  /* :30-32 */
  if (v == la) return FPO_OK;
  if (v != 0 && k > (n - 1) / v) return FPO_OK;                                     /* :34 */
  const size_t sh = (size_t)(v * k), inner = n - sh;
  size_t lu = la - v < inner ? la - v : inner;
  const uint32_t* u = a + v;
  while (lu > 0 && u[lu - 1] == 0) --lu;                                            /* trim, :39 */
  if (lu == 1) { out[sh] = fpo_pow(f, u[0], k); return FPO_OK; }                   /* :40-44 */
  if (k > 64 && lu == 2 && inner <= f->p) {                                        /* :45-60 */
    const uint32_t e = (uint32_t)(k % f->p);
    const size_t cnt = inner < (size_t)e + 1 ? inner : (size_t)e + 1;
    uint32_t* den = 0;
    int rc = inv_series(f, cnt - 1, &den);
    if (rc) return rc;
    out[sh] = fpo_pow(f, u[0], k);
    const uint32_t ratio = fpo_mul(f, u[1], fpo_minv(f, u[0]));
    uint32_t factor = fpo_mul(f, ratio, fpo_from_int(f, e));
    for (size_t j = 1; j < cnt; ++j) {
      out[sh + j] = fpo_mul(f, fpo_mul(f, out[sh + j - 1], factor), den[j]);
      factor = fpo_sub(f, factor, ratio);
    }
    free(den);
    return FPO_OK;
  }
  if (k <= 64 || inner > f->p) {                                                    /* :62-73 */
    uint32_t* r = (uint32_t*)malloc(4);
    uint32_t* base = (uint32_t*)malloc(lu * 4);
    if (!r || !base) { free(r); free(base); return FPO_E_NOMEM; }
    r[0] = fpo_from_int(f, 1);
    size_t lr = 1, lb = lu;
    memcpy(base, u, lu * 4);
    for (uint64_t e = k; e; e >>= 1) {
      if (e & 1) {
        uint32_t* t = 0; size_t lt = 0;
        int rc = conv_trunc(f, r, lr, base, lb, inner, &t, &lt);
        if (rc) { free(r); free(base); return rc; }
        free(r); r = t; lr = lt;
      }
      if (e > 1) {
        uint32_t* t = 0; size_t lt = 0;
        int rc = conv_trunc(f, base, lb, base, lb, inner, &t, &lt);
        if (rc) { free(r); free(base); return rc; }
        free(base); base = t; lb = lt;
      }
    }
    memcpy(out + sh, r, (lr < inner ? lr : inner) * 4);
    free(r); free(base);
    return FPO_OK;
  }
  /* exp(k log(a / c)) * c^k (:75-86) */
  const uint32_t c = u[0];
  uint32_t* nrm = (uint32_t*)malloc(lu * 4);
  uint32_t* l = (uint32_t*)malloc(inner * 4);
  uint32_t* r = (uint32_t*)malloc(inner * 4);
  int rc = (!nrm || !l || !r) ? FPO_E_NOMEM : FPO_OK;
  if (!rc) {
    const uint32_t ci = fpo_minv(f, c);
    for (size_t i = 0; i < lu; ++i) nrm[i] = fpo_mul(f, u[i], ci);
    rc = fpo_log(f, nrm, lu, inner, l);
  }
  if (!rc) {
    const uint32_t kk = fpo_from_int(f, k);
    for (size_t i = 0; i < inner; ++i) l[i] = fpo_mul(f, l[i], kk);
    rc = fpo_exp(f, l, inner, inner, r);
  }
  if (!rc) {
    const uint32_t ck = fpo_pow(f, c, k);
    for (size_t i = 0; i < inner; ++i) out[sh + i] = fpo_mul(f, r[i], ck);
  }
  free(nrm); free(l); free(r);
  return rc;
}

/* ------------------------------------------------------------------------
 * poly::divmod (detail/poly/division.hpp:12-104): a = b q + r, deg r < deg b,
 * both trimmed.  q holds at least la limbs, r at least la limbs.
 * ---------------------------------------------------------------------- */
int fpo_divmod(const fpo_field* f, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
               uint32_t* q, size_t* qlen, uint32_t* r, size_t* rlen) {
  size_t an = la, bn = lb;
  while (an && a[an - 1] == 0) --an;
  while (bn && b[bn - 1] == 0) --bn;
  *qlen = *rlen = 0;
  if (bn == 0) return FPO_E_DOMAIN;                                                 /* :17 */
  if (an < bn) {                                                                    /* :18-22 */
    memcpy(r, a, an * 4);
    *rlen = an;
    return FPO_OK;
  }
  if (bn == 1) {                                                                    /* :23-28 */
    const uint32_t bi = fpo_minv(f, b[0]);
    for (size_t i = 0; i < an; ++i) q[i] = fpo_mul(f, a[i], bi);
    *qlen = an;
    return FPO_OK;
  }
  const size_t dn = an - bn + 1, rn = bn - 1;
  if (bn <= 16 || dn <= 32) {                                                       /* long division :29-52 */
    uint32_t* w = (uint32_t*)malloc(an * 4);
    if (!w) return FPO_E_NOMEM;
    memcpy(w, a, an * 4);
    const uint32_t li = fpo_minv(f, b[bn - 1]);
    for (size_t i = dn; i-- > 0;) {
      const uint32_t fac = fpo_mul(f, w[i + bn - 1], li);
      q[i] = fac;
      if (fac == 0) continue;
      for (size_t j = 0; j < bn - 1; ++j) w[i + j] = fpo_sub(f, w[i + j], fpo_mul(f, fac, b[j]));
    }
    size_t lr = rn;
    while (lr && w[lr - 1] == 0) --lr;
    memcpy(r, w, lr * 4);
    *qlen = dn;
    *rlen = lr;
    free(w);
    return FPO_OK;
  }
  /* q = rev(rev(a) inv(rev(b)) mod x^dn) (:53-62) */
  const size_t lrb = bn < dn ? bn : dn;
  uint32_t* ra = (uint32_t*)malloc(dn * 4);
  uint32_t* rb = (uint32_t*)malloc(lrb * 4);
  uint32_t* ib = (uint32_t*)malloc(dn * 4);
  int rc = (!ra || !rb || !ib) ? FPO_E_NOMEM : FPO_OK;
  uint32_t* rq = 0;
  size_t lq = 0;
  if (!rc) {
    for (size_t i = 0; i < dn; ++i) ra[i] = a[an - 1 - i];
    for (size_t i = 0; i < lrb; ++i) rb[i] = b[bn - 1 - i];
    rc = fpo_inv(f, rb, lrb, dn, ib);
  }
  if (!rc) rc = conv_trunc(f, ra, dn, ib, dn, dn, &rq, &lq);
  if (!rc) {
    for (size_t i = 0; i < dn; ++i) q[i] = (dn - 1 - i) < lq ? rq[dn - 1 - i] : 0;  /* resize, reverse */
    *qlen = dn;
  }
  /* r = (a - b q) mod x^rn, trimmed (:63-75) */
  uint32_t* pr = 0;
  size_t lp = 0;
  if (!rc) rc = conv_trunc(f, b, lb, q, dn, rn, &pr, &lp);
  if (!rc) {
    for (size_t i = 0; i < rn; ++i) r[i] = fpo_sub(f, a[i], i < lp ? pr[i] : 0);
    size_t lr = rn;
    while (lr && r[lr - 1] == 0) --lr;
    *rlen = lr;
  }
  free(ra); free(rb); free(ib); free(rq); free(pr);
  return rc;
}
