/* Sanitizer run of the oracle (test infrastructure): every oracle routine on
 * small and edge-case inputs, built with -fsanitize=address,undefined by
 * `make -C oracle sanitize` and run by tests/test_sanitizers.py.  Results are
 * checked against the independent O(n^2) product and the series identities
 * (inv(a) * a = 1, exp(log a) = a, sqrt(a)^2 = a, b q + r = a), so the run
 * fails on a wrong result as well as on a sanitizer report.  Exit status 0
 * when everything holds. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fpx_oracle.h"

static int failures = 0;
#define CHECK(c, what)                                        \
  do {                                                        \
    if (!(c)) {                                               \
      fprintf(stderr, "FAIL %s (line %d)\n", what, __LINE__); \
      ++failures;                                             \
    }                                                         \
  } while (0)

static uint32_t* alloc_u32(size_t n) {
  uint32_t* p = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
  if (!p) abort();
  return p;
}

static void check_field(uint32_t p) {
  fpo_field f;
  CHECK(fpo_field_init(&f, p, 0) == FPO_OK, "field init");
  fpo_mt64 g;
  fpo_mt64_seed(&g, 7 + p);
  for (size_t la = 0; la <= 70; la += 7)
    for (size_t lb = 0; lb <= 70; lb += 11) {
      uint32_t *a = alloc_u32(la), *b = alloc_u32(lb), *c = alloc_u32(la + lb), *d = alloc_u32(la + lb);
      fpo_rnd(&f, &g, a, la);
      fpo_rnd(&f, &g, b, lb);
      size_t n = 0;
      CHECK(fpo_conv(&f, a, la, b, lb, (size_t)-1, c, &n) == FPO_OK, "conv");
      CHECK(n == (la && lb ? la + lb - 1 : 0), "conv length");
      CHECK(fpo_conv_naive(&f, a, la, b, lb, n, d) == FPO_OK, "naive");
      CHECK(memcmp(c, d, n * 4) == 0, "conv == naive");
      if (la && lb) {
        CHECK(fpo_conv_crt(&f, a, la, b, lb, n, 16, c) == FPO_OK, "crt blocks");
        CHECK(memcmp(c, d, n * 4) == 0, "crt == naive");
      }
      free(a); free(b); free(c); free(d);
    }
  for (size_t n = 1; n <= 300; n = n * 3 + 1) {
    uint32_t *a = alloc_u32(n), *x = alloc_u32(n), *y = alloc_u32(n), *z = alloc_u32(2 * n);
    fpo_rnd(&f, &g, a, n);
    a[0] = fpo_from_int(&f, 1);
    size_t m = 0;
    CHECK(fpo_inv(&f, a, n, n, x) == FPO_OK, "inv");
    CHECK(fpo_conv(&f, a, n, x, n, n, z, &m) == FPO_OK && z[0] == fpo_from_int(&f, 1), "a * inv(a) = 1");
    for (size_t i = 1; i < m; ++i) CHECK(z[i] == 0, "a * inv(a) = 1");
    if (n <= p) {
      CHECK(fpo_log(&f, a, n, n, x) == FPO_OK, "log");
      CHECK(fpo_exp(&f, x, n, n, y) == FPO_OK, "exp");
      CHECK(memcmp(y, a, n * 4) == 0, "exp(log a) = a");
    }
    CHECK(fpo_conv(&f, a, n, a, n, n, z, &m) == FPO_OK, "square");
    CHECK(fpo_sqrt_series(&f, z, n, n, x) == FPO_OK, "sqrt");
    CHECK(fpo_conv(&f, x, n, x, n, n, y, &m) == FPO_OK && memcmp(y, z, n * 4) == 0, "sqrt(a)^2 = a");
    CHECK(fpo_poly_pow(&f, a, n, 3, n, x) == FPO_OK, "pow");
    CHECK(fpo_conv(&f, a, n, a, n, n, y, &m) == FPO_OK && fpo_conv(&f, y, n, a, n, n, z, &m) == FPO_OK &&
              memcmp(x, z, n * 4) == 0,
          "pow(a, 3) = a a a");
    size_t lb = n / 2 + 1, lq = 0, lr = 0;
    uint32_t *q = alloc_u32(n), *r = alloc_u32(n);
    CHECK(fpo_divmod(&f, a, n, a, lb, q, &lq, r, &lr) == FPO_OK, "divmod");
    free(q); free(r);
    free(a); free(x); free(y); free(z);
  }
  /* transforms at the sizes the field supports, forward then inverse */
  for (uint32_t n = 2; n <= 4096 && (p - 1) % n == 0; n *= 2) {
    uint32_t *a = alloc_u32(n), *b = alloc_u32(n);
    fpo_rnd(&f, &g, a, n);
    memcpy(b, a, n * 4);
    CHECK(fpo_ntt_forward(&f, b, n) == FPO_OK && fpo_ntt_inverse(&f, b, n) == FPO_OK, "ntt");
    CHECK(memcmp(a, b, n * 4) == 0, "inverse(forward(a)) = a");
    free(a); free(b);
  }
  /* domain errors */
  uint32_t zero[4] = {0, 0, 0, 0}, out[8];
  CHECK(fpo_inv(&f, zero, 4, 4, out) == FPO_E_DOMAIN, "inv(a0 = 0) -> domain");
  CHECK(fpo_log(&f, zero, 4, 4, out) == FPO_E_DOMAIN, "log(a0 = 0) -> domain");
}

int main(void) {
  const uint32_t primes[] = {998244353u, 1000000007u, 167772161u, 469762049u, 754974721u, 1224736769u, 3u, 65537u};
  for (size_t i = 0; i < sizeof primes / sizeof primes[0]; ++i) check_field(primes[i]);
  if (failures) {
    fprintf(stderr, "%d failure(s)\n", failures);
    return 1;
  }
  printf("oracle sanitizer run: ok\n");
  return 0;
}
