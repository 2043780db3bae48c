// Driver of the ASan + UBSan build of the library's host code
// (tests/cpp/build_san.sh: the HIP device code is built as usual, the host
// code with -fsanitize=address,undefined).  Without a GPU it runs the error
// paths of the C ABI; with one it runs the host runtime end to end: products
// on every path (direct, tail repair, native four-step, split, three-prime),
// the Newton series with graph capture and replay, the device arena, merged
// batches, the server and the no-allocation product, each result compared
// with an O(n^2) product or a series identity computed here.  Exit status 0
// when every check holds and the sanitizers report nothing.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "fpxg.h"

static int failures = 0;
#define CHECK(c, what)                                                  \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "FAIL %s (line %d)\n", what, __LINE__);      \
      ++failures;                                                       \
    }                                                                   \
  } while (0)

namespace {

// Montgomery limbs of p (x * 2^32 mod p) and their plain products.
struct Field {
  uint32_t p;
  uint32_t from(uint64_t x) const { return (uint32_t)(((x % p) << 32) % p); }
  uint32_t to(uint32_t limb) const {
    uint64_t r = 1, b = ((uint64_t)1 << 32) % p, e = p - 2;  // (2^32)^-1
    for (; e; e >>= 1, b = b * b % p)
      if (e & 1) r = r * b % p;
    return (uint32_t)((uint64_t)limb * r % p);
  }
};

std::vector<uint32_t> naive(const Field& F, const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
  std::vector<uint64_t> acc(a.size() + b.size() - 1, 0);
  std::vector<uint32_t> ap(a.size()), bp(b.size());
  for (size_t i = 0; i < a.size(); ++i) ap[i] = F.to(a[i]);
  for (size_t j = 0; j < b.size(); ++j) bp[j] = F.to(b[j]);
  for (size_t i = 0; i < a.size(); ++i)
    for (size_t j = 0; j < b.size(); ++j) acc[i + j] = (acc[i + j] + (uint64_t)ap[i] * bp[j]) % F.p;
  std::vector<uint32_t> r(acc.size());
  for (size_t k = 0; k < r.size(); ++k) r[k] = F.from(acc[k]);
  return r;
}

std::vector<uint32_t> rnd(const Field& F, size_t n, std::mt19937_64& g) {
  std::vector<uint32_t> a(n);
  for (auto& x : a) x = F.from(g() % F.p);
  return a;
}

void products(uint32_t p, std::mt19937_64& g) {
  const Field F{p};
  const size_t shapes[][2] = {{1, 1}, {5000, 60}, {3000, 2000}, {(1u << 12) + 3, 1u << 12}, {20000, 17000}};
  for (auto& s : shapes) {
    auto a = rnd(F, s[0], g), b = rnd(F, s[1], g);
    std::vector<uint32_t> out(s[0] + s[1] - 1);
    size_t got = 0;
    CHECK(fpxg_conv_u32(p, a.data(), a.size(), b.data(), b.size(), SIZE_MAX, out.data(), out.size(), &got) == 0,
          "conv");
    CHECK(got == out.size() && out == naive(F, a, b), "conv == O(n^2) product");
    // the same through device buffers and the no-allocation entry point
    uint32_t* da = (uint32_t*)fpxg_dev_alloc(a.size() * 4);
    uint32_t* db = (uint32_t*)fpxg_dev_alloc(b.size() * 4);
    uint32_t* dc = (uint32_t*)fpxg_dev_alloc(out.size() * 4);
    CHECK(da && db && dc, "dev_alloc");
    fpxg_memcpy_h2d(da, a.data(), a.size() * 4);
    fpxg_memcpy_h2d(db, b.data(), b.size() * 4);
    CHECK(fpxg_conv_dev_noalloc(p, da, a.size(), db, b.size(), SIZE_MAX, dc, out.size(), &got) == 0, "noalloc");
    std::vector<uint32_t> back(out.size());
    fpxg_memcpy_d2h(back.data(), dc, back.size() * 4);
    CHECK(back == out, "noalloc == conv");
    fpxg_dev_free(da);
    fpxg_dev_free(db);
    fpxg_dev_free(dc);
  }
}

void series(uint32_t p, std::mt19937_64& g) {
  const Field F{p};
  for (size_t n : {size_t(7), size_t(600), size_t(5000), size_t(70000)}) {
    auto a = rnd(F, n, g);
    a[0] = F.from(1);
    std::vector<uint32_t> inv(n), lg(n), ex(n), prod(2 * n);
    size_t got = 0;
    for (int rep = 0; rep < 3; ++rep) {  // the second and third calls capture and replay graphs
      CHECK(fpxg_series_u32(FPXG_INV, p, a.data(), n, n, inv.data()) == 0, "inv");
      CHECK(fpxg_conv_u32(p, a.data(), n, inv.data(), n, n, prod.data(), prod.size(), &got) == 0, "conv");
      bool unit = prod[0] == F.from(1);
      for (size_t i = 1; i < n; ++i) unit = unit && prod[i] == 0;
      CHECK(unit, "a * inv(a) = 1 mod x^n");
    }
    CHECK(fpxg_series_u32(FPXG_LOG, p, a.data(), n, n, lg.data()) == 0, "log");
    CHECK(fpxg_series_u32(FPXG_EXP, p, lg.data(), n, n, ex.data()) == 0, "exp");
    CHECK(ex == a, "exp(log a) = a");
    CHECK(fpxg_conv_u32(p, a.data(), n, a.data(), n, n, prod.data(), prod.size(), &got) == 0, "square");
    std::vector<uint32_t> rt(n), sq(n);
    CHECK(fpxg_series_u32(FPXG_SQRT, p, prod.data(), n, n, rt.data()) == 0, "sqrt");
    CHECK(fpxg_conv_u32(p, rt.data(), n, rt.data(), n, n, sq.data(), n, &got) == 0 &&
              std::equal(sq.begin(), sq.end(), prod.begin()),
          "sqrt(a)^2 = a");
  }
}

void batch_and_server(std::mt19937_64& g) {
  const uint32_t ps[2] = {998244353u, 1000000007u};
  std::vector<std::vector<uint32_t>> ins, outs;
  std::vector<fpxg_req> reqs;
  for (int i = 0; i < 24; ++i) {
    const Field F{ps[i % 2]};
    const size_t la = 50 + (size_t)(g() % 3000), lb = 40 + (size_t)(g() % 2000);
    ins.push_back(rnd(F, la, g));
    ins.push_back(rnd(F, lb, g));
    outs.emplace_back(la + lb - 1);
  }
  for (int i = 0; i < 24; ++i) {
    fpxg_req q{};
    q.op = FPXG_CONV; q.p = ps[i % 2];
    q.a = ins[2 * i].data(); q.la = ins[2 * i].size();
    q.b = ins[2 * i + 1].data(); q.lb = ins[2 * i + 1].size();
    q.n = SIZE_MAX; q.out = outs[i].data(); q.out_cap = outs[i].size();
    reqs.push_back(q);
  }
  CHECK(fpxg_batch_u32(reqs.data(), reqs.size()) == 0, "batch");
  for (int i = 0; i < 24; ++i)
    CHECK(reqs[i].status == 0 && outs[i] == naive(Field{ps[i % 2]}, ins[2 * i], ins[2 * i + 1]), "batch request");
  fpxg_server* s = fpxg_server_create(0, 0);
  CHECK(s != nullptr, "server_create");
  if (!s) return;
  std::vector<std::vector<uint32_t>> souts(24);
  std::vector<uint64_t> ids(24);
  for (int i = 0; i < 24; ++i) {
    souts[i].assign(outs[i].size(), 0);
    reqs[i].out = souts[i].data();
    CHECK(fpxg_server_submit(s, &reqs[i], &ids[i]) == 0, "server_submit");
  }
  for (int i = 0; i < 24; ++i) {
    fpxg_result r{};
    CHECK(fpxg_server_wait(s, ids[i], &r) == 0 && r.status == 0 && souts[i] == outs[i], "served == batch");
  }
  fpxg_req bad{};
  bad.op = FPXG_INV; bad.p = 998244353u; bad.la = 5; bad.n = 4;  // null input with a length
  uint64_t id = 0;
  CHECK(fpxg_server_submit(s, &bad, &id) == FPXG_E_ARG, "server rejects a null input");
  fpxg_server_destroy(s);
}

}  // namespace

int main() {
  CHECK(std::strlen(fpxg_strerror(FPXG_E_DOMAIN)) > 0 && std::strlen(fpxg_strerror(99)) > 0, "strerror");
  const int rc = fpxg_init(0, 0);
  if (rc != 0) {
    // No device: every entry point reports FPXG_E_DEVICE and writes nothing.
    uint32_t a[4] = {1, 2, 3, 4}, out[8] = {7, 7, 7, 7, 7, 7, 7, 7};
    size_t got = 123;
    CHECK(rc == FPXG_E_DEVICE, "init without a device");
    CHECK(fpxg_conv_u32(998244353u, a, 4, a, 4, SIZE_MAX, out, 8, &got) == FPXG_E_DEVICE && out[0] == 7, "conv");
    CHECK(fpxg_series_u32(FPXG_INV, 998244353u, a, 4, 4, out) == FPXG_E_DEVICE && out[0] == 7, "series");
    CHECK(fpxg_server_create(0, 0) == nullptr, "server without a device");
    CHECK(std::strlen(fpxg_last_error()) > 0, "last_error");
    std::printf("%s: sanitizer run without a device: %d failure(s)\n", failures ? "FAIL" : "ok", failures);
    return failures ? 1 : 0;
  }
  std::mt19937_64 g(20261204);
  products(998244353u, g);
  products(1000000007u, g);
  products(65537u, g);  // small largest transform: the split product
  series(998244353u, g);
  series(1000000007u, g);
  batch_and_server(g);
  fpxg_stats st{};
  CHECK(fpxg_arena_stats(&st) == 0 && st.system_allocations >= 1, "arena stats");
  CHECK(fpxg_release_scratch() == 0, "release_scratch");
  fpxg_shutdown();
  std::printf("%s: sanitizer run on the device: %d failure(s)\n", failures ? "FAIL" : "ok", failures);
  return failures ? 1 : 0;
}
