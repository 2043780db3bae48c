// C++ drop-in checks: the fpx:: facade (include/fastpoly_amd/fastpoly.hpp)
// used exactly as reference code uses fastpoly, on the reference benchmark's
// inputs (std::mt19937_64, bench/bench.cpp:216-220) against its published
// FNV-1a checksums (bench/bench.cpp:152-159).  Built into a shared object by
// tests/cpp/build.sh; tests/test_facade.py calls fpx_facade_tests on the GPU.
#include <array>
#include <barrier>
#include <bit>
#include <memory>
#include <thread>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>

#include "fastpoly_amd/fastpoly.hpp"

namespace {

int failures = 0;
std::string log_text;

void check(bool ok, const char* what) {
  char line[256];
  std::snprintf(line, sizeof line, "%s %s\n", ok ? "ok  " : "FAIL", what);
  log_text += line;
  if (!ok) ++failures;
}

template <class V>
unsigned long long checksum(const V& a) {
  unsigned long long h = 14695981039346656037ull;
  for (const auto& x : a) {
    h ^= x.raw_val();
    h *= 1099511628211ull;
  }
  return h ^ a.size();
}

template <class M>
std::vector<M> rnd(size_t n, std::mt19937_64& rng) {
  std::vector<M> a(n);
  for (auto& x : a) x = M::from_int(rng() % M::mod);
  return a;
}

template <class F>
bool throws_domain(F&& f) {
  try { f(); } catch (const fpx::poly::domain_error&) { return true; } catch (...) { return false; }
  return false;
}
template <class F>
bool throws_size(F&& f) {
  try { f(); } catch (const fpx::ntt_size_error&) { return true; } catch (...) { return false; }
  return false;
}

// Eight threads ask for cold plans of three sizes at once; each size is built
// once and every thread gets that one plan (tests/test_ntt.cpp:338-378 of the
// reference).  Mont<998244353, 243> is a field no other test touches, so its
// cache starts cold.
void cold_plans_concurrently() {
  using Field = fpx::Mont<998244353u, 243u>;
  using Plan = fpx::NttPlan<Field>;
  constexpr size_t threads = 8;
  constexpr std::array<uint32_t, threads> sizes{256, 512, 2048, 256, 512, 2048, 256, 512};
  std::array<std::shared_ptr<const Plan>, threads> plans;
  std::array<bool, threads> correct{};
  std::barrier start(static_cast<std::ptrdiff_t>(threads));
  std::array<std::thread, threads> workers;
  for (size_t t = 0; t < threads; ++t) {
    workers[t] = std::thread([&, t] {
      start.arrive_and_wait();
      try {
        const uint32_t n = sizes[t];
        if (t & 1) plans[t] = Plan::get(n);
        const auto& borrowed = Plan::get_ref(n);
        if (!(t & 1)) plans[t] = Plan::get(n);
        std::vector<uint32_t> a(n);
        for (uint32_t i = 0; i < n; ++i) a[i] = Field::from_int(uint64_t(i + 1) * (t + 1)).raw_val();
        const auto original = a;
        borrowed.forward_lazy(a.data());
        borrowed.inverse(a.data());
        correct[t] = a == original && plans[t] == Plan::get(n) && &borrowed == plans[t].get();
      } catch (...) {
        correct[t] = false;
      }
    });
  }
  for (auto& w : workers) w.join();
  bool all = true, shared = true;
  for (size_t t = 0; t < threads; ++t) {
    all = all && correct[t] && plans[t] == Plan::get(sizes[t]) && plans[t].get() == &Plan::get_ref(sizes[t]);
    for (size_t u = 0; u < t; ++u) shared = shared && ((plans[t] == plans[u]) == (sizes[t] == sizes[u]));
  }
  check(all, "8 threads, cold plans: one plan per size, transforms round-trip");
  check(shared, "8 threads, cold plans: equal sizes share a plan, different sizes do not");
}

void run_all() {
  using M = fpx::mod998244353;
  {
    const size_t n = 65536;
    std::mt19937_64 rng(2 + n);
    const fpx::Poly<M> a(rnd<M>(n, rng)), b(rnd<M>(n, rng));
    check(checksum(a * b) == 0x841755dfc72581beull, "Poly operator* 65536 (bench conv checksum)");
    check(checksum(a.mul(b, n)) == 0x6548fcfc353988f8ull, "Poly::mul(b, n) 65536 (conv-trunc checksum)");
    check(checksum(fpx::poly::conv(a, a)) == 0x7069bfe1ae6aad95ull, "conv(a, a) (square checksum)");
  }
  {
    const size_t n = 65536;
    std::mt19937_64 rng(3 + n);
    fpx::Poly<M> a(rnd<M>(n, rng));
    a[0] = M::from_int(1);
    check(checksum(a.inv(n)) == 0x81c1f8a79506b8d2ull, "inv 65536");
    check(checksum(a.log(n)) == 0x605ffdc345fa9039ull, "log 65536");
    fpx::Poly<M> e = a;
    e[0] = M();
    check(checksum(e.exp(n)) == 0xff4d4e75dae0ae43ull, "exp 65536");
    const fpx::Poly<M> sq = a.mul(a, n);
    check(checksum(sq.sqrt(n)) == 0xd32ea1d3f38db84aull, "sqrt 65536");
    check(checksum(a.pow(1000003, n)) == 0x6e2ecc1757af26c1ull, "pow(1000003) 65536");
    const auto one = a.mul(a.inv(n), n);
    bool unit = one[0] == M::from_int(1);
    for (size_t i = 1; i < one.size(); ++i) unit = unit && one[i].is_zero();
    check(unit, "a * inv(a) == 1 mod x^n");
  }
  {
    using Q = fpx::mod1000000007;
    const size_t n = 16384;
    std::mt19937_64 rng(2 + n);
    const auto a = rnd<Q>(n, rng), b = rnd<Q>(n, rng);
    check(checksum(fpx::poly::conv(a, b)) == 0x80dc26cc95895f00ull, "conv mod 1e9+7 16384 (CRT)");
    std::mt19937_64 r3(3 + n);
    auto s = rnd<Q>(n, r3);
    s[0] = Q::from_int(1);
    check(checksum(fpx::poly::inv(s, n)) == 0x0b6d10016dec1f5bull, "inv mod 1e9+7 16384");
  }
  {
    const uint32_t n = 65536;
    std::mt19937_64 rng(1 + n);
    const auto in = rnd<M>(n, rng);
    std::vector<uint32_t> a(n);
    for (uint32_t i = 0; i < n; ++i) a[i] = in[i].raw_val();
    const auto& plan = fpx::NttPlan<M>::get_ref(n);
    plan.forward(a.data());
    std::vector<M> fa(n);
    for (uint32_t i = 0; i < n; ++i) fa[i] = M::raw(a[i]);
    check(checksum(fa) == 0xb83db2bcacb85d37ull, "NttPlan::forward 65536 (ntt-forward checksum)");
    plan.inverse(a.data());
    bool same = true;
    for (uint32_t i = 0; i < n; ++i) same = same && a[i] == in[i].raw_val();
    check(same, "NttPlan::inverse roundtrip");
  }
  {
    const size_t n = 65536;
    std::mt19937_64 rng(4 + n);
    const fpx::Poly<M> a(rnd<M>(n, rng)), b(rnd<M>(n / 2, rng));
    const auto qr = a.divmod(b);
    const unsigned long long h = checksum(qr.first) ^ std::rotl(checksum(qr.second), 17);
    check(h == 0x2762041dae10ec54ull, "divmod 65536 (bench checksum)");
    check(checksum(a.div(b)) == checksum(qr.first) && checksum(a.mod(b)) == checksum(qr.second), "div / mod");
    check(throws_domain([&] { a.divmod(std::vector<M>{M(), M()}); }), "divmod by zero -> domain_error");
  }
  check(throws_domain([] { fpx::poly::inv(std::vector<M>{M(), M(1)}, 4); }), "inv(a0 == 0) -> domain_error");
  check(throws_domain([] { fpx::poly::log(std::vector<M>{M(2)}, 4); }), "log(a0 != 1) -> domain_error");
  check(throws_domain([] { fpx::poly::exp(std::vector<M>{M(1)}, 4); }), "exp(a0 != 0) -> domain_error");
  check(throws_domain([] { fpx::poly::sqrt(std::vector<M>{M(), M(1)}, 4); }), "sqrt odd valuation -> domain_error");
  check(throws_size([] { fpx::NttPlan<M>::get(3); }), "NttPlan::get(3) -> ntt_size_error");
  check(throws_size([] { fpx::NttPlan<M>::get(1u << 24); }), "NttPlan::get(2^24) -> ntt_size_error");
  check(fpx::poly::conv(std::vector<M>{}, std::vector<M>{M(1)}).empty(), "conv with an empty operand");
  const auto st = fpx::scratch_memory_stats();
  check(st.system_allocations >= 1, "scratch_memory_stats reports the device arena");
  {
    // Plan identity (tests/test_ntt.cpp:197-203 of the reference).
    const auto owning = fpx::NttPlan<M>::get(1024);
    check(owning == fpx::NttPlan<M>::get(1024), "NttPlan::get(n) == NttPlan::get(n)");
    check(&fpx::NttPlan<M>::get_ref(1024) == owning.get(), "&NttPlan::get_ref(n) == NttPlan::get(n).get()");
    check(fpx::NttPlan<M>::get_ref(1024).size() == 1024u, "NttPlan::get_ref(n).size()");
  }
  cold_plans_concurrently();
}

}  // namespace

extern "C" int fpx_facade_tests(char* report, size_t cap) {
  failures = 0;
  log_text.clear();
  try {
    run_all();
  } catch (const std::exception& e) {
    log_text += std::string("FAIL exception: ") + e.what() + "\n";
    ++failures;
  }
  if (report && cap) {
    std::strncpy(report, log_text.c_str(), cap - 1);
    report[cap - 1] = 0;
  }
  return failures;
}

// Run in a process whose arena was capped at a few MiB (fpxg_init): a call
// whose scratch does not fit raises std::bad_alloc (the reference's
// allocation failure) and leaves the caller's data as it was; a call that
// fits still works afterwards.
extern "C" int fpx_facade_nomem_tests(char* report, size_t cap) {
  failures = 0;
  log_text.clear();
  try {
    using M = fpx::mod998244353;
    std::mt19937_64 rng(99);
    const fpx::Poly<M> big(rnd<M>(size_t(1) << 21, rng));
    bool threw = false;
    try {
      (void)(big * big);
    } catch (const std::bad_alloc&) {
      threw = true;
    }
    check(threw, "conv past the arena cap -> std::bad_alloc");
    threw = false;
    fpx::Poly<M> inv_in = big;
    inv_in[0] = M::from_int(1);
    try {
      (void)inv_in.inv(size_t(1) << 21);
    } catch (const std::bad_alloc&) {
      threw = true;
    }
    check(threw, "inv past the arena cap -> std::bad_alloc");
    const fpx::Poly<M> x(rnd<M>(300, rng)), y(rnd<M>(200, rng));
    const auto z = x * y;
    bool ok = z.size() == 499;
    for (size_t k = 0; ok && k < z.size(); k += 37) {
      M s{};
      for (size_t i = 0; i <= k; ++i)
        if (i < x.size() && k - i < y.size()) s += x[i] * y[k - i];
      ok = s == z[k];
    }
    check(ok, "a product that fits still works after the failures");
  } catch (const std::exception& e) {
    log_text += std::string("FAIL exception: ") + e.what() + "\n";
    ++failures;
  }
  if (report && cap) {
    std::strncpy(report, log_text.c_str(), cap - 1);
    report[cap - 1] = 0;
  }
  return failures;
}
