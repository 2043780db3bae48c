// fastpoly_amd — C++20 drop-in for fastpoly's NTT convolution API, computed on
// an MI355X through the fpxg C ABI (include/fpxg.h, libfpxg.so).
//
// Same names and contracts as the reference headers:
//   fpx::Mont<Mod, Root>                       include/fastpoly/modint.hpp:57-180
//   fpx::poly::conv / inv / log / exp / sqrt   detail/poly/convolution.hpp:145,
//                                              inverse.hpp:163, exponential.hpp:85/107,
//                                              square_root.hpp:12
//   fpx::poly::pow / divmod / div / mod        power.hpp:17, division.hpp:79-100
//   fpx::Poly<M>                               include/fastpoly/poly.hpp:42-112
//   fpx::NttPlan<M>::get / get_ref / forward...  include/fastpoly/ntt.hpp:130-192
//   fpx::release_scratch_memory / scratch_memory_stats   memory.hpp:251-258
// Limbs are passed zero-copy (sizeof(Mont) == 4); results are bit-identical.
// Errors: poly::domain_error, ntt_size_error, std::length_error,
// std::bad_alloc and std::runtime_error, raised in the reference's order.
#ifndef FASTPOLY_AMD_FASTPOLY_HPP
#define FASTPOLY_AMD_FASTPOLY_HPP

#include <algorithm>
#include <array>
#include <bit>
#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <new>
#include <ostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "fpxg.h"

namespace fpx {

// ---------------------------------------------------------------------------
// Field element: Montgomery limb x*2^32 mod Mod (the reference representation).
// ---------------------------------------------------------------------------
namespace detail {
constexpr uint32_t neg_inverse_2_32(uint32_t m) {
  uint32_t x = 1;
  for (int i = 0; i < 5; ++i) x *= 2u - m * x;
  return 0u - x;
}
constexpr uint64_t pow_plain(uint64_t a, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  a %= m;
  while (e) {
    if (e & 1) r = r * a % m;
    a = a * a % m;
    e >>= 1;
  }
  return r;
}
// Smallest primitive root: the discovery rule of modint.hpp:28-52.
constexpr uint32_t smallest_root(uint32_t p) {
  if (p == 2) return 1;
  uint32_t f[12] = {};
  int k = 0;
  uint32_t rest = p - 1;
  for (uint32_t d = 2; uint64_t(d) * d <= rest; ++d)
    if (rest % d == 0) {
      f[k++] = d;
      while (rest % d == 0) rest /= d;
    }
  if (rest > 1) f[k++] = rest;
  for (uint32_t g = 2; g < p; ++g) {
    bool ok = true;
    for (int i = 0; i < k && ok; ++i) ok = pow_plain(g, (p - 1) / f[i], p) != 1;
    if (ok) return g;
  }
  return 0;
}
}  // namespace detail

template <uint32_t Mod, uint32_t Root = 0>
class Mont {
  static_assert(Mod >= 2 && Mod < (1u << 31), "modulus must be a prime in [2, 2^31)");

 public:
  static constexpr uint32_t mod = Mod;
  static constexpr uint32_t primitive_root = Root != 0 ? Root : detail::smallest_root(Mod);
  static constexpr uint32_t ninv = detail::neg_inverse_2_32(Mod);
  static constexpr uint32_t one = Mod == 2 ? 1u : uint32_t((uint64_t(1) << 32) % Mod);
  static constexpr uint32_t r2 = uint32_t(uint64_t(one) * one % Mod);
  static constexpr uint32_t inv2 = (Mod + 1) / 2;

  constexpr Mont() = default;
  template <class I, class = std::enable_if_t<std::is_integral_v<I> && !std::is_same_v<I, bool>>>
  constexpr Mont(I x) : v_(enter(uint64_t(x))) {}

  static constexpr Mont from_int(uint64_t x) { return raw(enter(x)); }
  static constexpr Mont raw(uint32_t limb) {
    Mont m;
    m.v_ = limb;
    return m;
  }
  static constexpr uint32_t reduce(uint64_t t) {
    if constexpr (Mod == 2) {
      return uint32_t(t & 1u);
    } else {
      const uint32_t q = uint32_t(t) * ninv;
      const uint64_t u = (t + uint64_t(q) * Mod) >> 32;
      return uint32_t(u >= Mod ? u - Mod : u);
    }
  }
  constexpr uint32_t val() const { return reduce(v_); }
  constexpr uint32_t raw_val() const { return v_; }

  friend constexpr Mont operator+(Mont a, Mont b) {
    const uint32_t s = a.v_ + b.v_;
    return raw(s >= Mod ? s - Mod : s);
  }
  friend constexpr Mont operator-(Mont a, Mont b) { return raw(a.v_ >= b.v_ ? a.v_ - b.v_ : a.v_ + Mod - b.v_); }
  friend constexpr Mont operator-(Mont a) { return raw(a.v_ ? Mod - a.v_ : 0); }
  friend constexpr Mont operator*(Mont a, Mont b) { return raw(reduce(uint64_t(a.v_) * b.v_)); }
  friend constexpr Mont operator/(Mont a, Mont b) { return a * b.inv(); }
  constexpr Mont& operator+=(Mont b) { return *this = *this + b; }
  constexpr Mont& operator-=(Mont b) { return *this = *this - b; }
  constexpr Mont& operator*=(Mont b) { return *this = *this * b; }
  constexpr Mont& operator/=(Mont b) { return *this = *this / b; }
  friend constexpr bool operator==(Mont a, Mont b) { return a.v_ == b.v_; }
  friend constexpr bool operator!=(Mont a, Mont b) { return a.v_ != b.v_; }
  constexpr bool is_zero() const { return v_ == 0; }
  constexpr bool operator!() const { return v_ == 0; }
  constexpr explicit operator bool() const { return v_ != 0; }

  constexpr Mont pow(uint64_t e) const {
    Mont r = from_int(1), b = *this;
    for (; e; e >>= 1, b *= b)
      if (e & 1) r *= b;
    return r;
  }
  constexpr Mont inv() const {
    if constexpr (Mod == 2) return *this;
    else return pow(Mod - 2);
  }
  // Square root with the reference's deterministic choice: Euler's criterion,
  // a^((p+1)/4) for p = 3 mod 4, otherwise Tonelli-Shanks seeded by the
  // first non-residue z >= 2.
  bool sqrt(Mont& out) const {
    if constexpr (Mod == 2) {
      out = *this;
      return true;
    } else {
      const Mont o = from_int(1);
      if (v_ == 0) { out = *this; return true; }
      if (pow((Mod - 1) / 2) != o) return false;
      if (Mod % 4 == 3) { out = pow((uint64_t(Mod) + 1) / 4); return true; }
      uint64_t q = Mod - 1;
      int s = 0;
      for (; q % 2 == 0; q /= 2) ++s;
      Mont z = from_int(2);
      while (z.pow((Mod - 1) / 2) != -o) z += o;
      Mont c = z.pow(q), x = pow((q + 1) / 2), t = pow(q);
      for (int m = s; t != o;) {
        int i = 0;
        for (Mont t2 = t; t2 != o; t2 *= t2) ++i;
        const Mont b = c.pow(uint64_t(1) << (m - i - 1));
        x *= b;
        c = b * b;
        t *= c;
        m = i;
      }
      out = x;
      return true;
    }
  }
  std::string str() const { return std::to_string(val()); }

 private:
  static constexpr uint32_t enter(uint64_t x) { return uint32_t(uint64_t(uint32_t(x % Mod)) * one % Mod); }
  uint32_t v_ = 0;
};

template <uint32_t M, uint32_t R>
std::ostream& operator<<(std::ostream& os, const Mont<M, R>& a) {
  return os << a.val();
}

using mod998244353 = Mont<998244353u, 3>;
using mod1004535809 = Mont<1004535809u, 3>;
using mod469762049 = Mont<469762049u, 3>;
using mod167772161 = Mont<167772161u, 3>;
using mod754974721 = Mont<754974721u, 11>;
using mod1224736769 = Mont<1224736769u, 3>;
template <uint32_t Mod>
using ModInt = Mont<Mod>;
using mod1000000007 = Mont<1000000007u>;

class ntt_size_error : public std::runtime_error {
 public:
  explicit ntt_size_error(const std::string& s) : std::runtime_error(s) {}
};

namespace poly {
class domain_error : public std::runtime_error {
 public:
  explicit domain_error(const std::string& s) : std::runtime_error(s) {}
};
template <class M>
using vec = std::vector<M>;
}  // namespace poly

namespace detail {
// Map a C-ABI status onto the reference's exception types.
inline void raise_status(int rc, const char* what) {
  if (rc == FPXG_OK) return;
  const std::string msg = std::string(what) + ": " + fpxg_last_error();
  switch (rc) {
    case FPXG_E_DOMAIN: throw poly::domain_error(msg);
    case FPXG_E_SIZE: throw ntt_size_error(msg);
    case FPXG_E_NOMEM: throw std::bad_alloc();
    default: throw std::runtime_error(msg);
  }
}
template <class M>
const uint32_t* limbs(const std::vector<M>& v) {
  static_assert(sizeof(M) == sizeof(uint32_t), "Mont must be a single 32-bit limb");
  return reinterpret_cast<const uint32_t*>(v.data());
}
template <class M>
uint32_t* limbs(std::vector<M>& v) {
  return reinterpret_cast<uint32_t*>(v.data());
}
}  // namespace detail

// ---------------------------------------------------------------------------
// Polynomial operations on the device.
// ---------------------------------------------------------------------------
namespace poly {

template <class M>
vec<M> conv(const vec<M>& a, const vec<M>& b, size_t n = SIZE_MAX) {
  if (a.empty() || b.empty() || n == 0) return {};
  const size_t la = std::min(a.size(), n), lb = std::min(b.size(), n);
  if (la > SIZE_MAX - (lb - 1)) throw std::length_error("poly::conv: result length overflows size_t");
  vec<M> out(std::min(la + lb - 1, n));
  size_t got = 0;
  detail::raise_status(fpxg_conv_u32(M::mod, detail::limbs(a), a.size(), detail::limbs(b), b.size(), n,
                                     detail::limbs(out), out.size(), &got),
                       "poly::conv");
  out.resize(got);
  return out;
}

namespace detail_series {
template <class M>
vec<M> run(int op, const vec<M>& a, size_t n, const char* what) {
  if (n == 0) return {};
  vec<M> out(n);
  fpx::detail::raise_status(fpxg_series_u32(op, M::mod, fpx::detail::limbs(a), a.size(), n, fpx::detail::limbs(out)),
                            what);
  return out;
}
}  // namespace detail_series

template <class M>
vec<M> inv(const vec<M>& a, size_t n) { return detail_series::run(FPXG_INV, a, n, "poly::inv"); }
template <class M>
vec<M> log(const vec<M>& a, size_t n) { return detail_series::run(FPXG_LOG, a, n, "poly::log"); }
template <class M>
vec<M> exp(const vec<M>& a, size_t n) { return detail_series::run(FPXG_EXP, a, n, "poly::exp"); }
template <class M>
vec<M> sqrt(const vec<M>& a, size_t n) { return detail_series::run(FPXG_SQRT, a, n, "poly::sqrt"); }

template <class M>
void trim(vec<M>& a) {
  while (!a.empty() && a.back().is_zero()) a.pop_back();
}

/// a^k mod x^n (detail/poly/power.hpp:17-87).
template <class M>
vec<M> pow(const vec<M>& a, uint64_t k, size_t n) {
  if (n == 0) return {};
  vec<M> out(n);
  fpx::detail::raise_status(
      fpxg_pow_u32(M::mod, fpx::detail::limbs(a), a.size(), k, n, fpx::detail::limbs(out)), "poly::pow");
  return out;
}

/// Euclidean division a = b q + r, deg r < deg b, both trimmed (division.hpp:79-100).
template <class M>
std::pair<vec<M>, vec<M>> divmod(const vec<M>& a, const vec<M>& b) {
  vec<M> q(std::max<size_t>(a.size(), 1)), r(std::max<size_t>(a.size(), 1));
  size_t lq = 0, lr = 0;
  fpx::detail::raise_status(fpxg_divmod_u32(M::mod, fpx::detail::limbs(a), a.size(), fpx::detail::limbs(b),
                                            b.size(), fpx::detail::limbs(q), q.size(), &lq, fpx::detail::limbs(r),
                                            r.size(), &lr),
                            "poly::divmod");
  q.resize(lq);
  r.resize(lr);
  return {std::move(q), std::move(r)};
}
template <class M>
vec<M> div(const vec<M>& a, const vec<M>& b) { return divmod(a, b).first; }
template <class M>
vec<M> mod(const vec<M>& a, const vec<M>& b) { return divmod(a, b).second; }

}  // namespace poly

// ---------------------------------------------------------------------------
// Poly<M>: std::vector<M> with method syntax (poly.hpp:42-112, the NTT path).
// ---------------------------------------------------------------------------
template <class M>
class Poly : public std::vector<M> {
 public:
  using base = std::vector<M>;
  using value_type = M;
  Poly() = default;
  Poly(const base& v) : base(v) {}
  Poly(base&& v) : base(std::move(v)) {}
  Poly(std::initializer_list<M> il) : base(il) {}
  explicit Poly(size_t n, M v = M()) : base(n, v) {}
  size_t degree() const { return this->empty() ? 0 : this->size() - 1; }
  bool is_zero() const {
    for (const auto& x : *this)
      if (!x.is_zero()) return false;
    return true;
  }
  Poly& trim() { poly::trim(*this); return *this; }
  Poly operator*(const base& b) const { return poly::conv<M>(*this, b); }
  Poly& operator*=(const base& b) { return *this = poly::conv<M>(*this, b); }
  Poly mul(const base& b, size_t n) const { return poly::conv<M>(*this, b, n); }
  Poly inv(size_t n) const { return poly::inv<M>(*this, n); }
  Poly log(size_t n) const { return poly::log<M>(*this, n); }
  Poly exp(size_t n) const { return poly::exp<M>(*this, n); }
  Poly sqrt(size_t n) const { return poly::sqrt<M>(*this, n); }
  Poly pow(uint64_t k, size_t n) const { return poly::pow<M>(*this, k, n); }
  std::pair<Poly, Poly> divmod(const base& b) const {
    auto r = poly::divmod<M>(*this, b);
    return {Poly(std::move(r.first)), Poly(std::move(r.second))};
  }
  Poly div(const base& b) const { return poly::div<M>(*this, b); }
  Poly mod(const base& b) const { return poly::mod<M>(*this, b); }
  std::vector<uint32_t> to_ints() const {
    std::vector<uint32_t> r(this->size());
    for (size_t i = 0; i < this->size(); ++i) r[i] = (*this)[i].val();
    return r;
  }
  static Poly from_ints(const std::vector<uint32_t>& v) {
    Poly r(v.size());
    for (size_t i = 0; i < v.size(); ++i) r[i] = M::from_int(v[i]);
    return r;
  }
};

// ---------------------------------------------------------------------------
// NttPlan<M>: device-backed transforms of one size (ntt.hpp:130-192).
// ---------------------------------------------------------------------------
template <class M>
class NttPlan {
 public:
  // One plan per size for the whole process, built once even when many
  // threads ask for a cold size together: get(n) == get(n) and
  // &get_ref(n) == get(n).get(), as the reference's cache guarantees
  // (ntt.hpp:130-135, 250-276).
  static std::shared_ptr<const NttPlan> get(uint32_t n) { return cached(n); }
  static const NttPlan& get_ref(uint32_t n) { return *cached(n); }
  uint32_t size() const { return n_; }
  int log_size() const {
    int k = 0;
    while ((uint32_t(1) << k) < n_) ++k;
    return k;
  }
  void forward(uint32_t* a) const { run(a, FPXG_NTT_FORWARD, 0); }
  void forward_lazy(uint32_t* a) const { run(a, FPXG_NTT_FORWARD, 0); }
  void forward_prefix(uint32_t* a, size_t count) const { run(a, FPXG_NTT_FORWARD_PREFIX, count); }
  void inverse(uint32_t* a) const { run(a, FPXG_NTT_INVERSE, 0); }
  void inverse_prefix(uint32_t* a, size_t count) const { run(a, FPXG_NTT_INVERSE_PREFIX, count); }
  void inverse_highhalf(uint32_t* a) const { run(a, FPXG_NTT_INVERSE_HIGHHALF, 0); }

 private:
  explicit NttPlan(uint32_t n) : n_(n) { validate(n); }
  static void validate(uint32_t n) {
    if (n < 2 || (n & (n - 1)) != 0) throw ntt_size_error("NTT size must be a power of two");
    uint64_t t = M::mod - 1;
    int k = 0;
    while (t && t % 2 == 0) { t /= 2; ++k; }
    if (M::mod == 2 || (uint64_t(1) << k) < n) throw ntt_size_error("NTT size exceeds 2^v2(mod-1) for this modulus");
  }
  static const std::shared_ptr<const NttPlan>& cached(uint32_t n) {
    validate(n);
    struct Slot {
      std::once_flag built;
      std::shared_ptr<const NttPlan> plan;
    };
    static std::array<Slot, 32> slots;
    Slot& s = slots[std::countr_zero(n)];
    std::call_once(s.built, [&s, n] { s.plan = std::shared_ptr<const NttPlan>(new NttPlan(n)); });
    return s.plan;
  }
  void run(uint32_t* a, int variant, size_t count) const {
    detail::raise_status(fpxg_ntt_u32(M::mod, M::primitive_root, n_, variant, count, a), "NttPlan");
  }
  uint32_t n_;
};

struct ScratchMemoryStats {
  size_t cached_bytes = 0, cached_blocks = 0, system_allocations = 0, cache_hits = 0;
};
inline void release_scratch_memory() { detail::raise_status(fpxg_release_scratch(), "release_scratch_memory"); }
inline ScratchMemoryStats scratch_memory_stats() {
  fpxg_stats s{};
  detail::raise_status(fpxg_arena_stats(&s), "scratch_memory_stats");
  return {s.cached_bytes, s.cached_blocks, s.system_allocations, s.cache_hits};
}
inline const char* simd_backend() { return "hip-gfx950"; }

}  // namespace fpx

#endif  // FASTPOLY_AMD_FASTPOLY_HPP
