// C-ABI shim over the reference fastpoly headers, compiled IN PLACE from
// /root/reference/include by oracle/Makefile (target `ref`).  Test
// infrastructure only: it validates the C restatement in fpx_oracle.c and is
// the "reference" CPU baseline of bench.py.  No reference source is copied;
// this file only instantiates the public API of fastpoly.hpp
// (detail/poly/convolution.hpp:145, inverse.hpp:163, exponential.hpp:85/107,
// square_root.hpp:12, power.hpp:17, division.hpp:79, ntt.hpp:130-192) for the
// moduli the tests use.
#include <cstdint>
#include <cstring>
#include <exception>
#include <new>
#include <vector>

#include "fastpoly/fastpoly.hpp"

namespace {

template <class M>
std::vector<M> load(const uint32_t* a, size_t n) {
  std::vector<M> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = M::raw(a[i]);
  return v;
}
template <class M>
void store(const std::vector<M>& v, uint32_t* out) {
  for (size_t i = 0; i < v.size(); ++i) out[i] = v[i].raw_val();
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const fpx::poly::domain_error&) {
    return 1;
  } catch (const fpx::ntt_size_error&) {
    return 2;
  } catch (const std::length_error&) {
    return 2;
  } catch (const std::bad_alloc&) {
    return 3;
  } catch (...) {
    return 5;
  }
}

template <class M>
int conv_t(const uint32_t* a, size_t la, const uint32_t* b, size_t lb, size_t n, uint32_t* out,
           size_t* out_len) {
  return guarded([&] {
    auto va = load<M>(a, la), vb = load<M>(b, lb);
    auto r = fpx::poly::conv(va, vb, n);
    store(r, out);
    *out_len = r.size();
  });
}

template <class M>
int series_t(int op, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  return guarded([&] {
    auto va = load<M>(a, la);
    std::vector<M> r;
    switch (op) {
      case 0: r = fpx::poly::inv(va, n); break;
      case 1: r = fpx::poly::log(va, n); break;
      case 2: r = fpx::poly::exp(va, n); break;
      case 3: r = fpx::poly::sqrt(va, n); break;
      default: throw std::bad_alloc();
    }
    store(r, out);
  });
}

template <class M>
int pow_t(const uint32_t* a, size_t la, uint64_t k, size_t n, uint32_t* out) {
  return guarded([&] {
    auto r = fpx::poly::pow(load<M>(a, la), k, n);
    store(r, out);
  });
}

template <class M>
int divmod_t(const uint32_t* a, size_t la, const uint32_t* b, size_t lb, uint32_t* q, size_t* lq, uint32_t* r,
             size_t* lr) {
  return guarded([&] {
    auto qr = fpx::poly::divmod(load<M>(a, la), load<M>(b, lb));
    store(qr.first, q);
    store(qr.second, r);
    *lq = qr.first.size();
    *lr = qr.second.size();
  });
}

template <class M>
int msqrt_t(uint32_t limb, uint32_t* out) {
  M r;
  if (!M::raw(limb).sqrt(r)) return 1;
  *out = r.raw_val();
  return 0;
}

template <class M>
int ntt_t(int dir, uint32_t n, uint32_t* data) {
  return guarded([&] {
    const auto& plan = fpx::NttPlan<M>::get_ref(n);
    if (dir == 0) plan.forward(data);
    else plan.inverse(data);
  });
}

#define FPREF_DISPATCH(P, CALL)                                   \
  switch (P) {                                                    \
    case 998244353u: { using M = fpx::mod998244353; return CALL; } \
    case 1000000007u: { using M = fpx::mod1000000007; return CALL; } \
    case 167772161u: { using M = fpx::mod167772161; return CALL; } \
    case 469762049u: { using M = fpx::mod469762049; return CALL; } \
    case 754974721u: { using M = fpx::mod754974721; return CALL; } \
    case 1004535809u: { using M = fpx::mod1004535809; return CALL; } \
    case 1224736769u: { using M = fpx::mod1224736769; return CALL; } \
    case 100003u: { using M = fpx::ModInt<100003u>; return CALL; } \
    case 3u: { using M = fpx::ModInt<3u>; return CALL; } \
    case 2u: { using M = fpx::ModInt<2u>; return CALL; } \
    default: return 4;                                            \
  }

}  // namespace

extern "C" {

const char* fpref_backend(void) { return fpx::simd_backend(); }

int fpref_conv(uint32_t p, const uint32_t* a, size_t la, const uint32_t* b, size_t lb, size_t n,
               uint32_t* out, size_t* out_len) {
  FPREF_DISPATCH(p, (conv_t<M>(a, la, b, lb, n, out, out_len)))
}

int fpref_series(int op, uint32_t p, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  FPREF_DISPATCH(p, (series_t<M>(op, a, la, n, out)))
}

int fpref_pow(uint32_t p, const uint32_t* a, size_t la, uint64_t k, size_t n, uint32_t* out) {
  FPREF_DISPATCH(p, (pow_t<M>(a, la, k, n, out)))
}

int fpref_divmod(uint32_t p, const uint32_t* a, size_t la, const uint32_t* b, size_t lb, uint32_t* q, size_t* lq,
                 uint32_t* r, size_t* lr) {
  FPREF_DISPATCH(p, (divmod_t<M>(a, la, b, lb, q, lq, r, lr)))
}

int fpref_msqrt(uint32_t p, uint32_t limb, uint32_t* out) {
  FPREF_DISPATCH(p, (msqrt_t<M>(limb, out)))
}

int fpref_ntt(uint32_t p, int dir, uint32_t n, uint32_t* data) {
  FPREF_DISPATCH(p, (ntt_t<M>(dir, n, data)))
}

}  // extern "C"
