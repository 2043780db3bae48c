// fastpoly_amd device layer: Montgomery arithmetic on the integer VALU and the
// LDS-resident radix-16 NTT engine shared by every transform kernel.
//
// Representation (include/fastpoly/modint.hpp:54-98): a coefficient is one
// uint32 Montgomery limb x*2^32 mod p.  Products use one 32x32->64
// v_mad_u64_u32, one v_mul_lo_u32 and a second v_mad_u64_u32 (REDC), which
// measured faster on gfx950 than the Shoup/Q32 form of the x86 reference
// (simd.hpp:131-238) and needs single-word twiddles.
//
// Lazy convention (simd.hpp:19-38): for p < 2^30 every intermediate lives in
// [0, 2p); a REDC of a < 4p times a canonical twiddle lands in [0, 2p) without
// the final subtraction.  For 2^30 <= p < 2^31 the LAZY=false instantiation
// keeps everything canonical.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fpxg {

// Device twiddle tables of one prime (built once per field by the runtime).
//   tw_lo[k] = w^k, tw_hi[k] = w^(k << split) with w of order 2^logmax, and
//   the *_inv copies with inverse powers: w^x for any x < 2^logmax costs one
//   REDC (the four-step twiddle of the column pass);
//   planes_fwd / planes_inv: per-phase twiddle planes of blocks up to 2^14
//   points (layout in ntt2.hpp).  All canonical Montgomery limbs.
struct FieldDev {
  uint32_t p, ninv, r2, one;
  uint32_t logmax, split;
  const uint32_t* tw_lo;
  const uint32_t* tw_hi;
  const uint32_t* tw_lo_inv;
  const uint32_t* tw_hi_inv;
  const uint32_t* planes_fwd;
  const uint32_t* planes_inv;
};

constexpr int kPlaneLog = 14;  // twiddle planes cover blocks up to 2^14 points

template <bool LAZY>
struct Arith {
  uint32_t p, ninv, R;  // R = 2p (lazy) or p
  __device__ __forceinline__ explicit Arith(const FieldDev& f)
      : p(f.p), ninv(f.ninv), R(LAZY ? 2 * f.p : f.p) {}
  // REDC(a*b): a < 2^32 with b canonical, or both < 2p (lazy).
  __device__ __forceinline__ uint32_t mul(uint32_t a, uint32_t b) const {
    const uint64_t t = (uint64_t)a * b;
    const uint32_t m = (uint32_t)t * ninv;
    uint32_t u = (uint32_t)((t + (uint64_t)m * p) >> 32);
    if constexpr (!LAZY) u = min(u, u - p);
    return u;
  }
  __device__ __forceinline__ uint32_t add(uint32_t a, uint32_t b) const {
    const uint32_t s = a + b;
    return min(s, s - R);
  }
  __device__ __forceinline__ uint32_t sub(uint32_t a, uint32_t b) const {
    const uint32_t d = a - b + R;
    return min(d, d - R);
  }
  __device__ __forceinline__ uint32_t canon(uint32_t a) const {
    if constexpr (LAZY) return min(a, a - p);
    return a;
  }
  // From the inverse transforms' range [0, 4p) (lazy) to [0, 2p).
  __device__ __forceinline__ uint32_t half(uint32_t a) const {
    if constexpr (LAZY) return min(a, a - R);
    return a;
  }
};

// Signed residues for the forward transforms of primes p < 2^28 (8p < 2^31).
// A register holds the bit pattern of an int32 x with |x| < b*p standing for
// x mod p; sums and differences are one instruction each.  mul is a signed
// REDC: for |a| * w < 2^31 * p (w >= 0) the product a*w (v_mad_i64_i32),
// m = lo * ninv and a*w + m*p (v_mad_i64_i32 again, m read as signed) is a
// multiple of 2^32 below 2^32 * p in magnitude, so its high word u has
// |u| < p and u = a*w*2^-32 mod p.  The same three instructions as the
// unsigned REDC, but differences need no bias and sums no reduction.
struct ArithS {
  uint32_t p, ninv, one;
  __device__ __forceinline__ explicit ArithS(const FieldDev& f) : p(f.p), ninv(f.ninv), one(f.one) {}
  __device__ __forceinline__ uint32_t mul(uint32_t a, uint32_t w) const {
    const int64_t t = (int64_t)(int32_t)a * (int64_t)(int32_t)w;
    const uint32_t m = (uint32_t)t * ninv;
    return (uint32_t)((uint64_t)(t + (int64_t)(int32_t)m * (int64_t)(int32_t)p) >> 32);
  }
};

// Full reduction for arbitrary odd p < 2^31 and t < p*2^32 (Garner, scalars):
// (t + m*p) < 2^64, the quotient is < 2p, one conditional subtraction.
__device__ __forceinline__ uint32_t redc_full(uint64_t t, uint32_t p, uint32_t ninv) {
  const uint32_t m = (uint32_t)t * ninv;
  const uint32_t u = (uint32_t)((t + (uint64_t)m * p) >> 32);
  return u >= p ? u - p : u;
}
__device__ __forceinline__ uint32_t mulmod_full(uint32_t a, uint32_t b, uint32_t p, uint32_t ninv) {
  return redc_full((uint64_t)a * b, p, ninv);
}

// A record that is the same for the whole workgroup, word by word into
// scalar registers: tests on its fields become scalar branches.
template <class Rec>
__device__ __forceinline__ void make_uniform(Rec& r) {
  static_assert(sizeof(Rec) % 4 == 0, "whole words");
  uint32_t* w = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
  for (unsigned i = 0; i < sizeof(Rec) / 4; ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
}

// Padded LDS index: one spare word every 32 breaks power-of-two strides.
__device__ __forceinline__ uint32_t pidx(uint32_t i) { return i + (i >> 5); }
__host__ __device__ constexpr uint32_t lds_words(uint32_t n) { return n + (n >> 5) + 1; }

__host__ __device__ constexpr uint32_t bitrev_c(uint32_t x, int bits) {
  uint32_t r = 0;
  for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}
__device__ __forceinline__ uint32_t bitrev(uint32_t x, int bits) {
  return bits == 0 ? 0u : (__builtin_bitreverse32(x) >> (32 - bits));
}

// w^x for 0 <= x < 2^logmax from the two-level table (one extra REDC).
template <bool LAZY>
__device__ __forceinline__ uint32_t twiddle(const Arith<LAZY>& A, const uint32_t* lo,
                                            const uint32_t* hi, uint32_t split, uint32_t x) {
  typedef __attribute__((address_space(1))) const uint32_t g32;
  return A.mul(*(g32*)(hi + (x >> split)), *(g32*)(lo + (x & ((1u << split) - 1))));
}

}  // namespace fpxg
