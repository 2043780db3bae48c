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

// Device twiddle tables of one prime (built once per field, see runtime.cpp).
//   lvl_fwd[(L>>1) + m] = w_L^m        for 2 <= L <= 2^kLvlLog, m < L/2
//   lvl_inv[(L>>1) + m] = w_L^-m
//   tw_lo[k] = w^k, tw_hi[k] = w^(k << split) with w of order 2^logmax;
//   the *_inv planes hold the inverse powers.  All canonical Montgomery limbs.
struct FieldDev {
  uint32_t p, ninv, r2, one;
  uint32_t logmax, split;
  const uint32_t* lvl_fwd;
  const uint32_t* lvl_inv;
  const uint32_t* tw_lo;
  const uint32_t* tw_hi;
  const uint32_t* tw_lo_inv;
  const uint32_t* tw_hi_inv;
};

constexpr int kLvlLog = 14;  // per-level tables cover blocks up to 2^14 points

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

// Padded LDS index: one spare word every 32 breaks power-of-two strides.
__device__ __forceinline__ uint32_t pidx(uint32_t i) { return i + (i >> 5); }
__host__ __device__ constexpr uint32_t lds_words(uint32_t n) { return n + (n >> 5) + 1; }

__device__ __forceinline__ uint32_t bitrev(uint32_t x, int bits) {
  return bits == 0 ? 0u : (__builtin_bitreverse32(x) >> (32 - bits));
}

// w^x for 0 <= x < 2^logmax from the two-level table (one extra REDC).
template <bool LAZY>
__device__ __forceinline__ uint32_t twiddle(const Arith<LAZY>& A, const uint32_t* lo,
                                            const uint32_t* hi, uint32_t split, uint32_t x) {
  return A.mul(hi[x >> split], lo[x & ((1u << split) - 1)]);
}

// ---------------------------------------------------------------------------
// LDS engine.  A tile is C = 2^logC columns x 2^S rows, row-major in LDS
// (element (row i, column c) at pidx(i*C + c)).  Each column is an
// independent 2^S-point transform whose level l (0 = outermost) pairs rows at
// distance 2^(S-l-1) inside blocks of 2^(S-l) rows with twiddle w_{2^(S-l)}^m.
// A radix-2^r phase covers levels [s, s+r): a thread holds the 2^r rows
// hi*2^(S-s) + k*2^(S-s-r) + lo of one column in registers.
// ---------------------------------------------------------------------------

template <int r, bool LAZY>
__device__ __forceinline__ void dif_phase(uint32_t* X, int S, int logC, int s, const uint32_t* lvl,
                                          const Arith<LAZY>& A) {
  const int lo_bits = S - s - r;
  const uint32_t groups = 1u << (S - r + logC);
  const uint32_t cmask = (1u << logC) - 1, lomask = (1u << lo_bits) - 1;
  for (uint32_t g = threadIdx.x; g < groups; g += blockDim.x) {
    const uint32_t c = g & cmask, gg = g >> logC;
    const uint32_t lo = gg & lomask, hi = gg >> lo_bits;
    const uint32_t base = (hi << (S - s)) | lo;
    uint32_t x[1 << r];
#pragma unroll
    for (int k = 0; k < (1 << r); ++k) x[k] = X[pidx(((base + ((uint32_t)k << lo_bits)) << logC) | c)];
#pragma unroll
    for (int q = 0; q < r; ++q) {
      const int h = 1 << (r - q - 1);
      const uint32_t* tw = lvl + (1u << (S - s - q - 1));
#pragma unroll
      for (int k = 0; k < (1 << r); ++k) {
        if (k & h) continue;
        const uint32_t w = tw[((uint32_t)(k & (h - 1)) << lo_bits) | lo];
        const uint32_t a0 = x[k], a1 = x[k + h];
        x[k] = A.add(a0, a1);
        x[k + h] = A.mul(a0 - a1 + A.R, w);
      }
    }
#pragma unroll
    for (int k = 0; k < (1 << r); ++k) X[pidx(((base + ((uint32_t)k << lo_bits)) << logC) | c)] = x[k];
  }
}

template <int r, bool LAZY>
__device__ __forceinline__ void dit_phase(uint32_t* X, int S, int logC, int s, const uint32_t* lvl,
                                          const Arith<LAZY>& A) {
  const int lo_bits = S - s - r;
  const uint32_t groups = 1u << (S - r + logC);
  const uint32_t cmask = (1u << logC) - 1, lomask = (1u << lo_bits) - 1;
  for (uint32_t g = threadIdx.x; g < groups; g += blockDim.x) {
    const uint32_t c = g & cmask, gg = g >> logC;
    const uint32_t lo = gg & lomask, hi = gg >> lo_bits;
    const uint32_t base = (hi << (S - s)) | lo;
    uint32_t x[1 << r];
#pragma unroll
    for (int k = 0; k < (1 << r); ++k) x[k] = X[pidx(((base + ((uint32_t)k << lo_bits)) << logC) | c)];
#pragma unroll
    for (int q = r - 1; q >= 0; --q) {
      const int h = 1 << (r - q - 1);
      const uint32_t* tw = lvl + (1u << (S - s - q - 1));
#pragma unroll
      for (int k = 0; k < (1 << r); ++k) {
        if (k & h) continue;
        const uint32_t w = tw[((uint32_t)(k & (h - 1)) << lo_bits) | lo];
        const uint32_t t = A.mul(x[k + h], w);
        const uint32_t a0 = x[k];
        x[k] = A.add(a0, t);
        x[k + h] = A.sub(a0, t);
      }
    }
#pragma unroll
    for (int k = 0; k < (1 << r); ++k) X[pidx(((base + ((uint32_t)k << lo_bits)) << logC) | c)] = x[k];
  }
}

// All S levels, radix-16 phases then one shorter phase; barriers between.
template <bool LAZY>
__device__ __forceinline__ void dif_all(uint32_t* X, int S, int logC, const uint32_t* lvl,
                                        const Arith<LAZY>& A) {
  for (int s = 0; s < S; s += 4) {
    const int r = S - s < 4 ? S - s : 4;
    switch (r) {
      case 4: dif_phase<4, LAZY>(X, S, logC, s, lvl, A); break;
      case 3: dif_phase<3, LAZY>(X, S, logC, s, lvl, A); break;
      case 2: dif_phase<2, LAZY>(X, S, logC, s, lvl, A); break;
      default: dif_phase<1, LAZY>(X, S, logC, s, lvl, A); break;
    }
    __syncthreads();
  }
}

template <bool LAZY>
__device__ __forceinline__ void dit_all(uint32_t* X, int S, int logC, const uint32_t* lvl,
                                        const Arith<LAZY>& A) {
  int last = ((S - 1) / 4) * 4;
  for (int s = last; s >= 0; s -= 4) {
    const int r = S - s < 4 ? S - s : 4;
    switch (r) {
      case 4: dit_phase<4, LAZY>(X, S, logC, s, lvl, A); break;
      case 3: dit_phase<3, LAZY>(X, S, logC, s, lvl, A); break;
      case 2: dit_phase<2, LAZY>(X, S, logC, s, lvl, A); break;
      default: dit_phase<1, LAZY>(X, S, logC, s, lvl, A); break;
    }
    __syncthreads();
  }
}

}  // namespace fpxg
