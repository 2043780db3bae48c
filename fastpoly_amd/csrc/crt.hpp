// Three-prime CRT reconstruction kernel (restore_crt, crt_convolution.hpp:88-127).
#pragma once

#include "device.hpp"

namespace fpxg {

// Garner reconstruction of restore_crt (crt_convolution.hpp:88-127).  c_k are
// canonical Montgomery residues of the exact integer product of the raw target
// limbs modulo the auxiliary primes; the integer x1 + p1 x2 + p1 p2 x3 carries
// two target Montgomery factors and one REDC leaves the product's limb.
struct GarnerConst {
  uint32_t P[3], Pninv[3];
  uint32_t inv12;    // Montgomery (mod P2) of p1^-1
  uint32_t p1_in3;   // Montgomery (mod P3) of p1
  uint32_t inv123;   // Montgomery (mod P3) of (p1 p2)^-1
  uint32_t p, ninv;  // target
  uint32_t p1_t, p12_t;  // p1 mod p, p1 p2 mod p (plain)
  uint32_t primes, accumulate, neg;
};

// The target limb from the residues c1, c2, c3 (those past g.primes are
// ignored) and, when accumulating, the output's old limb.
__device__ __forceinline__ uint32_t garner_value(uint32_t c1, uint32_t c2, uint32_t c3, uint32_t old,
                                                 const GarnerConst& g) {
  {
    const uint32_t x1 = redc_full(c1, g.P[0], g.Pninv[0]);
    uint32_t x2 = 0, x3 = 0;
    if (g.primes >= 2) {
      const uint32_t r2 = redc_full(c2, g.P[1], g.Pninv[1]);
      const uint32_t d2 = r2 >= x1 ? r2 - x1 : r2 + g.P[1] - x1;
      x2 = mulmod_full(d2, g.inv12, g.P[1], g.Pninv[1]);
      if (g.primes == 3) {
        const uint32_t r3 = redc_full(c3, g.P[2], g.Pninv[2]);
        uint32_t x12 = x1 + mulmod_full(x2, g.p1_in3, g.P[2], g.Pninv[2]);
        x12 = x12 >= g.P[2] ? x12 - g.P[2] : x12;
        const uint32_t d3 = r3 >= x12 ? r3 - x12 : r3 + g.P[2] - x12;
        x3 = mulmod_full(d3, g.inv123, g.P[2], g.Pninv[2]);
      }
    }
    uint32_t v;
    if (g.p == 2) {
      v = (x1 + x2 + x3) & 1u;
      if (g.accumulate) v ^= old;
    } else {
      v = redc_full(x1, g.p, g.ninv);
      uint32_t t = redc_full((uint64_t)x2 * g.p1_t, g.p, g.ninv);
      v = v + t >= g.p ? v + t - g.p : v + t;
      t = redc_full((uint64_t)x3 * g.p12_t, g.p, g.ninv);
      v = v + t >= g.p ? v + t - g.p : v + t;
      if (g.neg) v = v == 0 ? 0u : g.p - v;
      if (g.accumulate) v = v + old >= g.p ? v + old - g.p : v + old;
    }
    return v;
  }
}

__device__ __forceinline__ void garner_one(const uint32_t* __restrict__ c1, const uint32_t* __restrict__ c2,
                                           const uint32_t* __restrict__ c3, uint32_t* __restrict__ out, uint32_t i,
                                           const GarnerConst& g) {
  out[i] = garner_value(c1[i], g.primes >= 2 ? c2[i] : 0u, g.primes == 3 ? c3[i] : 0u, g.accumulate ? out[i] : 0u, g);
}

// Garner fused into the last auxiliary prime's inverse column pass: the
// residues of the other primes, read at the output position.
struct GarnerArgs {
  const uint32_t* r[2];
  GarnerConst g;
};

__global__ void k_garner(const uint32_t* __restrict__ c1, const uint32_t* __restrict__ c2,
                         const uint32_t* __restrict__ c3, uint32_t* __restrict__ out, uint32_t n,
                         GarnerConst g) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    garner_one(c1, c2, c3, out, i, g);
}

// One record per CRT product window; the launch covers every record through a
// chunk table (chunk -> record, first element), 1024 elements per chunk.
struct GarnerRec {
  const uint32_t* c1;
  const uint32_t* c2;
  const uint32_t* c3;
  uint32_t* out;
  uint32_t n;
  GarnerConst g;
};
constexpr uint32_t kChunk = 1024;

__global__ void __launch_bounds__(256) k_garner_batch(const GarnerRec* __restrict__ recs, const uint2* __restrict__ chunks) {
  const uint2 ch = chunks[blockIdx.x];
  GarnerRec R = recs[ch.x];
  make_uniform(R);
  const uint32_t start = __builtin_amdgcn_readfirstlane(ch.y);
  const uint32_t end = min(R.n, start + kChunk);
  for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) garner_one(R.c1, R.c2, R.c3, R.out, i, R.g);
}

}  // namespace fpxg
