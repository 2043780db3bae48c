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

// Garner's steps one prime at a time, so that a kernel holding the residues
// of one prime after another keeps only a running state.  x_k are the plain
// residues (REDC of the Montgomery residues c_k), x = x1 + p1 x2 + p1 p2 x3.
// x2 = (r2 - x1) / p1 mod P2.
__device__ __forceinline__ uint32_t garner_x2(uint32_t x1, uint32_t c2, const GarnerConst& g) {
  const uint32_t r2 = redc_full(c2, g.P[1], g.Pninv[1]);
  const uint32_t d2 = r2 >= x1 ? r2 - x1 : r2 + g.P[1] - x1;
  return mulmod_full(d2, g.inv12, g.P[1], g.Pninv[1]);
}
// x1 + p1 x2 mod P3.
__device__ __forceinline__ uint32_t garner_x12(uint32_t x1, uint32_t x2, const GarnerConst& g) {
  const uint32_t x12 = x1 + mulmod_full(x2, g.p1_in3, g.P[2], g.Pninv[2]);
  return x12 >= g.P[2] ? x12 - g.P[2] : x12;
}
// x3 = (r3 - x12) / (p1 p2) mod P3.
__device__ __forceinline__ uint32_t garner_x3(uint32_t x12, uint32_t c3, const GarnerConst& g) {
  const uint32_t r3 = redc_full(c3, g.P[2], g.Pninv[2]);
  const uint32_t d3 = r3 >= x12 ? r3 - x12 : r3 + g.P[2] - x12;
  return mulmod_full(d3, g.inv123, g.P[2], g.Pninv[2]);
}
// The target's share of x1: REDC(x1) (one of the product's two Montgomery
// factors comes off), or its parity for p = 2.
__device__ __forceinline__ uint32_t garner_t1(uint32_t x1, const GarnerConst& g) {
  return g.p == 2 ? (x1 & 1u) : redc_full(x1, g.p, g.ninv);
}
// v + REDC(x m) mod p, m = p1 or p1 p2 mod p (both odd: parity for p = 2).
__device__ __forceinline__ uint32_t garner_tadd(uint32_t v, uint32_t x, uint32_t m, const GarnerConst& g) {
  if (g.p == 2) return v ^ (x & 1u);
  const uint32_t t = redc_full((uint64_t)x * m, g.p, g.ninv);
  return v + t >= g.p ? v + t - g.p : v + t;
}
// Negation and accumulation into the output's old limb.
__device__ __forceinline__ uint32_t garner_finish(uint32_t v, uint32_t old, const GarnerConst& g) {
  if (g.p == 2) return g.accumulate ? v ^ old : v;
  if (g.neg) v = v == 0 ? 0u : g.p - v;
  if (g.accumulate) v = v + old >= g.p ? v + old - g.p : v + old;
  return v;
}

// The target limb from the residues c1, c2, c3 (those past g.primes are
// ignored) and, when accumulating, the output's old limb.
__device__ __forceinline__ uint32_t garner_value(uint32_t c1, uint32_t c2, uint32_t c3, uint32_t old,
                                                 const GarnerConst& g) {
  const uint32_t x1 = redc_full(c1, g.P[0], g.Pninv[0]);
  uint32_t v = garner_t1(x1, g);
  if (g.primes >= 2) {
    const uint32_t x2 = garner_x2(x1, c2, g);
    v = garner_tadd(v, x2, g.p1_t, g);
    if (g.primes == 3) v = garner_tadd(v, garner_x3(garner_x12(x1, x2, g), c3, g), g.p12_t, g);
  }
  return garner_finish(v, old, g);
}

__device__ __forceinline__ void garner_one(const uint32_t* __restrict__ c1, const uint32_t* __restrict__ c2,
                                           const uint32_t* __restrict__ c3, uint32_t* __restrict__ out, uint32_t i,
                                           const GarnerConst& g) {
  out[i] = garner_value(c1[i], g.primes >= 2 ? c2[i] : 0u, g.primes == 3 ? c3[i] : 0u, g.accumulate ? out[i] : 0u, g);
}

// One CRT product's inverse column pass over all its primes in one
// workgroup per tile (k_cols2_crt): the primes' row-pass outputs, the output
// window of target limbs and Garner's constants.  The primes' field tables
// are the same for every product and travel once per launch (CrtFields).
struct CrtColJob {
  const uint32_t* src[3];
  uint32_t* dst;
  uint32_t out_lo, out_hi, out_shift;
  uint32_t inv_n[3];
  GarnerConst g;
  uint32_t pad;
};
static_assert(sizeof(CrtColJob) % 8 == 0, "by-value jobs follow each other in the kernel arguments");

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

}  // namespace fpxg
