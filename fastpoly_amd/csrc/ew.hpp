// Small device kernels around the transforms: elementwise records, the
// sequential seed/tail recurrences, the input scan, and their host emitters
// (immediate launch, or a record appended to the active batch recorder).
#pragma once

#include "crt_kernels.hpp"

namespace fpxg {

__device__ __forceinline__ uint32_t fs_mul(const FieldScalar& f, uint32_t a, uint32_t b) {
  if (f.p == 2) return a & b & 1u;
  return mulmod_full(a, b, f.p, f.ninv);
}
__device__ __forceinline__ uint32_t fs_sub(const FieldScalar& f, uint32_t a, uint32_t b) {
  return a >= b ? a - b : a - b + f.p;
}
__device__ __forceinline__ uint32_t fs_from_int(const FieldScalar& f, uint32_t x) {
  if (f.p == 2) return x & 1u;
  return mulmod_full(x % f.p, f.r2, f.p, f.ninv);
}
__device__ uint32_t fs_pow(const FieldScalar& f, uint32_t a, uint32_t e) {
  uint32_t r = f.one, b = a;
  for (; e; e >>= 1) {
    if (e & 1) r = fs_mul(f, r, b);
    b = fs_mul(f, b, b);
  }
  return r;
}
__device__ __forceinline__ uint32_t fs_inv(const FieldScalar& f, uint32_t a) {
  return f.p == 2 ? a : fs_pow(f, a, f.p - 2);
}

// ---------------------------------------------------------------------------
// Sequential recurrences (seeds and <= 8-coefficient tails), one workgroup per
// record: new values are sequential, each dot product is spread over the block.
//   INV : v[i] = -(sum_{j=1}^{min(i,lu-1)} u[j] v[i-j]) * v[0]     (inverse.hpp:68-77)
//   EXP : v[i] = (sum_{j=1}^{min(i,lu)} u[j-1] v[i-j]) / i         (exponential.hpp:158-163)
//   SQRT: v[i] = (w[i] - sum_{j=1}^{i-1} v[j] v[i-j]) * k          (square_root.hpp:46-52)
// For INV with lo == 0, v[0] = u[0]^-1; for EXP/SQRT v[0] = init.
// ---------------------------------------------------------------------------
template <int TPB>
__device__ __forceinline__ unsigned long long block_sum(unsigned long long v, unsigned long long* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if constexpr (TPB == 64) {
    return v;
  } else {
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long t = 0;
#pragma unroll
    for (int w = 0; w < TPB / 64; ++w) t += red[w];
    __syncthreads();
    return t;
  }
}

template <int TPB>
__global__ void __launch_bounds__(TPB) k_recur(const RecurRec* __restrict__ recs, RecurRec r0) {
  __shared__ uint32_t nv[kRecMax];
  __shared__ uint32_t recip[kRecMax];
  __shared__ uint32_t us[kRecMax], ws[kRecMax];
  __shared__ unsigned long long red[TPB / 64];
  __shared__ uint32_t v0s;
  RecurRec R = recs ? recs[blockIdx.x] : r0;
  // The record is the workgroup's: its words are made wave-uniform, so the
  // kind and field tests below are scalar branches, not exec-mask juggling
  // inside the serial loops.
  make_uniform(R);
  const FieldScalar f = R.f;
  const uint32_t lo = R.lo, hi = R.hi;
  uint32_t* v = R.v;
  // A seed (hi <= kRecMax): every index read is below hi, so u, w and the
  // earlier values are staged in LDS once and each new coefficient costs no
  // global round trip.  Longer recurrences read them from global memory.
  const bool small = hi <= (uint32_t)kRecMax;
  const uint32_t base = small ? 0u : R.lo;
  if (small) {
    for (uint32_t t = threadIdx.x; t < hi; t += TPB) {
      us[t] = t < R.lu ? R.u[t] : 0u;
      ws[t] = t < R.lw ? R.w[t] : 0u;
      if (t < lo) nv[t] = v[t];
    }
  }
  if (R.kind == REC_EXP && threadIdx.x < hi - lo) {
    const uint32_t i = lo + threadIdx.x;
    recip[threadIdx.x] = i == 0 ? 0u : fs_inv(f, fs_from_int(f, i));
  }
  if (threadIdx.x == 0) {
    uint32_t v0;
    if (lo == 0) {
      v0 = R.kind == REC_INV ? fs_inv(f, R.u[0]) : R.init;
      nv[0] = v0;
      v[0] = v0;
    } else {
      v0 = v[0];
    }
    v0s = v0;
  }
  __syncthreads();
  const uint32_t v0 = v0s;
  if constexpr (TPB == 64) {
    if (small) {
      // Online form of a seed: lane i owns coefficient i and adds the term of
      // each earlier coefficient as soon as it is known, so a new coefficient
      // costs one finishing product, one lane broadcast and one product per
      // lane, with no cross-lane reduction.
      const uint32_t i = threadIdx.x;
      const uint32_t k0 = lo == 0 ? 1u : lo;  // first coefficient computed here
      // Contribution of v_k to lane i's sum (k < i).  SQRT adds the pair
      // {k, i - k} once its larger member k is known (twice unless equal).
      auto term = [&](uint32_t k, uint32_t vk) -> uint32_t {
        const uint32_t d = i - k;
        if (R.kind == REC_INV) return (d >= 1 && d + 1 <= R.lu) ? fs_mul(f, us[d], vk) : 0u;
        if (R.kind == REC_EXP) return (d >= 1 && d <= R.lu) ? fs_mul(f, us[d - 1], vk) : 0u;
        if (k == 0 || d == 0 || d > k) return 0u;
        const uint32_t t = fs_mul(f, vk, nv[d]);
        if (d == k) return t;
        const uint32_t t2 = t + t;
        return t2 >= f.p ? t2 - f.p : t2;
      };
      uint32_t acc = 0;
      auto add = [&](uint32_t t) {
        acc += t;
        acc = acc >= f.p ? acc - f.p : acc;
      };
      if (i < hi)
        for (uint32_t k = 0; k < k0 && k < i; ++k) add(term(k, nv[k]));
      for (uint32_t k = k0; k < hi; ++k) {
        uint32_t cand;
        if (R.kind == REC_INV) cand = fs_mul(f, acc == 0 ? 0u : f.p - acc, v0);
        else if (R.kind == REC_EXP) cand = fs_mul(f, acc, (i >= lo && i < hi) ? recip[i - lo] : 0u);
        else cand = fs_mul(f, fs_sub(f, (i < hi && i < R.lw) ? ws[i] : 0u, acc), R.k);
        const uint32_t vk = (uint32_t)__builtin_amdgcn_readlane((int)cand, (int)k);
        if (i == k) nv[k] = vk;
        if (i > k && i < hi) add(term(k, vk));
      }
      for (uint32_t t = k0 + i; t < hi; t += TPB) v[t] = nv[t];
      return;
    }
  }
  auto val = [&](uint32_t idx) -> uint32_t { return idx < base ? v[idx] : nv[idx - base]; };
  auto uat = [&](uint32_t idx) -> uint32_t { return small ? us[idx] : R.u[idx]; };
  for (uint32_t i = lo == 0 ? 1 : lo; i < hi; ++i) {
    uint32_t jmax;
    if (R.kind == REC_INV) jmax = min(i, R.lu ? R.lu - 1 : 0);
    else if (R.kind == REC_EXP) jmax = min(i, R.lu);
    else jmax = i - 1;
    uint32_t sum;
    if constexpr (TPB == 64) {
      // One wave: modular partial sums and a butterfly, 32-bit throughout.
      uint32_t acc = 0;
      for (uint32_t j = 1 + threadIdx.x; j <= jmax; j += TPB) {
        const uint32_t x = R.kind == REC_EXP ? uat(j - 1) : (R.kind == REC_INV ? uat(j) : val(j));
        acc += fs_mul(f, x, val(i - j));
        acc = acc >= f.p ? acc - f.p : acc;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        acc += __shfl_xor(acc, o, 64);
        acc = acc >= f.p ? acc - f.p : acc;
      }
      sum = acc;
    } else {
      unsigned long long s = 0;
      for (uint32_t j = 1 + threadIdx.x; j <= jmax; j += TPB) {
        const uint32_t x = R.kind == REC_EXP ? uat(j - 1) : (R.kind == REC_INV ? uat(j) : val(j));
        s += fs_mul(f, x, val(i - j));
      }
      s = block_sum<TPB>(s, red);
      sum = (uint32_t)(s % f.p);
    }
    if (threadIdx.x == 0) {
      uint32_t r;
      if (R.kind == REC_INV) r = fs_mul(f, sum == 0 ? 0u : f.p - sum, v0);
      else if (R.kind == REC_EXP) r = fs_mul(f, sum, recip[i - lo]);
      else r = fs_mul(f, fs_sub(f, i < R.lw ? (small ? ws[i] : R.w[i]) : 0u, sum), R.k);
      nv[i - base] = r;
    }
    __syncthreads();
  }
  // The new values are written once at the end: no store is waited on per
  // coefficient by the barriers above.
  for (uint32_t i = (lo == 0 ? 1 : lo) + threadIdx.x; i < hi; i += TPB) v[i] = nv[i - base];
}

__device__ __forceinline__ uint32_t ew_value(const EwRec& R, uint32_t i) {
  const FieldScalar& f = R.f;
  switch (R.op) {
    case EW_FILL: return R.c;
    case EW_NEGATE: { const uint32_t x = R.src[i]; return x == 0 ? 0u : f.p - x; }
    case EW_SCALE: return fs_mul(f, R.src[i], R.c);
    case EW_MUL_RECIP: return fs_mul(f, R.src[i], fs_inv(f, fs_from_int(f, R.c + i)));
    case EW_DERIV: return fs_mul(f, R.src[i + 1], fs_from_int(f, i + 1));
    case EW_THETA: return fs_mul(f, R.src[i], fs_from_int(f, i));
    case EW_INTEGRAL:  // src2: a table of 1/i, if given
      return (i >= 1 && i - 1 < R.avail) ? fs_mul(f, R.src[i - 1], R.src2 ? R.src2[i] : fs_inv(f, fs_from_int(f, i)))
                                         : 0u;
    case EW_REVERSE: return R.src[R.n - 1 - i];
    case EW_RECIP: return R.c + i == 0 ? 0u : fs_inv(f, fs_from_int(f, R.c + i));
    case EW_MULADD: {
      const uint32_t m = fs_mul(f, R.src[i], R.src2[i]);
      if (!R.c) return m;
      const uint32_t s = R.dst[i] + m;
      return s >= f.p ? s - f.p : s;
    }
    default: return fs_sub(f, i < R.avail ? R.src[i] : 0u, R.src2[i]);
  }
}

// The ops built on 1 / (c + i) (EW_RECIP, EW_MUL_RECIP; EW_INTEGRAL with
// c = 0) over four consecutive elements with one Fermat inverse between them
// (Montgomery's batch inversion: prefix products, one inverse, products back
// down).  A factor c + i = 0 mod p is left out of the products and gives 0,
// as fs_inv(0) does.
__device__ __forceinline__ void ew_recip4(const EwRec& R, uint32_t i0, uint32_t end) {
  const FieldScalar& f = R.f;
  const uint32_t c0 = R.op == EW_INTEGRAL ? 0u : R.c;
  uint32_t x[4], pre[4], acc = f.one;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    x[j] = i0 + j < end ? fs_from_int(f, c0 + i0 + j) : 0u;
    pre[j] = acc;
    if (x[j]) acc = fs_mul(f, acc, x[j]);
  }
  uint32_t inv = fs_inv(f, acc);
#pragma unroll
  for (int j = 3; j >= 0; --j) {
    const uint32_t i = i0 + j;
    const uint32_t r = x[j] ? fs_mul(f, inv, pre[j]) : 0u;
    if (x[j]) inv = fs_mul(f, inv, x[j]);
    if (i >= end) continue;
    uint32_t v = r;
    if (R.op == EW_MUL_RECIP) v = fs_mul(f, R.src[i], r);
    else if (R.op == EW_INTEGRAL) v = (i >= 1 && i - 1 < R.avail) ? fs_mul(f, R.src[i - 1], r) : 0u;
    R.dst[i] = v;
  }
}

// Elementwise records; one launch serves any number of them through a chunk
// table of (record, first element) pairs, kChunk elements per chunk.
__global__ void __launch_bounds__(256) k_ew(const EwRec* __restrict__ recs, const uint2* __restrict__ chunks, EwRec r0) {
  uint32_t start = blockIdx.x * kChunk;
  EwRec R = r0;
  if (chunks) {
    const uint2 ch = chunks[blockIdx.x];
    R = recs[ch.x];
    start = ch.y;
  }
  make_uniform(R);
  start = __builtin_amdgcn_readfirstlane(start);
  const uint32_t end = min(R.n, start + kChunk);
  if (R.op == EW_RECIP || R.op == EW_MUL_RECIP || (R.op == EW_INTEGRAL && !R.src2)) {  // uniform branch
    for (uint32_t i0 = start + 4 * threadIdx.x; i0 < end; i0 += 4 * blockDim.x) ew_recip4(R, i0, end);
    return;
  }
  for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x) R.dst[i] = ew_value(R, i);
}

// ---------------------------------------------------------------------------
// Direct product with a short operand (convolution.hpp:33-95: the reference's
// tiled direct product for a short side <= 40): out[o] = sum_j s[j] l[o - j]
// for o < outc, with ns <= kDirectShort (256).  A workgroup owns kDirectTile
// consecutive outputs; the long operand's window [base - (ns - 1), base +
// tile) and the short operand are staged in LDS (zeros outside the inputs),
// and each thread slides a four-word window down the long operand (one LDS
// read per term for four outputs).  Two products are summed before each
// Montgomery reduction (2 p^2 < p 2^32), four for p < 2^30, so any odd prime
// p < 2^31 works: no transform and no auxiliary primes.
// ---------------------------------------------------------------------------
constexpr uint32_t kDirectShort = 256;
constexpr uint32_t kDirectTile = 1024;

// Outputs [base, base + kDirectTile) of one direct product (256 threads).
__device__ __forceinline__ void conv_direct_tile(const uint32_t* __restrict__ sh, uint32_t ns,
                                                 const uint32_t* __restrict__ lg, uint32_t nl,
                                                 uint32_t* __restrict__ out, uint32_t outc, uint32_t p,
                                                 uint32_t ninv, uint32_t base) {
  __shared__ uint32_t S[kDirectShort];
  __shared__ uint32_t W[kDirectTile + kDirectShort];
  const uint32_t t = threadIdx.x;
  const int64_t w0 = (int64_t)base - (int64_t)(ns - 1);  // long-operand index of W[0]
  for (uint32_t i = t; i < ns; i += 256) S[i] = sh[i];
  for (uint32_t i = t; i < kDirectTile + ns - 1; i += 256) {
    const int64_t k = w0 + (int64_t)i;
    W[i] = (k >= 0 && k < (int64_t)nl) ? lg[k] : 0u;
  }
  __syncthreads();
  // Outputs o = base + 4t + q; term j reads W[4t + q - j + ns - 1].
  uint32_t w[4];
  const uint32_t top = 4 * t + ns - 1;
#pragma unroll
  for (int q = 0; q < 4; ++q) w[q] = W[top + q];
  uint64_t a64[4] = {0, 0, 0, 0};
  uint32_t acc[4] = {0, 0, 0, 0};
  auto fold = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t r = redc_full(a64[q], p, ninv);
      const uint32_t z = acc[q] + r;
      acc[q] = z >= p ? z - p : z;
      a64[q] = 0;
    }
  };
  // Two terms an iteration, their LDS reads issued together: term j uses the
  // window w, term j + 1 the window shifted by one (n1 = W[top - j - 1]).
  // The 64-bit sums are reduced every iteration, or every second one for
  // p < 2^30 (four products below p 2^32).
  const bool pairs = p < (1u << 30);  // uniform
  uint32_t j = 0, it = 0;
  for (; j + 1 < ns; j += 2, ++it) {
    const uint64_t s0 = S[j], s1 = S[j + 1];
    const uint32_t n1 = W[top - j - 1];
    const uint32_t n2 = j + 2 < ns ? W[top - j - 2] : 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) a64[q] += s0 * w[q];
    a64[0] += s1 * n1;
#pragma unroll
    for (int q = 1; q < 4; ++q) a64[q] += s1 * w[q - 1];
    if (!pairs || (it & 1)) fold();
    w[3] = w[1];
    w[2] = w[0];
    w[1] = n1;
    w[0] = n2;
  }
  if (j < ns) {
    const uint64_t s0 = S[j];
#pragma unroll
    for (int q = 0; q < 4; ++q) a64[q] += s0 * w[q];
  }
  fold();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t o = base + 4 * t + q;
    if (o < outc) out[o] = acc[q];
  }
}

__global__ void __launch_bounds__(256) k_conv_direct(const uint32_t* __restrict__ sh, uint32_t ns,
                                                     const uint32_t* __restrict__ lg, uint32_t nl,
                                                     uint32_t* __restrict__ out, uint32_t outc, uint32_t p,
                                                     uint32_t ninv) {
  conv_direct_tile(sh, ns, lg, nl, out, outc, p, ninv, blockIdx.x * kDirectTile);
}

// Merged schedules: chunk (record, first output) per workgroup.
__global__ void __launch_bounds__(256) k_conv_direct_batch(const DirectRec* __restrict__ recs,
                                                           const uint2* __restrict__ chunks) {
  const uint2 ch = chunks[blockIdx.x];
  DirectRec R = recs[ch.x];
  make_uniform(R);
  conv_direct_tile(R.sh, R.ns, R.lg, R.nl, R.out, R.outc, R.p, R.ninv, __builtin_amdgcn_readfirstlane(ch.y));
}

// Tail repair (convolution.hpp:102-137): out[0, min(outc, N)) holds the
// N-point cyclic product of a and b, whose true length is N + tail (tail <=
// 40, la, lb <= N).  Thread i < tail computes the high coefficient h = (a
// b)[N + i] directly (tail - i terms), stores it at N + i and removes its
// wraparound from position i.
__device__ __forceinline__ void tail_fix(const uint32_t* __restrict__ a, uint32_t la, const uint32_t* __restrict__ b,
                                         uint32_t lb, uint32_t* __restrict__ out, uint32_t outc, uint32_t N,
                                         uint32_t tail, uint32_t p, uint32_t ninv) {
  const uint32_t i = threadIdx.x;
  if (i >= tail) return;
  const uint32_t s = N + i;  // a_j b_(s - j) with s - j < lb and j < la
  uint32_t h = 0;
  for (uint32_t j = s - (lb - 1); j < la; ++j) {
    const uint32_t t = mulmod_full(a[j], b[s - j], p, ninv);
    const uint32_t z = h + t;
    h = z >= p ? z - p : z;
  }
  if (s < outc) out[s] = h;
  if (i < outc) {
    const uint32_t v = out[i];
    out[i] = v >= h ? v - h : v - h + p;
  }
}

__global__ void __launch_bounds__(64) k_tail_fix(const uint32_t* __restrict__ a, uint32_t la,
                                                 const uint32_t* __restrict__ b, uint32_t lb,
                                                 uint32_t* __restrict__ out, uint32_t outc, uint32_t N,
                                                 uint32_t tail, uint32_t p, uint32_t ninv) {
  tail_fix(a, la, b, lb, out, outc, N, tail, p, ninv);
}

// Merged schedules: one workgroup a record.
__global__ void __launch_bounds__(64) k_tail_fix_batch(const TailRec* __restrict__ recs) {
  TailRec R = recs[blockIdx.x];
  make_uniform(R);
  tail_fix(R.a, R.la, R.b, R.lb, R.out, R.outc, R.N, R.tail, R.p, R.ninv);
}

// Characteristic two (square_root.hpp:15-25): (sum b_i x^i)^2 = sum b_i x^(2i).
// out[i/2] = a[i] for even i; any nonzero odd coefficient sets *bad.
__global__ void k_frobenius_sqrt(const uint32_t* a, size_t cnt, uint32_t* out, unsigned int* bad) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x) {
    if (i & 1) {
      if (a[i] != 0) atomicOr(bad, 1u);
    } else {
      out[i / 2] = a[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Input scan: for a[0, n): the first two nonzero indices, one past the last,
// a[0] and a[first].  One workgroup per record.
// ---------------------------------------------------------------------------
struct ScanRec {
  const uint32_t* a;
  uint32_t n, pad;
};
struct ScanOut {
  uint32_t first, second, last1, a0, afirst, pad[3];
};

struct ScanPart {
  uint32_t m1, m2, last, pad;
};

// Up to two records by value (one series input, or a product's operands):
// such a scan needs no upload of its records.
struct ScanPack {
  ScanRec r[2];
};
__device__ __forceinline__ ScanRec scan_rec(const ScanRec* recs, const ScanPack& pk, uint32_t i) {
  if (recs) return recs[i];
  return i == 0 ? pk.r[0] : pk.r[1];
}

// (first, second) nonzero indices of two disjoint ranges, merged.
__device__ __forceinline__ void scan_merge(uint32_t& a1, uint32_t& a2, uint32_t b1, uint32_t b2) {
  const uint32_t m2 = a1 < b1 ? min(a2, b1) : (b1 < a1 ? min(a1, b2) : min(a2, b2));
  a1 = min(a1, b1);
  a2 = m2;
}

// Stage 1: block x of record y scans slice x of B (indices are record-relative;
// the sentinel "none" is the record's n).
__global__ void __launch_bounds__(256) k_scan_part(const ScanRec* __restrict__ recs, ScanPack pk,
                                                  ScanPart* __restrict__ parts, uint32_t B) {
  __shared__ uint32_t m1s[256], m2s[256], lss[256];
  const ScanRec R = scan_rec(recs, pk, blockIdx.y);
  const uint32_t per = (uint32_t)(((uint64_t)R.n + B - 1) / B);
  const uint32_t lo = (uint32_t)min((uint64_t)blockIdx.x * per, (uint64_t)R.n);
  const uint32_t hi = (uint32_t)min((uint64_t)lo + per, (uint64_t)R.n);
  uint32_t m1 = R.n, m2 = R.n, last = 0;
  // Eight words a thread a round, all loads issued before any test (one
  // memory round trip a round); indices rise within a thread.
  constexpr uint32_t K = 8;
  for (uint32_t c0 = lo; c0 < hi; c0 += 256 * K) {
    uint32_t v[K];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
      const uint32_t i = c0 + threadIdx.x + k * 256;
      v[k] = i < hi ? R.a[i] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
      const uint32_t i = c0 + threadIdx.x + k * 256;
      if (v[k] != 0) {
        if (i < m1) { m2 = m1; m1 = i; } else if (i < m2) { m2 = i; }
        last = i + 1;
      }
    }
  }
  m1s[threadIdx.x] = m1; m2s[threadIdx.x] = m2; lss[threadIdx.x] = last;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      uint32_t a1 = m1s[threadIdx.x], a2 = m2s[threadIdx.x];
      scan_merge(a1, a2, m1s[threadIdx.x + st], m2s[threadIdx.x + st]);
      m1s[threadIdx.x] = a1;
      m2s[threadIdx.x] = a2;
      lss[threadIdx.x] = max(lss[threadIdx.x], lss[threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) parts[(size_t)blockIdx.y * B + blockIdx.x] = {m1s[0], m2s[0], lss[0], 0u};
}

// Stage 2: one workgroup per record merges its B <= 256 partials.
// token != 0: out is host memory the caller polls; the fields are written
// first and out[i].pad[0] = token last, after a system-scope fence, so a
// host that sees the token sees the fields.
__global__ void __launch_bounds__(256) k_scan_final(const ScanRec* __restrict__ recs, ScanPack pk,
                                                   const ScanPart* __restrict__ parts, uint32_t B,
                                                   ScanOut* __restrict__ out, uint32_t token) {
  __shared__ uint32_t m1s[256], m2s[256], lss[256];
  const ScanRec R = scan_rec(recs, pk, blockIdx.x);
  ScanPart q{R.n, R.n, 0u, 0u};
  for (uint32_t t = threadIdx.x; t < B; t += 256) {
    const ScanPart r = parts[(size_t)blockIdx.x * B + t];
    scan_merge(q.m1, q.m2, r.m1, r.m2);
    q.last = max(q.last, r.last);
  }
  m1s[threadIdx.x] = q.m1; m2s[threadIdx.x] = q.m2; lss[threadIdx.x] = q.last;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      uint32_t a1 = m1s[threadIdx.x], a2 = m2s[threadIdx.x];
      scan_merge(a1, a2, m1s[threadIdx.x + st], m2s[threadIdx.x + st]);
      m1s[threadIdx.x] = a1;
      m2s[threadIdx.x] = a2;
      lss[threadIdx.x] = max(lss[threadIdx.x], lss[threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ScanOut o{};
    o.first = m1s[0];
    o.second = m2s[0];
    o.last1 = lss[0];
    o.a0 = R.n ? R.a[0] : 0u;
    o.afirst = o.first < R.n ? R.a[o.first] : 0u;
    if (token) {
      volatile uint32_t* w = reinterpret_cast<volatile uint32_t*>(out + blockIdx.x);
      w[0] = o.first; w[1] = o.second; w[2] = o.last1; w[3] = o.a0; w[4] = o.afirst;
      __threadfence_system();
      w[5] = token;
    } else {
      out[blockIdx.x] = o;
    }
  }
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------

static FieldScalar scalar_of(const HostField& F) { return {F.p, F.ninv, F.r2, F.one}; }

static void emit_ew(Context& c, const HostField& T, int op, uint32_t* dst, const uint32_t* src, const uint32_t* src2,
                    size_t n, uint32_t avail, uint32_t cst) {
  if (n == 0) return;
  EwRec r{};
  r.op = op; r.f = scalar_of(T); r.dst = dst; r.src = src; r.src2 = src2;
  r.n = (uint32_t)n; r.avail = avail; r.c = cst;
  if (c.rec) {
    c.rec->push_ew(r);
    return;
  }
  const unsigned blocks = (unsigned)((n + kChunk - 1) / kChunk);
  hipLaunchKernelGGL(k_ew, dim3(blocks), dim3(256), 0, c.stream, (const EwRec*)nullptr, (const uint2*)nullptr, r);
  FPXG_HIP(hipGetLastError());
}

// A device table of 1/i for i < n (entry 0 = 0) over T's field, kept for the
// process: computed once per modulus and grown (never shrunk) by immediate
// calls, read by exp's Newton steps and log's integral.  Returns null where
// none of at least n entries exists and none may be made here (recording:
// the table would be computed before the launches recorded ahead of it run)
// or n is above 2^24 (64 MiB).
static const uint32_t* recip_table(Context& c, const HostField& T, size_t n) {
  auto it = c.recip_tabs.find(T.p);
  if (it != c.recip_tabs.end() && it->second.second >= n) return it->second.first;
  if (c.rec || n > ((size_t)1 << 24)) return nullptr;
  size_t len = (size_t)1 << 16;
  while (len < n) len <<= 1;
  uint32_t* t = nullptr;
  FPXG_HIP(hipMalloc((void**)&t, len * 4));
  if (it != c.recip_tabs.end()) c.recip_old.push_back(it->second.first);
  c.recip_tabs[T.p] = {t, len};
  emit_ew(c, T, EW_RECIP, t, nullptr, nullptr, len, 0, 0);
  return t;
}

// Seeds (dot products of at most ~128 terms) run on one wave; longer tails on
// a 1024-thread workgroup.
static void recur(Context& c, const HostField& T, int kind, const uint32_t* u, size_t lu, const uint32_t* w,
                  size_t lw, uint32_t* v, size_t lo, size_t hi, uint32_t init, uint32_t k) {
  if (hi <= lo) return;
  RecurRec r{};
  r.kind = kind; r.f = scalar_of(T); r.u = u; r.lu = (uint32_t)lu; r.w = w; r.lw = (uint32_t)lw; r.v = v;
  r.lo = (uint32_t)lo; r.hi = (uint32_t)hi; r.init = init; r.k = k;
  if (c.rec) {
    c.rec->push_recur(r);
    return;
  }
  if (hi <= 128) hipLaunchKernelGGL(k_recur<64>, dim3(1), dim3(64), 0, c.stream, (const RecurRec*)nullptr, r);
  else hipLaunchKernelGGL(k_recur<1024>, dim3(1), dim3(1024), 0, c.stream, (const RecurRec*)nullptr, r);
  FPXG_HIP(hipGetLastError());
}

static void emit_garner(Context& c, const GarnerRec& g) {
  if (g.n == 0) return;
  if (c.rec) {
    c.rec->push_garner(g);
    return;
  }
  hipLaunchKernelGGL(k_garner, dim3(grid_for(g.n)), dim3(256), 0, c.stream, g.c1, g.c2, g.c3, g.out, g.n, g.g);
  FPXG_HIP(hipGetLastError());
}

// Scan of several inputs with one read-back.
static std::vector<ScanOut> scan_inputs(Context& c, const std::vector<ScanRec>& recs) {
  std::vector<ScanOut> out(recs.size());
  if (recs.empty()) return out;
  uint32_t maxn = 0;
  for (const ScanRec& r : recs) maxn = std::max(maxn, r.n);
  // Slices of >= 1024 words (one round of loads a thread), at most 1024 per
  // record; the final kernel folds up to four slices a thread.
  const uint32_t B = std::max<uint32_t>(1, std::min<uint32_t>(1024, (maxn + 1023) / 1024));
  DevBuf dpart = c.buf(recs.size() * B * sizeof(ScanPart) / 4 + 1);
  if (recs.size() <= 2) {
    // The single calls' scan: records by value, results written straight
    // into pinned host memory (no upload, no download copy).
    if (!c.scan_host)
      FPXG_HIP(hipHostMalloc(&c.scan_host, 2 * sizeof(ScanOut), hipHostMallocCoherent));
    ScanPack pk{};
    for (size_t i = 0; i < recs.size(); ++i) pk.r[i] = recs[i];
    ScanOut* hout = static_cast<ScanOut*>(c.scan_host);
    hipLaunchKernelGGL(k_scan_part, dim3(B, (unsigned)recs.size()), dim3(256), 0, c.stream, (const ScanRec*)nullptr,
                       pk, (ScanPart*)dpart.get(), B);
    FPXG_HIP(hipGetLastError());
    if (++c.scan_token == 0) c.scan_token = 1;
    const uint32_t token = c.scan_token;
    hipLaunchKernelGGL(k_scan_final, dim3((unsigned)recs.size()), dim3(256), 0, c.stream, (const ScanRec*)nullptr, pk,
                       (const ScanPart*)dpart.get(), B, hout, token);
    FPXG_HIP(hipGetLastError());
    // The host needs the five words, not an idle stream: poll the tokens for
    // up to 200 us (a stream synchronisation costs ~6 us more than the poll),
    // then fall back to synchronising, which also reports any device error.
    volatile const uint32_t* hw = reinterpret_cast<volatile const uint32_t*>(hout);
    auto arrived = [&] {
      for (size_t i = 0; i < recs.size(); ++i)
        if (hw[i * (sizeof(ScanOut) / 4) + 5] != token) return false;
      return true;
    };
    const auto t0 = std::chrono::steady_clock::now();
    while (!arrived()) {
      if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(200)) {
        FPXG_HIP(hipStreamSynchronize(c.stream));
        if (!arrived()) throw Error(E_DEVICE, "input scan did not complete");
        break;
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    for (size_t i = 0; i < recs.size(); ++i) {
      const size_t b = i * (sizeof(ScanOut) / 4);
      out[i].first = hw[b]; out[i].second = hw[b + 1]; out[i].last1 = hw[b + 2];
      out[i].a0 = hw[b + 3]; out[i].afirst = hw[b + 4];
    }
    return out;
  }
  DevBuf dr = c.buf(recs.size() * sizeof(ScanRec) / 4 + 1);
  DevBuf dout = c.buf(recs.size() * sizeof(ScanOut) / 4 + 1);
  FPXG_HIP(hipMemcpyAsync(dr.get(), recs.data(), recs.size() * sizeof(ScanRec), hipMemcpyHostToDevice, c.stream));
  for (size_t s = 0; s < recs.size(); s += 65535) {
    const size_t cnt = std::min<size_t>(65535, recs.size() - s);
    const ScanRec* rp = (const ScanRec*)dr.get() + s;
    hipLaunchKernelGGL(k_scan_part, dim3(B, (unsigned)cnt), dim3(256), 0, c.stream, rp, ScanPack{},
                       (ScanPart*)dpart.get() + s * B, B);
    FPXG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_final, dim3((unsigned)cnt), dim3(256), 0, c.stream, rp, ScanPack{},
                       (const ScanPart*)dpart.get() + s * B, B, (ScanOut*)dout.get() + s, 0u);
    FPXG_HIP(hipGetLastError());
  }
  FPXG_HIP(hipMemcpyAsync(out.data(), dout.get(), recs.size() * sizeof(ScanOut), hipMemcpyDeviceToHost, c.stream));
  FPXG_HIP(hipStreamSynchronize(c.stream));
  return out;
}

static uint32_t read_word(Context& c, const uint32_t* p) {
  uint32_t h = 0;
  FPXG_HIP(hipMemcpyAsync(&h, p, 4, hipMemcpyDeviceToHost, c.stream));
  FPXG_HIP(hipStreamSynchronize(c.stream));
  return h;
}

}  // namespace fpxg
