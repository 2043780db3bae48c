// fastpoly_amd transform kernels for gfx950.
//
// An N = 2^L point transform (forward = DIF, natural order in, bit-reversed
// out, exactly the order of NttPlan<M>::forward, ntt.hpp:283-308) is split as
// N = R * M with R = 2^S, M = 2^b:
//   column pass  (k_cols): the top S levels as R-point transforms down the
//                columns of an R x M row-major view, followed by the
//                four-step twiddle w_N^(c * bitrev_S(row)); tiles of C
//                adjacent columns keep every global access a 4C-byte segment;
//   row pass     (k_rows): the bottom b levels on each contiguous M-word row.
// Small transforms (L <= b) are a single row pass.  The inverse runs the
// mirror DIT (ntt.hpp:195-229) row pass first, then the column pass with the
// conjugate twiddle and the 1/N normalisation folded into its pre-multiply.
// Every kernel takes an array of jobs (gridDim.y) so one launch serves a whole
// batch of same-shape transforms.
#pragma once

#include "device.hpp"

namespace fpxg {

struct ColJob {
  const uint32_t* src;
  uint32_t* dst;
  uint32_t in_count;   // forward: index >= in_count reads as zero (never loaded)
  uint32_t out_lo, out_hi, out_shift;  // inverse: store j in [lo,hi) at dst[j-shift]
  uint32_t inv_n;      // inverse: Montgomery 1/N
  uint32_t lift;       // forward: src are raw limbs of another field, lift by r2
  uint32_t canon;      // store canonical residues
  FieldDev f;
};

struct RowJob {
  const uint32_t* src_a;
  const uint32_t* src_b;
  uint32_t* dst;
  uint32_t* dst_b;     // MODE_STORE_B: spectrum of b written here
  uint32_t in_a, in_b; // zero-fill limits (array indices)
  uint32_t out_lo, out_hi, out_shift;
  uint32_t inv_n;      // MODE_SCALE: multiply by 1/N at the store
  uint32_t lift;
  uint32_t canon;
  FieldDev f;
};

enum : int {
  MODE_FWD_A = 1,    // forward DIF of a's row
  MODE_B_RAW = 2,    // load b's row, forward it, multiply
  MODE_B_SPEC = 4,   // multiply by the stored spectrum row of b
  MODE_INV = 8,      // inverse DIT after the product
  MODE_SCALE = 16,   // single-pass inverse: 1/N at the store
  MODE_STORE_B = 32  // with B_RAW: also store b's spectrum row
};

constexpr int kThreads = 256;
constexpr int kMaxStash = 64;  // b-spectrum words held per thread (M <= 2^14)

// Map the flat block id so that consecutive tiles land on one XCD: the
// hardware deals workgroups round-robin over the 8 XCDs, each with its own L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t bid, uint32_t tiles) {
  if (tiles < 64 || (tiles & 7)) return bid;
  return (bid & 7) * (tiles >> 3) + (bid >> 3);
}

template <bool LAZY, bool INV>
__global__ void __launch_bounds__(kThreads) k_cols(ColJob j0, const ColJob* __restrict__ jobs, int L, int S,
                                                   int logC) {
  extern __shared__ uint32_t X[];
  const ColJob& J = jobs ? jobs[blockIdx.y] : j0;
  const Arith<LAZY> A(J.f);
  const int b = L - S;
  const uint32_t M = 1u << b, C = 1u << logC, R = 1u << S;
  const uint32_t tiles = M >> logC;
  const uint32_t c0 = xcd_tile(blockIdx.x, tiles) << logC;
  const uint32_t shift = J.f.logmax - L;  // w_N = w_max^(2^shift)
  const uint32_t total = C * R;
  if constexpr (!INV) {
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
      const uint32_t c = e & (C - 1), row = e >> logC;
      const uint32_t gi = row * M + c0 + c;
      uint32_t v = 0;
      if (gi < J.in_count) {
        v = J.src[gi];
        if (J.lift) v = A.mul(v, J.f.r2);
      }
      X[pidx(e)] = v;
    }
    __syncthreads();
    dif_all<LAZY>(X, S, logC, J.f.lvl_fwd, A);
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
      const uint32_t c = e & (C - 1), row = e >> logC;
      const uint32_t gi = row * M + c0 + c;
      const uint32_t x = ((c0 + c) * bitrev(row, S)) << shift;
      uint32_t v = A.mul(X[pidx(e)], twiddle<LAZY>(A, J.f.tw_lo, J.f.tw_hi, J.f.split, x));
      if (J.canon) v = A.canon(v);
      J.dst[gi] = v;
    }
  } else {
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
      const uint32_t c = e & (C - 1), row = e >> logC;
      const uint32_t gi = row * M + c0 + c;
      const uint32_t x = ((c0 + c) * bitrev(row, S)) << shift;
      const uint32_t tw = A.mul(twiddle<LAZY>(A, J.f.tw_lo_inv, J.f.tw_hi_inv, J.f.split, x), J.inv_n);
      X[pidx(e)] = A.mul(J.src[gi], tw);
    }
    __syncthreads();
    dit_all<LAZY>(X, S, logC, J.f.lvl_inv, A);
    for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
      const uint32_t c = e & (C - 1), row = e >> logC;
      const uint32_t gi = row * M + c0 + c;
      if (gi >= J.out_lo && gi < J.out_hi) J.dst[gi - J.out_shift] = A.canon(X[pidx(e)]);
    }
  }
}

template <bool LAZY, int MODE>
__global__ void __launch_bounds__(kThreads) k_rows(RowJob j0, const RowJob* __restrict__ jobs, int logM) {
  extern __shared__ uint32_t X[];
  const RowJob& J = jobs ? jobs[blockIdx.y] : j0;
  const Arith<LAZY> A(J.f);
  const uint32_t M = 1u << logM;
  const uint32_t base = blockIdx.x << logM;
  uint32_t stash[kMaxStash];

  if constexpr ((MODE & MODE_B_RAW) != 0) {
    for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
      const uint32_t gi = base + i;
      uint32_t v = 0;
      if (gi < J.in_b) {
        v = J.src_b[gi];
        if (J.lift) v = A.mul(v, J.f.r2);
      }
      X[pidx(i)] = v;
    }
    __syncthreads();
    dif_all<LAZY>(X, logM, 0, J.f.lvl_fwd, A);
#pragma unroll
    for (int j = 0; j < kMaxStash; ++j) {
      const uint32_t i = threadIdx.x + (uint32_t)j * blockDim.x;
      if (i < M) {
        stash[j] = X[pidx(i)];
        if constexpr ((MODE & MODE_STORE_B) != 0) J.dst_b[base + i] = stash[j];
      }
    }
    __syncthreads();
  }

  for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
    const uint32_t gi = base + i;
    uint32_t v = 0;
    if (gi < J.in_a) {
      v = J.src_a[gi];
      if (J.lift) v = A.mul(v, J.f.r2);
    }
    X[pidx(i)] = v;
  }
  __syncthreads();
  if constexpr ((MODE & MODE_FWD_A) != 0) dif_all<LAZY>(X, logM, 0, J.f.lvl_fwd, A);

  if constexpr ((MODE & (MODE_B_RAW | MODE_B_SPEC)) != 0) {
#pragma unroll
    for (int j = 0; j < kMaxStash; ++j) {
      const uint32_t i = threadIdx.x + (uint32_t)j * blockDim.x;
      if (i < M) {
        uint32_t w;
        if constexpr ((MODE & MODE_B_RAW) != 0) w = stash[j];
        else w = J.src_b[base + i];
        X[pidx(i)] = A.mul(X[pidx(i)], w);
      }
    }
    __syncthreads();
  }
  if constexpr ((MODE & MODE_INV) != 0) dit_all<LAZY>(X, logM, 0, J.f.lvl_inv, A);

  for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
    const uint32_t gi = base + i;
    if (gi >= J.out_lo && gi < J.out_hi) {
      uint32_t v = X[pidx(i)];
      if constexpr ((MODE & MODE_SCALE) != 0) v = A.mul(v, J.inv_n);
      if (J.canon) v = A.canon(v);
      J.dst[gi - J.out_shift] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Elementwise kernels.
// ---------------------------------------------------------------------------

template <bool LAZY>
__global__ void k_pointwise(uint32_t* dst, const uint32_t* a, const uint32_t* b, uint32_t n, FieldDev f) {
  const Arith<LAZY> A(f);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] = A.mul(a[i], b[i]);
}

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
  uint32_t primes, accumulate;
};

__global__ void k_garner(const uint32_t* __restrict__ c1, const uint32_t* __restrict__ c2,
                         const uint32_t* __restrict__ c3, uint32_t* __restrict__ out, uint32_t n,
                         GarnerConst g) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t x1 = redc_full(c1[i], g.P[0], g.Pninv[0]);
    uint32_t x2 = 0, x3 = 0;
    if (g.primes >= 2) {
      const uint32_t r2 = redc_full(c2[i], g.P[1], g.Pninv[1]);
      const uint32_t d2 = r2 >= x1 ? r2 - x1 : r2 + g.P[1] - x1;
      x2 = mulmod_full(d2, g.inv12, g.P[1], g.Pninv[1]);
      if (g.primes == 3) {
        const uint32_t r3 = redc_full(c3[i], g.P[2], g.Pninv[2]);
        uint32_t x12 = x1 + mulmod_full(x2, g.p1_in3, g.P[2], g.Pninv[2]);
        x12 = x12 >= g.P[2] ? x12 - g.P[2] : x12;
        const uint32_t d3 = r3 >= x12 ? r3 - x12 : r3 + g.P[2] - x12;
        x3 = mulmod_full(d3, g.inv123, g.P[2], g.Pninv[2]);
      }
    }
    uint32_t v;
    if (g.p == 2) {
      v = (x1 + x2 + x3) & 1u;
      if (g.accumulate) v ^= out[i];
    } else {
      v = redc_full(x1, g.p, g.ninv);
      uint32_t t = redc_full((uint64_t)x2 * g.p1_t, g.p, g.ninv);
      v = v + t >= g.p ? v + t - g.p : v + t;
      t = redc_full((uint64_t)x3 * g.p12_t, g.p, g.ninv);
      v = v + t >= g.p ? v + t - g.p : v + t;
      if (g.accumulate) {
        const uint32_t o = out[i];
        v = v + o >= g.p ? v + o - g.p : v + o;
      }
    }
    out[i] = v;
  }
}

}  // namespace fpxg
