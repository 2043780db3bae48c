// Compile-time-shaped NTT passes for gfx950.
//
// A block of 2^LOGB points is transformed in radix-16 register phases: each
// thread owns one 16-point group per phase (2^(4-r) groups of a shorter
// radix-2^r phase).  The first phase is loaded straight from HBM into
// registers, the last phase is stored straight from registers, and only the
// phase transitions go through LDS.  Every phase reads its twiddles from a
// precomputed "plane" laid out in the order the thread consumes them, with
// 16-byte vector loads.  Because the last forward phase and the first inverse
// phase own the same elements, a pointwise product between them never leaves
// the registers.
//
// Level q of a phase starting at level s of a 2^LOGB block pairs elements at
// distance 2^(LOGB-s-q-1) inside sub-blocks of 2^(LOGB-s-q) with twiddle
// w_L^m, L = 2^(LOGB-s-q), m the upper element's offset: the decimation-in-
// frequency order of NttPlan<M>::forward (ntt.hpp:283-308), read in
// bit-reversed order, and its exact DIT inverse.
#pragma once

#include "crt.hpp"
#include "device.hpp"

#include <type_traits>

namespace fpxg {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const uint32_t gu32_c;
typedef __attribute__((address_space(1))) const u32x4 gu4_c;
typedef __attribute__((address_space(1))) uint32_t gu32;

__device__ __forceinline__ uint32_t gload(const uint32_t* p) { return *(gu32_c*)p; }

// Raw buffer loads: the base lives in scalar registers and each load adds the
// lane's 32-bit byte offset and a uniform byte offset, so a run of loads at
// fixed distances costs no per-load 64-bit address arithmetic.  Offsets at or
// past `bytes` read 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint32_t bload(__amdgpu_buffer_rsrc_t r, uint32_t lane_bytes, uint32_t uniform_bytes) {
  return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)lane_bytes, (int)uniform_bytes, 0);
}
__device__ __forceinline__ void gstore(uint32_t* p, uint32_t v) { *(gu32*)p = v; }
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t lane_bytes,
                                       uint32_t uniform_bytes) {
  __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)lane_bytes, (int)uniform_bytes, 0);
}

// Twiddle planes.  plane(logB, r) starts at word plane_off(logB, r) and holds,
// for lo < 2^(logB-r), the 2^r - 1 twiddles of group lo (level by level,
// q = 0..r-1: w_{2^(logB-q)}^((k << (logB-r)) | lo), k < 2^(r-q-1)) in words
// [lo*2^r, lo*2^r + 2^r - 1).  Planes exist for 1 <= r <= min(4, logB) <= 14.
__host__ __device__ constexpr uint32_t plane_off(int logB, int r) {
  return (uint32_t)(r - 1) * 32768u + (1u << logB) - (1u << r);
}
constexpr uint32_t kPlaneWords = 4u * 32768u;

// Loads a group's 2^r - 1 twiddles and its padding word tw[2^r - 1] (16-byte
// loads).  keep_pad() marks the padding word as read where the twiddles are
// used: a register the compiler believes dead is reused at once, and the
// reuse must then wait for the load in flight to land.
template <int r>
__device__ __forceinline__ void load_tw(uint32_t* tw, const uint32_t* plane, uint32_t lo) {
  const uint32_t* p = plane + (lo << r);
  if constexpr (r >= 2) {
#pragma unroll
    for (int v = 0; v < (1 << r) / 4; ++v) {
      const u32x4 q = *(gu4_c*)(p + 4 * v);
      tw[4 * v] = q.x; tw[4 * v + 1] = q.y; tw[4 * v + 2] = q.z; tw[4 * v + 3] = q.w;
    }
  } else {
    tw[0] = gload(p);
  }
}
template <int r>
__device__ __forceinline__ void keep_pad(const uint32_t* tw) {
  if constexpr (r >= 2) asm volatile("" ::"v"(tw[(1 << r) - 1]));
}

// LO0: the group's offset lo is 0 (last phase of a block), so every level's
// first twiddle is w^0 = 1 and its products are skipped.
template <int r, bool LAZY, bool LO0 = false>
__device__ __forceinline__ void dif_block(uint32_t* x, const uint32_t* tw, const Arith<LAZY>& A) {
  int idx = 0;
#pragma unroll
  for (int q = 0; q < r; ++q) {
    const int h = 1 << (r - q - 1);
#pragma unroll
    for (int k = 0; k < (1 << r); ++k) {
      if (k & h) continue;
      const uint32_t a0 = x[k], a1 = x[k + h];
      x[k] = A.add(a0, a1);
      if (LO0 && (k & (h - 1)) == 0) x[k + h] = A.sub(a0, a1);
      else x[k + h] = A.mul(a0 - a1 + A.R, tw[idx + (k & (h - 1))]);
    }
    idx += h;
  }
}

// dif_block on signed residues (ArithS, p < 2^28).  Every register's bound b
// (|x| < b*p, b in {1, 2, 4, 8}) is a compile-time constant after unrolling:
// inputs have b = 2 (any value in [0, 2p)), a sum or difference has 2b, a
// product 1.  A pair with b = 8 is first brought to b = 1 by a product with
// the Montgomery one, so no sum or difference reaches 16p > 2^31 and every
// product's factor stays below 2^31.  A butterfly is 5 instructions (add,
// subtract, product) against dif_block's 8.  Outputs have b <= BOUT.
template <int r, bool LO0, int BOUT>
__device__ __forceinline__ void dif_block_s(uint32_t* x, const uint32_t* tw, const ArithS& S) {
  int bd[1 << r];
#pragma unroll
  for (int k = 0; k < (1 << r); ++k) bd[k] = 2;
  int idx = 0;
#pragma unroll
  for (int q = 0; q < r; ++q) {
    const int h = 1 << (r - q - 1);
#pragma unroll
    for (int k = 0; k < (1 << r); ++k) {
      if (k & h) continue;
      int b = bd[k];  // == bd[k + h]: both come through the same sums and products
      uint32_t a0 = x[k], a1 = x[k + h];
      if (b > 4) {
        a0 = S.mul(a0, S.one);
        a1 = S.mul(a1, S.one);
        b = 1;
      }
      x[k] = a0 + a1;
      if (LO0 && (k & (h - 1)) == 0) {
        x[k + h] = a0 - a1;
        bd[k + h] = 2 * b;
      } else {
        x[k + h] = S.mul(a0 - a1, tw[idx + (k & (h - 1))]);
        bd[k + h] = 1;
      }
      bd[k] = 2 * b;
    }
    idx += h;
  }
#pragma unroll
  for (int k = 0; k < (1 << r); ++k)
    if (bd[k] > BOUT) x[k] = S.mul(x[k], S.one);
}

// One forward group with either arithmetic: unsigned lazy/canonical limbs or
// signed residues (outputs bounded by BOUT * p).
template <int r, bool LO0, int BOUT, bool LAZY>
__device__ __forceinline__ void dif_any(uint32_t* x, const uint32_t* tw, const Arith<LAZY>& A) {
  dif_block<r, LAZY, LO0>(x, tw, A);
}
template <int r, bool LO0, int BOUT, bool LAZY>
__device__ __forceinline__ void dif_any(uint32_t* x, const uint32_t* tw, const ArithS& S) {
  dif_block_s<r, LO0, BOUT>(x, tw, S);
}

template <int r, bool LAZY, bool LO0 = false>
__device__ __forceinline__ void dit_block(uint32_t* x, const uint32_t* tw, const Arith<LAZY>& A) {
#pragma unroll
  for (int q = r - 1; q >= 0; --q) {
    const int h = 1 << (r - q - 1);
    const int idx = (1 << r) - (1 << (r - q));
#pragma unroll
    for (int k = 0; k < (1 << r); ++k) {
      if (k & h) continue;
      if constexpr (LAZY) {
        // Values live in [0, 4p) between levels (4p < 2^32): reduce the upper
        // input to [0, 2p), take w * lower in [0, 2p) (the product accepts
        // any factor below 2^32), and leave both outputs unreduced.
        const uint32_t a0r = x[k];
        const uint32_t a0 = min(a0r, a0r - A.R);
        const uint32_t b = x[k + h];
        const uint32_t t = (LO0 && (k & (h - 1)) == 0) ? min(b, b - A.R) : A.mul(b, tw[idx + (k & (h - 1))]);
        x[k] = a0 + t;
        x[k + h] = a0 - t + A.R;
      } else {
        const uint32_t t = (LO0 && (k & (h - 1)) == 0) ? x[k + h] : A.mul(x[k + h], tw[idx + (k & (h - 1))]);
        const uint32_t a0 = x[k];
        x[k] = A.add(a0, t);
        x[k + h] = A.sub(a0, t);
      }
    }
  }
}

// Phase schedule of a 2^LOGB block whose threads own 2^E elements each
// (E = 4 unless stated): radix-2^E phases with the one shorter phase
// (LOGB mod E levels) second, so that the first and last phases stay radix
// 2^E whenever LOGB >= 2E.  Fewer elements a thread (E = 2, 3) put more
// threads on one block: the one-workgroup steps of mid-size Newton chains run
// on 4-16 waves instead of 1-8.
template <int LOGB, int E = 4>
struct Sched {
  static constexpr int a = LOGB / E, b = LOGB % E;
  static constexpr int nph = a + (b ? 1 : 0);
  static constexpr int short_at = a >= 1 ? 1 : 0;
  static constexpr int r_of(int p) { return (b && p == short_at) ? b : E; }
  static constexpr int s_of(int p) {
    int s = 0;
    for (int i = 0; i < p; ++i) s += r_of(i);
    return s;
  }
  // threads per unit (one 2^E-point group each in a radix-2^E phase)
  static constexpr int GT = LOGB >= E ? (1 << (LOGB - E)) : 1;
};

// Element index within a unit for group id gid, register k of phase p.
template <int LOGB, int P, int E = 4>
__device__ __forceinline__ uint32_t elem(uint32_t gid, int k) {
  constexpr int s = Sched<LOGB, E>::s_of(P), r = Sched<LOGB, E>::r_of(P);
  constexpr int lo_bits = LOGB - s - r;
  const uint32_t lo = gid & ((1u << lo_bits) - 1), hi = gid >> lo_bits;
  return (hi << (LOGB - s)) | ((uint32_t)k << lo_bits) | lo;
}
template <int LOGB, int P, int E = 4>
__device__ __forceinline__ uint32_t group_lo(uint32_t gid) {
  constexpr int s = Sched<LOGB, E>::s_of(P), r = Sched<LOGB, E>::r_of(P);
  constexpr int lo_bits = LOGB - s - r;
  return gid & ((1u << lo_bits) - 1);
}
// Number of groups a thread owns in phase P, and its u-th group id.  Threads
// of a unit are numbered g < GT; a wave holds WG consecutive g (64 / C of a
// column tile, 64 of a long row, or the whole unit when it fits a wave).  In a
// phase with NG > 1 groups per thread, the WG threads of a wave take WG * NG
// consecutive group ids, so a wave's groups cover whole sub-blocks whenever
// possible and the phase transitions around them stay inside the wave.
template <int LOGB, int P, int LOGC = 0, int E = 4>
struct PhaseGroups {
  static constexpr int r = Sched<LOGB, E>::r_of(P);
  static constexpr int NG = LOGB >= E ? (1 << (E - r)) : 1;
  static constexpr int GT = Sched<LOGB, E>::GT;
  static constexpr int WG = (64 >> LOGC) < GT ? (64 >> LOGC) : GT;
  __device__ static __forceinline__ uint32_t gid(uint32_t g, int u) {
    if constexpr (NG == 1) return g;
    else return (g / WG) * (uint32_t)(WG * NG) + (uint32_t)u * WG + (g % WG);
  }
};

// Whether the exchange between phases P-1 and P stays inside each wave: true
// when a wave's groups of phase P-1 cover whole 2^(LOGB - s(P-1)) sub-blocks
// (then its phase-P groups are the same elements), or the unit fits a wave.
template <int LOGB, int P, int LOGC, int E = 4>
struct WaveLocal {
  using G = PhaseGroups<LOGB, P - 1, LOGC, E>;
  static constexpr int lb = LOGB - Sched<LOGB, E>::s_of(P - 1) - Sched<LOGB, E>::r_of(P - 1);
  static constexpr bool value = (Sched<LOGB, E>::GT << LOGC) <= 64 || (G::WG * G::NG >= (1 << lb));
};

// Orders one wave's LDS accesses (a wave's LDS operations execute in order;
// this only keeps the compiler from moving them across the exchange).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Register <-> LDS for phase P.  The unit's element e lives at padded word
// pidx(e << LOGC | c) (LOGC = 0 for rows, with a per-row base).  For a group
// (hi, lo) of phase P the padded address is base(gid) + off(k) with off a
// compile-time function of k whenever the block spans >= 32 words, so every
// LDS access uses an immediate offset.
template <int LOGB, int P, int LOGC, int E = 4>
struct LdsMap {
  static constexpr int s = Sched<LOGB, E>::s_of(P), r = Sched<LOGB, E>::r_of(P);
  static constexpr int lo_bits = LOGB - s - r;
  static constexpr int span = LOGB - s + LOGC;  // log2 words of one hi-block
  static constexpr bool affine = span >= 5;
  // unpadded word of (hi, k, lo, c)
  __device__ static __forceinline__ uint32_t word(uint32_t gid, int k, uint32_t c) {
    return (elem<LOGB, P, E>(gid, k) << LOGC) | c;
  }
  __device__ static __forceinline__ uint32_t base(uint32_t gid, uint32_t c) { return pidx(word(gid, 0, c)); }
  // pidx(word(gid, k, c)) - base(gid, c), independent of gid and c when affine
  static constexpr uint32_t off(int k) {
    const uint32_t kw = (uint32_t)k << (lo_bits + LOGC);
    if (lo_bits + LOGC >= 5) return kw + (kw >> 5);
    // k*2^(lo_bits+LOGC) + (lo<<LOGC|c) stays below the next multiple of 32
    return kw + (kw >> 5);
  }
};

template <int LOGB, int P, int LOGC, int E = 4>
__device__ __forceinline__ void phase_to_lds(uint32_t* X, const uint32_t* x, uint32_t g, uint32_t c) {
  using LM = LdsMap<LOGB, P, LOGC, E>;
  constexpr int r = LM::r;
  constexpr int NG = PhaseGroups<LOGB, P, LOGC, E>::NG;
#pragma unroll
  for (int u = 0; u < NG; ++u) {
    const uint32_t gid = PhaseGroups<LOGB, P, LOGC, E>::gid(g, u);
    if constexpr (LM::affine) {
      uint32_t* b = X + LM::base(gid, c);
#pragma unroll
      for (int k = 0; k < (1 << r); ++k) b[LM::off(k)] = x[u * (1 << r) + k];
    } else {
#pragma unroll
      for (int k = 0; k < (1 << r); ++k) X[pidx(LM::word(gid, k, c))] = x[u * (1 << r) + k];
    }
  }
}
template <int LOGB, int P, int LOGC, int E = 4>
__device__ __forceinline__ void phase_from_lds(uint32_t* x, const uint32_t* X, uint32_t g, uint32_t c) {
  using LM = LdsMap<LOGB, P, LOGC, E>;
  constexpr int r = LM::r;
  constexpr int NG = PhaseGroups<LOGB, P, LOGC, E>::NG;
#pragma unroll
  for (int u = 0; u < NG; ++u) {
    const uint32_t gid = PhaseGroups<LOGB, P, LOGC, E>::gid(g, u);
    if constexpr (LM::affine) {
      const uint32_t* b = X + LM::base(gid, c);
#pragma unroll
      for (int k = 0; k < (1 << r); ++k) x[u * (1 << r) + k] = b[LM::off(k)];
    } else {
#pragma unroll
      for (int k = 0; k < (1 << r); ++k) x[u * (1 << r) + k] = X[pidx(LM::word(gid, k, c))];
    }
  }
}

// Twiddles of phase P for thread g: the 2^r-word plane groups of its NG
// groups, 16 words in all (some unused for r = 1).  Loading them is separate
// from using them so that every load can be issued a phase (or a whole
// memory round trip) before the butterflies that need it.
template <int LOGB, int P, int LOGC = 0, int E = 4>
__device__ __forceinline__ void tw_load(uint32_t* T, const uint32_t* planes, uint32_t g) {
  constexpr int r = Sched<LOGB, E>::r_of(P);
  const uint32_t* plane = planes + plane_off(LOGB - Sched<LOGB, E>::s_of(P), r);
#pragma unroll
  for (int u = 0; u < PhaseGroups<LOGB, P, LOGC, E>::NG; ++u)
    load_tw<r>(T + u * (1 << r), plane, group_lo<LOGB, P, E>(PhaseGroups<LOGB, P, LOGC, E>::gid(g, u)));
}
template <int LOGB, int P, bool INV, bool LAZY, int LOGC = 0, int BOUT = 2, int E = 4, class AR = Arith<LAZY>>
__device__ __forceinline__ void tw_apply(uint32_t* x, const uint32_t* T, const AR& A) {
  constexpr int r = Sched<LOGB, E>::r_of(P);
  constexpr bool lo0 = LOGB - Sched<LOGB, E>::s_of(P) - r == 0;
#pragma unroll
  for (int u = 0; u < PhaseGroups<LOGB, P, LOGC, E>::NG; ++u) {
    if constexpr (INV) dit_block<r, LAZY, lo0>(x + u * (1 << r), T + u * (1 << r), A);
    else dif_any<r, lo0, BOUT, LAZY>(x + u * (1 << r), T + u * (1 << r), A);
    keep_pad<r>(T + u * (1 << r));
  }
}

// Register exchange between phases P-1 (registers) and P (registers) of a
// forward transform, through LDS.
template <int LOGB, int P, int LOGC, int E = 4>
__device__ __forceinline__ void fwd_exchange(uint32_t* x, uint32_t* X, uint32_t g, uint32_t c) {
  if constexpr (WaveLocal<LOGB, P, LOGC, E>::value) {
    wave_sync();
    phase_to_lds<LOGB, P - 1, LOGC, E>(X, x, g, c);
    wave_sync();
    phase_from_lds<LOGB, P, LOGC, E>(x, X, g, c);
  } else {
    if constexpr (P > 1) __syncthreads();  // earlier exchanges may still be reading
    phase_to_lds<LOGB, P - 1, LOGC, E>(X, x, g, c);
    __syncthreads();
    phase_from_lds<LOGB, P, LOGC, E>(x, X, g, c);
  }
}

// Phases [P, END) of a forward transform whose registers already hold phase
// P-1's output (P >= 1); T holds phase P's twiddles (loads issued earlier).
// Each step issues the next phase's twiddle loads, exchanges through LDS and
// computes, so a twiddle load overlaps a whole phase.  All threads of the
// workgroup must call it.  With signed residues (AR = ArithS) phases hand
// over values below 2p in magnitude and the last one's outputs are below
// BLAST * p.
// PF = false (kernels held to 64 registers): no prefetch; each phase loads
// its own twiddles after its exchange and T is unused.
template <int LOGB, int P, int END, int LOGC, bool LAZY, int BLAST = 2, bool PF = true, int E = 4,
          class AR = Arith<LAZY>>
__device__ __forceinline__ void fwd_phases(uint32_t* x, uint32_t* X, uint32_t g, uint32_t c, const uint32_t* planes,
                                           const AR& A, const uint32_t* T) {
  if constexpr (P < END) {
    uint32_t Tn[16];
    if constexpr (PF && P + 1 < END) tw_load<LOGB, P + 1, LOGC, E>(Tn, planes, g);
    fwd_exchange<LOGB, P, LOGC, E>(x, X, g, c);
    if constexpr (PF) {
      tw_apply<LOGB, P, false, LAZY, LOGC, P + 1 == END ? BLAST : 2, E>(x, T, A);
    } else {
      uint32_t Tp[16];
      tw_load<LOGB, P, LOGC, E>(Tp, planes, g);
      tw_apply<LOGB, P, false, LAZY, LOGC, P + 1 == END ? BLAST : 2, E>(x, Tp, A);
    }
    fwd_phases<LOGB, P + 1, END, LOGC, LAZY, BLAST, PF, E>(x, X, g, c, planes, A, Tn);
  }
}
// A whole forward transform of registers holding phase 0's input; T0 holds
// phase 0's twiddles (tw_load<LOGB, 0, LOGC>, best issued with the data
// loads).  BOUT0: output bound of a one-phase transform (signed residues).
template <int LOGB, int LOGC, bool LAZY, int BLAST = 2, int BOUT0 = 2, bool PF = true, int E = 4,
          class AR = Arith<LAZY>>
__device__ __forceinline__ void fwd_transform(uint32_t* x, uint32_t* X, uint32_t g, uint32_t c, const uint32_t* planes,
                                              const AR& A, const uint32_t* T0) {
  constexpr int NPH = Sched<LOGB, E>::nph;
  uint32_t T1[16];
  if constexpr (PF && NPH > 1) tw_load<LOGB, 1, LOGC, E>(T1, planes, g);
  tw_apply<LOGB, 0, false, LAZY, LOGC, NPH == 1 ? BOUT0 : 2, E>(x, T0, A);
  fwd_phases<LOGB, 1, NPH, LOGC, LAZY, BLAST, PF, E>(x, X, g, c, planes, A, T1);
}

// Inverse phases P, P-1, ..., 0 where registers hold phase P+1's inverse
// output and T holds phase P's twiddles.
template <int LOGB, int P, int LOGC, bool LAZY, bool PF = true, int E = 4>
__device__ __forceinline__ void inv_phases(uint32_t* x, uint32_t* X, uint32_t g, uint32_t c, const uint32_t* planes,
                                           const Arith<LAZY>& A, const uint32_t* T) {
  if constexpr (P >= 0) {
    uint32_t Tn[16];
    if constexpr (PF && P >= 1) tw_load<LOGB, P - 1, LOGC, E>(Tn, planes, g);
    if constexpr (WaveLocal<LOGB, P + 1, LOGC, E>::value) {
      wave_sync();
      phase_to_lds<LOGB, P + 1, LOGC, E>(X, x, g, c);
      wave_sync();
      phase_from_lds<LOGB, P, LOGC, E>(x, X, g, c);
    } else {
      __syncthreads();  // a forward pass or an earlier exchange may still be reading
      phase_to_lds<LOGB, P + 1, LOGC, E>(X, x, g, c);
      __syncthreads();
      phase_from_lds<LOGB, P, LOGC, E>(x, X, g, c);
    }
    if constexpr (PF) {
      tw_apply<LOGB, P, true, LAZY, LOGC, 2, E>(x, T, A);
    } else {
      uint32_t Tp[16];
      tw_load<LOGB, P, LOGC, E>(Tp, planes, g);
      tw_apply<LOGB, P, true, LAZY, LOGC, 2, E>(x, Tp, A);
    }
    inv_phases<LOGB, P - 1, LOGC, LAZY, PF, E>(x, X, g, c, planes, A, Tn);
  }
}
// A whole inverse transform of registers holding the spectrum in the last
// phase's grouping; TL holds that phase's inverse twiddles (issued earlier).
template <int LOGB, int LOGC, bool LAZY, bool PF = true, int E = 4>
__device__ __forceinline__ void inv_transform(uint32_t* x, uint32_t* X, uint32_t g, uint32_t c, const uint32_t* planes,
                                              const Arith<LAZY>& A, const uint32_t* TL) {
  constexpr int PL = Sched<LOGB, E>::nph - 1;
  uint32_t Tn[16];
  if constexpr (PF && PL >= 1) tw_load<LOGB, PL - 1, LOGC, E>(Tn, planes, g);
  tw_apply<LOGB, PL, true, LAZY, LOGC, 2, E>(x, TL, A);
  inv_phases<LOGB, PL - 1, LOGC, LAZY, PF, E>(x, X, g, c, planes, A, Tn);
}

// Up to K jobs passed by value at the start of the kernel arguments: a launch
// of a few independent jobs of one shape (both operands of a product, the
// three auxiliary primes of a CRT product) needs no device job array.
template <class Job, int K>
struct JobPack {
  Job j[K];
};

// The workgroup's job: jobs[blockIdx.y] when a job array is given, else the
// by-value job number blockIdx.y of the JobPack at offset 0 of the kernel
// arguments.  Both are read through the constant address space, so the job
// lands in scalar registers with one round of scalar loads.
template <class Job>
__device__ __forceinline__ Job load_job(const Job& j0, const Job* jobs) {
#if defined(__HIP_DEVICE_COMPILE__)
  (void)j0;
  typedef __attribute__((address_space(4))) const uint32_t cword;
  static_assert(sizeof(Job) % 8 == 0, "by-value jobs follow each other in the kernel arguments");
  const cword* jp = jobs ? (const cword*)(uintptr_t)(jobs + blockIdx.y)
                         : (const cword*)__builtin_amdgcn_kernarg_segment_ptr() + blockIdx.y * (sizeof(Job) / 4);
  static_assert(sizeof(Job) % 4 == 0, "jobs are whole words");
  Job J;
  uint32_t* w = reinterpret_cast<uint32_t*>(&J);
#pragma unroll
  for (unsigned i = 0; i < sizeof(Job) / 4; ++i) w[i] = jp[i];
  return J;
#else
  return jobs ? jobs[0] : j0;  // host pass only: never executed
#endif
}

// ---------------------------------------------------------------------------
// Row pass: U rows of 2^LOGM contiguous words per workgroup.
// ---------------------------------------------------------------------------
template <int LOGM, int E = 4>
struct RowShape {
  static constexpr int GT = Sched<LOGM, E>::GT;
  static constexpr int U = GT >= 64 ? 1 : 64 / GT;
  static constexpr int T = GT * U;
  static constexpr uint32_t M = 1u << LOGM;
  static constexpr uint32_t lds_words_per_row = lds_words(M);
};

struct RowJob2 {
  const uint32_t* src_a;
  const uint32_t* src_b;
  uint32_t* dst;
  uint32_t nrows;            // rows of 2^LOGM in this job
  uint32_t in_a, in_b;       // zero-fill limits (array indices)
  uint32_t out_lo, out_hi, out_shift;
  uint32_t inv_n;            // flags & R2_SCALE: multiply by inv_n at the store
  uint32_t flags;
  const uint32_t* epi;       // operand of R2_SUBFROM / R2_MULTAB, indexed by output position
  uint32_t epi_avail, aux;   // aux: R2M_EXPSTEP's window shift
  const uint32_t* src_c;     // third operand of the Newton-step modes
  FieldDev f;
};
// By-value jobs per row launch: one for each of the three auxiliary primes.
constexpr int kRowPack = 3;
enum : uint32_t { R2_LIFT = 1, R2_CANON = 2, R2_SCALE = 4, R2_NEG = 8, R2_ACC = 16, R2_SUBFROM = 32, R2_MULTAB = 64,
                  R2_PREFIX = 128, R2_ACCLO = 256, R2_TWIST = 512, R2_COSET = 1024 };
// R2_TWIST (forward column pass only): input coefficient t times g^t, g the
// field's coset generator (so the spectrum is that of the coset g w^k).
// R2_COSET (inverse column pass only): the output v at position t < 2^L is
// the untwisted coset product's coefficient d = v g^-t; with c1 = epi[t]
// (the cyclic product's coefficient) and acc_end = 1/(g^N - 1), the high
// coefficient h = (d - c1) / (g^N - 1) goes to dst[N + t] (t < epi_avail)
// and the low one c1 - h to dst[t].
// R2_PREFIX (chained column pass only): the operand's positions below
// epi_avail come from epi, and its window is also stored to epi.
// R2_ACCLO (column inverse only): like R2_ACC for output positions below
// acc_end (relative to out_shift), a plain store from there on.
// R2_ACC (column inverse only): dst[i] = dst[i] + result, both canonical.
// Inverse-pass epilogues on the canonical result v at output position t (the
// row pass needs R2_CANON with them): R2_SUBFROM v = (t < epi_avail ? epi[t]
// : 0) - v; R2_MULTAB v = v * epi[t] (Montgomery limbs).
__device__ __forceinline__ uint32_t out_epilogue(uint32_t flags, const uint32_t* epi, uint32_t avail, uint32_t t,
                                                 uint32_t v, uint32_t p, uint32_t ninv) {
  if (flags & R2_SUBFROM) {
    const uint32_t s = t < avail ? gload(epi + t) : 0u;
    v = s >= v ? s - v : s - v + p;
  }
  if (flags & R2_MULTAB) v = mulmod_full(v, gload(epi + t), p, ninv);
  return v;
}
enum : int { R2M_FWD = 1, R2M_INV = 2, R2M_CONV = 3, R2M_SPEC = 4, R2M_SPEC_ONLY = 5, R2M_INVSTEP = 6,
             R2M_EXPSTEP = 7, R2M_SQRTSTEP = 8, R2M_NINV = 9, R2M_NEXP = 10, R2M_NSQRT = 11, R2M_FWDS = 12,
             R2M_SPEC_SUM2 = 13 };
// R2M_FWDS: R2M_FWD on signed residues (ArithS, p < 2^28), chosen at launch
// time for a recorded R2M_FWD; never recorded itself.
// R2M_NINV / R2M_NEXP / R2M_NSQRT: the first coefficients of a series in one
// workgroup (k_newton_small), recorded and merged like the row modes.
// R2M_INVSTEP (one Newton step of the power-series inverse, inverse.hpp:18-65,
// for a 2^LOGM = 2k point row): with c = b[0, k) and a = a[0, in_a),
// e = (a * c)[k, k + out_hi) and out = (e * c)[0, out_hi) (negated with
// R2_NEG); c's spectrum is computed once and e never leaves the registers.
// R2M_EXPSTEP (one doubling m -> m + h of exp, exponential.hpp:176-200, for a
// 2^LOGM = 2m point row, h = out_hi): b = src_b[0, in_b = m), a' = src_a[0,
// in_a), c = src_c[0, h): r = (a' * b)[aux = m - 1, aux + h), t = (r * c)[0, h)
// times epi[i] (1 / (m + i)), out = (t * b)[0, h).
// R2M_SQRTSTEP (square_root.hpp:86-100, 2^LOGM = 2m): b = src_b[0, m), c =
// src_c[0, h): res = (i < epi_avail ? epi[i] : 0) - (b * b)[m + i], out =
// (res * c)[0, h) * aux (aux = 1/2 as a Montgomery limb).
// R2M_FWD:  forward a                      R2M_INV:       inverse a
// R2M_CONV: forward a, forward b, a*b, inverse
// R2M_SPEC: forward a, a*spectrum(b), inverse
// R2M_SPEC_ONLY: a (already a spectrum) * spectrum(b), inverse
// R2M_SPEC_SUM2: a * b + c * d (four spectra: src_a, src_b, src_c, epi), inverse

// Target waves per SIMD (register budget 512 / waves).  Eight waves (64
// registers) lets two 1024-thread workgroups share a CU; the fused modes hold
// two operands in registers and get a larger budget.
#ifndef FPXG_ROW_WPE
#define FPXG_ROW_WPE 8
#endif
#ifndef FPXG_ROW_WPE_FUSED
#define FPXG_ROW_WPE_FUSED 4
#endif
#ifndef FPXG_COL_WPE
#define FPXG_COL_WPE 8
#endif
template <int MODE>
struct RowWaves {
  static constexpr int value = (MODE == R2M_FWD || MODE == R2M_FWDS || MODE == R2M_INV) ? FPXG_ROW_WPE
                               : (MODE == R2M_INVSTEP || MODE == R2M_EXPSTEP || MODE == R2M_SQRTSTEP) ? 2
                                                                                             : FPXG_ROW_WPE_FUSED;
};

// E: log2 of the elements a thread owns (4; 2 or 3 for the one-row launches
// of mid-size Newton steps and products, which then run on 4x or 2x the
// waves).
template <int LOGM, bool LAZY, int MODE, int E = 4>
__global__ void __launch_bounds__((RowShape<LOGM, E>::T)) __attribute__((amdgpu_waves_per_eu(RowWaves<MODE>::value)))
k_rows2(JobPack<RowJob2, kRowPack> jp, const RowJob2* __restrict__ jobs) {
  using RS = RowShape<LOGM, E>;
  using SC = Sched<LOGM, E>;
  constexpr int NE = 1 << E;  // elements a thread owns
  constexpr int NPH = SC::nph;
  extern __shared__ uint32_t lds[];
  const RowJob2 J = load_job(jp.j[0], jobs);
  const Arith<LAZY> A(J.f);
  const uint32_t flags = __builtin_amdgcn_readfirstlane(J.flags);  // uniform: scalar branches
  const bool lift = (flags & R2_LIFT) != 0;
  const uint32_t t = threadIdx.x;
  const uint32_t g = t % RS::GT, ur = t / RS::GT;
  const uint32_t row = blockIdx.x * RS::U + ur;
  if (RS::U == 1 && row >= J.nrows) return;  // whole workgroup idle (uniform): no barrier is skipped
  const bool active = RS::U == 1 || row < J.nrows;
  const uint32_t base = row << LOGM;
  uint32_t* X = lds + ur * RS::lds_words_per_row;
  uint32_t x[16];
  // The first forward phase's and the first inverse phase's twiddles, issued
  // before any data load so that their latency hides behind it.
  constexpr int PL = NPH - 1;  // last forward phase = first inverse phase
  constexpr bool kHasFwd = MODE != R2M_INV && MODE != R2M_SPEC_ONLY && MODE != R2M_SPEC_SUM2;
  constexpr bool kHasInv = MODE != R2M_FWD && MODE != R2M_FWDS;
  // Modes held to 64 registers (the plain forward and inverse passes, which
  // stream the largest transforms) load each phase's twiddles where they are
  // used instead.
  constexpr bool kPF = RowWaves<MODE>::value < 8;
  uint32_t TF0[16], TI[16];
  if constexpr (kPF && kHasFwd) tw_load<LOGM, 0, 0, E>(TF0, J.f.planes_fwd, g);
  if constexpr (kPF && kHasInv) tw_load<LOGM, PL, 0, E>(TI, J.f.planes_inv, g);

  // Branch-free first-phase load: indices past in_count (or of an inactive
  // row) read a clamped valid word and are then replaced by zero.  load_issue
  // only issues the loads (every index clamped, so they overlap); load_fix
  // waits for them and applies the lift and the zero fill past in_count.  A
  // mode with several operands issues all of them before its first transform,
  // so they cost one memory round trip together.
  //
  // The whole row is input: no clamps, no zero fill, and (radix-16 first
  // phase) register k of thread g is word k * GT + g of the row, so each load
  // is a uniform base plus the thread's 32-bit offset (no 64-bit address
  // arithmetic per load).
  constexpr int r0 = SC::r_of(0);
  constexpr int NE0 = PhaseGroups<LOGM, 0, 0, E>::NG * (1 << r0);
  auto whole_row = [&](uint32_t in_count) {
    return RS::U == 1 && PhaseGroups<LOGM, 0, 0, E>::NG == 1 && base + RS::M <= in_count;
  };
  auto load_issue = [&](const uint32_t* src, uint32_t in_count, uint32_t* dst_regs) {
    if (in_count == 0) return;
    if (whole_row(in_count)) {
      const __amdgpu_buffer_rsrc_t rs = buf_rsrc(src + base, RS::M * 4);
#pragma unroll
      for (int k = 0; k < (1 << r0); ++k) dst_regs[k] = bload(rs, g * 4, (uint32_t)k * RS::GT * 4);
      return;
    }
    const uint32_t last = in_count - 1;
#pragma unroll
    for (int u = 0; u < PhaseGroups<LOGM, 0, 0, E>::NG; ++u)
#pragma unroll
      for (int k = 0; k < (1 << r0); ++k)
        dst_regs[u * (1 << r0) + k] =
            gload(src + min(base + elem<LOGM, 0, E>(PhaseGroups<LOGM, 0, 0, E>::gid(g, u), k), last));
  };
  auto load_fix = [&](uint32_t in_count, uint32_t* dst_regs) {
    if (in_count == 0) {
#pragma unroll
      for (int i = 0; i < NE0; ++i) dst_regs[i] = 0u;
      return;
    }
    if (lift) {
#pragma unroll
      for (int i = 0; i < NE0; ++i) dst_regs[i] = A.mul(dst_regs[i], J.f.r2);
    }
    if (whole_row(in_count)) return;
#pragma unroll
    for (int u = 0; u < PhaseGroups<LOGM, 0, 0, E>::NG; ++u)
#pragma unroll
      for (int k = 0; k < (1 << r0); ++k) {
        const uint32_t gi = base + elem<LOGM, 0, E>(PhaseGroups<LOGM, 0, 0, E>::gid(g, u), k);
        if (!(active && gi < in_count)) dst_regs[u * (1 << r0) + k] = 0u;
      }
  };
  auto load_first = [&](const uint32_t* src, uint32_t in_count, uint32_t* dst_regs) {
    load_issue(src, in_count, dst_regs);
    load_fix(in_count, dst_regs);
  };
  // Forward transform of registers holding phase 0's input, ending with phase
  // NPH-1's groups in regs; `forward` loads the input first.
  static_assert(MODE != R2M_FWDS || LAZY, "signed residues: lazy primes only");
  const ArithS Sg(J.f);
  auto forward_regs = [&](uint32_t* regs) {
    if constexpr (!kPF) tw_load<LOGM, 0, 0, E>(TF0, J.f.planes_fwd, g);
    if constexpr (MODE == R2M_FWDS) {
      // outputs below 4p in magnitude; the store brings them to [0, 2p)
      fwd_transform<LOGM, 0, LAZY, 4, 4, kPF, E>(regs, X, g, 0u, J.f.planes_fwd, Sg, TF0);
    } else {
      fwd_transform<LOGM, 0, LAZY, 2, 2, kPF, E>(regs, X, g, 0u, J.f.planes_fwd, A, TF0);
    }
  };
  auto forward = [&](const uint32_t* src, uint32_t in_count, uint32_t* regs) {
    load_first(src, in_count, regs);
    forward_regs(regs);
  };

  constexpr int rl = SC::r_of(PL);
  constexpr int NGL = PhaseGroups<LOGM, PL, 0, E>::NG;

  // Canonical outputs: min(v, v - p) with p, or a no-op min(v, v) with 0.
  const uint32_t canon_p = (flags & R2_CANON) ? J.f.p : 0u;
  if constexpr (MODE == R2M_FWD || MODE == R2M_FWDS) {
    forward(J.src_a, J.in_a, x);
#pragma unroll
    for (int u = 0; u < NGL; ++u)
#pragma unroll
      for (int k = 0; k < (1 << rl); ++k) {
        const uint32_t gi = base + elem<LOGM, PL, E>(PhaseGroups<LOGM, PL, 0, E>::gid(g, u), k);
        const uint32_t v = MODE == R2M_FWDS ? Sg.mul(x[u * (1 << rl) + k], Sg.one) + Sg.p : x[u * (1 << rl) + k];
        if (active) gstore(J.dst + gi, min(v, v - canon_p));
      }
    return;
  } else {
    if constexpr (MODE == R2M_CONV) {
      uint32_t y[16];
      load_issue(J.src_b, J.in_b, y);
      load_issue(J.src_a, J.in_a, x);
      load_fix(J.in_b, y);
      forward_regs(y);
      __syncthreads();
      load_fix(J.in_a, x);
      forward_regs(x);
#pragma unroll
      for (int i = 0; i < NGL * (1 << rl); ++i) x[i] = A.mul(x[i], y[i]);
    } else if constexpr (MODE == R2M_INVSTEP) {
      static_assert(LOGM >= E && PhaseGroups<LOGM, 0, 0, E>::NG == 1, "the half shift below needs a radix-2^E first phase");
      uint32_t y[16];
      load_issue(J.src_b, J.in_b, y);  // c
      load_issue(J.src_a, J.in_a, x);  // a
      load_fix(J.in_b, y);
      forward_regs(y);
      __syncthreads();
      load_fix(J.in_a, x);
      forward_regs(x);
#pragma unroll
      for (int i = 0; i < NGL * (1 << rl); ++i) x[i] = A.mul(x[i], y[i]);
      inv_transform<LOGM, 0, LAZY, true, E>(x, X, g, 0u, J.f.planes_inv, A, TI);
      // Register k of thread g holds position k * GT + g: the upper half
      // (k >= NE/2) is e's [0, k); keep e's first out_hi positions, zero the rest.
#pragma unroll
      for (int i = 0; i < NE / 2; ++i) {
        const uint32_t pos = (uint32_t)i * RowShape<LOGM, E>::GT + g;
        const uint32_t v = A.mul(x[i + NE / 2], J.inv_n);
        x[i] = pos < J.out_hi ? v : 0u;
        x[i + NE / 2] = 0u;
      }
      __syncthreads();  // the exchanges above are done with LDS
      forward_regs(x);
#pragma unroll
      for (int i = 0; i < NGL * (1 << rl); ++i) x[i] = A.mul(x[i], y[i]);
    } else if constexpr (MODE == R2M_EXPSTEP || MODE == R2M_SQRTSTEP) {
      static_assert(LOGM >= E && PhaseGroups<LOGM, 0, 0, E>::NG == 1, "positions k * GT + g need a radix-2^E first phase");
      constexpr uint32_t GT = RowShape<LOGM, E>::GT;
      const uint32_t h = J.out_hi;
      uint32_t yb[16], yc[16], ep[16];
      load_issue(J.src_b, J.in_b, yb);  // b[0, m)
      load_issue(J.src_c, h, yc);       // c[0, h)
      if constexpr (MODE == R2M_EXPSTEP) load_issue(J.src_a, J.in_a, x);  // a'
      // The epilogue operand at positions i * GT + g (EXPSTEP: 1/(m + pos)
      // for pos < h; SQRTSTEP: a[v + m + pos] for pos < epi_avail), clamped.
      {
        const uint32_t ecount = MODE == R2M_EXPSTEP ? h : J.epi_avail;
        if (ecount) {
#pragma unroll
          for (int i = 0; i < NE; ++i) ep[i] = gload(J.epi + min((uint32_t)i * GT + g, ecount - 1));
        } else {
#pragma unroll
          for (int i = 0; i < NE; ++i) ep[i] = 0u;
        }
      }
      load_fix(J.in_b, yb);
      forward_regs(yb);
      __syncthreads();
      load_fix(h, yc);
      forward_regs(yc);
      __syncthreads();
      // Inverse of regs into phase 0's layout (register k of thread g holds
      // position k * GT + g), times 1/N: values in [0, 2p).
      auto inverse_scaled = [&](uint32_t* r) {
        inv_transform<LOGM, 0, LAZY, true, E>(r, X, g, 0u, J.f.planes_inv, A, TI);
#pragma unroll
        for (int i = 0; i < NE; ++i) r[i] = A.mul(r[i], J.inv_n);
      };
      if constexpr (MODE == R2M_EXPSTEP) {
        load_fix(J.in_a, x);
        forward_regs(x);
#pragma unroll
        for (int i = 0; i < NE; ++i) x[i] = A.mul(x[i], yb[i]);
        inverse_scaled(x);
        // r = window [aux, aux + h): a shift by aux positions through LDS
        // (aux + h <= 2^LOGM - 1, inside the row's words).
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NE; ++i) X[pidx((uint32_t)i * GT + g)] = x[i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NE; ++i) {
          const uint32_t pos = (uint32_t)i * GT + g;
          x[i] = pos < h ? X[pidx(pos + J.aux)] : 0u;
        }
        __syncthreads();
        forward_regs(x);
#pragma unroll
        for (int i = 0; i < NE; ++i) x[i] = A.mul(x[i], yc[i]);
        inverse_scaled(x);
        // t = (r * c)[0, h) / (m + i)
#pragma unroll
        for (int i = 0; i < NE; ++i) {
          const uint32_t pos = (uint32_t)i * GT + g;
          x[i] = pos < h ? A.mul(x[i], ep[i]) : 0u;
        }
        __syncthreads();
        forward_regs(x);
#pragma unroll
        for (int i = 0; i < NE; ++i) x[i] = A.mul(x[i], yb[i]);
      } else {
#pragma unroll
        for (int i = 0; i < NE; ++i) x[i] = A.mul(yb[i], yb[i]);
        inverse_scaled(x);
        // (b * b)[m + i] is the upper half, registers NE/2 .. NE-1; res = a - it.
        const uint32_t p = J.f.p;
#pragma unroll
        for (int i = 0; i < NE / 2; ++i) {
          const uint32_t pos = (uint32_t)i * GT + g;
          const uint32_t v = A.canon(x[i + NE / 2]);
          const uint32_t s = pos < J.epi_avail ? ep[i] : 0u;
          const uint32_t r = s >= v ? s - v : s - v + p;
          x[i] = pos < h ? r : 0u;
          x[i + NE / 2] = 0u;
        }
        __syncthreads();
        forward_regs(x);
#pragma unroll
        for (int i = 0; i < NE; ++i) x[i] = A.mul(A.mul(x[i], yc[i]), J.aux);  // times aux = 1/2
      }
    } else if constexpr (MODE == R2M_SPEC) {
      // The spectrum (phase PL's grouping) is loaded with the operand.
      uint32_t w[16];
      load_issue(J.src_a, J.in_a, x);
#pragma unroll
      for (int u = 0; u < NGL; ++u)
#pragma unroll
        for (int k = 0; k < (1 << rl); ++k) {
          const uint32_t gi = base + elem<LOGM, PL, E>(PhaseGroups<LOGM, PL, 0, E>::gid(g, u), k);
          w[u * (1 << rl) + k] = gload(J.src_b + (active ? gi : 0u));
        }
      load_fix(J.in_a, x);
      forward_regs(x);
#pragma unroll
      for (int i = 0; i < NGL * (1 << rl); ++i) x[i] = A.mul(x[i], w[i]);
    } else {  // R2M_INV / R2M_SPEC_ONLY / R2M_SPEC_SUM2: load the spectrum in phase PL's grouping
#pragma unroll
      for (int u = 0; u < NGL; ++u)
#pragma unroll
        for (int k = 0; k < (1 << rl); ++k) {
          const uint32_t gi = active ? base + elem<LOGM, PL, E>(PhaseGroups<LOGM, PL, 0, E>::gid(g, u), k) : 0u;
          uint32_t v = gload(J.src_a + gi);
          if constexpr (MODE == R2M_SPEC_ONLY) v = A.mul(v, gload(J.src_b + gi));
          if constexpr (MODE == R2M_SPEC_SUM2)
            v = A.add(A.mul(v, gload(J.src_b + gi)), A.mul(gload(J.src_c + gi), gload(J.epi + gi)));
          x[u * (1 << rl) + k] = v;
        }
    }
    // Inverse: phase PL on registers, then PL-1 .. 0 through LDS.
    if constexpr (!kPF) tw_load<LOGM, PL, 0, E>(TI, J.f.planes_inv, g);
    inv_transform<LOGM, 0, LAZY, kPF, E>(x, X, g, 0u, J.f.planes_inv, A, TI);
    // Inverse outputs are in [0, 4p) (lazy); the scale brings them to [0, 2p).
    if (flags & R2_SCALE) {
#pragma unroll
      for (int i = 0; i < NE0; ++i) x[i] = A.mul(x[i], J.inv_n);
    } else if (flags & (R2_NEG | R2_CANON)) {
#pragma unroll
      for (int i = 0; i < NE0; ++i) x[i] = A.half(x[i]);
    }
    if (flags & R2_NEG) {
#pragma unroll
      for (int i = 0; i < NE0; ++i) x[i] = x[i] == 0 ? 0u : A.R - x[i];
    }
#pragma unroll
    for (int u = 0; u < PhaseGroups<LOGM, 0, 0, E>::NG; ++u)
#pragma unroll
      for (int k = 0; k < (1 << r0); ++k) {
        const uint32_t gi = base + elem<LOGM, 0, E>(PhaseGroups<LOGM, 0, 0, E>::gid(g, u), k);
        const uint32_t v = x[u * (1 << r0) + k];
        if (active && gi >= J.out_lo && gi < J.out_hi) {
          uint32_t w = min(v, v - canon_p);
          if (flags & (R2_SUBFROM | R2_MULTAB))
            w = out_epilogue(flags, J.epi, J.epi_avail, gi - J.out_shift, w, J.f.p, J.f.ninv);
          gstore(J.dst + (gi - J.out_shift), w);
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Persistent forward row pass for many rows of 2^14 words (the second pass of
// large transforms).  One 1024-thread workgroup per CU walks rows blockIdx.x,
// + gridDim.x, ...; while it transforms a row, the next row streams into a
// second LDS buffer by global_load_lds_dwordx4 (no registers), and every
// phase's twiddles stay in registers for all rows.  The only wait for that
// prefetch is the vmcnt(0) right before the row's stores: by then the
// prefetch (issued after the first phase) and the previous row's stores
// (issued a whole row earlier) are done, so the wait is short and relies on
// no ordering between loads and stores.  Barriers inside use LDS-only waits
// (s_waitcnt lgkmcnt(0) + s_barrier), not __syncthreads, whose workgroup
// fence would wait for the prefetch.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt left at their maxima
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0)

template <int LOGB, int P>
__device__ __forceinline__ void load_phase_tw(uint32_t* T, const uint32_t* planes, uint32_t g) {
  constexpr int r = Sched<LOGB>::r_of(P);
  const uint32_t* plane = planes + plane_off(LOGB - Sched<LOGB>::s_of(P), r);
#pragma unroll
  for (int u = 0; u < PhaseGroups<LOGB, P>::NG; ++u)
    load_tw<r>(T + u * (1 << r), plane, group_lo<LOGB, P>(PhaseGroups<LOGB, P>::gid(g, u)));
}
template <int LOGB, int P, bool LAZY, int BOUT = 2, class AR = Arith<LAZY>>
__device__ __forceinline__ void compute_with(uint32_t* x, const uint32_t* T, const AR& A) {
  constexpr int r = Sched<LOGB>::r_of(P);
  constexpr bool lo0 = LOGB - Sched<LOGB>::s_of(P) - r == 0;
#pragma unroll
  for (int u = 0; u < PhaseGroups<LOGB, P>::NG; ++u) dif_any<r, lo0, BOUT, LAZY>(x + u * (1 << r), T + u * (1 << r), A);
}

constexpr int kDmaRowLog = 14;
// By-value jobs of one persistent row launch (the pieces of a split product
// share one launch).
constexpr int kDmaPack = 8;

// SGN: signed residues (ArithS) for p < 2^28; the row's values are converted
// back to [0, 2p) by one product with the Montgomery one and + p.
template <bool LAZY, bool SGN>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4)))
k_rows_fwd_dma(JobPack<RowJob2, kDmaPack> jp, uint32_t njobs) {
  constexpr int LOGM = kDmaRowLog;
  static_assert(!SGN || LAZY, "signed residues: lazy primes only");
  using SC = Sched<LOGM>;
  static_assert(SC::nph == 4 && SC::r_of(0) == 4 && SC::r_of(1) == 2 && SC::r_of(2) == 4 && SC::r_of(3) == 4,
                "phase shape of a 2^14 row");
  static_assert(!WaveLocal<LOGM, 1, 0>::value && WaveLocal<LOGM, 2, 0>::value && WaveLocal<LOGM, 3, 0>::value,
                "one workgroup-wide exchange, two inside waves");
  static_assert(RowShape<LOGM>::T == 1024 && RowShape<LOGM>::U == 1, "one row a workgroup");
  __shared__ uint32_t LA[1u << LOGM];          // prefetched row, unpadded
  __shared__ uint32_t LE[lds_words(1u << LOGM)];  // phase exchanges, padded
  // njobs jobs of J.nrows rows each (one field, flags and size; their own
  // src / dst): rows are numbered across the jobs.
  const RowJob2 J = load_job(jp.j[0], (const RowJob2*)nullptr);
  const uint32_t total = J.nrows * njobs;
  auto row_src = [&](uint32_t r) -> const uint32_t* {
    return jp.j[r / J.nrows].src_a + ((size_t)(r % J.nrows) << LOGM);
  };
  auto row_dst = [&](uint32_t r) -> uint32_t* { return jp.j[r / J.nrows].dst + ((size_t)(r % J.nrows) << LOGM); };
  const Arith<LAZY> A(J.f);
  const ArithS Sg(J.f);
  auto comp = [&](auto ph, uint32_t* x, const uint32_t* T) {
    constexpr int P = decltype(ph)::value;
    if constexpr (SGN) compute_with<LOGM, P, LAZY, P == 3 ? 4 : 2>(x, T, Sg);
    else compute_with<LOGM, P, LAZY>(x, T, A);
  };
  const uint32_t g = threadIdx.x;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(g >> 6), ln = g & 63;
  const uint32_t canon_p = (J.flags & R2_CANON) ? J.f.p : 0u;
  // Row r into LA: 64 chunks of 1 KiB (64 lanes x 16 bytes), four per wave.
  auto prefetch = [&](uint32_t r) {
    const uint32_t* src = row_src(r);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t chunk = wv * 4 + i;
      __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void*)(src + chunk * 256 + ln * 4),
                                       (__attribute__((address_space(3))) void*)(LA + chunk * 256), 16, 0, 0);
    }
  };
  uint32_t T0[16], T1[16], T2[16], T3[16];
  load_phase_tw<LOGM, 0>(T0, J.f.planes_fwd, g);
  load_phase_tw<LOGM, 1>(T1, J.f.planes_fwd, g);
  load_phase_tw<LOGM, 2>(T2, J.f.planes_fwd, g);
  load_phase_tw<LOGM, 3>(T3, J.f.planes_fwd, g);
  uint32_t row = blockIdx.x;
  prefetch(row);
  wait_vmem();
  for (;;) {
    lds_barrier();  // every wave's part of this row has landed (each waited before its stores)
    uint32_t x[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = LA[elem<LOGM, 0>(g, k)];
    comp(std::integral_constant<int, 0>(), x, T0);
    const uint32_t next = row + gridDim.x;
    lds_barrier();  // LA consumed by every wave; the previous row's LE reads are done
    // Always issued (the last row re-reads a valid row) so the code stays straight-line.
    prefetch(min(next, total - 1));
    phase_to_lds<LOGM, 0, 0>(LE, x, g, 0);
    lds_barrier();
    phase_from_lds<LOGM, 1, 0>(x, LE, g, 0);
    comp(std::integral_constant<int, 1>(), x, T1);
    wave_sync();
    phase_to_lds<LOGM, 1, 0>(LE, x, g, 0);
    wave_sync();
    phase_from_lds<LOGM, 2, 0>(x, LE, g, 0);
    comp(std::integral_constant<int, 2>(), x, T2);
    wave_sync();
    phase_to_lds<LOGM, 2, 0>(LE, x, g, 0);
    wave_sync();
    phase_from_lds<LOGM, 3, 0>(x, LE, g, 0);
    comp(std::integral_constant<int, 3>(), x, T3);
    // Coalesced stores: the wave's outputs are the 1024 consecutive words
    // elem = g * 16 + k of its own phase-3 slice of LE (the 2 -> 3 exchange
    // stays inside the wave), so they go back through that slice and leave
    // as 16 stores of 64 consecutive words.
    static_assert(LOGM - SC::s_of(3) - SC::r_of(3) == 0, "last phase groups are contiguous");
    wave_sync();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint32_t v = SGN ? Sg.mul(x[k], Sg.one) + Sg.p : x[k];
      LE[pidx(elem<LOGM, 3>(g, k))] = min(v, v - canon_p);
    }
    wave_sync();
    wait_vmem();  // the prefetch of the next row and this wave's previous stores
    uint32_t* const out = row_dst(row);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const uint32_t e = wv * 1024 + j * 64 + ln;
      gstore(out + e, LE[pidx(e)]);
    }
    row = next;
    if (row >= total) break;
  }
  wait_vmem();  // the redundant last prefetch lands before the workgroup's LDS is released
}

constexpr uint32_t kMidTPO = 4;  // threads an output of a direct middle product

// sum_j X[j] Y[o + off - j] over 0 <= j < nx, 0 <= o + off - j < ny, for an
// output o < h: the terms with j = par mod kMidTPO, added to those of the
// output's other kMidTPO - 1 lanes (adjacent lanes of one wave).  Canonical
// limbs; products are summed in 64 bits before each Montgomery reduction,
// four at a time for p < 2^30 (4 p^2 < p 2^32), else two (2 p^2 < p 2^32).
__device__ __forceinline__ uint32_t mid_product(const uint32_t* X, uint32_t nx, const uint32_t* Y, uint32_t ny,
                                                uint32_t off, uint32_t o, uint32_t par, uint32_t h, uint32_t p,
                                                uint32_t ninv) {
  constexpr uint32_t D = kMidTPO;
  uint32_t acc = 0;
  const uint32_t s = o + off;
  if (o < h && nx && ny) {
    const uint32_t jlo = s >= ny ? s - (ny - 1) : 0u;
    const uint32_t jhi = min(nx - 1, s);  // inclusive
    uint32_t j = jlo + ((par - jlo) & (D - 1));  // the first j >= jlo with j = par mod D
    if (p < (1u << 30)) {  // uniform: four products below p 2^32 (4 p < 2^32), one reduction
      for (; j + 3 * D <= jhi; j += 4 * D) {
        const uint64_t t = (uint64_t)X[j] * Y[s - j] + (uint64_t)X[j + D] * Y[s - j - D] +
                           (uint64_t)X[j + 2 * D] * Y[s - j - 2 * D] + (uint64_t)X[j + 3 * D] * Y[s - j - 3 * D];
        const uint32_t z = acc + redc_full(t, p, ninv);
        acc = z >= p ? z - p : z;
      }
    }
    for (; j + D <= jhi; j += 2 * D) {
      const uint32_t r = redc_full((uint64_t)X[j] * Y[s - j] + (uint64_t)X[j + D] * Y[s - j - D], p, ninv);
      const uint32_t z = acc + r;
      acc = z >= p ? z - p : z;
    }
    if (j <= jhi) {
      const uint32_t z = acc + redc_full((uint64_t)X[j] * Y[s - j], p, ninv);
      acc = z >= p ? z - p : z;
    }
  }
#pragma unroll
  for (uint32_t d = 1; d < D; d <<= 1) {
    const uint32_t z = acc + (uint32_t)__shfl_xor((int)acc, (int)d, 64);
    acc = z >= p ? z - p : z;
  }
  return acc;
}

// ---------------------------------------------------------------------------
// Newton steps with m = 2^LM <= 256 (transforms of 2^6..2^9 points): the
// step's window products, computed directly.  Every window the
// NTT step modes read is free of cyclic wraparound, so each is a plain middle
// product out[o] = sum_j X[j] * Y[o + off - j]; at these sizes 4m threads
// (four per output, j split mod 4) finish one in a few hundred instructions,
// where the one-workgroup NTT step is a chain of radix-16 blocks on a few
// waves.  Same job fields and results as the R2M_INVSTEP / R2M_EXPSTEP /
// R2M_SQRTSTEP row modes, which serve the larger steps:
//   INVSTEP : e = (a c)[k + i], out[i] = -(e c)[i]        (inverse.hpp:18-65)
//   EXPSTEP : r = (a' b)[m - 1 + i], t = (r c)[i] / (m + i),
//             out[i] = (t b)[i]                           (exponential.hpp:172-204)
//   SQRTSTEP: res = a[v + m + i] - (b b)[m + i],
//             out[i] = (res c)[i] / 2                     (square_root.hpp:86-100)
// for i < out_hi.  Values are canonical limbs; two products are summed
// before each reduction (2 p^2 < p 2^32 for p < 2^31).  Only Montgomery
// arithmetic mod p is used, so any odd prime p < 2^31 is served, NTT prime or
// not.  One workgroup a job: the job by value, or jobs[blockIdx.y].
// ---------------------------------------------------------------------------
template <int MODE, int LM>
__global__ void __launch_bounds__(kMidTPO << LM) k_step_direct(RowJob2 j0, const RowJob2* __restrict__ jobs) {
  constexpr uint32_t M = 1u << LM;
  const RowJob2 J = load_job(j0, jobs);
  __shared__ uint32_t A[2 * M], B[M], Cc[M], W[M];
  const uint32_t t = threadIdx.x, p = J.f.p, ninv = J.f.ninv;
  const uint32_t h = J.out_hi;  // outputs, <= M
  const uint32_t o = t / kMidTPO, par = t % kMidTPO;
  auto mid = [&](const uint32_t* X, uint32_t nx, const uint32_t* Y, uint32_t ny, uint32_t off) -> uint32_t {
    return mid_product(X, nx, Y, ny, off, o, par, h, p, ninv);
  };
  auto stage = [&](uint32_t* dst, const uint32_t* src, uint32_t count, uint32_t cap) {
    for (uint32_t i = t; i < cap; i += kMidTPO * M) dst[i] = i < count ? src[i] : 0u;
  };
  if constexpr (MODE == R2M_INVSTEP) {
    const uint32_t k = J.in_b;
    stage(A, J.src_a, min(J.in_a, 2 * M), 2 * M);
    stage(B, J.src_b, k, M);
    __syncthreads();
    const uint32_t e = mid(A, min(J.in_a, 2 * M), B, k, k);
    if (par == 0 && o < h) W[o] = e;
    __syncthreads();
    const uint32_t v = mid(W, h, B, k, 0);
    if (par == 0 && o < h) J.dst[o] = v == 0 ? 0u : p - v;
  } else if constexpr (MODE == R2M_EXPSTEP) {
    const uint32_t m = J.in_b;
    stage(A, J.src_a, min(J.in_a, 2 * M), 2 * M);
    stage(B, J.src_b, m, M);
    stage(Cc, J.src_c, h, M);
    __syncthreads();
    const uint32_t r = mid(A, min(J.in_a, 2 * M), B, m, J.aux);
    if (par == 0 && o < h) W[o] = r;
    __syncthreads();
    uint32_t tt = mid(W, h, Cc, h, 0);
    __syncthreads();  // W is rewritten below
    if (par == 0 && o < h) W[o] = mulmod_full(tt, J.epi[o], p, ninv);
    __syncthreads();
    const uint32_t v = mid(W, h, B, m, 0);
    if (par == 0 && o < h) J.dst[o] = v;
  } else {
    static_assert(MODE == R2M_SQRTSTEP, "direct Newton steps: INVSTEP, EXPSTEP, SQRTSTEP");
    const uint32_t m = J.in_b;
    stage(B, J.src_b, m, M);
    stage(Cc, J.src_c, h, M);
    __syncthreads();
    const uint32_t sq = mid(B, m, B, m, m);
    if (par == 0 && o < h) {
      const uint32_t a = o < J.epi_avail ? J.epi[o] : 0u;
      W[o] = a >= sq ? a - sq : a - sq + p;
    }
    __syncthreads();
    const uint32_t v = mid(W, h, Cc, h, 0);
    if (par == 0 && o < h) J.dst[o] = mulmod_full(v, J.aux, p, ninv);
  }
}

// ---------------------------------------------------------------------------
// The first N <= 512 coefficients of a Newton series in one workgroup, from
// the constant term up: doublings m -> w = min(2m, N) whose products are the
// same direct middle products as k_step_direct, every operand in LDS.  It
// stands in for the reference's 32-coefficient seed recurrences and first
// doubling steps (inverse.hpp:68-77 and 18-65, exponential.hpp:158-204,
// square_root.hpp:46-100): every output coefficient is unique, so the limbs
// are the same.  Job fields (out_hi = N):
//   R2M_NINV : c = 1/a mod x^N; src_a = a (in_a words, a[0] != 0), dst = c.
//   R2M_NEXP : b = exp(a) mod x^N; src_a = a' (in_a words), epi = 1/i table,
//              dst = b; src_c (an output here) <- c = 1/b mod x^N.
//   R2M_NSQRT: b = sqrt(a) mod x^N with b[0] = aux; src_a = a (in_a words),
//              inv_n = 1/aux, dst = b; src_c (an output) <- c = 1/b mod x^N.
// Each doubling keeps c = 1/b to the precision of the next step's half
// (extend_c, the inverse step with b in place of a).  kSmallThreads =
// 1024 threads, four an output of each product (each product has at most
// N/2 outputs).
// ---------------------------------------------------------------------------
constexpr uint32_t kSmallSeries = 512;
constexpr int kSmallSeriesLog = 9;  // the size key its launches are recorded under
constexpr uint32_t kSmallThreads = kSmallSeries / 2 * kMidTPO;

template <int MODE>
__global__ void __launch_bounds__(kSmallThreads) k_newton_small(RowJob2 j0, const RowJob2* __restrict__ jobs) {
  constexpr uint32_t NM = kSmallSeries;
  const RowJob2 J = load_job(j0, jobs);
  // RT (exp): the reciprocal table 1/(m + i) for every step, loaded with the
  // input so that no step waits on a global load.
  __shared__ uint32_t A[NM], B[NM], C[NM], E[NM / 2], RT[MODE == R2M_NEXP ? NM : 1];
  const uint32_t t = threadIdx.x, p = J.f.p, ninv = J.f.ninv;
  const uint32_t N = min(J.out_hi, NM), na = min(J.in_a, N);
  const uint32_t o = t / kMidTPO, par = t % kMidTPO;
  for (uint32_t i = t; i < NM; i += NM) {
    A[i] = i < na ? J.src_a[i] : 0u;
    B[i] = 0u;
    C[i] = 0u;
    if constexpr (MODE == R2M_NEXP) RT[i] = i < N ? J.epi[i] : 0u;
  }
  if (t == 0) {
    if constexpr (MODE == R2M_NINV) {
      // 1/a[0] = a[0]^(p-2) (Montgomery limbs), read from the device: the
      // schedule stays independent of the input's values.
      uint32_t r = J.f.one, x = na ? J.src_a[0] : 0u;
      for (uint32_t e = p - 2; e; e >>= 1) {
        if (e & 1) r = mulmod_full(r, x, p, ninv);
        x = mulmod_full(x, x, p, ninv);
      }
      C[0] = r;
    } else if constexpr (MODE == R2M_NEXP) {
      B[0] = J.f.one;
      C[0] = J.f.one;
    } else {
      B[0] = J.aux;
      C[0] = J.inv_n;
    }
  }
  __syncthreads();
  auto mp = [&](const uint32_t* X, uint32_t nx, const uint32_t* Y, uint32_t ny, uint32_t off, uint32_t h) {
    return mid_product(X, nx, Y, ny, off, o, par, h, p, ninv);
  };
  // c[k, w) from c[0, k) = 1/x mod x^k, x = X[0, nx): e = (x c)[k + i],
  // c[k + i] = -(e c)[i] (inverse.hpp:18-65).  Reads C[0, k) only, writes C[k, w).
  auto extend_c = [&](const uint32_t* X, uint32_t nx, uint32_t k, uint32_t w) {
    const uint32_t hh = w - k;
    const uint32_t e = mp(X, min(nx, w), C, k, k, hh);
    if (par == 0 && o < hh) E[o] = e;
    __syncthreads();
    const uint32_t v = mp(E, hh, C, k, 0, hh);
    if (par == 0 && o < hh) C[k + o] = v == 0 ? 0u : p - v;
    __syncthreads();
  };
  uint32_t cp = 1;  // precision of c
  for (uint32_t m = 1; m < N; m *= 2) {
    const uint32_t w = min(2 * m, N), h = w - m;
    if constexpr (MODE == R2M_NINV) {
      extend_c(A, na, m, w);
    } else {
      if (cp < h) {
        extend_c(B, m, cp, h);
        cp = h;
      }
      if constexpr (MODE == R2M_NEXP) {
        // r = (a' b)[m - 1 + i]; t = (r c)[i] / (m + i); b[m + i] = (t b)[i]
        const uint32_t r = mp(A, min(na, w - 1), B, m, m - 1, h);
        if (par == 0 && o < h) E[o] = r;
        __syncthreads();
        const uint32_t rc = mp(E, h, C, h, 0, h);
        __syncthreads();  // E is rewritten below
        if (par == 0 && o < h) E[o] = mulmod_full(rc, RT[m + o], p, ninv);
        __syncthreads();
        const uint32_t v = mp(E, h, B, m, 0, h);
        if (par == 0 && o < h) B[m + o] = v;
      } else {
        // res = a[m + i] - (b b)[m + i]; b[m + i] = (res c)[i] / 2
        const uint32_t sq = mp(B, m, B, m, m, h);
        if (par == 0 && o < h) {
          const uint32_t av = A[m + o];
          E[o] = av >= sq ? av - sq : av - sq + p;
        }
        __syncthreads();
        const uint32_t v = mp(E, h, C, h, 0, h);
        const uint32_t half = mulmod_full((p + 1) / 2, J.f.r2, p, ninv);
        if (par == 0 && o < h) B[m + o] = mulmod_full(v, half, p, ninv);
      }
      __syncthreads();
    }
  }
  if constexpr (MODE != R2M_NINV) {
    // c to all N coefficients: the first Newton step after this launch needs
    // 1/b to its full length, and b[0, N) is final here.
    if (cp < N) {
      extend_c(B, N, cp, N);
      cp = N;
    }
  }
  if constexpr (MODE == R2M_NINV) {
    for (uint32_t i = t; i < N; i += NM) J.dst[i] = C[i];
  } else {
    uint32_t* c_out = const_cast<uint32_t*>(J.src_c);
    for (uint32_t i = t; i < N; i += NM) J.dst[i] = B[i];
    for (uint32_t i = t; i < cp; i += NM) c_out[i] = C[i];
  }
}

// ---------------------------------------------------------------------------
// Column pass: a tile of C = 2^LOGC adjacent columns x 2^S rows of the
// 2^S x M row-major view, with the four-step twiddle w_N^(c * bitrev_S(row))
// applied after the forward levels (before the inverse ones).
// ---------------------------------------------------------------------------
struct ColJob2 {
  const uint32_t* src;
  uint32_t* dst;
  uint32_t in_count;
  uint32_t out_lo, out_hi, out_shift;
  uint32_t inv_n;
  uint32_t flags;  // R2_LIFT, R2_CANON, R2_NEG, R2_ACC, R2_SUBFROM, R2_MULTAB
  const uint32_t* epi;
  uint32_t epi_avail, acc_end;  // acc_end: R2_ACCLO
  FieldDev f;
};

// By-value jobs per column launch: both operands of a product for each of
// the three auxiliary primes.
constexpr int kColPack = 6;

// g^e (inverse: g^-e) for e < 2^logmax, g the field's coset generator: the
// coset tables follow the inverse twiddle planes in the field's table block,
// split like the root tables (lo[e & mask], hi[e >> split]).
template <class Ar>
__device__ __forceinline__ uint32_t coset_pow(const FieldDev& f, const Ar& A, uint32_t e, bool inv) {
  typedef __attribute__((address_space(1))) const uint32_t g32;
  const uint32_t lo_n = 1u << f.split, hi_n = 1u << (f.logmax - f.split);
  const uint32_t* ct = f.planes_inv + kPlaneWords + (inv ? lo_n : 0u);
  const uint32_t* ch = f.planes_inv + kPlaneWords + 2 * lo_n + (inv ? hi_n : 0u);
  return A.mul(*(g32*)(ch + (e >> f.split)), *(g32*)(ct + (e & (lo_n - 1))));
}

template <int S, int LOGC, int E = 4>
struct ColShape {
  static constexpr int GT = Sched<S, E>::GT;
  static constexpr int C = 1 << LOGC;
  static constexpr int T = GT * C;
  static constexpr uint32_t words = lds_words((1u << S) << LOGC);
};

// SGN (forward, p < 2^28): the phases on signed residues (ArithS); the
// four-step twiddle product (factor < 2p, values below 4p in magnitude)
// brings them back, + p gives the usual [0, 2p).
// Column passes of up to 2^9 levels serve mid-size transforms, whose few
// workgroups are latency-bound: 128 registers, room to prefetch every
// phase's twiddles a phase ahead.  Longer columns (the streaming passes of
// the largest transforms) keep 64 registers for two workgroups a CU.
template <int S>
struct ColTune {
  static constexpr bool PF = S <= 9;
  static constexpr int waves = S <= 9 ? 4 : FPXG_COL_WPE;
};

// E: log2 of the elements a thread owns (4; 2 for the by-value launches of
// mid-size columns, whose threads then issue a quarter of the loads).
// CH (inverse only; the chained pass): the inverse column levels of one
// product, then, in the same tile, the forward column levels of the next
// product's operand taken from it: e[t] = v[t + out_shift] where t +
// out_shift lies in the window [out_lo, out_hi) (out_shift is 0, a quarter,
// half or minus a quarter of the transform, modulo 2^32: a register move,
// since register k of the first phase holds row k * 2^(S-E) + g), negated
// with R2_NEG, subtracted from epi[t] (0 from epi_avail on) with R2_SUBFROM,
// times epi[t] with R2_MULTAB; outside the window e[t] = epi[t]
// for t < epi_avail with R2_PREFIX (which also stores the window's e[t] to
// epi[t], positions at or above epi_avail), zero otherwise.  Its forward
// column spectrum (with the forward four-step twiddle, lazy limbs) goes to
// dst.  A Newton step's two dependent products then share a launch and a
// memory round trip.
template <int S, int LOGC, bool LAZY, bool INV, bool SGN = false, int E = 4, bool CH = false>
__global__ void __launch_bounds__((ColShape<S, LOGC, E>::T)) __attribute__((amdgpu_waves_per_eu(ColTune<S>::waves)))
k_cols2(JobPack<ColJob2, kColPack> jp, const ColJob2* __restrict__ jobs, int L) {
  using SC = Sched<S, E>;
  static_assert(!SGN || (LAZY && !INV), "signed residues: lazy forward passes only");
  static_assert(!CH || (INV && S >= E && PhaseGroups<S, 0, LOGC, E>::NG == 1), "chained pass: inverse, one group a phase-0 thread");
  constexpr int NPH = SC::nph;
  constexpr int C = 1 << LOGC;
  extern __shared__ uint32_t X[];
  const ColJob2 J = load_job(jp.j[0], jobs);
  const Arith<LAZY> A(J.f);
  const uint32_t flags = __builtin_amdgcn_readfirstlane(J.flags);
  const bool lift = (flags & R2_LIFT) != 0;
  const int b = L - S;
  const uint32_t M = 1u << b;
  const uint32_t canon_p = (flags & R2_CANON) ? J.f.p : 0u;
  const uint32_t tiles = M >> LOGC;
  uint32_t tile = blockIdx.x;
  if (tiles >= 64 && (tiles & 7) == 0) tile = (tile & 7) * (tiles >> 3) + (tile >> 3);
  const uint32_t c = threadIdx.x % C, g = threadIdx.x / C;
  const uint32_t col = (tile << LOGC) + c;
  const uint32_t shift = J.f.logmax - L;
  uint32_t x[16];
  constexpr int PL = NPH - 1;
  constexpr int rl = SC::r_of(PL);
  constexpr int NGL = PhaseGroups<S, PL, LOGC, E>::NG;
  typedef __attribute__((address_space(1))) const uint32_t g32;
  const uint32_t split = J.f.split, smask = (1u << split) - 1;
  // Four-step twiddle w^(col * bitrev_S(e)) (inverse: w^-(...)).  In the last
  // forward phase e = gid * 2^rl + k, so the exponent is col *
  // bitrev_(S-rl)(gid) + j * col * 2^(S-rl) with j = bitrev_rl(k): one table
  // lookup (two words, w^x = hi[x >> split] * lo[x & mask]) for each group's
  // first factor and one for the ratio D, then a running product.  The
  // lookups are issued here, before the data, so they cost no round trip of
  // their own.
  const uint32_t* twl = INV ? J.f.tw_lo_inv : J.f.tw_lo;
  const uint32_t* twh = INV ? J.f.tw_hi_inv : J.f.tw_hi;
  // (Passes held to 64 registers issue them where they are used.)
  uint32_t d_hi, d_lo, q_hi[NGL], q_lo[NGL];
  auto fourstep_load = [&] {
    const uint32_t xd = (col << (S - rl)) << shift;
    d_hi = *(g32*)(twh + (xd >> split));
    d_lo = *(g32*)(twl + (xd & smask));
#pragma unroll
    for (int u = 0; u < NGL; ++u) {
      const uint32_t xq = (col * bitrev(PhaseGroups<S, PL, LOGC, E>::gid(g, u) << rl, S)) << shift;
      q_hi[u] = *(g32*)(twh + (xq >> split));
      q_lo[u] = *(g32*)(twl + (xq & smask));
    }
  };
  if constexpr (ColTune<S>::PF || INV) fourstep_load();
  // The chained pass's forward four-step factors (forward tables), issued
  // with the inverse's.
  uint32_t fd_hi = 0, fd_lo = 0, fq_hi[CH ? NGL : 1], fq_lo[CH ? NGL : 1];
  if constexpr (CH) {
    const uint32_t xd = (col << (S - rl)) << shift;
    fd_hi = *(g32*)(J.f.tw_hi + (xd >> split));
    fd_lo = *(g32*)(J.f.tw_lo + (xd & smask));
#pragma unroll
    for (int u = 0; u < NGL; ++u) {
      const uint32_t xq = (col * bitrev(PhaseGroups<S, PL, LOGC, E>::gid(g, u) << rl, S)) << shift;
      fq_hi[u] = *(g32*)(J.f.tw_hi + (xq >> split));
      fq_lo[u] = *(g32*)(J.f.tw_lo + (xq & smask));
    }
  }
  if constexpr (!INV) {
    constexpr int r0 = SC::r_of(0);
    constexpr int NE0 = PhaseGroups<S, 0, LOGC, E>::NG * (1 << r0);
    uint32_t T0[16];
    if constexpr (ColTune<S>::PF) tw_load<S, 0, LOGC, E>(T0, J.f.planes_fwd, g);  // with the data
    if (J.in_count == 0) {
#pragma unroll
      for (int i = 0; i < NE0; ++i) x[i] = 0u;
    } else {
      // Buffer loads whose range ends at in_count: the zero tail reads 0
      // without clamps or compares.  Element (u, k) is at word
      // elem(gid_u, 0) * M + col + (k << lo_bits) * M.
      const uint32_t N = M << S;
      const __amdgpu_buffer_rsrc_t rs = buf_rsrc(J.src, (J.in_count < N ? J.in_count : N) * 4);
      constexpr int lo0 = S - r0;
#pragma unroll
      for (int u = 0; u < PhaseGroups<S, 0, LOGC, E>::NG; ++u) {
        const uint32_t lane = (elem<S, 0, E>(PhaseGroups<S, 0, LOGC, E>::gid(g, u), 0) * M + col) * 4;
#pragma unroll
        for (int k = 0; k < (1 << r0); ++k) x[u * (1 << r0) + k] = bload(rs, lane + (((uint32_t)k * M) << (lo0 + 2)), 0);
      }
      if (lift) {
#pragma unroll
        for (int i = 0; i < NE0; ++i) x[i] = A.mul(x[i], J.f.r2);
      }
      if (flags & R2_TWIST) {
        // Element (u, k) is coefficient t = gid_u * M + col + k * (M << lo0):
        // g^t from the coset tables for k = 0, then a running product.
        const uint32_t Dg = coset_pow(J.f, A, M << lo0, false);
#pragma unroll
        for (int u = 0; u < PhaseGroups<S, 0, LOGC, E>::NG; ++u) {
          uint32_t q = coset_pow(J.f, A, PhaseGroups<S, 0, LOGC, E>::gid(g, u) * M + col, false);
#pragma unroll
          for (int k = 0; k < (1 << r0); ++k) {
            x[u * (1 << r0) + k] = A.mul(x[u * (1 << r0) + k], q);
            if (k + 1 < (1 << r0)) q = A.mul(q, Dg);
          }
        }
      }
    }
    const ArithS Sg(J.f);
    if constexpr (!ColTune<S>::PF) tw_load<S, 0, LOGC, E>(T0, J.f.planes_fwd, g);  // after the data: registers
    if constexpr (SGN) fwd_transform<S, LOGC, LAZY, 4, 4, ColTune<S>::PF, E>(x, X, g, c, J.f.planes_fwd, Sg, T0);
    else fwd_transform<S, LOGC, LAZY, 2, 2, ColTune<S>::PF, E>(x, X, g, c, J.f.planes_fwd, A, T0);
    if constexpr (!ColTune<S>::PF) fourstep_load();
    const uint32_t D = A.mul(d_hi, d_lo);
    // Stores at a per-lane offset plus a uniform one: element (u, k) of the
    // last phase is at word elem(gid_u, 0) * M + col + (k << lo_bits) * M.
    const __amdgpu_buffer_rsrc_t rd = buf_rsrc(J.dst, (M << S) * 4);
    constexpr int lol = S - Sched<S, E>::s_of(PL) - rl;
#pragma unroll
    for (int u = 0; u < NGL; ++u) {
      const uint32_t gid = PhaseGroups<S, PL, LOGC, E>::gid(g, u);
      const uint32_t lane = (elem<S, PL, E>(gid, 0) * M + col) * 4;
      uint32_t q = A.mul(q_hi[u], q_lo[u]);
#pragma unroll
      for (int j = 0; j < (1 << rl); ++j) {
        const int k = (int)bitrev_c(j, rl);
        const uint32_t v = SGN ? Sg.mul(x[u * (1 << rl) + k], q) + Sg.p : A.mul(x[u * (1 << rl) + k], q);
        bstore(rd, min(v, v - canon_p), lane, ((uint32_t)k * M) << (lol + 2));
        if (j + 1 < (1 << rl)) q = A.mul(q, D);
      }
    }
  } else {
    constexpr int r0 = SC::r_of(0);
    constexpr int NE0 = PhaseGroups<S, 0, LOGC, E>::NG * (1 << r0);
    const bool neg = (flags & R2_NEG) != 0;
    uint32_t TI[16];
    tw_load<S, PL, LOGC, E>(TI, J.f.planes_inv, g);
    // The spectrum, in the last phase's grouping.
#pragma unroll
    for (int u = 0; u < NGL; ++u)
#pragma unroll
      for (int k = 0; k < (1 << rl); ++k)
        x[u * (1 << rl) + k] = gload(J.src + elem<S, PL, E>(PhaseGroups<S, PL, LOGC, E>::gid(g, u), k) * M + col);
    // The epilogue operand of every output (epi for R2_SUBFROM / R2_MULTAB,
    // else the destination itself for R2_ACC / R2_ACCLO), issued with the
    // data at clamped positions; unused values are never read.
    const bool epi_op = (flags & (R2_SUBFROM | R2_MULTAB | R2_COSET)) != 0;
    const bool acc_op = (flags & (R2_ACC | R2_ACCLO)) != 0;
    const uint32_t win = J.out_hi > J.out_shift ? J.out_hi - J.out_shift : 0u;  // output words past out_shift
    const uint32_t elen = epi_op ? ((flags & R2_SUBFROM) ? J.epi_avail : win) : (acc_op ? win : 0u);
    // (Passes held to 64 registers load them where they are used.)
    constexpr bool kEvPF = ColTune<S>::PF;
    uint32_t ev[NE0];
    if constexpr (CH) {
      // The chained operand's words at its own positions t: the factors
      // epi[t] of its window (R2_MULTAB), the minuends epi[t] (R2_SUBFROM) or
      // its prefix epi[t] (R2_PREFIX).
      const uint32_t ecnt = (flags & (R2_PREFIX | R2_SUBFROM)) ? J.epi_avail : J.out_hi - J.out_shift;
      if ((flags & (R2_MULTAB | R2_PREFIX | R2_SUBFROM)) && ecnt != 0) {
#pragma unroll
        for (int k = 0; k < NE0; ++k) ev[k] = gload(J.epi + min(elem<S, 0, E>(g, k) * M + col, ecnt - 1));
      }
    } else if (kEvPF && elen != 0) {
      const uint32_t* esrc = epi_op ? J.epi : J.dst;
      const uint32_t elim = elen - 1;
#pragma unroll
      for (int u = 0; u < PhaseGroups<S, 0, LOGC, E>::NG; ++u)
#pragma unroll
        for (int k = 0; k < (1 << r0); ++k) {
          const uint32_t gi = elem<S, 0, E>(PhaseGroups<S, 0, LOGC, E>::gid(g, u), k) * M + col;
          const uint32_t t = gi >= J.out_shift ? gi - J.out_shift : 0u;
          ev[u * (1 << r0) + k] = gload(esrc + min(t, elim));
        }
    }
    // Inverse four-step twiddle times 1/N, by the same running product.
    const uint32_t D = A.mul(d_hi, d_lo);
#pragma unroll
    for (int u = 0; u < NGL; ++u) {
      uint32_t q = A.mul(A.mul(q_hi[u], q_lo[u]), J.inv_n);
#pragma unroll
      for (int j = 0; j < (1 << rl); ++j) {
        const int k = (int)bitrev_c(j, rl);
        x[u * (1 << rl) + k] = A.mul(x[u * (1 << rl) + k], q);
        if (j + 1 < (1 << rl)) q = A.mul(q, D);
      }
    }
    inv_transform<S, LOGC, LAZY, ColTune<S>::PF, E>(x, X, g, c, J.f.planes_inv, A, TI);
    if constexpr (CH) {
      constexpr int NE = 1 << E, HALF = NE / 2;
      static_assert(NE0 == NE, "phase 0 holds one full group a thread");
      uint32_t T0[16];
      tw_load<S, 0, LOGC, E>(T0, J.f.planes_fwd, g);
      // New position t takes old position t + out_shift when that lies in
      // the window [out_lo, out_hi): the shift is a whole number of quarter
      // transforms, so register k takes register k + DK.
      const uint32_t NN = M << S, sh = J.out_shift, wlo = J.out_lo, wn = J.out_hi - J.out_lo;
      const bool mul = (flags & R2_MULTAB) != 0, pre = (flags & R2_PREFIX) != 0;
      uint32_t* const pdst = const_cast<uint32_t*>(J.epi);
      uint32_t y[NE];
      auto take = [&](auto dkc) {
        constexpr int DK = decltype(dkc)::value;
#pragma unroll
        for (int k = 0; k < NE; ++k) {
          const uint32_t t = elem<S, 0, E>(g, k) * M + col;  // position in the new operand
          uint32_t v = 0u;
          if (k + DK >= 0 && k + DK < NE && t + sh - wlo < wn) {
            v = A.canon(A.half(x[min(max(k + DK, 0), NE - 1)]));
            if (flags & R2_NEG) v = v == 0 ? 0u : J.f.p - v;
            if (flags & R2_SUBFROM) {
              const uint32_t e = t < J.epi_avail ? ev[k] : 0u;
              v = e >= v ? e - v : e - v + J.f.p;
            }
            if (mul) v = mulmod_full(v, ev[k], J.f.p, J.f.ninv);
            if (pre) pdst[t] = v;
          } else if (pre && t < J.epi_avail) {
            v = ev[k];
          }
          y[k] = v;
        }
      };
      if (sh == 0) take(std::integral_constant<int, 0>());
      else if (sh == NN / 2) take(std::integral_constant<int, HALF>());
      else if (sh == NN / 4) take(std::integral_constant<int, NE / 4>());
      else take(std::integral_constant<int, -(NE / 4)>());
#pragma unroll
      for (int k = 0; k < NE; ++k) x[k] = y[k];
      __syncthreads();  // the inverse's last exchange has been read by every thread
      fwd_transform<S, LOGC, LAZY, 2, 2, ColTune<S>::PF, E>(x, X, g, c, J.f.planes_fwd, A, T0);
      const uint32_t D = A.mul(fd_hi, fd_lo);
      const __amdgpu_buffer_rsrc_t rd = buf_rsrc(J.dst, (M << S) * 4);
      constexpr int lol = S - Sched<S, E>::s_of(PL) - rl;
#pragma unroll
      for (int u = 0; u < NGL; ++u) {
        const uint32_t gid = PhaseGroups<S, PL, LOGC, E>::gid(g, u);
        const uint32_t lane = (elem<S, PL, E>(gid, 0) * M + col) * 4;
        uint32_t q = A.mul(fq_hi[u], fq_lo[u]);
#pragma unroll
        for (int j = 0; j < (1 << rl); ++j) {
          const int k = (int)bitrev_c(j, rl);
          bstore(rd, A.mul(x[u * (1 << rl) + k], q), lane, ((uint32_t)k * M) << (lol + 2));
          if (j + 1 < (1 << rl)) q = A.mul(q, D);
        }
      }
      return;
    }
    const uint32_t p = J.f.p;
    // R2_COSET: g^-gi by the forward twist's running product (position gi =
    // gid_u * M + col + k * (M << (S - r0))).
    const bool coset = (flags & R2_COSET) != 0;
    const uint32_t Dc = coset ? coset_pow(J.f, A, M << (S - r0), true) : 0u;
#pragma unroll
    for (int u = 0; u < PhaseGroups<S, 0, LOGC, E>::NG; ++u) {
      uint32_t cq = coset ? coset_pow(J.f, A, PhaseGroups<S, 0, LOGC, E>::gid(g, u) * M + col, true) : 0u;
#pragma unroll
      for (int k = 0; k < (1 << r0); ++k) {
        const uint32_t e = elem<S, 0, E>(PhaseGroups<S, 0, LOGC, E>::gid(g, u), k);
        const uint32_t gi = e * M + col;
        const uint32_t v = A.canon(A.half(x[u * (1 << r0) + k]));
        const uint32_t gq = A.canon(A.half(cq));
        if (coset && k + 1 < (1 << r0)) cq = A.mul(cq, Dc);
        if (gi >= J.out_lo && gi < J.out_hi) {
          const uint32_t t = gi - J.out_shift;
          uint32_t w = neg ? (v == 0 ? 0u : p - v) : v;
          if (flags & R2_SUBFROM) {
            const uint32_t sv = t < J.epi_avail ? (kEvPF ? ev[u * (1 << r0) + k] : gload(J.epi + t)) : 0u;
            w = sv >= w ? sv - w : sv - w + p;
          }
          if (flags & R2_MULTAB)
            w = mulmod_full(w, kEvPF ? ev[u * (1 << r0) + k] : gload(J.epi + t), p, J.f.ninv);
          if (coset) {
            const uint32_t c1 = kEvPF ? ev[u * (1 << r0) + k] : gload(J.epi + t);
            const uint32_t d = mulmod_full(w, gq, p, J.f.ninv);
            const uint32_t h = mulmod_full(d >= c1 ? d - c1 : d - c1 + p, J.acc_end, p, J.f.ninv);
            if (gi < J.epi_avail) gstore(J.dst + (M << S) + gi, h);
            w = c1 >= h ? c1 - h : c1 - h + p;
          }
          if ((flags & R2_ACC) || ((flags & R2_ACCLO) && t < J.acc_end)) {
            // the destination's old value: prefetched unless an epilogue operand took the registers
            const uint32_t old = (epi_op || !kEvPF) ? gload(J.dst + t) : ev[u * (1 << r0) + k];
            const uint32_t s2 = old + w;
            w = min(s2, s2 - p);
          }
          gstore(J.dst + t, w);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// The inverse column pass of a CRT product over all its primes, Garner in the
// epilogue (restore_crt, crt_convolution.hpp:88-127, fused): each workgroup
// takes its tile of every prime's row-pass output in turn, runs the inverse
// levels, and folds the canonical residues into Garner's running state (two
// words an element); after the last prime it writes the target limbs of its
// output window.  No residue array is written or read back, and there is no
// Garner launch.
// ---------------------------------------------------------------------------
struct CrtFields {
  FieldDev f[3];
};
// By-value jobs per launch (one product; two when products are merged).
constexpr int kCrtPack = 2;

// 128 registers (the Garner state is two words an element beside the data);
// twiddles are not loaded a phase ahead, which would spill.
template <int S>
struct ColTuneCrt {
  static constexpr bool PF = false;
  static constexpr int waves = 4;
};

template <int S, int LOGC, bool LAZY, int E = 4>
__global__ void __launch_bounds__((ColShape<S, LOGC, E>::T)) __attribute__((amdgpu_waves_per_eu(ColTuneCrt<S>::waves)))
k_cols2_crt(JobPack<CrtColJob, kCrtPack> jp, const CrtColJob* __restrict__ jobs, int L, CrtFields F) {
  using SC = Sched<S, E>;
  constexpr int NPH = SC::nph;
  constexpr int C = 1 << LOGC;
  extern __shared__ uint32_t X[];
  const CrtColJob J = load_job(jp.j[0], jobs);
  const uint32_t np = __builtin_amdgcn_readfirstlane(J.g.primes);
  const int b = L - S;
  const uint32_t M = 1u << b;
  const uint32_t tiles = M >> LOGC;
  uint32_t tile = blockIdx.x;
  if (tiles >= 64 && (tiles & 7) == 0) tile = (tile & 7) * (tiles >> 3) + (tile >> 3);
  const uint32_t c = threadIdx.x % C, g = threadIdx.x / C;
  const uint32_t col = (tile << LOGC) + c;
  constexpr int PL = NPH - 1;
  constexpr int rl = SC::r_of(PL);
  constexpr int NGL = PhaseGroups<S, PL, LOGC, E>::NG;
  constexpr int r0 = SC::r_of(0);
  constexpr int NG0 = PhaseGroups<S, 0, LOGC, E>::NG;
  constexpr int NE0 = NG0 * (1 << r0);
  typedef __attribute__((address_space(1))) const uint32_t g32;
  uint32_t s0[NE0], s1[NE0];  // Garner state: x1; then the target's share of x1 + p1 x2, and x1 + p1 x2 mod P3
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if ((uint32_t)k >= np) break;  // uniform
    if (k > 0) __syncthreads();    // the previous prime's last LDS reads are done
    const FieldDev& f = F.f[k];
    const Arith<LAZY> A(f);
    uint32_t TI[16], x[16];
    tw_load<S, PL, LOGC, E>(TI, f.planes_inv, g);
    // Four-step twiddle w^-(col * bitrev_S(e)) times 1/N by a running product
    // (as in k_cols2's inverse).
    const uint32_t shift = f.logmax - L, split = f.split, smask = (1u << split) - 1;
    const uint32_t xd = (col << (S - rl)) << shift;
    const uint32_t d_hi = *(g32*)(f.tw_hi_inv + (xd >> split));
    const uint32_t d_lo = *(g32*)(f.tw_lo_inv + (xd & smask));
    uint32_t q_hi[NGL], q_lo[NGL];
#pragma unroll
    for (int u = 0; u < NGL; ++u) {
      const uint32_t xq = (col * bitrev(PhaseGroups<S, PL, LOGC, E>::gid(g, u) << rl, S)) << shift;
      q_hi[u] = *(g32*)(f.tw_hi_inv + (xq >> split));
      q_lo[u] = *(g32*)(f.tw_lo_inv + (xq & smask));
    }
    {
      const __amdgpu_buffer_rsrc_t rs = buf_rsrc(J.src[k], (M << S) * 4);
      constexpr int lol = S - Sched<S, E>::s_of(PL) - rl;
#pragma unroll
      for (int u = 0; u < NGL; ++u) {
        const uint32_t lane = (elem<S, PL, E>(PhaseGroups<S, PL, LOGC, E>::gid(g, u), 0) * M + col) * 4;
#pragma unroll
        for (int kk = 0; kk < (1 << rl); ++kk)
          x[u * (1 << rl) + kk] = bload(rs, lane, ((uint32_t)kk * M) << (lol + 2));
      }
    }
    const uint32_t D = A.mul(d_hi, d_lo);
#pragma unroll
    for (int u = 0; u < NGL; ++u) {
      uint32_t q = A.mul(A.mul(q_hi[u], q_lo[u]), J.inv_n[k]);
#pragma unroll
      for (int j = 0; j < (1 << rl); ++j) {
        const int kk = (int)bitrev_c(j, rl);
        x[u * (1 << rl) + kk] = A.mul(x[u * (1 << rl) + kk], q);
        if (j + 1 < (1 << rl)) q = A.mul(q, D);
      }
    }
    inv_transform<S, LOGC, LAZY, ColTuneCrt<S>::PF, E>(x, X, g, c, f.planes_inv, A, TI);
#pragma unroll
    for (int i = 0; i < NE0; ++i) {
      const uint32_t v = A.canon(A.half(x[i]));  // Montgomery residue mod P_k
      if (k == 0) {
        s0[i] = redc_full(v, J.g.P[0], J.g.Pninv[0]);
      } else if (k == 1) {
        const uint32_t x1 = s0[i], x2 = garner_x2(x1, v, J.g);
        s1[i] = garner_x12(x1, x2, J.g);
        s0[i] = garner_tadd(garner_t1(x1, J.g), x2, J.g.p1_t, J.g);
      } else {
        s0[i] = garner_tadd(s0[i], garner_x3(s1[i], v, J.g), J.g.p12_t, J.g);
      }
    }
  }
  const bool acc = J.g.accumulate != 0;
#pragma unroll
  for (int u = 0; u < NG0; ++u)
#pragma unroll
    for (int kk = 0; kk < (1 << r0); ++kk) {
      const uint32_t gi = elem<S, 0, E>(PhaseGroups<S, 0, LOGC, E>::gid(g, u), kk) * M + col;
      if (gi < J.out_lo || gi >= J.out_hi) continue;
      const uint32_t o = gi - J.out_shift;
      uint32_t v = s0[u * (1 << r0) + kk];
      if (np == 1) v = garner_t1(v, J.g);
      gstore(J.dst + o, garner_finish(v, acc ? gload(J.dst + o) : 0u, J.g));
    }
}

}  // namespace fpxg
