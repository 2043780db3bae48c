// Kernel launches over the compile-time shapes of ntt2.hpp.
//
// The switch tables that instantiate every row, column and CRT-column kernel
// are spread over the launch_*.hip translation units (compiled in parallel);
// the drivers in fpxg.hip reach them through the part functions declared at
// the end of this file.  Each part also sets the LDS attributes of the large
// kernels it instantiates (a kernel's attribute is set through the
// translation unit that registers it).
#pragma once

#include "runtime.hpp"

namespace fpxg {

// Forward passes on signed residues (ArithS) need 8p < 2^31.
inline bool signed_ok(uint32_t p) { return p < (1u << 28); }

// Elements a thread owns in a one-row launch (log2): fewer than 16 for the
// fused products and Newton steps of 2^10 .. 2^13 points, whose single
// workgroup would otherwise run on 1-8 waves.
// (Measured by graph replay of the launch alone: at 2^12 points eight elements
// a thread beat four by 0.3-1.1 us, INVSTEP 11.6 -> 11.1 us, EXPSTEP 18.2 ->
// 17.2; below 2^12 four win.)
static constexpr int row_small_e(int logm, int mode) {
  const bool fused = mode == R2M_CONV || mode == R2M_INVSTEP || mode == R2M_EXPSTEP || mode == R2M_SQRTSTEP;
  if (!fused || logm < 10 || logm > 13) return 4;
  return logm <= 11 ? 2 : 3;
}

// Elements a thread (log2) of the fused product's row pass over many rows of
// 2^4 .. 2^12 words, up to 2^21 points a job: four (2), or eight (3) once the
// launch holds 2^21 points or more, where the exchanges of the four-element
// form start to cost more than its extra waves save.  Measured on the MI355X
// by graph replay of the pass alone: rows of 2^4 / 2^5 words 5.0 / 5.6 ->
// 2.6 / 3.0 us (sixteen elements before), two 2^20-point jobs of 2^12-word
// rows 15.8 -> 13.9 us.
static constexpr int row_multi_e(int logm, int mode) {
  return (mode == R2M_CONV && logm >= 4 && logm <= 12) ? 2 : 4;
}

// Whether njobs by-value forward jobs of nrows rows of 2^14 words each can run
// as the persistent prefetching pass: enough rows for every CU twice, no
// lift, whole rows of input, 16-byte aligned rows (LDS DMA loads and 16-byte
// stores), and at most kDmaPack jobs.
static bool rows_dma_ok(const Context& c, const RowJob2* j, uint32_t nrows, uint32_t njobs) {
  bool dma = c.row_dma && njobs >= 1 && njobs <= (uint32_t)kDmaPack && (size_t)nrows * njobs >= 2u * (size_t)c.cus;
  for (uint32_t i = 0; dma && i < njobs; ++i)
    dma = !(j[i].flags & R2_LIFT) && (size_t)j[i].in_a >= ((size_t)nrows << kDmaRowLog) &&
          ((uintptr_t)j[i].src_a & 15) == 0 && ((uintptr_t)j[i].dst & 15) == 0;
  return dma;
}

template <bool LAZY, bool SGN>
static void rows_dma_go(Context& c, const RowJob2* j, uint32_t nrows, uint32_t njobs) {
  JobPack<RowJob2, kDmaPack> jp;
  for (int i = 0; i < kDmaPack; ++i) {
    jp.j[i] = j[(uint32_t)i < njobs ? i : 0];
    jp.j[i].nrows = nrows;
  }
  hipLaunchKernelGGL((k_rows_fwd_dma<LAZY, SGN>), dim3((uint32_t)c.cus), dim3(RowShape<kDmaRowLog>::T), 0, c.stream,
                     jp, njobs);
}

// A row launch: njobs jobs from the device array dj, or (dj null) the njobs
// by-value jobs j[0 .. njobs).
template <int LOGM, bool LAZY, int MODE>
static void rows2_launch(Context& c, const RowJob2* j, uint32_t nrows, uint32_t njobs, const RowJob2* dj) {
  using RS = RowShape<LOGM>;
  const size_t lds = (size_t)RS::U * RS::lds_words_per_row * 4;
  // Many whole rows of 2^14 words, forward: the persistent prefetching pass.
  if constexpr (LOGM == kDmaRowLog && (MODE == R2M_FWD || MODE == R2M_FWDS)) {
    if (!dj && rows_dma_ok(c, j, nrows, njobs)) {
      rows_dma_go<LAZY, MODE == R2M_FWDS>(c, j, nrows, njobs);
      return;
    }
  }
  if constexpr (MODE == R2M_NINV || MODE == R2M_NEXP || MODE == R2M_NSQRT) {
    if (!dj && njobs != 1) throw Error(E_DEVICE, "small-series launches take one job by value");
    hipLaunchKernelGGL((k_newton_small<MODE>), dim3(1, njobs), dim3(kSmallThreads), 0, c.stream, j[0], dj);
  } else if constexpr ((MODE == R2M_INVSTEP || MODE == R2M_EXPSTEP || MODE == R2M_SQRTSTEP) && LOGM <= 9) {
    // Small Newton steps: their window products computed directly (one job
    // by value, or a device job array).
    if (!dj && njobs != 1) throw Error(E_DEVICE, "direct Newton steps take one job by value");
    hipLaunchKernelGGL((k_step_direct<MODE, LOGM - 1>), dim3(1, njobs), dim3(kMidTPO << (LOGM - 1)), 0, c.stream,
                       j[0], dj);
  } else {
    JobPack<RowJob2, kRowPack> pack;
    for (int i = 0; i < kRowPack; ++i) pack.j[i] = j[(!dj && (uint32_t)i < njobs) ? i : 0];
    // A lone row of a mid-size product or Newton step: one workgroup, so
    // fewer elements a thread put it on more waves.
    constexpr int ES = row_small_e(LOGM, MODE);
    if constexpr (ES != 4) {
      if (!dj && njobs == 1 && nrows == 1) {
        using RSS = RowShape<LOGM, ES>;
        hipLaunchKernelGGL((k_rows2<LOGM, LAZY, MODE, ES>), dim3(1, 1), dim3(RSS::T),
                           (size_t)RSS::U * RSS::lds_words_per_row * 4, c.stream, pack, dj);
        return;
      }
    }
    // The row pass of a four-step product of up to 2^21 points (rows of
    // 2^4 .. 2^12 words): four or eight elements a thread (row_multi_e), so
    // each thread issues a quarter or half of the loads (the pass is
    // latency-bound at these sizes).
    constexpr int EM = row_multi_e(LOGM, MODE);
    if constexpr (EM != 4) {
      if (!dj && nrows > 1 && ((size_t)nrows << LOGM) <= ((size_t)1 << 21)) {
        if constexpr (LOGM >= 8) {
          if (((size_t)nrows << LOGM) * njobs >= ((size_t)1 << 21)) {
            using RS3 = RowShape<LOGM, 3>;
            hipLaunchKernelGGL((k_rows2<LOGM, LAZY, MODE, 3>), dim3((nrows + RS3::U - 1) / RS3::U, njobs),
                               dim3(RS3::T), (size_t)RS3::U * RS3::lds_words_per_row * 4, c.stream, pack, dj);
            return;
          }
        }
        using RSM = RowShape<LOGM, EM>;
        hipLaunchKernelGGL((k_rows2<LOGM, LAZY, MODE, EM>), dim3((nrows + RSM::U - 1) / RSM::U, njobs),
                           dim3(RSM::T), (size_t)RSM::U * RSM::lds_words_per_row * 4, c.stream, pack, dj);
        return;
      }
    }
    dim3 grid((nrows + RS::U - 1) / RS::U, njobs), block(RS::T);
    hipLaunchKernelGGL((k_rows2<LOGM, LAZY, MODE>), grid, block, lds, c.stream, pack, dj);
  }
}

template <int LOGM, bool LAZY>
static void rows2_mode(Context& c, int mode, const RowJob2* j, uint32_t nrows, uint32_t njobs, const RowJob2* dj) {
  switch (mode) {
    case R2M_FWD:
      if constexpr (LAZY) {
        // Signed residues when every by-value job's prime allows them.
        bool sgn = !dj;
        for (uint32_t i = 0; sgn && i < njobs; ++i) sgn = signed_ok(j[i].f.p);
        if (sgn) {
          rows2_launch<LOGM, LAZY, R2M_FWDS>(c, j, nrows, njobs, dj);
          break;
        }
      }
      rows2_launch<LOGM, LAZY, R2M_FWD>(c, j, nrows, njobs, dj);
      break;
    case R2M_INV: rows2_launch<LOGM, LAZY, R2M_INV>(c, j, nrows, njobs, dj); break;
    case R2M_CONV: rows2_launch<LOGM, LAZY, R2M_CONV>(c, j, nrows, njobs, dj); break;
    case R2M_SPEC: rows2_launch<LOGM, LAZY, R2M_SPEC>(c, j, nrows, njobs, dj); break;
    case R2M_SPEC_ONLY: rows2_launch<LOGM, LAZY, R2M_SPEC_ONLY>(c, j, nrows, njobs, dj); break;
    case R2M_SPEC_SUM2: rows2_launch<LOGM, LAZY, R2M_SPEC_SUM2>(c, j, nrows, njobs, dj); break;
    case R2M_INVSTEP:
    case R2M_EXPSTEP:
    case R2M_SQRTSTEP:
      if constexpr (LOGM >= 6 && LOGM <= 13) {
        if (mode == R2M_INVSTEP) rows2_launch<LOGM, LAZY, R2M_INVSTEP>(c, j, nrows, njobs, dj);
        else if (mode == R2M_EXPSTEP) rows2_launch<LOGM, LAZY, R2M_EXPSTEP>(c, j, nrows, njobs, dj);
        else rows2_launch<LOGM, LAZY, R2M_SQRTSTEP>(c, j, nrows, njobs, dj);
        break;
      }
      throw Error(E_DEVICE, "Newton-step row pass outside 2^6 .. 2^13 points");
    case R2M_NINV:
    case R2M_NEXP:
    case R2M_NSQRT:
      if constexpr (LOGM == kSmallSeriesLog) {
        if (mode == R2M_NINV) rows2_launch<LOGM, LAZY, R2M_NINV>(c, j, nrows, njobs, dj);
        else if (mode == R2M_NEXP) rows2_launch<LOGM, LAZY, R2M_NEXP>(c, j, nrows, njobs, dj);
        else rows2_launch<LOGM, LAZY, R2M_NSQRT>(c, j, nrows, njobs, dj);
        break;
      }
      throw Error(E_DEVICE, "small-series launch recorded with the wrong size key");
    default: throw Error(E_DEVICE, "unsupported row-pass mode");
  }
}



// Column tiles: 16 columns up to 2^10 levels, 8 for 2^11, 4 for 2^12.  A
// 2^10-level pass with fewer than 512 tiles of 16 uses tiles of 8, so a
// mid-size transform still spreads over the CUs.
static constexpr int col_tile_log(int S) { return S <= 10 ? 4 : (S <= 11 ? 3 : 2); }
inline int tile_log_cols(int S, int L) {
  if (S == 10 && L - S - 4 < 9) return 3;
  return col_tile_log(S);
}

// Tile width (log2 of the columns a tile) of the by-value column launches of
// up to 2^20 points (2^5 .. 2^10 levels, four elements a thread): these
// launches are latency-bound, and a workgroup's arithmetic is serialised over
// its waves, so narrow tiles on many CUs finish first.  One column a tile
// while rows have at most 2^7 words (a tile's reads are then at most 512
// bytes apart), four at 2^8 .. 2^10 words, eight for the forward pass at
// 2^10; at least 64 and at most 1024 threads a tile.  Measured on the MI355X
// by graph replay of the pass alone, 2^13 .. 2^20 points: 2.1 .. 7 us a
// launch against 2.2 .. 9 us with the widest tile that fits 1024 threads.
static int small_col_tile_log(int S, int b, bool inv) {
  int lc = b <= 7 ? 0 : (b == 10 && !inv ? 3 : 2);
  lc = std::max(lc, 8 - S);
  lc = std::min(lc, 12 - S);
  return lc;
}

template <int S, int LC, bool LAZY, bool INV, bool SGN, bool CH>
static void cols2_small_go(Context& c, const JobPack<ColJob2, kColPack>& pack, const PassPlan& P, uint32_t njobs) {
  using CS2 = ColShape<S, LC, 2>;
  hipLaunchKernelGGL((k_cols2<S, LC, LAZY, INV, SGN, 2, CH>), dim3(1u << (P.b - LC), njobs), dim3(CS2::T),
                     (size_t)CS2::words * 4, c.stream, pack, (const ColJob2*)nullptr, P.L);
}

// Dispatches the tile width lc (8 - S <= lc <= 12 - S, lc <= P.b); false if
// none applies.
template <int S, int LC, bool LAZY, bool INV, bool SGN, bool CH>
static bool cols2_small(Context& c, const JobPack<ColJob2, kColPack>& pack, const PassPlan& P, uint32_t njobs,
                        int lc) {
  if constexpr (LC > 4) {
    return false;
  } else {
    if constexpr (LC >= 8 - S && LC <= 12 - S) {
      if (lc == LC && P.b >= LC) {
        cols2_small_go<S, LC, LAZY, INV, SGN, CH>(c, pack, P, njobs);
        return true;
      }
    }
    return cols2_small<S, LC + 1, LAZY, INV, SGN, CH>(c, pack, P, njobs, lc);
  }
}

// njobs jobs from the device array dj, or (dj null) the njobs by-value jobs
// j[0 .. njobs).
template <int S, int LOGC, bool LAZY, bool INV, bool SGN, bool CH = false>
static void cols2_launch_c(Context& c, const ColJob2* j, const PassPlan& P, uint32_t njobs, const ColJob2* dj) {
  using CS = ColShape<S, LOGC>;
  const size_t lds = (size_t)CS::words * 4;
  JobPack<ColJob2, kColPack> pack;
  for (int i = 0; i < kColPack; ++i) pack.j[i] = j[(!dj && (uint32_t)i < njobs) ? i : 0];
  // Columns of lone calls and Newton steps up to 2^20 points (2^5 .. 2^10
  // levels): four elements a thread in narrow tiles (small_col_tile_log),
  // except the 2^20-point forward pass of 2^10 levels, which keeps sixteen
  // elements a thread in tiles of eight columns.  Merged batches and larger
  // transforms keep the wide tiles: their launches are throughput-bound.
  if constexpr (S >= 5 && S <= 10) {
    if (!dj && P.L <= 20 && !(S == 10 && P.b == 10 && !INV)) {
      if (cols2_small<S, 0, LAZY, INV, SGN, CH>(c, pack, P, njobs, small_col_tile_log(S, P.b, INV))) return;
      constexpr int LC = LOGC < 10 - (S - 2) ? LOGC : 10 - (S - 2);
      using CS2 = ColShape<S, LC, 2>;
      hipLaunchKernelGGL((k_cols2<S, LC, LAZY, INV, SGN, 2, CH>), dim3(1u << (P.b - LC), njobs), dim3(CS2::T),
                         (size_t)CS2::words * 4, c.stream, pack, dj, P.L);
      return;
    }
  }
  dim3 grid(1u << (P.b - LOGC), njobs), block(CS::T);
  hipLaunchKernelGGL((k_cols2<S, LOGC, LAZY, INV, SGN, 4, CH>), grid, block, lds, c.stream, pack, dj, P.L);
}
template <int S, bool LAZY, bool INV, bool SGN, bool CH = false>
static void cols2_launch(Context& c, const ColJob2* j, const PassPlan& P, uint32_t njobs, const ColJob2* dj) {
  if constexpr (S == 10) {
    if (P.logC == 3) {
      cols2_launch_c<S, 3, LAZY, INV, SGN, CH>(c, j, P, njobs, dj);
      return;
    }
  }
  cols2_launch_c<S, col_tile_log(S), LAZY, INV, SGN, CH>(c, j, P, njobs, dj);
}



// The inverse column pass of CRT products with Garner fused (k_cols2_crt):
// tiles as in launch_cols; the auxiliary primes are lazy (all below 2^30).
template <int S, int LOGC, int E>
static void colcrt_go(Context& c, const CrtColJob* j, const PassPlan& P, uint32_t njobs, const CrtColJob* dj,
                      const CrtFields& F) {
  using CS = ColShape<S, LOGC, E>;
  JobPack<CrtColJob, kCrtPack> pack;
  for (int i = 0; i < kCrtPack; ++i) pack.j[i] = j[(!dj && (uint32_t)i < njobs) ? i : 0];
  hipLaunchKernelGGL((k_cols2_crt<S, LOGC, true, E>), dim3(1u << (P.b - LOGC), njobs), dim3(CS::T),
                     (size_t)CS::words * 4, c.stream, pack, dj, P.L, F);
}
template <int S, int LOGC>
static void colcrt_launch_c(Context& c, const CrtColJob* j, const PassPlan& P, uint32_t njobs, const CrtColJob* dj,
                            const CrtFields& F) {
  if constexpr (S >= 5 && S <= 10) {
    if (!dj && P.L <= 20) {  // four elements a thread, as in cols2_launch_c
      constexpr int LC = LOGC < 10 - (S - 2) ? LOGC : 10 - (S - 2);
      colcrt_go<S, LC, 2>(c, j, P, njobs, dj, F);
      return;
    }
  }
  colcrt_go<S, LOGC, 4>(c, j, P, njobs, dj, F);
}
template <int S>
static void colcrt_launch(Context& c, const CrtColJob* j, const PassPlan& P, uint32_t njobs, const CrtColJob* dj,
                          const CrtFields& F) {
  if constexpr (S == 10) {
    if (P.logC == 3) {
      colcrt_launch_c<S, 3>(c, j, P, njobs, dj, F);
      return;
    }
  }
  colcrt_launch_c<S, col_tile_log(S)>(c, j, P, njobs, dj, F);
}




template <int LOGM, bool LAZY, int MODE>
static bool rows2_attr() {
  using RS = RowShape<LOGM>;
  return hipFuncSetAttribute((const void*)k_rows2<LOGM, LAZY, MODE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(RS::U * RS::lds_words_per_row * 4)) == hipSuccess;
}
template <int S, bool LAZY, bool INV, bool SGN = false, bool CH = false>
static bool cols2_attr() {
  constexpr int LOGC = col_tile_log(S);
  return hipFuncSetAttribute((const void*)k_cols2<S, LOGC, LAZY, INV, SGN, 4, CH>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)(ColShape<S, LOGC>::words * 4)) ==
         hipSuccess;
}

template <int S>
static bool colcrt_attr() {
  constexpr int LOGC = col_tile_log(S);
  return hipFuncSetAttribute((const void*)k_cols2_crt<S, LOGC, true, 4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(ColShape<S, LOGC>::words * 4)) == hipSuccess;
}

// ---- the parts (one translation unit each) ----
// Row passes: 2^logM-word rows, lazy (p < 2^30) or canonical limbs.
void rows2_part_lazy_lo(Context& c, int logM, int mode, const RowJob2* j, uint32_t nrows, uint32_t njobs,
                        const RowJob2* dj);  // logM 1..11
void rows2_part_lazy_12(Context& c, int logM, int mode, const RowJob2* j, uint32_t nrows, uint32_t njobs,
                        const RowJob2* dj);
void rows2_part_lazy_13(Context& c, int logM, int mode, const RowJob2* j, uint32_t nrows, uint32_t njobs,
                        const RowJob2* dj);
void rows2_part_lazy_14(Context& c, int logM, int mode, const RowJob2* j, uint32_t nrows, uint32_t njobs,
                        const RowJob2* dj);
// The forward row pass of up to kDmaPack transforms of nrows rows of 2^14
// words each (lazy limbs) as one persistent launch; false (nothing launched)
// where rows_dma_ok says no.
bool rows_dma_part_lazy(Context& c, bool sgn, const RowJob2* j, uint32_t nrows, uint32_t njobs);
void rows2_part_canon(Context& c, int logM, int mode, const RowJob2* j, uint32_t nrows, uint32_t njobs,
                      const RowJob2* dj);  // logM 1..14
// Column passes of 2^S levels.
void cols2_part_lazy_fwd(Context& c, bool sgn, const ColJob2* j, const PassPlan& P, uint32_t njobs,
                         const ColJob2* dj);
void cols2_part_lazy_inv(Context& c, const ColJob2* j, const PassPlan& P, uint32_t njobs, const ColJob2* dj);
// The chained pass (k_cols2 with CH: inverse levels, window, forward levels)
// of 2^5 .. 2^12 column levels, lazy limbs.
void cols2_part_lazy_chain(Context& c, const ColJob2* j, const PassPlan& P, uint32_t njobs, const ColJob2* dj);
void cols2_part_canon(Context& c, bool inv, const ColJob2* j, const PassPlan& P, uint32_t njobs,
                      const ColJob2* dj);
void colcrt_part(Context& c, const CrtColJob* j, const PassPlan& P, uint32_t njobs, const CrtColJob* dj,
                 const CrtFields& F);
// LDS attributes of the 2^13- and 2^14-word kernels of each part; false when
// the device refuses one (the caller then keeps to 2^13).
bool rows2_attrs_lazy_13();
bool rows2_attrs_lazy_14();
bool rows2_attrs_canon();
bool cols2_attrs_lazy_fwd();
bool cols2_attrs_lazy_inv();
bool cols2_attrs_canon();
bool colcrt_attrs();

}  // namespace fpxg
