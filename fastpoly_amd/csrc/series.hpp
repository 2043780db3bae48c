// fastpoly_amd Newton series on the device: inv, log, exp, sqrt.
//
// Control flow follows the reference (detail/poly/inverse.hpp:18-188,
// exponential.hpp:85-207, square_root.hpp:12-112, calculus.hpp:17-116): a
// 32-coefficient recurrence seed, doubling steps m -> min(2m, n) whose
// products are cyclic transforms of 2m points read only where no wraparound
// lands (the 2m middle-product trick), a persistent inverse prefix c for exp
// and sqrt, and <= 8-coefficient tails filled by the recurrence.  All series
// data stays in device memory.  Every output coefficient is unique, so the
// reference's special cases for short inputs (inverse_linear, inverse_short,
// exp_linear) are served by the same Newton path.
//
// A call runs in two halves: one scan of the input (a single launch and one
// read-back of five scalars) settles every domain check and special case, then
// the Newton schedule is emitted.  The schedule reads no device data on the
// host, so the same code records it for the merged batch executor.
//
// Products go through an Engine: the target prime's own NTT when it supports
// the 2m-point transform, otherwise three auxiliary primes (167772161,
// 469762049, 754974721) with Garner reconstruction (crt_convolution.hpp).
#pragma once

namespace fpxg {

// Input words a series call reads from a host array a[0, la): min(la, n),
// except sqrt, which finds the valuation v over the whole input and then reads
// a[v + i] for i < n - v/2, i.e. up to n + v/2 > n words (square_root.hpp:26-44).
static size_t series_input_words(int op, const uint32_t* a, size_t la, size_t n) {
  if (op != FPXG_SQRT) return std::min(la, n);
  size_t v = 0;
  while (v < la && a[v] == 0) ++v;
  if (v == la) return std::min(la, n);
  return std::min(la, std::max(v + 1, n + v / 2));
}

// The same for device inputs, whose valuation is found by the device scan: the
// whole input for sqrt (the scan's positions are 32-bit words).
static size_t series_input_words_dev(int op, size_t la, size_t n) {
  if (op != FPXG_SQRT) return std::min(la, n);
  if (la >= (size_t)UINT32_MAX)
    throw Error(E_SIZE, "poly::sqrt: inputs of 2^32 - 1 or more coefficients are not supported");
  return la;
}

static bool series_native(const HostField& T, int L) {
  return T.p != 2 && T.maxlog >= 6 && L <= std::min(T.maxlog, 26);
}

// A Newton step's products: 2^L-point cyclic products over the target
// prime's own transform when it supports them (`native_conv`), else through
// the three auxiliary primes (`crt_conv`), whose transforms stop at 2^24.
struct Engine {
  int L;
  bool native;
  Engine(Context&, const HostField& t, int l) : L(l), native(series_native(t, l)) {
    if (!native && L > 24) throw Error(E_SIZE, "series step beyond the auxiliary primes' 2^24-point transforms");
  }
};

// Native fields: out[i - lo] = (x * y)[i] for i in [lo, hi) of the 2^L-point
// cyclic product of x[0, cx) and y[0, cy), both target limbs, with no stored
// spectra: one launch for a single-row transform (two forwards, product and
// inverse in one workgroup), three for the four-step ones (both column passes
// in one launch, the fused row pass, the column inverse into the window).  A
// Newton step's products are latency-bound chains of small launches, so fewer
// launches beat reusing a spectrum.
// Epilogue of a window product: the result is multiplied by `scale` (a
// Montgomery limb, folded into 1/N; 0 for none), negated, then, with
// epi_flags R2_SUBFROM or R2_MULTAB, combined with epi[t] at output position t
// (see out_epilogue), which saves a Newton step its elementwise launch.
struct Epi {
  uint32_t scale = 0;
  uint32_t flags = 0;
  const uint32_t* src = nullptr;
  uint32_t avail = 0;
};

static void native_conv(Context& c, const HostField& T, int L, const uint32_t* x, size_t cx, const uint32_t* y,
                        size_t cy, size_t lo, size_t hi, uint32_t* out, bool neg = false, const Epi& ep = Epi()) {
  const uint32_t N = 1u << L;
  const PassPlan P = c.conv_plan(L);
  const uint32_t inv_n = ep.scale ? T.mul(inv_pow2(T, L), ep.scale) : inv_pow2(T, L);
  const uint32_t negf = (neg ? R2_NEG : 0u) | ep.flags;
  RowJob2 rj{};
  rj.f = T.dev;
  rj.epi = ep.src;
  rj.epi_avail = ep.avail;
  if (P.S == 0) {
    rj.src_a = x; rj.src_b = y; rj.dst = out; rj.nrows = 1; rj.in_a = (uint32_t)cx; rj.in_b = (uint32_t)cy;
    rj.out_lo = (uint32_t)lo; rj.out_hi = (uint32_t)hi; rj.out_shift = (uint32_t)lo; rj.inv_n = inv_n;
    rj.flags = R2_SCALE | R2_CANON | negf;
    launch_rows(c, T.lazy, R2M_CONV, rj, L, 1);
    return;
  }
  const bool same = x == y && cx == cy;
  DevBuf X = c.buf(N), Y;
  ColJob2 jx{};
  jx.src = x; jx.dst = X.get(); jx.in_count = (uint32_t)cx; jx.f = T.dev;
  if (same) {
    launch_cols(c, T.lazy, false, jx, P);
  } else {
    Y = c.buf(N);
    ColJob2 jy = jx;
    jy.src = y; jy.dst = Y.get(); jy.in_count = (uint32_t)cy;
    launch_cols_pair(c, T.lazy, false, jx, jy, P);
  }
  rj.src_a = X.get(); rj.src_b = same ? X.get() : Y.get(); rj.dst = X.get(); rj.nrows = 1u << P.S;
  rj.in_a = N; rj.in_b = N; rj.out_lo = 0; rj.out_hi = N;
  launch_rows(c, T.lazy, R2M_CONV, rj, P.b, rj.nrows);
  ColJob2 ij{};
  ij.src = X.get(); ij.dst = out; ij.out_lo = (uint32_t)lo; ij.out_hi = (uint32_t)hi; ij.out_shift = (uint32_t)lo;
  ij.inv_n = inv_n; ij.flags = negf; ij.epi = ep.src; ij.epi_avail = ep.avail; ij.f = T.dev;
  launch_cols(c, T.lazy, true, ij, P);
}

// One window product of native_conv_multi: out[i - lo] = (x * y)[i] for i in
// [lo, hi) of the 2^L-point cyclic product, negated on request, with the
// epilogue ep.
struct WinProd {
  const uint32_t* x;
  size_t cx;
  const uint32_t* y;
  size_t cy;
  size_t lo, hi;
  uint32_t* out;
  bool neg;
  Epi ep;
};

// k <= 3 independent window products of one 2^L-point size (with a column
// pass) in three launches shared by all of them: the forward column pass of
// every distinct operand, the fused row pass of every product, the inverse
// column pass into every window.  A Newton step's products are chains of
// latency-bound launches, so two products that do not depend on each other
// cost the launches of one.  An operand used by several products (same
// pointer and length) is transformed once.
static void native_conv_multi(Context& c, const HostField& T, int L, const WinProd* w, int k) {
  const PassPlan P = c.conv_plan(L);
  if (P.S == 0 || k < 1 || k > kRowPack) throw Error(E_DEVICE, "native_conv_multi: shape");
  const uint32_t N = 1u << L;
  const uint32_t* srcs[2 * kRowPack];
  size_t cnts[2 * kRowPack];
  DevBuf spec[2 * kRowPack], dst[kRowPack];
  int nsrc = 0, ia[kRowPack], ib[kRowPack];
  auto operand = [&](const uint32_t* p, size_t n) {
    for (int i = 0; i < nsrc; ++i)
      if (srcs[i] == p && cnts[i] == n) return i;
    srcs[nsrc] = p;
    cnts[nsrc] = n;
    return nsrc++;
  };
  for (int i = 0; i < k; ++i) {
    ia[i] = operand(w[i].x, w[i].cx);
    ib[i] = operand(w[i].y, w[i].cy);
  }
  ColJob2 cj[2 * kRowPack];
  for (int i = 0; i < nsrc; ++i) {
    spec[i] = c.buf(N);
    cj[i] = ColJob2{};
    cj[i].src = srcs[i]; cj[i].dst = spec[i].get(); cj[i].in_count = (uint32_t)cnts[i]; cj[i].f = T.dev;
  }
  launch_cols_n(c, T.lazy, false, cj, P, (uint32_t)nsrc, nullptr);
  RowJob2 rj[kRowPack];
  for (int i = 0; i < k; ++i) {
    dst[i] = c.buf(N);
    rj[i] = RowJob2{};
    rj[i].f = T.dev;
    rj[i].src_a = spec[ia[i]].get(); rj[i].src_b = spec[ib[i]].get(); rj[i].dst = dst[i].get();
    rj[i].in_a = N; rj[i].in_b = N; rj[i].out_lo = 0; rj[i].out_hi = N;
  }
  launch_rows_group(c, T.lazy, R2M_CONV, rj, k, P.b, 1u << P.S);
  ColJob2 ij[kRowPack];
  for (int i = 0; i < k; ++i) {
    const Epi& ep = w[i].ep;
    ij[i] = ColJob2{};
    ij[i].src = dst[i].get(); ij[i].dst = w[i].out;
    ij[i].out_lo = (uint32_t)w[i].lo; ij[i].out_hi = (uint32_t)w[i].hi; ij[i].out_shift = (uint32_t)w[i].lo;
    ij[i].inv_n = ep.scale ? T.mul(inv_pow2(T, L), ep.scale) : inv_pow2(T, L);
    ij[i].flags = (w[i].neg ? R2_NEG : 0u) | ep.flags; ij[i].epi = ep.src; ij[i].epi_avail = ep.avail;
    ij[i].f = T.dev;
  }
  launch_cols_n(c, T.lazy, true, ij, P, (uint32_t)k, nullptr);
}

// A full doubling step of 2^L = 2m points that one launch serves: the direct
// window products up to 2^9 points (any odd prime), the one-workgroup NTT
// step modes up to 2^13 (NTT primes).
static bool one_launch_step(const Context& c, const HostField& T, const Engine& E, size_t m) {
  if (E.L < 6 || ((size_t)1 << E.L) != 2 * m || T.p == 2) return false;
  return E.L <= 9 || (E.native && E.L <= std::min(c.row_log_max, c.step_log_max));
}

// Fields without their own transform: out[i - lo] = (x * y)[i] for i in
// [lo, hi) of the 2^L-point cyclic product through the three auxiliary primes,
// each pass one launch for all three, then Garner (negated on request), in the
// inverse column pass where crt_cols_fused says so.  No
// spectrum is stored for reuse: fewer launches a step beat the saved
// transforms at every size measured (exp mod 1e9+7 at 2^20: 1.42 ms, against
// 1.50 ms with stored spectra).
static void crt_conv(Context& c, const HostField& T, int L, const uint32_t* x, size_t cx, const uint32_t* y,
                     size_t cy, size_t lo, size_t hi, uint32_t* out, bool neg = false) {
  const bool fused = crt_cols_fused(c, L);
  DevBuf res[3];
  const HostField* Fs[3];
  uint32_t* outs[3];
  for (int i = 0; i < 3; ++i) {
    if (!fused) res[i] = c.buf(hi - lo);
    Fs[i] = &c.field(kCrtP[i]);
    outs[i] = res[i].get();
  }
  GarnerFuse gf{out, garner_const(c, T, 3, false)};
  gf.g.neg = neg;
  if (fused) {
    if (!conv_core_fields(c, Fs, 3, x, cx, y, cy, outs, hi - lo, true, L, lo, &gf))
      throw Error(E_DEVICE, "CRT product: fused column pass not taken");
    return;
  }
  conv_core_fields(c, Fs, 3, x, cx, y, cy, outs, hi - lo, true, L, lo);
  GarnerRec g{};
  g.c1 = res[0].get(); g.c2 = res[1].get(); g.c3 = res[2].get(); g.out = out; g.n = (uint32_t)(hi - lo);
  g.g = gf.g;
  emit_garner(c, g);
}

// extend_inverse (inverse.hpp:18-65): c[0, k) is 1/a mod x^k; fill c[k, wanted).
static void extend_inverse(Context& c, const HostField& T, const uint32_t* a, size_t la, uint32_t* cc, size_t k,
                           size_t wanted) {
  const size_t hlen = wanted - k;
  if (hlen <= 8) {
    recur(c, T, REC_INV, a, la, nullptr, 0, cc, k, wanted, 0, 0);
    return;
  }
  Engine E(c, T, ceil_log2(2 * k));
  if (one_launch_step(c, T, E, k)) {
    // Both products of the step in one workgroup (R2M_INVSTEP).
    RowJob2 rj{};
    rj.f = T.dev;
    rj.src_a = a; rj.in_a = (uint32_t)std::min(la, wanted);
    rj.src_b = cc; rj.in_b = (uint32_t)k;
    rj.dst = cc + k; rj.nrows = 1; rj.out_lo = 0; rj.out_hi = (uint32_t)hlen; rj.out_shift = 0;
    rj.inv_n = inv_pow2(T, E.L);
    rj.flags = R2_SCALE | R2_CANON | R2_NEG;
    launch_rows(c, T.lazy, R2M_INVSTEP, rj, E.L, 1);
    return;
  }
  if (E.native && T.lazy && ((size_t)1 << E.L) == 2 * k && c.conv_plan(E.L).S >= 5) {
    // Four-step products: the first product's inverse column pass hands e to
    // the second product's forward column pass in the same launch (the
    // chained pass), and c's column spectrum from the first product's
    // forward pass serves the second: five launches instead of six.
    const PassPlan P = c.conv_plan(E.L);
    const uint32_t N = 1u << E.L;
    const uint32_t inv_n = inv_pow2(T, E.L);
    DevBuf X = c.buf(N), Y = c.buf(N), Z = c.buf(N);
    ColJob2 jx{};
    jx.src = a; jx.dst = X.get(); jx.in_count = (uint32_t)std::min(la, wanted); jx.f = T.dev;
    ColJob2 jy = jx;
    jy.src = cc; jy.dst = Y.get(); jy.in_count = (uint32_t)k;
    launch_cols_pair(c, T.lazy, false, jx, jy, P);
    RowJob2 rj{};
    rj.f = T.dev; rj.nrows = 1u << P.S; rj.in_a = N; rj.in_b = N; rj.out_lo = 0; rj.out_hi = N;
    rj.src_a = X.get(); rj.src_b = Y.get(); rj.dst = X.get();
    launch_rows(c, T.lazy, R2M_CONV, rj, P.b, rj.nrows);
    // e = -(a c)[k, wanted), moved down by k, then its forward column levels.
    ColJob2 ch{};
    ch.src = X.get(); ch.dst = Z.get(); ch.out_lo = (uint32_t)k; ch.out_hi = (uint32_t)wanted;
    ch.out_shift = (uint32_t)k; ch.inv_n = inv_n; ch.flags = R2_NEG; ch.f = T.dev;
    launch_cols(c, T.lazy, kColChain, ch, P);
    rj.src_a = Z.get(); rj.src_b = Y.get(); rj.dst = Z.get();
    launch_rows(c, T.lazy, R2M_CONV, rj, P.b, rj.nrows);
    ColJob2 ij{};
    ij.src = Z.get(); ij.dst = cc + k; ij.out_lo = 0; ij.out_hi = (uint32_t)hlen; ij.out_shift = 0;
    ij.inv_n = inv_n; ij.f = T.dev;
    launch_cols(c, T.lazy, kColInv, ij, P);
    return;
  }
  DevBuf e = c.buf(hlen);
  if (E.native) {
    native_conv(c, T, E.L, a, std::min(la, wanted), cc, k, k, wanted, e.get(), true);
    native_conv(c, T, E.L, e.get(), hlen, cc, k, 0, hlen, cc + k);
    return;
  }
  crt_conv(c, T, E.L, a, std::min(la, wanted), cc, k, k, wanted, e.get(), true);
  crt_conv(c, T, E.L, e.get(), hlen, cc, k, 0, hlen, cc + k);
}

// Whether extending c from k to `wanted` coefficients (extend_inverse), a
// step of at least 2^10 points, can run its first product at 2^L points in the
// launches of an independent product of that size (native_conv_multi): at
// 2^L >= wanted + k points that product has no wraparound at all.  Its second
// product stays at 2k points (one launch up to 2^13, where the one-workgroup
// inverse step did both products in one slower launch).
static bool cext_mergeable(Context& c, const HostField& T, size_t k, size_t wanted, int L) {
  const int Li = ceil_log2(2 * k);
  if (wanted <= k + 8 || Li < 10 || !series_native(T, Li) || !series_native(T, L)) return false;
  return c.conv_plan(L).S > 0 && ((size_t)1 << L) >= wanted + k;
}

// The second half of that extension: c[k, wanted) = -(e * c)[0, wanted - k)
// (inverse.hpp:18-65), e = (b * c)[k, wanted) computed beforehand.
static void cext_finish(Context& c, const HostField& T, const uint32_t* e, uint32_t* cc, size_t k, size_t wanted) {
  native_conv(c, T, ceil_log2(2 * k), e, wanted - k, cc, k, 0, wanted - k, cc + k);
}

// The first min(n, 512) coefficients of a small-series mode in one launch
// (k_newton_small); the caller fills the mode's fields.
static void small_series(Context& c, const HostField& T, int mode, RowJob2& rj, size_t n0) {
  rj.f = T.dev;
  rj.nrows = 1;
  rj.out_hi = (uint32_t)n0;
  launch_rows(c, T.lazy, mode, rj, kSmallSeriesLog, 1);
}

static void inv_newton(Context& c, const HostField& T, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  size_t m0;
  if (T.p != 2) {
    m0 = std::min<size_t>(n, kSmallSeries);
    RowJob2 rj{};
    rj.src_a = a; rj.in_a = (uint32_t)std::min(la, m0); rj.dst = out;
    small_series(c, T, R2M_NINV, rj, m0);
  } else {
    m0 = std::min<size_t>(n, 32);
    recur(c, T, REC_INV, a, la, nullptr, 0, out, 0, m0, 0, 0);
  }
  for (size_t m = m0; m < n; m <<= 1) extend_inverse(c, T, a, la, out, m, std::min(2 * m, n));
}

// One series call: its input and output (device pointers).
struct SeriesReq {
  int op;
  const uint32_t* a;
  size_t la, n;
  uint32_t* out;
};

static ScanRec scan_rec_of(const SeriesReq& q) { return {q.a, (uint32_t)q.la, 0}; }

// Emits the schedule of one series call given its scan; throws the reference's
// domain errors in the reference's order.
// Whether exp's or sqrt's doubling from m to 2m coefficients (c known to
// m/2) runs as exp_fused_doubling / sqrt_fused_doubling: a native lazy field,
// 2m points with a column pass the chained form serves, a step too large for
// one launch.
static bool exp_fused_ok(Context& c, const HostField& T, const Engine& E, size_t m, size_t hlen, size_t csize) {
  return E.native && T.lazy && hlen == m && 2 * csize == m && ((size_t)1 << E.L) == 2 * m &&
         !one_launch_step(c, T, E, m) && c.conv_plan(E.L).S >= 5;
}

// One doubling of exp's Newton iteration (exponential.hpp:158-204, with the
// inverse's extension inverse.hpp:18-65) at N = 2m points, b known to m and
// c = 1/b to m/2 coefficients, in nine launches that pass spectra from one
// product to the next through the chained column pass (F: forward transform):
//   B = F(b), C0 = F(c[0, m/2)), H = F(th)                     one launch
//   E = B C0, R = H B                                          one launch
//   F(e), e = E[m/2, m) moved down a quarter;
//   F(r), r = R[m, 2m) moved down a half                       one launch
//   G = F(e) C0
//   F(c), c[m/2, m) = -G[0, m/2) moved up a quarter (stored to cc) under
//   c[0, m/2) (read from cc)
//   T = F(r) F(c)
//   F(t), t = T[0, m) / (m + i)
//   U = F(t) B
//   b[m, 2m) = U[0, m)
// Every window is free of wraparound at N points: b c0 has fewer than 3m/2
// terms, e c0 and r c fewer than 2m, and (th b)'s wrapped terms land below m.
static void exp_fused_doubling(Context& c, const HostField& T, int L, const uint32_t* th, size_t cth, uint32_t* b,
                               size_t m, uint32_t* cc, const uint32_t* recip) {
  const PassPlan P = c.conv_plan(L);
  const uint32_t N = 1u << L, inv_n = inv_pow2(T, L), h = (uint32_t)(m / 2), mm = (uint32_t)m;
  DevBuf Bs = c.buf(N), C0 = c.buf(N), Hs = c.buf(N), X = c.buf(N), Y = c.buf(N), Z = c.buf(N);
  ColJob2 fj[3];
  for (ColJob2& j : fj) {
    j = ColJob2{};
    j.f = T.dev;
  }
  fj[0].src = b; fj[0].dst = Bs.get(); fj[0].in_count = mm;
  fj[1].src = cc; fj[1].dst = C0.get(); fj[1].in_count = h;
  fj[2].src = th; fj[2].dst = Hs.get(); fj[2].in_count = (uint32_t)cth;
  launch_cols_n(c, T.lazy, kColFwd, fj, P, 3, nullptr);
  RowJob2 rj[2];
  for (RowJob2& j : rj) {
    j = RowJob2{};
    j.f = T.dev; j.nrows = 1u << P.S; j.in_a = N; j.in_b = N; j.out_lo = 0; j.out_hi = N;
  }
  rj[0].src_a = Bs.get(); rj[0].src_b = C0.get(); rj[0].dst = X.get();
  rj[1].src_a = Hs.get(); rj[1].src_b = Bs.get(); rj[1].dst = Hs.get();
  launch_rows_group(c, T.lazy, R2M_CONV, rj, 2, P.b, 1u << P.S);
  auto chain = [&](const uint32_t* src, uint32_t* dst, uint32_t lo, uint32_t hi, uint32_t shift) {
    ColJob2 j{};
    j.f = T.dev; j.inv_n = inv_n;
    j.src = src; j.dst = dst; j.out_lo = lo; j.out_hi = hi; j.out_shift = shift;
    return j;
  };
  const ColJob2 ch[2] = {chain(X.get(), Y.get(), h, mm, h), chain(Hs.get(), Z.get(), mm, N, mm)};
  launch_cols_n(c, T.lazy, kColChain, ch, P, 2, nullptr);
  auto rows = [&](uint32_t* a, const uint32_t* bb) {
    RowJob2 j = rj[0];
    j.src_a = a; j.src_b = bb; j.dst = a;
    launch_rows(c, T.lazy, R2M_CONV, j, P.b, 1u << P.S);
  };
  rows(Y.get(), C0.get());
  ColJob2 cj = chain(Y.get(), X.get(), 0, h, 0u - h);
  cj.flags = R2_NEG | R2_PREFIX; cj.epi = cc; cj.epi_avail = h;
  launch_cols(c, T.lazy, kColChain, cj, P);
  rows(Z.get(), X.get());
  ColJob2 tj = chain(Z.get(), Y.get(), 0, mm, 0);
  tj.flags = R2_MULTAB; tj.epi = recip;
  launch_cols(c, T.lazy, kColChain, tj, P);
  rows(Y.get(), Bs.get());
  ColJob2 ij = chain(Y.get(), b + m, 0, mm, 0);
  launch_cols(c, T.lazy, kColInv, ij, P);
}

// sqrt's doubling from m to 2m coefficients (square_root.hpp:46-100, with the
// inverse's extension inverse.hpp:18-65) at N = 2m points, c = 1/b known to
// m/2, in seven launches (as exp_fused_doubling):
//   B = F(b), C0 = F(c[0, m/2))                                one launch
//   E = B C0, Q = B B                                          one launch
//   F(e), e = E[m/2, m) moved down a quarter;
//   F(res), res = a[v + m + i] - Q[m + i] moved down a half    one launch
//   G = F(e) C0
//   F(c), c[m/2, m) = -G[0, m/2) moved up a quarter (stored to cc) under
//   c[0, m/2) (read from cc)
//   T = F(res) F(c)
//   b[m, 2m) = T[0, m) / 2
// amin = a + v + m with avail words; b b has fewer than 2m terms.
static void sqrt_fused_doubling(Context& c, const HostField& T, int L, uint32_t* b, size_t m, uint32_t* cc,
                                const uint32_t* amin, uint32_t avail, uint32_t half_inv) {
  const PassPlan P = c.conv_plan(L);
  const uint32_t N = 1u << L, inv_n = inv_pow2(T, L), h = (uint32_t)(m / 2), mm = (uint32_t)m;
  DevBuf Bs = c.buf(N), C0 = c.buf(N), X = c.buf(N), Y = c.buf(N), Z = c.buf(N);
  ColJob2 fj[2];
  for (ColJob2& j : fj) {
    j = ColJob2{};
    j.f = T.dev;
  }
  fj[0].src = b; fj[0].dst = Bs.get(); fj[0].in_count = mm;
  fj[1].src = cc; fj[1].dst = C0.get(); fj[1].in_count = h;
  launch_cols_n(c, T.lazy, kColFwd, fj, P, 2, nullptr);
  RowJob2 rj[2];
  for (RowJob2& j : rj) {
    j = RowJob2{};
    j.f = T.dev; j.nrows = 1u << P.S; j.in_a = N; j.in_b = N; j.out_lo = 0; j.out_hi = N;
  }
  rj[0].src_a = Bs.get(); rj[0].src_b = C0.get(); rj[0].dst = X.get();
  rj[1].src_a = Bs.get(); rj[1].src_b = Bs.get(); rj[1].dst = Z.get();
  launch_rows_group(c, T.lazy, R2M_CONV, rj, 2, P.b, 1u << P.S);
  auto chain = [&](const uint32_t* src, uint32_t* dst, uint32_t lo, uint32_t hi, uint32_t shift) {
    ColJob2 j{};
    j.f = T.dev; j.inv_n = inv_n;
    j.src = src; j.dst = dst; j.out_lo = lo; j.out_hi = hi; j.out_shift = shift;
    return j;
  };
  ColJob2 ch[2] = {chain(X.get(), Y.get(), h, mm, h), chain(Z.get(), Z.get(), mm, N, mm)};
  ch[1].flags = R2_SUBFROM; ch[1].epi = amin; ch[1].epi_avail = avail;
  launch_cols_n(c, T.lazy, kColChain, ch, P, 2, nullptr);
  auto rows = [&](uint32_t* a, const uint32_t* bb) {
    RowJob2 j = rj[0];
    j.src_a = a; j.src_b = bb; j.dst = a;
    launch_rows(c, T.lazy, R2M_CONV, j, P.b, 1u << P.S);
  };
  rows(Y.get(), C0.get());
  ColJob2 cj = chain(Y.get(), X.get(), 0, h, 0u - h);
  cj.flags = R2_NEG | R2_PREFIX; cj.epi = cc; cj.epi_avail = h;
  launch_cols(c, T.lazy, kColChain, cj, P);
  rows(Z.get(), X.get());
  ColJob2 ij = chain(Z.get(), b + m, 0, mm, 0);
  ij.inv_n = T.mul(inv_n, half_inv);
  launch_cols(c, T.lazy, kColInv, ij, P);
}

static void series_emit(Context& c, const HostField& T, const SeriesReq& q, const ScanOut& s) {
  const uint32_t* a = q.a;
  const size_t la = q.la, n = q.n;
  uint32_t* out = q.out;
  const uint32_t one = T.from_int(1);
  switch (q.op) {
    case FPXG_INV: {  // poly::inv (inverse.hpp:163-188)
      if (la == 0 || s.a0 == 0) throw Error(E_DOMAIN, "poly::inv: constant term is zero");
      const size_t cnt = std::max<size_t>(1, std::min<size_t>(s.last1, la));
      if (cnt == 1) {
        emit_ew(c, T, EW_FILL, out, nullptr, nullptr, n, 0, 0);
        recur(c, T, REC_INV, a, 1, nullptr, 0, out, 0, 1, 0, 0);
        return;
      }
      inv_newton(c, T, a, cnt, n, out);
      return;
    }
    case FPXG_LOG: {  // poly::log (exponential.hpp:85-98)
      if (la == 0 || s.a0 != one) throw Error(E_DOMAIN, "poly::log: constant term must be 1");
      if (n > T.p) throw Error(E_DOMAIN, "poly::log: required denominators must be invertible (n <= mod)");
      if (n == 1) {
        emit_ew(c, T, EW_FILL, out, nullptr, nullptr, 1, 0, 0);
        return;
      }
      const size_t cnt = std::min(la, n);
      const size_t ld = cnt > 1 ? cnt - 1 : 0;
      DevBuf d = c.buf(std::max<size_t>(ld, 1));
      emit_ew(c, T, EW_DERIV, d.get(), a, nullptr, ld, 0, 0);
      DevBuf ia = c.buf(n - 1);
      const size_t icnt = std::max<size_t>(1, std::min<size_t>(s.last1, std::min(la, n - 1)));
      if (icnt == 1) {
        emit_ew(c, T, EW_FILL, ia.get(), nullptr, nullptr, n - 1, 0, 0);
        recur(c, T, REC_INV, a, 1, nullptr, 0, ia.get(), 0, 1, 0, 0);
      } else {
        inv_newton(c, T, a, icnt, n - 1, ia.get());
      }
      size_t lp = 0;
      DevBuf pr = c.buf(n - 1);
      if (ld) {
        lp = std::min(ld + (n - 1) - 1, n - 1);
        conv_dev(c, T, d.get(), ld, ia.get(), n - 1, pr.get(), lp);
      }
      emit_ew(c, T, EW_INTEGRAL, out, pr.get(), recip_table(c, T, n), n, (uint32_t)lp, 0);
      return;
    }
    case FPXG_EXP: {  // poly::exp (exponential.hpp:107-207)
      if (la != 0 && s.a0 != 0) throw Error(E_DOMAIN, "poly::exp: constant term must be 0");
      if (n - 1 >= T.p) throw Error(E_DOMAIN, "poly::inv_series: length must be < mod");
      const size_t cnt = std::min(la, n);
      // a' = derivative_prefix(a, n) trimmed: its length is one less than one
      // past the last nonzero index below cnt (i < p makes every factor i invertible).
      const size_t last = std::min<size_t>(s.last1, cnt);
      const size_t lap = last > 1 ? last - 1 : 0;
      if (lap == 0) {
        emit_ew(c, T, EW_FILL, out, nullptr, nullptr, n, 0, 0);
        emit_ew(c, T, EW_FILL, out, nullptr, nullptr, 1, 0, one);
        return;
      }
      // (No zero fill of out: the small-series launch and the Newton steps
      // below write every coefficient of b = out[0, n).)
      // th = x a'(x) (th[i] = i a[i], lap + 1 words) and a' = th + 1: the
      // fused doublings take r = (a' b)[m - 1, 2m - 1) as (th b)[m, 2m),
      // a window at half the transform.
      DevBuf th = c.buf(lap + 1);
      emit_ew(c, T, EW_THETA, th.get(), a, nullptr, lap + 1, 0, 0);
      const uint32_t* ap = th.get() + 1;
      uint32_t* b = out;
      if (T.p == 2) {  // n <= 2 (n - 1 < p): the recurrence alone
        recur(c, T, REC_EXP, ap, lap, nullptr, 0, b, 0, n, one, 0);
        return;
      }
      // 1/i for the steps' t[i] /= m + i (exponential.hpp:194-195), read by the
      // small-series kernel, the direct steps (up to 2^9 points: indices
      // below 512) and, on native fields, the product epilogues.  (Other
      // fields read it only below 512.)
      // (The process-wide table where there is one, else a table of this call.)
      const size_t nr = series_native(T, 6) ? n : std::min<size_t>(n, std::max<size_t>(kSmallSeries, 512));
      DevBuf rbuf;
      const uint32_t* rtab = recip_table(c, T, nr);
      if (!rtab) {
        rbuf = c.buf(nr);
        emit_ew(c, T, EW_RECIP, rbuf.get(), nullptr, nullptr, nr, 0, 0);
        rtab = rbuf.get();
      }
      // The first min(n, 512) coefficients of b, and c = 1/b to as many.
      const size_t n0 = std::min<size_t>(n, kSmallSeries);
      DevBuf cbuf = c.buf(n);
      {
        RowJob2 rj{};
        rj.src_a = ap; rj.in_a = (uint32_t)std::min(lap, n0 - 1); rj.epi = rtab;
        rj.dst = b; rj.src_c = cbuf.get();
        small_series(c, T, R2M_NEXP, rj, n0);
      }
      if (n == n0) return;
      size_t csize = n0;  // the small-series launch leaves c = 1/b mod x^n0
      for (size_t m = n0; m < n; m *= 2) {
        const size_t m2 = std::min(2 * m, n), hlen = m2 - m;
        if (hlen <= 8) {
          recur(c, T, REC_EXP, ap, lap, nullptr, 0, b, m, m2, one, 0);
          continue;
        }
        Engine E(c, T, ceil_log2(2 * m));
        if (rtab && exp_fused_ok(c, T, E, m, hlen, csize)) {
          exp_fused_doubling(c, T, E.L, th.get(), std::min(lap, m2 - 1) + 1, b, m, cbuf.get(), rtab + m);
          csize = m;
          continue;
        }
        // The c extension's first product and r = (a' b)[m - 1, ...) depend
        // only on b and the old c: one set of launches for both.
        const bool merged = csize < hlen && E.native && rtab && !one_launch_step(c, T, E, m) &&
                            cext_mergeable(c, T, csize, hlen, E.L);
        DevBuf r;
        if (merged) {
          r = c.buf(hlen);
          DevBuf e = c.buf(hlen - csize);
          const WinProd w[2] = {{b, std::min(m, hlen), cbuf.get(), csize, csize, hlen, e.get(), true, Epi()},
                                {ap, std::min(lap, m2 - 1), b, m, m - 1, m - 1 + hlen, r.get(), false, Epi()}};
          native_conv_multi(c, T, E.L, w, 2);
          cext_finish(c, T, e.get(), cbuf.get(), csize, hlen);
          csize = hlen;
        } else if (csize < hlen) {
          extend_inverse(c, T, b, m, cbuf.get(), csize, hlen);
          csize = hlen;
        }
        if (rtab && one_launch_step(c, T, E, m)) {
          // The step's three products in one workgroup (R2M_EXPSTEP).
          RowJob2 rj{};
          rj.f = T.dev;
          rj.src_a = ap; rj.in_a = (uint32_t)std::min(lap, m2 - 1);
          rj.src_b = b; rj.in_b = (uint32_t)m;
          rj.src_c = cbuf.get(); rj.aux = (uint32_t)(m - 1); rj.epi = rtab + m;
          rj.dst = b + m; rj.nrows = 1; rj.out_lo = 0; rj.out_hi = (uint32_t)hlen; rj.out_shift = 0;
          rj.inv_n = inv_pow2(T, E.L);
          rj.flags = R2_SCALE | R2_CANON;
          launch_rows(c, T.lazy, R2M_EXPSTEP, rj, E.L, 1);
          continue;
        }
        if (!merged) r = c.buf(hlen);
        if (E.native && rtab && !merged)
          native_conv(c, T, E.L, ap, std::min(lap, m2 - 1), b, m, m - 1, m - 1 + hlen, r.get());
        if (E.native && rtab && T.lazy && ((size_t)1 << E.L) == 2 * m && c.conv_plan(E.L).S >= 5) {
          // t = (r c)[0, h) / (m + i) hands itself to out = (t b)[0, h)
          // through the chained pass; b's column spectrum rides in the first
          // product's column launch: five launches instead of six.
          const PassPlan P = c.conv_plan(E.L);
          const uint32_t N = 1u << E.L;
          const uint32_t inv_n = inv_pow2(T, E.L);
          DevBuf R = c.buf(N), Cs = c.buf(N), Bs = c.buf(N), Tz = c.buf(N);
          ColJob2 cj[3];
          for (ColJob2& j : cj) {
            j = ColJob2{};
            j.f = T.dev;
          }
          cj[0].src = r.get(); cj[0].dst = R.get(); cj[0].in_count = (uint32_t)hlen;
          cj[1].src = cbuf.get(); cj[1].dst = Cs.get(); cj[1].in_count = (uint32_t)hlen;
          cj[2].src = b; cj[2].dst = Bs.get(); cj[2].in_count = (uint32_t)m;
          launch_cols_n(c, T.lazy, kColFwd, cj, P, 3, nullptr);
          RowJob2 rj{};
          rj.f = T.dev; rj.nrows = 1u << P.S; rj.in_a = N; rj.in_b = N; rj.out_lo = 0; rj.out_hi = N;
          rj.src_a = R.get(); rj.src_b = Cs.get(); rj.dst = R.get();
          launch_rows(c, T.lazy, R2M_CONV, rj, P.b, rj.nrows);
          ColJob2 ch{};
          ch.src = R.get(); ch.dst = Tz.get(); ch.out_lo = 0; ch.out_hi = (uint32_t)hlen; ch.out_shift = 0;
          ch.inv_n = inv_n; ch.flags = R2_MULTAB; ch.epi = rtab + m; ch.f = T.dev;
          launch_cols(c, T.lazy, kColChain, ch, P);
          rj.src_a = Tz.get(); rj.src_b = Bs.get(); rj.dst = Tz.get();
          launch_rows(c, T.lazy, R2M_CONV, rj, P.b, rj.nrows);
          ColJob2 ij{};
          ij.src = Tz.get(); ij.dst = b + m; ij.out_lo = 0; ij.out_hi = (uint32_t)hlen; ij.out_shift = 0;
          ij.inv_n = inv_n; ij.f = T.dev;
          launch_cols(c, T.lazy, kColInv, ij, P);
          continue;
        }
        DevBuf t = c.buf(hlen);
        if (E.native && rtab) {
          Epi ep;
          ep.flags = R2_MULTAB;
          ep.src = rtab + m;
          native_conv(c, T, E.L, r.get(), hlen, cbuf.get(), hlen, 0, hlen, t.get(), false, ep);
          native_conv(c, T, E.L, t.get(), hlen, b, m, 0, hlen, b + m);
          continue;
        }
        crt_conv(c, T, E.L, ap, std::min(lap, m2 - 1), b, m, m - 1, m - 1 + hlen, r.get());
        crt_conv(c, T, E.L, r.get(), hlen, cbuf.get(), hlen, 0, hlen, t.get());
        emit_ew(c, T, EW_MUL_RECIP, t.get(), t.get(), nullptr, hlen, 0, (uint32_t)m);
        crt_conv(c, T, E.L, t.get(), hlen, b, m, 0, hlen, b + m);
      }
      return;
    }
    case FPXG_SQRT: {  // poly::sqrt (square_root.hpp:12-112)
      // The domain checks come before the output's zero fill: a failing call
      // writes nothing.
      const size_t v = s.first;
      uint32_t root0 = 0;
      if (v < la) {
        if (v & 1) throw Error(E_DOMAIN, "poly::sqrt: odd valuation, no square root");
        if (!T.sqrt(s.afirst, &root0)) throw Error(E_DOMAIN, "poly::sqrt: non-residue constant term");
      }
      const size_t half = v / 2;
      if (v >= la || half >= n) {  // the zero series, or a root vanishing below x^n
        emit_ew(c, T, EW_FILL, out, nullptr, nullptr, n, 0, 0);
        return;
      }
      const size_t inner = n - half, au = la - v;
      const size_t ucnt = std::min(au, inner);
      uint32_t* b = out + half;
      if (s.second >= v + ucnt) {  // no nonzero coefficient in (v, v + ucnt)
        emit_ew(c, T, EW_FILL, out, nullptr, nullptr, n, 0, 0);
        emit_ew(c, T, EW_FILL, b, nullptr, nullptr, 1, 0, root0);
        return;
      }
      // The Newton steps below write every coefficient of b = out[half, n);
      // only the valuation's zeros need a fill.
      emit_ew(c, T, EW_FILL, out, nullptr, nullptr, half, 0, 0);
      const uint32_t half_inv = T.from_int(((uint64_t)T.p + 1) / 2);
      const uint32_t tri = T.mul(T.inv(root0), half_inv);
      const uint32_t* av = a + v;
      // The first min(inner, 512) coefficients of b, and c = 1/b to as many.
      const size_t n0 = std::min<size_t>(inner, kSmallSeries);
      DevBuf cbuf = c.buf(inner);
      {
        RowJob2 rj{};
        rj.src_a = av; rj.in_a = (uint32_t)std::min(au, n0); rj.aux = root0; rj.inv_n = T.inv(root0);
        rj.dst = b; rj.src_c = cbuf.get();
        small_series(c, T, R2M_NSQRT, rj, n0);
      }
      if (inner == n0) return;
      size_t csize = n0;  // the small-series launch leaves c = 1/b mod x^n0
      for (size_t m = n0; m < inner; m *= 2) {
        const size_t m2 = std::min(2 * m, inner), hlen = m2 - m;
        if (hlen <= 8) {
          recur(c, T, REC_SQRT, nullptr, 0, av, au, b, m, m2, root0, tri);
          continue;
        }
        Engine E(c, T, ceil_log2(2 * m));
        if (exp_fused_ok(c, T, E, m, hlen, csize)) {
          sqrt_fused_doubling(c, T, E.L, b, m, cbuf.get(), av + std::min(m, au), (uint32_t)(au > m ? au - m : 0),
                              half_inv);
          csize = m;
          continue;
        }
        // res = a[v + m + i] - (b * b)[m + i]  (square_root.hpp:93-94)
        Epi sub;
        sub.flags = R2_SUBFROM;
        sub.src = av + std::min(m, au);
        sub.avail = (uint32_t)(au > m ? au - m : 0);
        // The c extension's first product and b * b depend only on b and the
        // old c: one set of launches for both.
        const bool merged = csize < hlen && E.native && !one_launch_step(c, T, E, m) &&
                            cext_mergeable(c, T, csize, hlen, E.L);
        DevBuf res_m;
        if (merged) {
          res_m = c.buf(hlen);
          DevBuf e = c.buf(hlen - csize);
          const WinProd w[2] = {{b, std::min(m, hlen), cbuf.get(), csize, csize, hlen, e.get(), true, Epi()},
                                {b, m, b, m, m, m2, res_m.get(), false, sub}};
          native_conv_multi(c, T, E.L, w, 2);
          cext_finish(c, T, e.get(), cbuf.get(), csize, hlen);
          csize = hlen;
        } else if (csize < hlen) {
          extend_inverse(c, T, b, m, cbuf.get(), csize, hlen);
          csize = hlen;
        }
        if (one_launch_step(c, T, E, m)) {
          // The step's two products in one workgroup (R2M_SQRTSTEP).
          RowJob2 rj{};
          rj.f = T.dev;
          rj.src_b = b; rj.in_b = (uint32_t)m;
          rj.src_c = cbuf.get(); rj.aux = half_inv;
          rj.epi = av + std::min(m, au); rj.epi_avail = (uint32_t)(au > m ? au - m : 0);
          rj.dst = b + m; rj.nrows = 1; rj.out_lo = 0; rj.out_hi = (uint32_t)hlen; rj.out_shift = 0;
          rj.inv_n = inv_pow2(T, E.L);
          rj.flags = R2_SCALE | R2_CANON;
          launch_rows(c, T.lazy, R2M_SQRTSTEP, rj, E.L, 1);
          continue;
        }
        DevBuf sq = c.buf(hlen), res = merged ? std::move(res_m) : c.buf(hlen);
        if (E.native) {
          if (!merged) native_conv(c, T, E.L, b, m, b, m, m, m2, res.get(), false, sub);
          Epi half;  // times 1/2 (square_root.hpp:96-97)
          half.scale = half_inv;
          native_conv(c, T, E.L, res.get(), hlen, cbuf.get(), hlen, 0, hlen, b + m, false, half);
          continue;
        }
        crt_conv(c, T, E.L, b, m, b, m, m, m2, sq.get());
        emit_ew(c, T, EW_SUB_FROM, res.get(), av + std::min(m, au), sq.get(), hlen,
                (uint32_t)(au > m ? au - m : 0), 0);
        crt_conv(c, T, E.L, res.get(), hlen, cbuf.get(), hlen, 0, hlen, b + m);
        emit_ew(c, T, EW_SCALE, b + m, b + m, nullptr, hlen, 0, half_inv);
      }
      return;
    }
    default:
      throw Error(E_ARG, "unknown series op");
  }
}

// Characteristic-two square root (square_root.hpp:15-25), immediate only.
// The result is built in a buffer of its own and copied out only after the
// check, so a failing call writes nothing.
static void sqrt_char2(Context& c, const HostField& T, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  DevBuf res = c.buf(n);
  emit_ew(c, T, EW_FILL, res.get(), nullptr, nullptr, n, 0, 0);
  const size_t cnt = std::min(la, n);
  DevBuf flag = c.buf(1);
  emit_ew(c, T, EW_FILL, flag.get(), nullptr, nullptr, 1, 0, 0);
  if (cnt) {
    hipLaunchKernelGGL(k_frobenius_sqrt, dim3(grid_for(cnt)), dim3(256), 0, c.stream, a, cnt, res.get(),
                       (unsigned int*)flag.get());
    FPXG_HIP(hipGetLastError());
  }
  if (read_word(c, flag.get())) throw Error(E_DOMAIN, "poly::sqrt: nonzero odd coefficient in characteristic 2");
  FPXG_HIP(hipMemcpyAsync(out, res.get(), n * 4, hipMemcpyDeviceToDevice, c.stream));
}

static void series_dev(Context& c, const HostField& T, int op, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  if (op < FPXG_INV || op > FPXG_SQRT) throw Error(E_ARG, "unknown series op");
  if (op == FPXG_SQRT && T.p == 2) {
    sqrt_char2(c, T, a, la, n, out);
    return;
  }
  const SeriesReq q{op, a, la, n, out};
  const std::vector<ScanOut> s = scan_inputs(c, {scan_rec_of(q)});
  series_emit(c, T, q, s[0]);
}

}  // namespace fpxg
