// Launch records shared by the immediate drivers and the merged batch
// executor: every elementwise op, recurrence and CRT reconstruction is a small
// POD record, so one launch can serve any number of them.
#pragma once

#include <stdint.h>

namespace fpxg {

struct FieldScalar {
  uint32_t p, ninv, r2, one;
};

// Sequential recurrences (seeds and <= 8-coefficient tails).
enum { REC_INV = 0, REC_EXP = 1, REC_SQRT = 2 };
constexpr int kRecMax = 64;
struct RecurRec {
  int kind;
  FieldScalar f;
  const uint32_t* u;
  uint32_t lu;
  const uint32_t* w;
  uint32_t lw;
  uint32_t* v;
  uint32_t lo, hi, init, k;
};

// Elementwise ops on Montgomery limbs.
enum {
  EW_FILL = 0,      // dst[i] = c
  EW_NEGATE = 1,    // dst[i] = -src[i]
  EW_SCALE = 2,     // dst[i] = src[i] * c (Montgomery limb c)
  EW_MUL_RECIP = 3, // dst[i] = src[i] / (c + i)                    (exponential.hpp:194-195)
  EW_DERIV = 4,     // dst[i] = src[i+1] * (i+1)                    (calculus.hpp:17-39)
  EW_INTEGRAL = 5,  // dst[i] = i in [1, avail] ? src[i-1] / i : 0  (calculus.hpp:97-116)
  EW_SUB_FROM = 6,  // dst[i] = (i < avail ? src[i] : 0) - src2[i]  (square_root.hpp:93-94)
  EW_REVERSE = 7,   // dst[i] = src[n - 1 - i]                      (division.hpp:53-58)
  EW_MULADD = 8,    // dst[i] = (c ? dst[i] : 0) + src[i] * src2[i] (spectra of split products)
  EW_RECIP = 9,     // dst[i] = 1 / (c + i), 0 for c + i == 0       (exponential.hpp:194-195 table)
  EW_THETA = 10     // dst[i] = src[i] * i: x a'(x), so dst + 1 is a' (calculus.hpp:17-39)
};
struct EwRec {
  int op;
  FieldScalar f;
  uint32_t* dst;
  const uint32_t* src;
  const uint32_t* src2;
  uint32_t n, avail, c;
};

// Products with a short side, computed directly (k_conv_direct): out[0, outc)
// of sh[0, ns) * lg[0, nl), ns <= kDirectShort.
struct DirectRec {
  const uint32_t* sh;
  const uint32_t* lg;
  uint32_t* out;
  uint32_t ns, nl, outc, p, ninv, pad;
};

// Tail repair of a product N + tail coefficients long (k_tail_fix): the high
// tail of a * b computed directly into out[N, N + tail) and its wraparound
// removed from out[0, tail).
struct TailRec {
  const uint32_t* a;
  const uint32_t* b;
  uint32_t* out;
  uint32_t la, lb, outc, N, tail, p, ninv, pad;
};

}  // namespace fpxg
