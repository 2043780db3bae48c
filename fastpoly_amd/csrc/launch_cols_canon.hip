// Column-pass launches, canonical limbs (2^30 <= p < 2^31), forward and
// inverse (see launch.hpp).
#include "launch.hpp"

namespace fpxg {

template <bool INV>
static void cols_canon(Context& c, const ColJob2* j, const PassPlan& P, uint32_t njobs, const ColJob2* dj) {
  switch (P.S) {
#define FPXG_C(S) \
  case S: cols2_launch<S, false, INV, false>(c, j, P, njobs, dj); break;
    FPXG_C(1) FPXG_C(2) FPXG_C(3) FPXG_C(4) FPXG_C(5) FPXG_C(6)
    FPXG_C(7) FPXG_C(8) FPXG_C(9) FPXG_C(10) FPXG_C(11) FPXG_C(12)
#undef FPXG_C
    default: throw Error(E_SIZE, "column pass size out of range");
  }
}

void cols2_part_canon(Context& c, bool inv, const ColJob2* j, const PassPlan& P, uint32_t njobs,
                      const ColJob2* dj) {
  if (inv) cols_canon<true>(c, j, P, njobs, dj);
  else cols_canon<false>(c, j, P, njobs, dj);
}

bool cols2_attrs_canon() {
  bool ok = true;
#define X(S) ok = cols2_attr<S, false, false>() && cols2_attr<S, false, true>() && ok;
  X(9) X(10) X(11) X(12)
#undef X
  return ok;
}

}  // namespace fpxg
