// Inverse column-pass launches, lazy limbs (see launch.hpp).
#include "launch.hpp"

namespace fpxg {

void cols2_part_lazy_inv(Context& c, const ColJob2* j, const PassPlan& P, uint32_t njobs, const ColJob2* dj) {
  switch (P.S) {
#define FPXG_C(S) \
  case S: cols2_launch<S, true, true, false>(c, j, P, njobs, dj); break;
    FPXG_C(1) FPXG_C(2) FPXG_C(3) FPXG_C(4) FPXG_C(5) FPXG_C(6)
    FPXG_C(7) FPXG_C(8) FPXG_C(9) FPXG_C(10) FPXG_C(11) FPXG_C(12)
#undef FPXG_C
    default: throw Error(E_SIZE, "column pass size out of range");
  }
}

void cols2_part_lazy_chain(Context& c, const ColJob2* j, const PassPlan& P, uint32_t njobs, const ColJob2* dj) {
  switch (P.S) {
#define FPXG_C(S) \
  case S: cols2_launch<S, true, true, false, true>(c, j, P, njobs, dj); break;
    FPXG_C(5) FPXG_C(6) FPXG_C(7) FPXG_C(8) FPXG_C(9) FPXG_C(10) FPXG_C(11) FPXG_C(12)
#undef FPXG_C
    default: throw Error(E_SIZE, "chained column pass size out of range");
  }
}

bool cols2_attrs_lazy_inv() {
  bool ok = true;
#define X(S) ok = cols2_attr<S, true, true>() && cols2_attr<S, true, true, false, true>() && ok;
  X(9) X(10) X(11) X(12)
#undef X
  return ok;
}

}  // namespace fpxg
