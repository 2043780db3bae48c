// Row-pass launches, canonical limbs (2^30 <= p < 2^31), rows of 2^1 .. 2^14
// words (see launch.hpp).
#include "launch.hpp"

namespace fpxg {

void rows2_part_canon(Context& c, int logM, int mode, const RowJob2* j, uint32_t nrows, uint32_t njobs,
                      const RowJob2* dj) {
  switch (logM) {
#define FPXG_R(L) \
  case L: rows2_mode<L, false>(c, mode, j, nrows, njobs, dj); break;
    FPXG_R(1) FPXG_R(2) FPXG_R(3) FPXG_R(4) FPXG_R(5) FPXG_R(6) FPXG_R(7)
    FPXG_R(8) FPXG_R(9) FPXG_R(10) FPXG_R(11) FPXG_R(12) FPXG_R(13) FPXG_R(14)
#undef FPXG_R
    default: throw Error(E_SIZE, "row pass size out of range");
  }
}

bool rows2_attrs_canon() {
  bool ok = true;
#define X(L) \
  ok = rows2_attr<L, false, R2M_FWD>() && rows2_attr<L, false, R2M_INV>() && rows2_attr<L, false, R2M_CONV>() && \
       rows2_attr<L, false, R2M_SPEC>() && rows2_attr<L, false, R2M_SPEC_ONLY>() &&                            \
       rows2_attr<L, false, R2M_SPEC_SUM2>() && ok;
  X(13) X(14)
#undef X
  return ok;
}

}  // namespace fpxg
