// Row-pass launches, lazy limbs, rows of 2^1 .. 2^11 words (see launch.hpp).
#include "launch.hpp"

namespace fpxg {

void rows2_part_lazy_lo(Context& c, int logM, int mode, const RowJob2* j, uint32_t nrows, uint32_t njobs,
                        const RowJob2* dj) {
  switch (logM) {
#define FPXG_R(L) \
  case L: rows2_mode<L, true>(c, mode, j, nrows, njobs, dj); break;
    FPXG_R(1) FPXG_R(2) FPXG_R(3) FPXG_R(4) FPXG_R(5) FPXG_R(6) FPXG_R(7) FPXG_R(8) FPXG_R(9) FPXG_R(10) FPXG_R(11)
#undef FPXG_R
    default: throw Error(E_SIZE, "row pass size out of range");
  }
}

}  // namespace fpxg
