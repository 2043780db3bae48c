// Row-pass launches, lazy limbs, rows of 2^12 words (see launch.hpp).
#include "launch.hpp"

namespace fpxg {

void rows2_part_lazy_12(Context& c, int logM, int mode, const RowJob2* j, uint32_t nrows, uint32_t njobs,
                        const RowJob2* dj) {
  if (logM != 12) throw Error(E_SIZE, "row pass size out of range");
  rows2_mode<12, true>(c, mode, j, nrows, njobs, dj);
}

}  // namespace fpxg
