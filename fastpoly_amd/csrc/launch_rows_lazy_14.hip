// Row-pass launches, lazy limbs, rows of 2^14 words (see launch.hpp).
#include "launch.hpp"

namespace fpxg {

void rows2_part_lazy_14(Context& c, int logM, int mode, const RowJob2* j, uint32_t nrows, uint32_t njobs,
                        const RowJob2* dj) {
  if (logM != 14) throw Error(E_SIZE, "row pass size out of range");
  rows2_mode<14, true>(c, mode, j, nrows, njobs, dj);
}

bool rows_dma_part_lazy(Context& c, bool sgn, const RowJob2* j, uint32_t nrows, uint32_t njobs) {
  if (!rows_dma_ok(c, j, nrows, njobs)) return false;
  if (sgn) rows_dma_go<true, true>(c, j, nrows, njobs);
  else rows_dma_go<true, false>(c, j, nrows, njobs);
  FPXG_HIP(hipGetLastError());
  return true;
}

bool rows2_attrs_lazy_14() {
  return rows2_attr<14, true, R2M_FWD>() && rows2_attr<14, true, R2M_INV>() && rows2_attr<14, true, R2M_CONV>() &&
         rows2_attr<14, true, R2M_SPEC>() && rows2_attr<14, true, R2M_SPEC_ONLY>() &&
         rows2_attr<14, true, R2M_SPEC_SUM2>() && rows2_attr<14, true, R2M_FWDS>();
}

}  // namespace fpxg
