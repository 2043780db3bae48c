// Inverse column passes of CRT products with Garner fused (see launch.hpp).
#include "launch.hpp"

namespace fpxg {

void colcrt_part(Context& c, const CrtColJob* j, const PassPlan& P, uint32_t njobs, const CrtColJob* dj,
                 const CrtFields& F) {
  switch (P.S) {
#define FPXG_C(S) \
  case S: colcrt_launch<S>(c, j, P, njobs, dj, F); break;
    FPXG_C(1) FPXG_C(2) FPXG_C(3) FPXG_C(4) FPXG_C(5) FPXG_C(6)
    FPXG_C(7) FPXG_C(8) FPXG_C(9) FPXG_C(10) FPXG_C(11) FPXG_C(12)
#undef FPXG_C
    default: throw Error(E_SIZE, "column pass size out of range");
  }
}

bool colcrt_attrs() { return colcrt_attr<10>() && colcrt_attr<11>() && colcrt_attr<12>(); }

}  // namespace fpxg
