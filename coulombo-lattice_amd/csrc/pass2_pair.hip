// Code object of the two-level pass kernels for one compile-time radix pair
// (R1V x R2V, set by the build for each pair in CLAT_P2_PAIRS): the general
// kernel and the compile-time column kernel.
#include "colpass.hpp"

#define CLAT_P2_FN2(a, b) clat_p2_##a##x##b
#define CLAT_P2_FN(a, b) CLAT_P2_FN2(a, b)
#define CLAT_C2_FN2(a, b) clat_c2_##a##x##b
#define CLAT_C2_FN(a, b) CLAT_C2_FN2(a, b)

namespace clat {
hipError_t CLAT_P2_FN(R1V, R2V)(int combo, int mode, bool line, const P2Launch& L) {
  return p2_launch<R1V, R2V>(combo, mode, line, L);
}
hipError_t CLAT_C2_FN(R1V, R2V)(int combo, int mode, int nsub, const ColLaunch& L) {
  return col_launch<R1V, R2V>(combo, mode, nsub, L);
}
}  // namespace clat
