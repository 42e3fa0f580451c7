// Code object of the two-level pass kernels for one compile-time radix pair
// (R1V x R2V, set by the build for each pair in CLAT_P2_PAIRS): the general
// kernel, the compile-time column kernel and the atom-path line kernel.
#include "linepass.hpp"

#define CLAT_P2_FN2(a, b) clat_p2_##a##x##b
#define CLAT_P2_FN(a, b) CLAT_P2_FN2(a, b)
#define CLAT_C2_FN2(a, b) clat_c2_##a##x##b
#define CLAT_C2_FN(a, b) CLAT_C2_FN2(a, b)
#define CLAT_L2_FN2(a, b) clat_l2_##a##x##b
#define CLAT_L2_FN(a, b) CLAT_L2_FN2(a, b)
#define CLAT_L2N_FN2(a, b) clat_l2_lines_##a##x##b
#define CLAT_L2N_FN(a, b) CLAT_L2N_FN2(a, b)

namespace clat {
hipError_t CLAT_P2_FN(R1V, R2V)(int combo, int mode, bool line, const P2Launch& L) {
  return p2_launch<R1V, R2V>(combo, mode, line, L);
}
hipError_t CLAT_C2_FN(R1V, R2V)(int combo, int mode, int nsub, const ColLaunch& L) {
  return col_launch<R1V, R2V>(combo, mode, nsub, L);
}
hipError_t CLAT_L2_FN(R1V, R2V)(int combo, int mode, const LineLaunch& L) {
  return line_launch<R1V, R2V>(combo, mode, L);
}
int CLAT_L2N_FN(R1V, R2V)(int combo) { return line_lines<R1V, R2V>(combo); }
}  // namespace clat
