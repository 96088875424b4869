// Column passes, split-row transforms and row synthesis: FP64 launcher
// instantiations, split by GLW_COL_SET into units the build compiles in
// parallel (0: column-inverse + synthesis, 1: column-forward + split rows).
#include "glw_launch_impl.hpp"

#ifndef GLW_COL_SET
#error "GLW_COL_SET must select a launcher set"
#endif

namespace glw {

#if GLW_COL_SET == 0

bool length_supported(int n) {
  switch (n) {
#define GLW_CASE(L) case L:
    GLW_LENGTHS(GLW_CASE)
#undef GLW_CASE
    return true;
    default:
      return false;
  }
}

int launch_synth(int gx, const SynthArgs& a, int nseq, const void* tw, hipStream_t st) {
  switch (gx) {
#define GLW_CASE(L)                                                                                            \
  case L: {                                                                                                    \
    constexpr int RY = (L <= 256) ? 4 : 2;                                                                     \
    if (a.gy % RY) return glw_report_error(GLW_EINVAL, "row count not divisible by the synthesis tile");       \
    hipLaunchKernelGGL((k_synth<double, L, RY>), dim3(a.gy / RY, nseq), dim3(NT), 0, st, a,                    \
                       (const cpx<double>*)tw);                                                                \
    break;                                                                                                     \
  }
    GLW_LENGTHS(GLW_CASE)
#undef GLW_CASE
    default:
      return glw_report_error(GLW_EUNSUPPORTED, "synthesis length " + std::to_string(gx) + " not instantiated");
  }
  GLW_HIP_CHECK(hipGetLastError());
  return GLW_OK;
}

template int launch_colinv<double>(int, const ColInvArgs&, int, const void*, const Exec&, hipStream_t);
#elif GLW_COL_SET == 1
template int launch_colfwd<double>(int, const ColFwdArgs&, int, const void*, const Exec&, hipStream_t);
template int launch_rowinv<double>(int, const RowArgs&, int, const void*, int, void*, hipStream_t);
template int launch_rowfwd<double>(int, const RowArgs&, int, const void*, int, const void*, hipStream_t);
#else
#error "unknown GLW_COL_SET"
#endif

}  // namespace glw
