// Launchers of the modal-path kernels, one switch over the supported grid
// lengths per kernel family.  GLW_MODAL_SET picks the families built by this
// translation unit (the units are compiled in parallel):
//   0: time pass, order-2 sources and eta20 forcing
//   1: time pass, order-3 forcing
//   2: modal pass, eta22 (+ pair fields)
//   3: modal pass, eta33 and eta20
#include "glw_modal.hpp"
#include "glw_modal_launch.hpp"

namespace glw {
namespace modal {

namespace {

template <int T, int MODE>
int t_go(const Args& a, hipStream_t st) {
  if (a.q_len % t_rows(T, MODE)) return 1;
  const int dyn = t_dyn_bytes<T, MODE>(a.span + 1);
  if (dyn && hipFuncSetAttribute((const void*)k_modal_t<T, MODE>, hipFuncAttributeMaxDynamicSharedMemorySize, dyn) !=
                 hipSuccess)
    return 2;
  hipLaunchKernelGGL((k_modal_t<T, MODE>), dim3(a.q_len / t_rows(T, MODE)), dim3(t_threads(T, MODE)), dyn, st, a);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

template <int Q, int MODE>
int q_go(const Args& a, int ncol, int r0, hipStream_t st) {
  if (a.q_len != Q * r0) return 1;
  hipLaunchKernelGGL((k_modal_q<Q, MODE>), dim3(ncol, r0), dim3(q_threads(Q)), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

template <int Q>
int p_go(const Args& a, int ncol, int r0, hipStream_t st) {
  if (a.q_len != Q * r0) return 1;
  hipLaunchKernelGGL((k_modal_pairs<Q>), dim3(ncol, r0, kPairFields), dim3(q_threads(Q)), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // namespace

#define GLW_T_CASE(L) \
  case L:             \
    return t_go<L, MODE>(a, st);
#define GLW_Q_CASE(L) \
  case L:             \
    return q_go<L, MODE>(a, ncol, r0, st);

#if GLW_MODAL_SET == 0 || GLW_MODAL_SET == 1
template <int MODE>
static int t_switch(const Args& a, hipStream_t st) {
  switch (a.t_len) {
    GLW_MODAL_T_LENGTHS(GLW_T_CASE)
    default: return 1;
  }
}
#endif
#if GLW_MODAL_SET >= 2
template <int MODE>
static int q_switch(const Args& a, int ncol, int r0, hipStream_t st) {
  if (r0 < 1 || a.q_len % r0) return 1;
  switch (a.q_len / r0) {
    GLW_MODAL_Q_LENGTHS(GLW_Q_CASE)
    default: return 1;
  }
}
#endif
#if GLW_MODAL_SET == 2
#define GLW_P_CASE(L) \
  case L:             \
    return p_go<L>(a, ncol, r0, st);
static int p_switch(const Args& a, int ncol, int r0, hipStream_t st) {
  if (r0 < 1 || a.q_len % r0) return 1;
  switch (a.q_len / r0) {
    GLW_MODAL_Q_LENGTHS(GLW_P_CASE)
    default: return 1;
  }
}
#endif

#if GLW_MODAL_SET == 0
int launch_t_order2(const Args& a, hipStream_t st) { return t_switch<MODE_O2>(a, st); }
int launch_t_eta20(const Args& a, hipStream_t st) { return t_switch<MODE_E20>(a, st); }
#elif GLW_MODAL_SET == 1
int launch_t_order3(const Args& a, hipStream_t st) { return t_switch<MODE_O3>(a, st); }
#elif GLW_MODAL_SET == 2
int launch_q_order2(const Args& a, int ncol, int r0, hipStream_t st) { return q_switch<MODE_O2>(a, ncol, r0, st); }
int launch_q_pairs(const Args& a, int ncol, int r0, hipStream_t st) { return p_switch(a, ncol, r0, st); }
#elif GLW_MODAL_SET == 3
int launch_q_order3(const Args& a, int ncol, int r0, hipStream_t st) { return q_switch<MODE_O3>(a, ncol, r0, st); }
int launch_q_eta20(const Args& a, int ncol, hipStream_t st) { return q_switch<MODE_E20>(a, ncol, 1, st); }
#endif

}  // namespace modal
}  // namespace glw
