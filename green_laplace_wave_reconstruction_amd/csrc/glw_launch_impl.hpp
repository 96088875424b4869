// Definitions of the launchers declared in glw_launch.hpp.  Included only by
// the glw_launch_*.hip translation units, which instantiate them.
#pragma once

#include <string>
#include <type_traits>

#include "glw_internal.hpp"
#include "glw_launch.hpp"

namespace glw {

#ifndef GLW_COL_LDS_KB
#define GLW_COL_LDS_KB 48
#endif
#ifndef GLW_ROW_LDS_KB
#define GLW_ROW_LDS_KB 64
#endif
constexpr int kLdsBudget = GLW_COL_LDS_KB * 1024;  // column tiles: keep several workgroups per CU
constexpr int kRowLds = GLW_ROW_LDS_KB * 1024;     // fused row kernel tile
constexpr int kLdsMax = 160 * 1024;    // gfx950 LDS per workgroup

// Widest power-of-two column tile (<= 32 columns) within kLdsBudget.
template <typename T, int L>
constexpr int col_width() {
  int cw = kLdsBudget / (col_stride<L>() * (int)sizeof(cpx<T>));
  int p = 1;
  while (p * 2 <= cw && p < 32) p *= 2;
  return p;
}

// LDS budget of a fused row tile for an operator: kRowLds unless the operator
// sets kRowLdsKb.  The eta44 third-order operator (31 fields of 64 points)
// takes 96 KB, two rows per workgroup: its products and forward transforms then
// keep more lanes busy (eta44 GL6/8/10 at 256^2: +4-6 % on MI355X), while the
// order-2 and order-3 operators are faster at 64 KB (-2.5 to -4 % at 96 KB).
template <class Op, class = void>
struct row_lds_of {
  static constexpr int value = kRowLds;
};
template <class Op>
struct row_lds_of<Op, std::void_t<decltype(Op::kRowLdsKb)>> {
  static constexpr int value = Op::kRowLdsKb * 1024;
};

// Rows per workgroup of the fused row kernel: as many as fit the operator's
// LDS budget (at most 4).
template <typename T, int NIN, int L, class Op = void>
constexpr int row_ry() {
  int per = NIN * seq_stride<L>() * (int)sizeof(cpx<T>);
  int ry = row_lds_of<Op>::value / per;
  if (ry >= 4) return 4;
  if (ry >= 2) return 2;
  return 1;
}

template <typename T, int NIN, int L, int RY>
constexpr bool row_fits() {
  return (size_t)(NIN * RY * seq_stride<L>() + (kTwiddlesInLdsRow ? L : 1)) * sizeof(cpx<T>) <= (size_t)kLdsMax;
}

// Rows per workgroup of the split-path transforms.
template <typename T, int L>
constexpr int split_ry() {
  int ry = kLdsBudget / (seq_stride<L>() * (int)sizeof(cpx<T>));
  return ry >= 4 ? 4 : (ry >= 2 ? 2 : 1);
}

// ---------------------------------------------------------------- columns
// Streaming stores per column pass (measured on MI355X at 512^2 order 3 with
// whole-line column-inverse tiles: plain stores instead made the inner and
// outer column-inverse passes 0.32 -> 0.40 and 0.44 -> 0.55 ms, and the
// column-forward passes with their accumulations 0.08 ms per step slower).
#ifndef GLW_STREAM_COLINV
#define GLW_STREAM_COLINV 1
#endif
#ifndef GLW_STREAM_COLFWD
#define GLW_STREAM_COLFWD 1
#endif
// Streaming stores for outputs of kStreamBytes or more per launch, and only when
// a workgroup's row segment (CW columns) fills whole 128-byte lines: narrower
// tiles (CW = 2 at 768 and 1024 points) leave partial lines that the L2 must
// merge across workgroups (streamed, k_colinv<1024, 2> ran 2x slower).
template <typename T, int CW>
constexpr bool whole_lines() {
  return (CW * sizeof(cpx<T>)) % 128 == 0;
}

template <typename T, int L, int CW, int PLAN, int NTH>
void colinv_go(const ColInvArgs& a, int ntask, const void* tw, hipStream_t st) {
  ColInvArgs b = a;
  b.stream = GLW_STREAM_COLINV && whole_lines<T, CW>() && a.ostride % 8 == 0 &&
             (size_t)ntask * a.nfield * L * a.ostride * sizeof(cpx<T>) >= kStreamBytes;
  hipLaunchKernelGGL((k_colinv<T, L, CW, PLAN, NTH>), dim3((a.wx + CW - 1) / CW, a.nfield, ntask), dim3(NTH), 0, st,
                     b, (const cpx<T>*)tw);
}

template <typename T, int L, int CW, int PLAN, int NTH>
void colfwd_go(const ColFwdArgs& a, int ntask, const void* tw, hipStream_t st) {
  ColFwdArgs b = a;
  const int pitch = a.opitch ? a.opitch : a.owx;
  b.stream = GLW_STREAM_COLFWD && whole_lines<double, CW>() &&
             (size_t)ntask * a.nsrc * a.owy * pitch * sizeof(cpx<double>) >= kStreamBytes;
  hipLaunchKernelGGL((k_colfwd<T, L, CW, PLAN, NTH>), dim3((a.owx + CW - 1) / CW, a.nsrc, ntask), dim3(NTH), 0, st,
                     b, (const cpx<T>*)tw);
}

// Column tiles: 0 = widest tile, 256 threads; 1 = half tile, 256 threads;
// 2 = widest tile, 128 threads (fewer idle lanes in the first and last stages).
template <typename T, int L, int PLAN, class F>
void col_tile(int tile, F&& go) {
  constexpr int CW0 = col_width<T, L>();
  constexpr int CW1 = CW0 > 1 ? CW0 / 2 : 1;
  if (tile == 1) go(std::integral_constant<int, CW1>(), std::integral_constant<int, NT>());
  else if (tile == 2) go(std::integral_constant<int, CW0>(), std::integral_constant<int, NT / 2>());
  else go(std::integral_constant<int, CW0>(), std::integral_constant<int, NT>());
}

template <typename T, int L>
void colinv_len(const ColInvArgs& a, int ntask, const void* tw, Sched s, hipStream_t st) {
  constexpr int P1 = canonical_plan<L, 1>();
  auto run = [&](auto plan) {
    col_tile<T, L, decltype(plan)::value>(s.tile, [&](auto cw, auto nth) {
      colinv_go<T, L, decltype(cw)::value, decltype(plan)::value, decltype(nth)::value>(a, ntask, tw, st);
    });
  };
  if (s.plan == 0) run(std::integral_constant<int, 0>());
  else run(std::integral_constant<int, P1>());
}

template <typename T, int L>
void colfwd_len(const ColFwdArgs& a, int ntask, const void* tw, Sched s, hipStream_t st) {
  constexpr int P1 = canonical_plan<L, 1>();
  auto run = [&](auto plan) {
    col_tile<T, L, decltype(plan)::value>(s.tile, [&](auto cw, auto nth) {
      colfwd_go<T, L, decltype(cw)::value, decltype(plan)::value, decltype(nth)::value>(a, ntask, tw, st);
    });
  };
  if (s.plan == 0) run(std::integral_constant<int, 0>());
  else run(std::integral_constant<int, P1>());
}

template <typename T>
int launch_colinv(int gy, const ColInvArgs& a, int ntask, const void* tw, const Exec& x, hipStream_t st) {
  switch (gy) {
#define GLW_CASE(L)                                 \
  case L:                                           \
    colinv_len<T, L>(a, ntask, tw, x.s, st);        \
    break;
    GLW_LENGTHS(GLW_CASE)
#undef GLW_CASE
    default:
      return glw_report_error(GLW_EUNSUPPORTED, "column length " + std::to_string(gy) + " not instantiated");
  }
  GLW_HIP_CHECK(hipGetLastError());
  return GLW_OK;
}

template <typename T>
int launch_colfwd(int gy, const ColFwdArgs& a, int ntask, const void* tw, const Exec& x, hipStream_t st) {
  switch (gy) {
#define GLW_CASE(L)                                 \
  case L:                                           \
    colfwd_len<T, L>(a, ntask, tw, x.s, st);        \
    break;
    GLW_LENGTHS(GLW_CASE)
#undef GLW_CASE
    default:
      return glw_report_error(GLW_EUNSUPPORTED, "column length " + std::to_string(gy) + " not instantiated");
  }
  GLW_HIP_CHECK(hipGetLastError());
  return GLW_OK;
}

// ---------------------------------------------------------------- split rows
template <typename T, int L, int PLAN>
void rowinv_go(const RowArgs& a, int ntask, const void* tw, void* phys, hipStream_t st) {
  constexpr int RY = split_ry<T, L>();
  hipLaunchKernelGGL((k_rowinv<T, L, RY, PLAN>), dim3(a.gy / RY, a.nin, ntask), dim3(NT), 0, st, a, (cpx<T>*)phys,
                     (const cpx<T>*)tw);
}

template <typename T, int L, int PLAN>
void rowfwd_go(const RowArgs& a, int ntask, const void* tw, const void* phys, hipStream_t st) {
  constexpr int RY = split_ry<T, L>();
  hipLaunchKernelGGL((k_rowfwd<T, L, RY, PLAN>), dim3(a.gy / RY, a.nout, ntask), dim3(NT), 0, st, a,
                     (const cpx<T>*)phys, (const cpx<T>*)tw);
}

template <typename T>
int launch_rowinv(int gx, const RowArgs& a, int ntask, const void* tw, int plan, void* phys, hipStream_t st) {
  switch (gx) {
#define GLW_CASE(L)                                                                                      \
  case L:                                                                                                \
    if (a.gy % split_ry<T, L>()) return glw_report_error(GLW_EINVAL, "row count not divisible by tile"); \
    plan == 0 ? rowinv_go<T, L, 0>(a, ntask, tw, phys, st)                                              \
              : rowinv_go<T, L, canonical_plan<L, 1>()>(a, ntask, tw, phys, st);                        \
    break;
    GLW_LENGTHS(GLW_CASE)
#undef GLW_CASE
    default:
      return glw_report_error(GLW_EUNSUPPORTED, "row length " + std::to_string(gx) + " not instantiated");
  }
  GLW_HIP_CHECK(hipGetLastError());
  return GLW_OK;
}

template <typename T>
int launch_rowfwd(int gx, const RowArgs& a, int ntask, const void* tw, int plan, const void* phys, hipStream_t st) {
  switch (gx) {
#define GLW_CASE(L)                                                                                      \
  case L:                                                                                                \
    if (a.gy % split_ry<T, L>()) return glw_report_error(GLW_EINVAL, "row count not divisible by tile"); \
    plan == 0 ? rowfwd_go<T, L, 0>(a, ntask, tw, phys, st)                                              \
              : rowfwd_go<T, L, canonical_plan<L, 1>()>(a, ntask, tw, phys, st);                        \
    break;
    GLW_LENGTHS(GLW_CASE)
#undef GLW_CASE
    default:
      return glw_report_error(GLW_EUNSUPPORTED, "row length " + std::to_string(gx) + " not instantiated");
  }
  GLW_HIP_CHECK(hipGetLastError());
  return GLW_OK;
}

// ---------------------------------------------------------------- fused rows
template <typename T, int NIN, int NOUT, class Op, int L, int RY, int PLAN>
int row_go(const RowArgs& a, int ntask, const void* tw, hipStream_t st) {
  if (a.gy % RY) return glw_report_error(GLW_EINVAL, "row count not divisible by the row tile");
  hipLaunchKernelGGL((k_row<T, L, NIN, NOUT, RY, Op, PLAN>), dim3(a.gy / RY, ntask), dim3(NT), 0, st, a,
                     (const cpx<T>*)tw);
  return GLW_OK;
}

template <typename T, int NIN, int NOUT, class Op>
int row_split(int gx, const RowArgs& a, int ntask, const void* tw, const Exec& x, hipStream_t st) {
  const size_t npts = (size_t)a.gy * gx;
  const size_t bin = (size_t)ntask * NIN * npts * sizeof(cpx<T>);
  const size_t bout = (size_t)ntask * NOUT * npts * sizeof(cpx<T>);
  if (!x.scratch) return glw_report_error(GLW_EINVAL, "split row path needs a scratch provider");
  void* ws = nullptr;
  int rc = x.scratch(x.ctx, bin + bout, &ws);
  if (rc) return rc;
  char* pin = (char*)ws;
  char* pout = pin + bin;
  if ((rc = launch_rowinv<T>(gx, a, ntask, tw, x.s.plan, pin, st))) return rc;
  hipLaunchKernelGGL((k_pointop<T, NIN, NOUT, Op>), dim3((unsigned)((npts + NT - 1) / NT), ntask), dim3(NT), 0, st,
                     (const cpx<T>*)pin, (cpx<T>*)pout, (int)npts);
  GLW_HIP_CHECK(hipGetLastError());
  return launch_rowfwd<T>(gx, a, ntask, tw, x.s.plan, pout, st);
}

template <typename T, int NIN, int NOUT, class Op, int L>
int row_len(const RowArgs& a, int ntask, const void* tw, const Exec& x, hipStream_t st) {
  static_assert(NOUT <= NIN, "the fused row kernel reuses the input tile for the outputs");
  constexpr int RY0 = row_ry<T, NIN, L, Op>();
  constexpr int P1 = canonical_plan<L, 1>();
  int tile = x.s.tile;
  if (tile == 0 && !row_fits<T, NIN, L, RY0>()) tile = 1;
  if (tile == 1 && !row_fits<T, NIN, L, 1>()) tile = 2;
  if (tile == 0) {
    if constexpr (row_fits<T, NIN, L, RY0>())
      return x.s.plan == 0 ? row_go<T, NIN, NOUT, Op, L, RY0, 0>(a, ntask, tw, st)
                           : row_go<T, NIN, NOUT, Op, L, RY0, P1>(a, ntask, tw, st);
  } else if (tile == 1) {
    if constexpr (row_fits<T, NIN, L, 1>())
      return x.s.plan == 0 ? row_go<T, NIN, NOUT, Op, L, 1, 0>(a, ntask, tw, st)
                           : row_go<T, NIN, NOUT, Op, L, 1, P1>(a, ntask, tw, st);
  }
  return row_split<T, NIN, NOUT, Op>(L, a, ntask, tw, x, st);
}

template <typename T, int NIN, int NOUT, class Op>
int launch_row(int gx, const RowArgs& a, int ntask, const void* tw, const Exec& x, hipStream_t st) {
  int rc;
  switch (gx) {
#define GLW_CASE(L)                                               \
  case L:                                                         \
    rc = row_len<T, NIN, NOUT, Op, L>(a, ntask, tw, x, st);       \
    break;
    GLW_LENGTHS(GLW_CASE)
#undef GLW_CASE
    default:
      return glw_report_error(GLW_EUNSUPPORTED, "row length " + std::to_string(gx) + " not instantiated");
  }
  if (rc) return rc;
  GLW_HIP_CHECK(hipGetLastError());
  return GLW_OK;
}

template <typename T, int NIN, int NG, class Op>
int launch_row_adjoint(int gx, const RowArgs& fwd, const RowArgs& gin, const RowArgs& gout, int ntask,
                       const void* tw, const Exec& x, hipStream_t st) {
  if (fwd.nin != NIN || gin.nin != NG || gout.nout != NIN || !gout.out_per_field)
    return glw_report_error(GLW_EINVAL, "row adjoint: field counts do not match the operator");
  const size_t npts = (size_t)fwd.gy * gx;
  const size_t fb = (size_t)ntask * npts * sizeof(cpx<T>);
  if (!x.scratch) return glw_report_error(GLW_EINVAL, "row adjoint needs a scratch provider");
  void* ws = nullptr;
  int rc = x.scratch(x.ctx, fb * (2 * NIN + NG), &ws);
  if (rc) return rc;
  char* pv = (char*)ws;
  char* pg = pv + fb * NIN;
  char* pgv = pg + fb * NG;
  if ((rc = launch_rowinv<T>(gx, fwd, ntask, tw, x.s.plan, pv, st))) return rc;
  if ((rc = launch_rowinv<T>(gx, gin, ntask, tw, x.s.plan, pg, st))) return rc;
  hipLaunchKernelGGL((k_pointadj<T, NIN, NG, Op>), dim3((unsigned)((npts + NT - 1) / NT), ntask), dim3(NT), 0, st,
                     (const cpx<T>*)pv, (const cpx<T>*)pg, (cpx<T>*)pgv, (int)npts);
  GLW_HIP_CHECK(hipGetLastError());
  return launch_rowfwd<T>(gx, gout, ntask, tw, x.s.plan, pgv, st);
}

}  // namespace glw
