// Row passes: launcher instantiations for one set of pointwise operators,
// chosen by GLW_ROW_SET at compile time (the build compiles this file once
// per set, in parallel).
#include "glw_launch_impl.hpp"
#include "glw_ops.hpp"

#ifndef GLW_ROW_SET
#error "GLW_ROW_SET must select an operator set"
#endif

#if GLW_ROW_SET == 0
#define GLW_ROW_OPS(X) X(12, 4, InnerOp)
#define GLW_ADJ_OPS(X) X(12, 4, InnerOp)
#elif GLW_ROW_SET == 1
#define GLW_ROW_OPS(X) X(21, 4, OuterOp)
#define GLW_ADJ_OPS(X) X(21, 2, OuterOp)
#elif GLW_ROW_SET == 2
#define GLW_ROW_OPS(X) X(8, 2, Eta22Op) X(2, 1, ContactOp)
#define GLW_ADJ_OPS(X) X(8, 2, Eta22Op)
#elif GLW_ROW_SET == 3
#define GLW_ROW_OPS(X) X(16, 8, Psi22Op)
#elif GLW_ROW_SET == 4
#define GLW_ROW_OPS(X) X(16, 6, Pair44Op)
#elif GLW_ROW_SET == 5
#define GLW_ROW_OPS(X) X(31, 4, Third44Op)
#elif GLW_ROW_SET == 6
#define GLW_ROW_OPS(X) X(45, 2, Quartic44Op)
#elif GLW_ROW_SET == 7
#define GLW_ROW_OPS(X) X(45, 2, Quartic31Op)
#else
#error "unknown GLW_ROW_SET"
#endif

namespace glw {

// The split-path transforms are instantiated once, in glw_launch_cols.hip.
extern template int launch_rowinv<double>(int, const RowArgs&, int, const void*, int, void*, hipStream_t);
extern template int launch_rowfwd<double>(int, const RowArgs&, int, const void*, int, const void*, hipStream_t);

// GL plans are FP64 only (FP32 transforms fail the 1e-5 bound: glw_plan_create refuses them).
#define GLW_INST(NIN, NOUT, OP) \
  template int launch_row<double, NIN, NOUT, OP>(int, const RowArgs&, int, const void*, const Exec&, hipStream_t);
GLW_ROW_OPS(GLW_INST)
#undef GLW_INST

// adjoints (FP64 only)
#ifdef GLW_ADJ_OPS
#define GLW_INST(NIN, NG, OP)                                                                               \
  template int launch_row_adjoint<double, NIN, NG, OP>(int, const RowArgs&, const RowArgs&, const RowArgs&, int, \
                                                       const void*, const Exec&, hipStream_t);
GLW_ADJ_OPS(GLW_INST)
#undef GLW_INST
#endif

}  // namespace glw
