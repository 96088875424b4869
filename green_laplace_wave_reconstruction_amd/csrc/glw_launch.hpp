// Launchers of the FFT-pass kernels, with a per-stage schedule.
//
// A schedule picks, for one pipeline stage, the radix sequence of the LDS
// FFTs (plan 0: two passes of composite in-register radices; plan 1: greedy
// radix 8/4/2/3) and the tile:
//   column stages  tile 0: widest column tile within the LDS budget,
//                  tile 1: half of it, tile 2: widest tile on 128 threads;
//   row stages     tile 0: fused row kernel with the most rows per
//                          workgroup that fit, tile 1: fused with one row,
//                  tile 2: split path (x-IFFT to HBM, pointwise products,
//                          x-FFT), which also serves tiles that do not fit LDS.
// Every schedule computes the same transforms; results agree to rounding.
// The launchers are defined in glw_launch_impl.hpp and instantiated in the
// glw_launch_*.hip translation units, which are compiled in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "glw_kernels.hpp"

namespace glw {

enum StageKind : int { STAGE_COL = 0, STAGE_ROW = 1, STAGE_OTHER = 2 };

struct Sched {
  int plan = 0;
  int tile = 0;
};

// Returns (in *ptr) device scratch of at least `bytes`, or an error code.
typedef int (*ScratchFn)(void* ctx, size_t bytes, void** ptr);

struct Exec {
  Sched s;
  ScratchFn scratch = nullptr;
  void* ctx = nullptr;
};

#define GLW_LENGTHS(X) \
  X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(144) X(192) X(256) X(288) X(384) X(512) X(576) X(768) X(1024) \
  X(1152) X(1536) X(2048)

bool length_supported(int n);

template <typename T>
int launch_colinv(int gy, const ColInvArgs& a, int ntask, const void* tw, const Exec& x, hipStream_t st);
template <typename T>
int launch_colfwd(int gy, const ColFwdArgs& a, int ntask, const void* tw, const Exec& x, hipStream_t st);
int launch_synth(int gx, const SynthArgs& a, int nseq, const void* tw, hipStream_t st);

// Split row path, first and last pass; `phys` holds [ntask][nin or nout][gy][gx].
template <typename T>
int launch_rowinv(int gx, const RowArgs& a, int ntask, const void* tw, int plan, void* phys, hipStream_t st);
template <typename T>
int launch_rowfwd(int gx, const RowArgs& a, int ntask, const void* tw, int plan, const void* phys, hipStream_t st);

template <typename T, int NIN, int NOUT, class Op>
int launch_row(int gx, const RowArgs& a, int ntask, const void* tw, const Exec& x, hipStream_t st);

// Adjoint of launch_row<T, NIN, *, Op> for the first NG outputs: `fwd` holds
// the forward inputs (as given to launch_row), `gin` the output-box
// gradients (nin = NG fields, boxes in fmx0/fwx), `gout` receives the input
// gradients (nout = NIN, out_per_field boxes).  Runs the split path.
template <typename T, int NIN, int NG, class Op>
int launch_row_adjoint(int gx, const RowArgs& fwd, const RowArgs& gin, const RowArgs& gout, int ntask,
                       const void* tw, const Exec& x, hipStream_t st);

}  // namespace glw
