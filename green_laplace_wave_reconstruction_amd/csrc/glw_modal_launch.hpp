// Launchers of the modal-path kernels (glw_modal_launch.hip).  Each returns
// 0, 1 for an unsupported length or launch shape, 2 for a launch error.
#pragma once

#include <hip/hip_runtime.h>

#include "glw_modal.hpp"

// Time-grid and modal-grid lengths with kernels (2^a 3^b, multiples of 16).
#define GLW_MODAL_T_LENGTHS(X) \
  X(16) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048) X(3072) X(4096)
#define GLW_MODAL_Q_LENGTHS(X)                                                                          \
  X(16) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048) \
  X(3072) X(4096) X(6144) X(8192)
// Modal lengths beyond one workgroup's LDS: R0 = 2, 4, 8 sub-transforms of 6144 or 8192 points.
#define GLW_MODAL_Q_SPLIT_LENGTHS(X) X(12288) X(16384) X(24576) X(32768) X(49152) X(65536)

namespace glw {
namespace modal {

int launch_t_order2(const Args& a, hipStream_t st);
int launch_t_order3(const Args& a, hipStream_t st);
int launch_t_eta20(const Args& a, hipStream_t st);
// r0 > 1: the split transform (a.q_len = r0 x an LDS length), one workgroup
// per column and sub-transform.
int launch_q_order2(const Args& a, int ncol, int r0, hipStream_t st);
// order 3 after launch_q_order2: the five pair fields from a.pair_scratch
int launch_q_pairs(const Args& a, int ncol, int r0, hipStream_t st);
int launch_q_order3(const Args& a, int ncol, int r0, hipStream_t st);
int launch_q_eta20(const Args& a, int ncol, hipStream_t st);

}  // namespace modal
}  // namespace glw
