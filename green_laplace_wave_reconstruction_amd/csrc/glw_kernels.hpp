// Generic kernels of the alias-free pseudo-spectral GL pipeline.
//
// Spectra live on "boxes": rectangles of signed FFT mode numbers
// (mx0..mx0+wx-1, my0..my0+wy-1) stored [wy][wx] row-major in FP64.  Physical
// fields live on a reduced working grid gy x gx that is at least as wide as
// every product's spectral support, so that bin (mx, my) of the original
// ny x nx grid maps one-to-one onto (my mod gy, mx mod gx) and the
// pseudo-spectral products are exact (no aliasing onto kept bins).
//
// Transform split: a 2-D inverse transform is a pruned y-pass over the box
// columns (k_colinv) followed by an x-pass per physical row; the x-pass, the
// pointwise products and the forward x-pass of the products are fused in one
// row kernel; the forward y-pass is pruned to the output box rows
// (k_colfwd), which also applies the support mask and records the support
// maximum |x| for the roundoff floor.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "glw_fft.hpp"

namespace glw {

constexpr int kMaxFields = 48;
constexpr int NT = 256;

// Workgroups are dealt to the 8 XCDs round robin, and each XCD has its own
// L2.  xcd_block renumbers dispatch slot w of an n-workgroup grid so that XCD
// k runs one consecutive range of logical workgroups (range sizes differ by
// at most one; a bijection on [0, n)).  Neighbouring logical workgroups --
// the column tiles and fields of one task, which read the same spectrum lines
// -- then share one L2 instead of each XCD fetching those lines again
// (order 3 at 512^2 on MI355X: outer column-inverse reads 406 -> 145 MB per
// dispatch, inner column-inverse 0.474 -> 0.432 ms).
#ifndef GLW_XCD_COL
#define GLW_XCD_COL 1
#endif
__device__ __forceinline__ int xcd_block(int w, int n) {
  const int q = n >> 3, r = n & 7, k = w & 7;
  return k * q + (k < r ? k : r) + (w >> 3);
}

// LDS layout of the column kernels' tiles.  Their lanes run across the CW
// columns (sequences) of a tile at one element, so the 16-element padding of
// lds_pad buys nothing there; what decides bank conflicts is the sequence
// stride S.  Measured on MI355X with ds_read_b128 / ds_write_b128 over 8
// sequences x consecutive elements: reads are served 32 lanes at a time over
// a 512-byte bank period, writes 8 lanes at a time over 128 bytes.  Odd S
// keeps the writes conflict-free; among odd S, S = 23 or 25 (mod 32) halves
// the read conflicts of other residues (1.55x the conflict-free time instead
// of 2.5x); no stride makes both conflict-free.  The column passes are
// store-bound, so this moved them by 1 % at most.  GLW_COL_SSTR_RES < 0 keeps
// the padded row layout (seq_stride).
#ifndef GLW_COL_SSTR_RES
#define GLW_COL_SSTR_RES 23
#endif
constexpr bool kColPad = GLW_COL_SSTR_RES < 0;
template <int L>
__host__ __device__ constexpr int col_stride() {
  if (GLW_COL_SSTR_RES < 0) return seq_stride<L>();
  int s = L;
  while (s % 32 != GLW_COL_SSTR_RES) ++s;
  return s;
}

// Columns the column-inverse pass writes per row: the box width w rounded up
// to whole 128-byte lines within the row pitch, when its tiles span whole
// lines and the last tile is at least half full (the padding columns get the
// loader's zeros; no consumer reads past a field's width).  A line written in
// part costs HBM time out of proportion: the inner column-inverse pass at 512^2
// (45 of 48 columns per row) took 0.42 ms with the box columns only and
// 0.33 ms with whole lines on MI355X.  Measured and not taken: a last tile
// holding 1 of 8 columns writing all 8 (89 -> 96 columns: 5 % slower), whole
// lines for the row kernels' output rows (89 -> 96 and 133 -> 136: 3-4 %
// slower, and the column-forward passes that read them 2-3 % slower) and for
// the column-forward spectra.  GLW_FULL_LINES=0 writes the box columns only.
#ifndef GLW_FULL_LINES
#define GLW_FULL_LINES 1
#endif
template <typename T, int CW>
__host__ __device__ inline int fill_width_cols(int w, int pitch) {
  constexpr int kLine = 128 / (int)sizeof(cpx<T>);
  const int part = w % kLine;
  if (!GLW_FULL_LINES || CW < kLine || (part != 0 && 2 * part < kLine)) return w;
  const int wl = (w + kLine - 1) / kLine * kLine;
  return wl < pitch ? wl : (w > pitch ? w : pitch);
}

// FFT stages read twiddles from the global table (a few KB, L1/L2-resident)
// rather than an LDS copy: the LDS goes to data tiles (more workgroups per CU).
// Measured on MI355X (GL order 3, 512^2, 64 instants): LDS copies in the
// column kernels 4.856 ms per step, in the row kernels 5.181 ms, neither 4.818 ms.
#ifndef GLW_TW_LDS_COL
#define GLW_TW_LDS_COL 0
#endif
#ifndef GLW_TW_LDS_ROW
#define GLW_TW_LDS_ROW 0
#endif
constexpr bool kTwiddlesInLdsCol = GLW_TW_LDS_COL, kTwiddlesInLdsRow = GLW_TW_LDS_ROW;
// Twiddle prefetch one stage ahead (global tables only) per kernel class, and
// whether the last transform of a kernel skips its final read-before-write
// barrier (the kernel ends there).
#ifndef GLW_PF_COL
#define GLW_PF_COL 1
#endif
#ifndef GLW_PF_ROW
#define GLW_PF_ROW 1
#endif
#ifndef GLW_PF_SYNTH
#define GLW_PF_SYNTH 0
#endif
#ifndef GLW_FINAL_NOBAR
#define GLW_FINAL_NOBAR 1
#endif
constexpr bool kPfCol = GLW_PF_COL && !kTwiddlesInLdsCol, kPfRow = GLW_PF_ROW && !kTwiddlesInLdsRow;
constexpr bool kPfSynth = GLW_PF_SYNTH && !kTwiddlesInLdsRow, kFinal = GLW_FINAL_NOBAR;

// Spectral multiplier codes (functions of qx, qy, nu, 1/nu, q^2, A on a bin).
enum MulCode : int {
  M_ONE = 0,
  M_I_QX = 1,       // i qx
  M_I_QY = 2,       // i qy
  M_QX_NU = 3,      // qx / nu
  M_QY_NU = 4,      // qy / nu
  M_MI_NU = 5,      // -i nu
  M_NU_QX = 6,      // nu qx
  M_NU_QY = 7,      // nu qy
  M_MI_Q2_NU = 8,   // -i q^2 / nu
  M_MI_Q2NU = 9,    // -i q^2 nu
  M_MA = 10,        // -A
  M_MQ2 = 11,       // -q^2
  M_NU = 12,        // nu
  M_NU2 = 13,       // nu^2
  M_NU3 = 14,       // nu^3
  M_QX = 15,        // qx
  M_QY = 16,        // qy
  M_Q2_NU = 17,     // q^2 / nu
  M_Q2 = 18,        // q^2
  M_A = 19,         // A
  M_MI_INU = 20,    // -i / nu  (0 where nu == 0)
  M_NU4 = 21,       // nu^4
  M_Q2NU = 22,      // q^2 nu
  M_NU2_QX = 23,    // nu^2 qx
  M_NU2_QY = 24,    // nu^2 qy
  M_MI_NU2_QX = 25, // -i nu^2 qx
  M_MI_NU2_QY = 26, // -i nu^2 qy
  M_MQ2A = 27,      // -q^2 A
  M_I_ANU = 28,     // i A nu
  M_QXQ2_NU = 29,   // qx q^2 / nu   (i qx q^2 applied to phi = -i s / nu)
  M_QYQ2_NU = 30,   // qy q^2 / nu
  M_MI_Q4_NU = 31,  // -i q^4 / nu   (q^4 applied to phi = -i s / nu)
  M_I_QXA = 32,     // i qx A
  M_I_QYA = 33,     // i qy A
  M_Q2A = 34,       // q^2 A
};
constexpr int kNumMulCodes = 35;

struct BoxTables {
  const double* qx;   // [wy][wx]
  const double* qy;
  const double* nu;
  const double* inu;  // 1 / max(nu, tiny)
  const double* q2;
  const double* a;    // q tanh q
  // [kNumMulCodes][wy*wx]: every multiplier code evaluated on the box
  // (apply_mul(code, 1)), so the column loads read one value instead of
  // branching on the code per element
  const cpx<double>* mul;
};

__device__ __forceinline__ cpx<double> apply_mul(int code, cpx<double> s, const BoxTables& t, int b) {
  double re = 0.0, im = 0.0;  // multiplier m = re + i im
  switch (code) {
    case M_ONE: re = 1.0; break;
    case M_I_QX: im = t.qx[b]; break;
    case M_I_QY: im = t.qy[b]; break;
    case M_QX_NU: re = t.qx[b] * t.inu[b]; break;
    case M_QY_NU: re = t.qy[b] * t.inu[b]; break;
    case M_MI_NU: im = -t.nu[b]; break;
    case M_NU_QX: re = t.nu[b] * t.qx[b]; break;
    case M_NU_QY: re = t.nu[b] * t.qy[b]; break;
    case M_MI_Q2_NU: im = -t.q2[b] * t.inu[b]; break;
    case M_MI_Q2NU: im = -t.q2[b] * t.nu[b]; break;
    case M_MA: re = -t.a[b]; break;
    case M_MQ2: re = -t.q2[b]; break;
    case M_NU: re = t.nu[b]; break;
    case M_NU2: re = t.nu[b] * t.nu[b]; break;
    case M_NU3: re = t.nu[b] * t.nu[b] * t.nu[b]; break;
    case M_QX: re = t.qx[b]; break;
    case M_QY: re = t.qy[b]; break;
    case M_Q2_NU: re = t.q2[b] * t.inu[b]; break;
    case M_Q2: re = t.q2[b]; break;
    case M_A: re = t.a[b]; break;
    case M_MI_INU: im = t.nu[b] > 0.0 ? -1.0 / t.nu[b] : 0.0; break;
    case M_NU4: re = t.nu[b] * t.nu[b] * t.nu[b] * t.nu[b]; break;
    case M_Q2NU: re = t.q2[b] * t.nu[b]; break;
    case M_NU2_QX: re = t.nu[b] * t.nu[b] * t.qx[b]; break;
    case M_NU2_QY: re = t.nu[b] * t.nu[b] * t.qy[b]; break;
    case M_MI_NU2_QX: im = -t.nu[b] * t.nu[b] * t.qx[b]; break;
    case M_MI_NU2_QY: im = -t.nu[b] * t.nu[b] * t.qy[b]; break;
    case M_MQ2A: re = -t.q2[b] * t.a[b]; break;
    case M_I_ANU: im = t.a[b] * t.nu[b]; break;
    case M_QXQ2_NU: re = t.nu[b] > 0.0 ? t.qx[b] * t.q2[b] / t.nu[b] : 0.0; break;
    case M_QYQ2_NU: re = t.nu[b] > 0.0 ? t.qy[b] * t.q2[b] / t.nu[b] : 0.0; break;
    case M_MI_Q4_NU: im = t.nu[b] > 0.0 ? -t.q2[b] * t.q2[b] / t.nu[b] : 0.0; break;
    case M_I_QXA: im = t.qx[b] * t.a[b]; break;
    case M_I_QYA: im = t.qy[b] * t.a[b]; break;
    case M_Q2A: re = t.q2[b] * t.a[b]; break;
    default: break;
  }
  return {re * s.x - im * s.y, re * s.y + im * s.x};
}

// Streaming store (global_store ... nt): the column passes write each
// intermediate once, in whole 128-byte lines, for the next pass to read once.
// Measured on MI355X (GL order 3, 512^2, 64 instants): the two column-inverse
// stages 0.575 -> 0.444 ms and 0.761 -> 0.607 ms.  The same hint on the row
// pass outputs changes nothing, and on box outputs whose rows are not
// line-aligned it is 1.6x slower (k_colfwd uses it only at a padded pitch).
// The launchers set it by output size (kStreamBytes).
// Column-pass outputs of at least this many bytes per launch use streaming
// stores; smaller ones stay cacheable for their consumer (an eta44 launch at
// 256^2, or the small sea states of a batch, slow down when their consumer
// misses the cache).
constexpr size_t kStreamBytes = size_t(64) << 20;
template <typename T>
__device__ __forceinline__ void store_stream(cpx<T>* p, cpx<T> v) {
  using V = T __attribute__((ext_vector_type(2)));
  const V w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<V*>(p));
}

__device__ __forceinline__ int wrap(int m, int n) {
  int r = m % n;
  return r < 0 ? r + n : r;
}

// ----------------------------------------------------------------------------
// Pruned inverse y-pass.
// ----------------------------------------------------------------------------
struct ColInvArgs {
  const cpx<double>* spec;  // [nspecset][nsrc][wy*wx]
  int nsrc;
  int sdiv;                 // spectrum set = task / sdiv
  int mx0, my0, wx, wy;     // input box
  BoxTables tab;
  const double* damp;       // [ndamp][wy*wx] or null; row = (task % dmod) * dmul + dadd
  int dmod, dmul, dadd;
  const uint64_t* fmax;     // bit patterns of support max |x| (or null), [nspecset][fmax_stride]
  int fmax_stride;          // 1: one maximum per spectrum set; nsrc: one per source
  int floor_src_mask;       // bit s set: apply the roundoff floor to source s
  double floor_level;
  int nfield;
  int fsrc[kMaxFields];
  int fcode[kMaxFields];
  double fscale[kMaxFields];
  void* out;                // cpx<T> [task][out_nfield][gy][ostride]
  int out_nfield, out_foff, ostride;
  bool stream = false;      // streaming stores (set by the launcher for large outputs)
};

// Box column (damped, floored, multiplied, scaled) as the first stage's input;
// element e of column c is box row wrap(e - my0, GY).
template <typename T>
struct ColInvLoad {
  static constexpr bool SEQ_FAST = true;  // neighbouring columns are contiguous
  const cpx<double>* spec;
  const double* damp;
  const cpx<double>* mul;  // this field's row of the multiplier table
  double level, scale;
  int col0, wx, wy, my0, gy;
  __device__ __forceinline__ cpx<T> value(int b) const {
    cpx<double> s = spec[b];
    const cpx<double> m = mul[b];
    const double d = damp ? damp[b] : 1.0;
    if (level >= 0.0 && hypot(s.x, s.y) < level) s = {0.0, 0.0};
    if (damp) s = {s.x * d, s.y * d};
    s = {m.x * s.x - m.y * s.y, m.x * s.y + m.y * s.x};  // the product apply_mul forms
    return {T(s.x * scale), T(s.y * scale)};
  }
  __device__ __forceinline__ cpx<T> operator()(int c, int e) const {
    const int col = col0 + c;
    const int r = wrap(e - my0, gy);
    if (col >= wx || r >= wy) return {T(0), T(0)};
    return value(r * wx + col);
  }
  // elements j + q STEP (q < R, R STEP = gy) of column c: one wrap per butterfly
  static constexpr bool BATCH = true;
  template <int R, int STEP>
  __device__ __forceinline__ void load_batch(int c, int j, cpx<T>* v) const {
    const int col = col0 + c;
    const int r0 = wrap(j - my0, gy);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int r = r0 + q * STEP;
      if (r >= gy) r -= gy;
      v[q] = (col < wx && r < wy) ? value(r * wx + col) : cpx<T>{T(0), T(0)};
    }
  }
};

template <typename T>
struct ColInvStore {
  static constexpr bool SEQ_FAST = true;
  cpx<T>* out;
  int ostride, col0;
  int wx;  // columns written: fill_width_cols of the box (the loader's zeros past it)
  bool stream;
  __device__ __forceinline__ void put(cpx<T>* p, cpx<T> v) {
    if (stream) store_stream(p, v);
    else *p = v;
  }
  __device__ __forceinline__ void operator()(int c, int e, cpx<T> v) {
    if (col0 + c < wx) put(out + (size_t)e * ostride + col0 + c, v);
  }
  static constexpr bool BATCH = true;
  template <int R, int STEP>
  __device__ __forceinline__ void store_batch(int c, int j, const cpx<T>* v) {
    if (col0 + c >= wx) return;
    cpx<T>* p = out + (size_t)j * ostride + col0 + c;
#pragma unroll
    for (int q = 0; q < R; ++q) put(p + (size_t)q * STEP * ostride, v[q]);
  }
};

template <typename T, int GY, int CW, int PLAN = 0, int NTH = NT>
__global__ __launch_bounds__(NTH) void k_colinv(ColInvArgs a, const cpx<T>* __restrict__ tw) {
  __shared__ cpx<T> buf[CW * col_stride<GY>()];
  __shared__ cpx<T> twl[kTwiddlesInLdsCol ? GY : 1];
  const cpx<T>* twp = tw;
  if constexpr (kTwiddlesInLdsCol) {  // read from the second stage on, after a barrier
    stage_twiddles<T, GY, NTH>(twl, tw);
    twp = twl;
  }
  int tile = blockIdx.x, f = blockIdx.y, task = blockIdx.z;
  if constexpr (GLW_XCD_COL) {
    int w = xcd_block(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                      gridDim.x * gridDim.y * gridDim.z);
    tile = w % gridDim.x;
    w /= gridDim.x;
    f = w % gridDim.y;
    task = w / gridDim.y;
  }
  const int set = task / a.sdiv;
  const int src = a.fsrc[f];
  const int box = a.wx * a.wy;
  ColInvLoad<T> ld;
  ld.spec = a.spec + ((size_t)set * a.nsrc + src) * box;
  ld.damp = a.damp ? a.damp + (size_t)((task % a.dmod) * a.dmul + a.dadd) * box : nullptr;
  ld.level = -1.0;
  if (a.fmax && ((a.floor_src_mask >> src) & 1)) {
    const size_t k = (size_t)set * a.fmax_stride + (a.fmax_stride == 1 ? 0 : src);
    ld.level = a.floor_level * __longlong_as_double((long long)a.fmax[k]);
  }
  ld.scale = a.fscale[f];
  ld.mul = a.tab.mul + (size_t)a.fcode[f] * box;
  ld.col0 = tile * CW;
  ld.wx = a.wx;
  ld.wy = a.wy;
  ld.my0 = a.my0;
  ld.gy = GY;
  ColInvStore<T> st;
  st.out = reinterpret_cast<cpx<T>*>(a.out) + ((size_t)task * a.out_nfield + a.out_foff + f) * GY * a.ostride;
  st.ostride = a.ostride;
  st.col0 = ld.col0;
  st.wx = fill_width_cols<T, CW>(a.wx, a.ostride);
  st.stream = a.stream;
  fft_io_pf<T, GY, CW, col_stride<GY>(), NTH, +1, PLAN, kPfCol, kFinal, kColPad>(buf, twp, ld, st);
}

// ----------------------------------------------------------------------------
// Pruned forward y-pass, support mask, scale and support max |x|.
// ----------------------------------------------------------------------------
struct ColFwdArgs {
  const void* in;           // cpx<T> [task][in_nfield][gy][istride], sources from field in_foff
  int nsrc, istride;
  int in_nfield = 0;        // fields per task in `in` (0: nsrc)
  int in_foff = 0;
  int omx0, owx, omy0, owy; // output box (columns = the row kernel's output box)
  const uint8_t* mask;      // [owy*owx]
  double scale;
  cpx<double>* out;         // [task][nsrc][owy][opitch]
  uint64_t* fmax;           // [task][nsrc] or null
  // Row pitch of the output boxes (0: owx).  A multiple of 8 makes every
  // 8-column tile of a box row one whole 128-byte line; at pitch owx the rows
  // are not line-aligned and two workgroups write parts of one line.
  int opitch = 0;
  // Streaming stores for line-aligned outputs too large to stay in the
  // last-level cache for their consumer (set by the launcher by size).
  bool stream = false;
};

template <int NTH = NT>
__device__ __forceinline__ void block_max_atomic(double v, uint64_t* dst) {
  // non-negative doubles order like their bit patterns
  __shared__ double red[NTH / 64];
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int i = 1; i < NTH / 64; ++i) m = fmax(m, red[i]);
    if (m > 0.0) atomicMax((unsigned long long*)dst, (unsigned long long)__double_as_longlong(m));
  }
}

template <typename T>
struct ColFwdLoad {
  static constexpr bool SEQ_FAST = true;
  const cpx<T>* in;
  int istride, col0, owx;
  __device__ __forceinline__ cpx<T> operator()(int c, int e) const {
    return (col0 + c < owx) ? in[(size_t)e * istride + col0 + c] : cpx<T>{T(0), T(0)};
  }
  static constexpr bool BATCH = true;
  template <int R, int STEP>
  __device__ __forceinline__ void load_batch(int c, int j, cpx<T>* v) const {
    const bool in_box = col0 + c < owx;
    const cpx<T>* p = in + (size_t)j * istride + col0 + c;
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = in_box ? p[(size_t)q * STEP * istride] : cpx<T>{T(0), T(0)};
  }
};

// Keeps the output box rows (element e is box row wrap(e - omy0, GY)), masks,
// scales and records the largest kept |x|.
template <typename T>
struct ColFwdStore {
  static constexpr bool SEQ_FAST = true;
  cpx<double>* out;
  const uint8_t* mask;
  double scale;
  double vmax;
  int col0, owx, owy, omy0, gy;
  int opitch;               // output row pitch
  bool stream;              // line-aligned rows: streaming stores
  __device__ __forceinline__ void put(int r, int col, cpx<T> v) {
    cpx<double> d = {double(v.x) * scale, double(v.y) * scale};
    if (!mask[r * owx + col]) d = {0.0, 0.0};
    else vmax = fmax(vmax, hypot(d.x, d.y));
    cpx<double>* p = out + (size_t)r * opitch + col;
    if (stream) store_stream(p, d);
    else *p = d;
  }
  __device__ __forceinline__ void operator()(int c, int e, cpx<T> v) {
    const int col = col0 + c;
    const int r = wrap(e - omy0, gy);
    if (col >= owx || r >= owy) return;
    put(r, col, v);
  }
  static constexpr bool BATCH = true;
  template <int R, int STEP>
  __device__ __forceinline__ void store_batch(int c, int j, const cpx<T>* v) {
    const int col = col0 + c;
    if (col >= owx) return;
    const int r0 = wrap(j - omy0, gy);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int r = r0 + q * STEP;
      if (r >= gy) r -= gy;
      if (r < owy) put(r, col, v[q]);
    }
  }
};

template <typename T, int GY, int CW, int PLAN = 0, int NTH = NT>
__global__ __launch_bounds__(NTH) void k_colfwd(ColFwdArgs a, const cpx<T>* __restrict__ tw) {
  __shared__ cpx<T> buf[CW * col_stride<GY>()];
  __shared__ cpx<T> twl[kTwiddlesInLdsCol ? GY : 1];
  const cpx<T>* twp = tw;
  if constexpr (kTwiddlesInLdsCol) {
    stage_twiddles<T, GY, NTH>(twl, tw);
    twp = twl;
  }
  const int s = blockIdx.y;
  const int task = blockIdx.z;
  const int nf = a.in_nfield ? a.in_nfield : a.nsrc;
  ColFwdLoad<T> ld;
  ld.in = reinterpret_cast<const cpx<T>*>(a.in) + ((size_t)task * nf + a.in_foff + s) * GY * a.istride;
  ld.istride = a.istride;
  ld.col0 = blockIdx.x * CW;
  ld.owx = a.owx;
  ColFwdStore<T> st;
  const int pitch = a.opitch ? a.opitch : a.owx;
  st.out = a.out + ((size_t)task * a.nsrc + s) * a.owy * pitch;
  st.opitch = pitch;
  st.stream = a.stream && pitch % 8 == 0;
  st.mask = a.mask;
  st.scale = a.scale;
  st.vmax = 0.0;
  st.col0 = ld.col0;
  st.owx = a.owx;
  st.owy = a.owy;
  st.omy0 = a.omy0;
  st.gy = GY;
  fft_io_pf<T, GY, CW, col_stride<GY>(), NTH, -1, PLAN, kPfCol, kFinal, kColPad>(buf, twp, ld, st);
  if (a.fmax) block_max_atomic<NTH>(st.vmax, a.fmax + (size_t)task * a.nsrc + s);
}

// ----------------------------------------------------------------------------
// Fused row kernel: x-IFFT of the input fields, pointwise products, x-FFT of
// the products, pruned to the output box columns.
// ----------------------------------------------------------------------------
struct RowArgs {
  const void* in;           // cpx<T> [task][nin][gy][istride]
  int nin, istride;
  int fmx0[kMaxFields];     // per-field input box first column (signed mode)
  int fwx[kMaxFields];      // per-field input box width
  void* out;                // cpx<T> [task][nout][gy][ostride]
  int nout, ostride;
  int omx0, owx;
  int gy;
  int out_per_field;        // k_rowfwd: output s is pruned to box (fmx0[s], fwx[s]) instead of (omx0, owx)
};

// Box row segments of sequence seq = field * RY + row as the first stage's
// input: element e is box column wrap(e - mx0[field], GX).
template <typename T, int GX, int RY>
struct RowLoad {
  static constexpr bool SEQ_FAST = false;  // elements of one row are contiguous
  const cpx<T>* in;                        // task base: [field][gy][istride]
  const int* mx0;                          // per field (LDS)
  const int* wx;
  int gy, y0, istride;
  __device__ __forceinline__ cpx<T> operator()(int seq, int e) const {
    const int f = seq / RY, r = seq - (seq / RY) * RY;
    const int j = wrap(e - mx0[f], GX);
    return j < wx[f] ? in[((size_t)f * gy + y0 + r) * istride + j] : cpx<T>{T(0), T(0)};
  }
  static constexpr bool BATCH = true;
  template <int R, int STEP>
  __device__ __forceinline__ void load_batch(int seq, int e0, cpx<T>* v) const {
    const int f = seq / RY, r = seq - (seq / RY) * RY;
    const int w = wx[f];
    const cpx<T>* row = in + ((size_t)f * gy + y0 + r) * istride;
    const int j0 = wrap(e0 - mx0[f], GX);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int j = j0 + q * STEP;
      if (j >= GX) j -= GX;
      v[q] = j < w ? row[j] : cpx<T>{T(0), T(0)};
    }
  }
};

// Output box columns of sequence seq = output * RY + row.
template <typename T, int GX, int RY>
struct RowStore {
  static constexpr bool SEQ_FAST = false;
  cpx<T>* out;                             // task base: [output][gy][ostride]
  int gy, y0, ostride, omx0, owx;
  __device__ __forceinline__ void operator()(int seq, int e, cpx<T> v) {
    const int s = seq / RY, r = seq - (seq / RY) * RY;
    const int j = wrap(e - omx0, GX);
    if (j < owx) out[((size_t)s * gy + y0 + r) * ostride + j] = v;
  }
  static constexpr bool BATCH = true;
  template <int R, int STEP>
  __device__ __forceinline__ void store_batch(int seq, int e0, const cpx<T>* v) {
    const int s = seq / RY, r = seq - (seq / RY) * RY;
    cpx<T>* row = out + ((size_t)s * gy + y0 + r) * ostride;
    const int j0 = wrap(e0 - omx0, GX);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int j = j0 + q * STEP;
      if (j >= GX) j -= GX;
      if (j < owx) row[j] = v[q];
    }
  }
};

template <typename T, int GX, int NIN, int NOUT, int RY, class Op, int PLAN = 0>
__global__ __launch_bounds__(NT) void k_row(RowArgs a, const cpx<T>* __restrict__ tw) {
  constexpr int SS = seq_stride<GX>();
  __shared__ cpx<T> buf[NIN * RY * SS];
  __shared__ cpx<T> twl[kTwiddlesInLdsRow ? GX : 1];
  __shared__ int fmx0[NIN], fwx[NIN];
  const cpx<T>* twp = tw;
  if constexpr (kTwiddlesInLdsRow) {
    stage_twiddles<T, GX, NT>(twl, tw);
    twp = twl;
  }
#pragma unroll
  for (int f = 0; f < NIN; ++f)
    if (threadIdx.x == f) {
      fmx0[f] = a.fmx0[f];
      fwx[f] = a.fwx[f];
    }
  __syncthreads();
  const int y0 = blockIdx.x * RY;
  const int task = blockIdx.y;
  RowLoad<T, GX, RY> ld;
  ld.in = reinterpret_cast<const cpx<T>*>(a.in) + (size_t)task * a.nin * a.gy * a.istride;
  ld.mx0 = fmx0;
  ld.wx = fwx;
  ld.gy = a.gy;
  ld.y0 = y0;
  ld.istride = a.istride;
  NoIO none;
  fft_io_pf<T, GX, NIN * RY, SS, NT, +1, PLAN, kPfRow>(buf, twp, ld, none);
  for (int i = threadIdx.x; i < RY * GX; i += NT) {
    const int r = i / GX;
    const int x = i - r * GX;
    cpx<T> v[NIN];
#pragma unroll
    for (int f = 0; f < NIN; ++f) v[f] = buf[lds_at<GX>(f * RY + r, x)];
    cpx<T> o[NOUT];
    Op::apply(v, o);
#pragma unroll
    for (int s = 0; s < NOUT; ++s) buf[lds_at<GX>(s * RY + r, x)] = o[s];
  }
  __syncthreads();
  RowStore<T, GX, RY> st;
  st.out = reinterpret_cast<cpx<T>*>(a.out) + (size_t)task * a.nout * a.gy * a.ostride;
  st.gy = a.gy;
  st.y0 = y0;
  st.ostride = a.ostride;
  st.omx0 = a.omx0;
  st.owx = a.owx;
  fft_io_pf<T, GX, NOUT * RY, SS, NT, -1, PLAN, kPfRow, kFinal>(buf, twp, none, st);
}

// ----------------------------------------------------------------------------
// Split row path, for row tiles that do not fit LDS (or when it is faster):
// x-IFFT of every input field to physical rows in HBM (k_rowinv), pointwise
// products per physical point (k_pointop), x-FFT of the products pruned to
// the output box columns (k_rowfwd).  Same results as k_row, element for
// element: the transforms and the products are the same operations.
// ----------------------------------------------------------------------------
// One field's box row segments (element e is box column wrap(e - mx0, GX)).
template <typename T, int GX>
struct BoxRowLoad {
  static constexpr bool SEQ_FAST = false;
  const cpx<T>* in;  // first row of the tile
  int istride, mx0, wx;
  __device__ __forceinline__ cpx<T> operator()(int r, int e) const {
    const int j = wrap(e - mx0, GX);
    return j < wx ? in[(size_t)r * istride + j] : cpx<T>{T(0), T(0)};
  }
  static constexpr bool BATCH = true;
  template <int R, int STEP>
  __device__ __forceinline__ void load_batch(int r, int e0, cpx<T>* v) const {
    const cpx<T>* row = in + (size_t)r * istride;
    const int j0 = wrap(e0 - mx0, GX);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int j = j0 + q * STEP;
      if (j >= GX) j -= GX;
      v[q] = j < wx ? row[j] : cpx<T>{T(0), T(0)};
    }
  }
};

// Full physical rows [row][GX].
template <typename T, typename U, int GX>
struct FullRowStore {
  static constexpr bool SEQ_FAST = false;
  U* out;            // first row of the tile
  __device__ __forceinline__ void operator()(int r, int e, cpx<T> v) {
    out[(size_t)r * GX + e] = {decltype(out->x)(v.x), decltype(out->y)(v.y)};
  }
};

template <typename T, int GX>
struct FullRowLoad {
  static constexpr bool SEQ_FAST = false;
  const cpx<T>* in;  // first row of the tile
  __device__ __forceinline__ cpx<T> operator()(int r, int e) const { return in[(size_t)r * GX + e]; }
};

// Output box columns of one output (element e is column wrap(e - omx0, GX)).
template <typename T, int GX>
struct BoxRowStore {
  static constexpr bool SEQ_FAST = false;
  cpx<T>* out;       // first row of the tile
  int ostride, omx0, owx;
  __device__ __forceinline__ void operator()(int r, int e, cpx<T> v) {
    const int j = wrap(e - omx0, GX);
    if (j < owx) out[(size_t)r * ostride + j] = v;
  }
  static constexpr bool BATCH = true;
  template <int R, int STEP>
  __device__ __forceinline__ void store_batch(int r, int e0, const cpx<T>* v) {
    cpx<T>* row = out + (size_t)r * ostride;
    const int j0 = wrap(e0 - omx0, GX);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int j = j0 + q * STEP;
      if (j >= GX) j -= GX;
      if (j < owx) row[j] = v[q];
    }
  }
};

template <typename T, int GX, int RY, int PLAN = 0>
__global__ __launch_bounds__(NT) void k_rowinv(RowArgs a, cpx<T>* __restrict__ phys, const cpx<T>* __restrict__ tw) {
  constexpr int SS = seq_stride<GX>();
  __shared__ cpx<T> buf[RY * SS];
  __shared__ cpx<T> twl[kTwiddlesInLdsRow ? GX : 1];
  const cpx<T>* twp = tw;
  if constexpr (kTwiddlesInLdsRow) {
    stage_twiddles<T, GX, NT>(twl, tw);
    twp = twl;
  }
  const int y0 = blockIdx.x * RY;
  const int f = blockIdx.y;
  const int task = blockIdx.z;
  BoxRowLoad<T, GX> ld;
  ld.in = reinterpret_cast<const cpx<T>*>(a.in) + (((size_t)task * a.nin + f) * a.gy + y0) * a.istride;
  ld.istride = a.istride;
  ld.mx0 = a.fmx0[f];
  ld.wx = a.fwx[f];
  FullRowStore<T, cpx<T>, GX> st;
  st.out = phys + (((size_t)task * a.nin + f) * a.gy + y0) * GX;
  fft_io_pf<T, GX, RY, SS, NT, +1, PLAN, kPfRow, kFinal>(buf, twp, ld, st);
}

// in: [task][NIN][npts], out: [task][NOUT][npts]
template <typename T, int NIN, int NOUT, class Op>
__global__ __launch_bounds__(NT) void k_pointop(const cpx<T>* __restrict__ in, cpx<T>* __restrict__ out, int npts) {
  const int p = blockIdx.x * NT + threadIdx.x;
  const int task = blockIdx.y;
  if (p >= npts) return;
  const cpx<T>* src = in + (size_t)task * NIN * npts + p;
  cpx<T> v[NIN];
#pragma unroll
  for (int f = 0; f < NIN; ++f) v[f] = src[(size_t)f * npts];
  cpx<T> o[NOUT];
  Op::apply(v, o);
  cpx<T>* dst = out + (size_t)task * NOUT * npts + p;
#pragma unroll
  for (int s = 0; s < NOUT; ++s) dst[(size_t)s * npts] = o[s];
}

template <typename T, int GX, int RY, int PLAN = 0>
__global__ __launch_bounds__(NT) void k_rowfwd(RowArgs a, const cpx<T>* __restrict__ phys, const cpx<T>* __restrict__ tw) {
  constexpr int SS = seq_stride<GX>();
  __shared__ cpx<T> buf[RY * SS];
  __shared__ cpx<T> twl[kTwiddlesInLdsRow ? GX : 1];
  const cpx<T>* twp = tw;
  if constexpr (kTwiddlesInLdsRow) {
    stage_twiddles<T, GX, NT>(twl, tw);
    twp = twl;
  }
  const int y0 = blockIdx.x * RY;
  const int s = blockIdx.y;
  const int task = blockIdx.z;
  FullRowLoad<T, GX> ld;
  ld.in = phys + (((size_t)task * a.nout + s) * a.gy + y0) * GX;
  BoxRowStore<T, GX> st;
  st.out = reinterpret_cast<cpx<T>*>(a.out) + (((size_t)task * a.nout + s) * a.gy + y0) * a.ostride;
  st.ostride = a.ostride;
  st.owx = a.out_per_field ? a.fwx[s] : a.owx;
  st.omx0 = a.out_per_field ? a.fmx0[s] : a.omx0;
  fft_io_pf<T, GX, RY, SS, NT, -1, PLAN, kPfRow, kFinal>(buf, twp, ld, st);
}

// Adjoint of the pointwise products: v [task][NIN][npts] (forward inputs),
// g [task][NG][npts] (output gradients) -> gv [task][NIN][npts].
template <typename T, int NIN, int NG, class Op>
__global__ __launch_bounds__(NT) void k_pointadj(const cpx<T>* __restrict__ v, const cpx<T>* __restrict__ g,
                                                 cpx<T>* __restrict__ gv, int npts) {
  static_assert(NG == Op::NADJ, "gradient count must match the operator's adjoint");
  const int p = blockIdx.x * NT + threadIdx.x;
  const int task = blockIdx.y;
  if (p >= npts) return;
  cpx<T> f[NIN], gg[NG], o[NIN];
#pragma unroll
  for (int i = 0; i < NIN; ++i) f[i] = v[((size_t)task * NIN + i) * npts + p];
#pragma unroll
  for (int s = 0; s < NG; ++s) gg[s] = g[((size_t)task * NG + s) * npts + p];
  Op::adjoint(f, gg, o);
#pragma unroll
  for (int i = 0; i < NIN; ++i) gv[((size_t)task * NIN + i) * npts + p] = o[i];
}

// ----------------------------------------------------------------------------
// Row synthesis: x-IFFT of box columns to full physical rows (outputs).
// ----------------------------------------------------------------------------
struct SynthArgs {
  const void* in;           // cpx<T>: sequence s starts at in + s * in_seq, rows of istride
  size_t in_seq;
  int istride, mx0, wx, gy;
  cpx<double>* out;         // [nseq][gy][gx]
};

template <typename T, int GX, int RY>
__global__ __launch_bounds__(NT) void k_synth(SynthArgs a, const cpx<T>* __restrict__ tw) {
  constexpr int SS = seq_stride<GX>();
  __shared__ cpx<T> buf[RY * SS];
  __shared__ cpx<T> twl[kTwiddlesInLdsRow ? GX : 1];
  const cpx<T>* twp = tw;
  if constexpr (kTwiddlesInLdsRow) {
    stage_twiddles<T, GX, NT>(twl, tw);
    twp = twl;
  }
  const int y0 = blockIdx.x * RY;
  const int seq = blockIdx.y;
  BoxRowLoad<T, GX> ld;
  ld.in = reinterpret_cast<const cpx<T>*>(a.in) + (size_t)seq * a.in_seq + (size_t)y0 * a.istride;
  ld.istride = a.istride;
  ld.mx0 = a.mx0;
  ld.wx = a.wx;
  FullRowStore<T, cpx<double>, GX> st;
  st.out = a.out + ((size_t)seq * a.gy + y0) * GX;
  fft_io_pf<T, GX, RY, SS, NT, +1, 0, kPfSynth, kFinal>(buf, twp, ld, st);
}

}  // namespace glw
