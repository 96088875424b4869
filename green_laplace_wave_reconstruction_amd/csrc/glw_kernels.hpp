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

struct BoxTables {
  const double* qx;   // [wy][wx]
  const double* qy;
  const double* nu;
  const double* inu;  // 1 / max(nu, tiny)
  const double* q2;
  const double* a;    // q tanh q
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
};

template <typename T, int GY, int CW, int PLAN = 0>
__global__ __launch_bounds__(NT) void k_colinv(ColInvArgs a, const cpx<T>* __restrict__ tw) {
  __shared__ cpx<T> buf[CW * seq_stride<GY>()];
  __shared__ cpx<T> twl[GY];
  stage_twiddles<T, GY, NT>(twl, tw);
  const int col0 = blockIdx.x * CW;
  const int f = blockIdx.y;
  const int task = blockIdx.z;
  const int set = task / a.sdiv;
  const int src = a.fsrc[f];
  const int code = a.fcode[f];
  const double scale = a.fscale[f];
  const int box = a.wx * a.wy;
  const cpx<double>* spec = a.spec + ((size_t)set * a.nsrc + src) * box;
  const double* damp = a.damp ? a.damp + (size_t)((task % a.dmod) * a.dmul + a.dadd) * box : nullptr;
  double level = -1.0;
  if (a.fmax && ((a.floor_src_mask >> src) & 1)) {
    const size_t k = (size_t)set * a.fmax_stride + (a.fmax_stride == 1 ? 0 : src);
    level = a.floor_level * __longlong_as_double((long long)a.fmax[k]);
  }
  for (int i = threadIdx.x; i < CW * seq_stride<GY>(); i += NT) buf[i] = {T(0), T(0)};
  __syncthreads();
  for (int i = threadIdx.x; i < CW * a.wy; i += NT) {
    const int r = i / CW;
    const int c = i - r * CW;
    const int col = col0 + c;
    if (col < a.wx) {
      const int b = r * a.wx + col;
      cpx<double> s = spec[b];
      if (level >= 0.0 && hypot(s.x, s.y) < level) s = {0.0, 0.0};
      if (damp) s = {s.x * damp[b], s.y * damp[b]};
      s = apply_mul(code, s, a.tab, b);
      buf[lds_at<GY>(c, wrap(a.my0 + r, GY))] = {T(s.x * scale), T(s.y * scale)};
    }
  }
  __syncthreads();
  fft_lds<T, GY, CW, seq_stride<GY>(), NT, +1, PLAN>(buf, twl);
  cpx<T>* out = reinterpret_cast<cpx<T>*>(a.out) +
                ((size_t)task * a.out_nfield + a.out_foff + f) * GY * a.ostride;
  for (int i = threadIdx.x; i < CW * GY; i += NT) {
    const int y = i / CW;
    const int c = i - y * CW;
    if (col0 + c < a.wx) out[(size_t)y * a.ostride + col0 + c] = buf[lds_at<GY>(c, y)];
  }
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
  cpx<double>* out;         // [task][nsrc][owy*owx]
  uint64_t* fmax;           // [task][nsrc] or null
};

__device__ __forceinline__ void block_max_atomic(double v, uint64_t* dst) {
  // non-negative doubles order like their bit patterns
  __shared__ double red[NT / 64];
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int i = 1; i < NT / 64; ++i) m = fmax(m, red[i]);
    if (m > 0.0) atomicMax((unsigned long long*)dst, (unsigned long long)__double_as_longlong(m));
  }
}

template <typename T, int GY, int CW, int PLAN = 0>
__global__ __launch_bounds__(NT) void k_colfwd(ColFwdArgs a, const cpx<T>* __restrict__ tw) {
  __shared__ cpx<T> buf[CW * seq_stride<GY>()];
  __shared__ cpx<T> twl[GY];
  stage_twiddles<T, GY, NT>(twl, tw);
  const int col0 = blockIdx.x * CW;
  const int s = blockIdx.y;
  const int task = blockIdx.z;
  const int nf = a.in_nfield ? a.in_nfield : a.nsrc;
  const cpx<T>* in = reinterpret_cast<const cpx<T>*>(a.in) + ((size_t)task * nf + a.in_foff + s) * GY * a.istride;
  for (int i = threadIdx.x; i < CW * GY; i += NT) {
    const int y = i / CW;
    const int c = i - y * CW;
    buf[lds_at<GY>(c, y)] = (col0 + c < a.owx) ? in[(size_t)y * a.istride + col0 + c] : cpx<T>{T(0), T(0)};
  }
  __syncthreads();
  fft_lds<T, GY, CW, seq_stride<GY>(), NT, -1, PLAN>(buf, twl);
  cpx<double>* out = a.out + ((size_t)task * a.nsrc + s) * a.owy * a.owx;
  double vmax = 0.0;
  for (int i = threadIdx.x; i < CW * a.owy; i += NT) {
    const int r = i / CW;
    const int c = i - r * CW;
    const int col = col0 + c;
    if (col < a.owx) {
      const int b = r * a.owx + col;
      cpx<T> v = buf[lds_at<GY>(c, wrap(a.omy0 + r, GY))];
      cpx<double> d = {double(v.x) * a.scale, double(v.y) * a.scale};
      if (!a.mask[b]) d = {0.0, 0.0};
      else vmax = fmax(vmax, hypot(d.x, d.y));
      out[b] = d;
    }
  }
  if (a.fmax) block_max_atomic(vmax, a.fmax + (size_t)task * a.nsrc + s);
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

template <typename T, int GX, int NIN, int NOUT, int RY, class Op, int PLAN = 0>
__global__ __launch_bounds__(NT) void k_row(RowArgs a, const cpx<T>* __restrict__ tw) {
  constexpr int SS = seq_stride<GX>();
  __shared__ cpx<T> buf[NIN * RY * SS];
  __shared__ cpx<T> twl[GX];
  stage_twiddles<T, GX, NT>(twl, tw);
  const int y0 = blockIdx.x * RY;
  const int task = blockIdx.y;
  const cpx<T>* in = reinterpret_cast<const cpx<T>*>(a.in) + (size_t)task * a.nin * a.gy * a.istride;
  for (int i = threadIdx.x; i < NIN * RY * SS; i += NT) buf[i] = {T(0), T(0)};
  __syncthreads();
#pragma unroll
  for (int f = 0; f < NIN; ++f) {
    const int wx = a.fwx[f], mx0 = a.fmx0[f];
    for (int i = threadIdx.x; i < RY * wx; i += NT) {
      const int r = i / wx;
      const int j = i - r * wx;
      buf[lds_at<GX>(f * RY + r, wrap(mx0 + j, GX))] = in[((size_t)f * a.gy + y0 + r) * a.istride + j];
    }
  }
  __syncthreads();
  fft_lds<T, GX, NIN * RY, SS, NT, +1, PLAN>(buf, twl);
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
  fft_lds<T, GX, NOUT * RY, SS, NT, -1, PLAN>(buf, twl);
  cpx<T>* out = reinterpret_cast<cpx<T>*>(a.out) + (size_t)task * a.nout * a.gy * a.ostride;
  for (int i = threadIdx.x; i < NOUT * RY * a.owx; i += NT) {
    const int sr = i / a.owx;
    const int j = i - sr * a.owx;
    const int s = sr / RY, r = sr - (sr / RY) * RY;
    out[((size_t)s * a.gy + y0 + r) * a.ostride + j] = buf[lds_at<GX>(s * RY + r, wrap(a.omx0 + j, GX))];
  }
}

// ----------------------------------------------------------------------------
// Split row path, for row tiles that do not fit LDS (or when it is faster):
// x-IFFT of every input field to physical rows in HBM (k_rowinv), pointwise
// products per physical point (k_pointop), x-FFT of the products pruned to
// the output box columns (k_rowfwd).  Same results as k_row, element for
// element: the transforms and the products are the same operations.
// ----------------------------------------------------------------------------
template <typename T, int GX, int RY, int PLAN = 0>
__global__ __launch_bounds__(NT) void k_rowinv(RowArgs a, cpx<T>* __restrict__ phys, const cpx<T>* __restrict__ tw) {
  constexpr int SS = seq_stride<GX>();
  __shared__ cpx<T> buf[RY * SS];
  __shared__ cpx<T> twl[GX];
  stage_twiddles<T, GX, NT>(twl, tw);
  const int y0 = blockIdx.x * RY;
  const int f = blockIdx.y;
  const int task = blockIdx.z;
  const int wx = a.fwx[f], mx0 = a.fmx0[f];
  const cpx<T>* in = reinterpret_cast<const cpx<T>*>(a.in) + ((size_t)task * a.nin + f) * a.gy * a.istride;
  for (int i = threadIdx.x; i < RY * SS; i += NT) buf[i] = {T(0), T(0)};
  __syncthreads();
  for (int i = threadIdx.x; i < RY * wx; i += NT) {
    const int r = i / wx;
    const int j = i - r * wx;
    buf[lds_at<GX>(r, wrap(mx0 + j, GX))] = in[(size_t)(y0 + r) * a.istride + j];
  }
  __syncthreads();
  fft_lds<T, GX, RY, SS, NT, +1, PLAN>(buf, twl);
  cpx<T>* out = phys + (((size_t)task * a.nin + f) * a.gy + y0) * GX;
  for (int i = threadIdx.x; i < RY * GX; i += NT) {
    const int r = i / GX;
    const int x = i - r * GX;
    out[(size_t)r * GX + x] = buf[lds_at<GX>(r, x)];
  }
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
  __shared__ cpx<T> twl[GX];
  stage_twiddles<T, GX, NT>(twl, tw);
  const int y0 = blockIdx.x * RY;
  const int s = blockIdx.y;
  const int task = blockIdx.z;
  const cpx<T>* src = phys + (((size_t)task * a.nout + s) * a.gy + y0) * GX;
  for (int i = threadIdx.x; i < RY * GX; i += NT) {
    const int r = i / GX;
    const int x = i - r * GX;
    buf[lds_at<GX>(r, x)] = src[(size_t)r * GX + x];
  }
  __syncthreads();
  fft_lds<T, GX, RY, SS, NT, -1, PLAN>(buf, twl);
  cpx<T>* out = reinterpret_cast<cpx<T>*>(a.out) + ((size_t)task * a.nout + s) * a.gy * a.ostride;
  const int owx = a.out_per_field ? a.fwx[s] : a.owx;
  const int omx0 = a.out_per_field ? a.fmx0[s] : a.omx0;
  for (int i = threadIdx.x; i < RY * owx; i += NT) {
    const int r = i / owx;
    const int j = i - r * owx;
    out[(size_t)(y0 + r) * a.ostride + j] = buf[lds_at<GX>(r, wrap(omx0 + j, GX))];
  }
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
  __shared__ cpx<T> twl[GX];
  stage_twiddles<T, GX, NT>(twl, tw);
  const int y0 = blockIdx.x * RY;
  const int seq = blockIdx.y;
  const cpx<T>* in = reinterpret_cast<const cpx<T>*>(a.in) + (size_t)seq * a.in_seq;
  for (int i = threadIdx.x; i < RY * SS; i += NT) buf[i] = {T(0), T(0)};
  __syncthreads();
  for (int i = threadIdx.x; i < RY * a.wx; i += NT) {
    const int r = i / a.wx;
    const int j = i - r * a.wx;
    buf[lds_at<GX>(r, wrap(a.mx0 + j, GX))] = in[(size_t)(y0 + r) * a.istride + j];
  }
  __syncthreads();
  fft_lds<T, GX, RY, SS, NT, +1>(buf, twl);
  cpx<double>* out = a.out + (size_t)seq * a.gy * GX;
  for (int i = threadIdx.x; i < RY * GX; i += NT) {
    const int r = i / GX;
    const int x = i - r * GX;
    cpx<T> v = buf[lds_at<GX>(r, x)];
    out[(size_t)(y0 + r) * GX + x] = {double(v.x), double(v.y)};
  }
}

}  // namespace glw
