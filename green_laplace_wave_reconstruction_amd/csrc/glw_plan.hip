// Plan, workspace, accumulation kernels and C ABI of the glw library.
//
// Build: hipcc --offload-arch=gfx950 -O3 -shared -fPIC
//        -Xclang -target-feature -Xclang -packed-fp32-ops
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "glw.h"
#include "glw_internal.hpp"
#include "glw_kernels.hpp"
#include "glw_launch.hpp"
#include "glw_ops.hpp"

using namespace glw;

namespace {

thread_local std::string g_err;

// Row stride of column-pass intermediates: box width rounded up to 8 complex
// values, so every 8-column tile starts on a 128-byte line.
inline int sw(int w) { return (w + 7) & ~7; }

// Output grids: any even length up to this; FFT lengths (2^a 3^b, 16..2048)
// synthesize through the FFT passes, the others through direct DFT passes.
constexpr int kMaxOutputLength = 16384;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define GLW_HIP(call)                                                                \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(GLW_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)

// ----------------------------------------------------------------------------
// FP64 elementwise kernels
// ----------------------------------------------------------------------------
__global__ void k_fill_u64(uint64_t* p, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// S_t = S0 exp(-i omega t)   (python/src/green_laplace/spatial.py:70-72;
// src/gl_spectral_surface.m:145-147)
__global__ void k_phase(const cpx<double>* s0, const double* omega, const double* times, int box,
                        cpx<double>* st) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  if (b >= box) return;
  double sn, cs;
  sincos(omega[b] * times[t], &sn, &cs);
  const cpx<double> s = s0[b];
  st[(size_t)t * box + b] = {s.x * cs + s.y * sn, s.y * cs - s.x * sn};
}

// ----------------------------------------------------------------------------
// Output grids whose length is not an FFT length (any even N): the synthesis
// passes as direct inverse DFTs over the box, with tables of
// exp(+2 pi i m n / N) built on the host for the box's signed modes m.  Only
// the final synthesis depends on N (the working grids are FFT lengths).
// ----------------------------------------------------------------------------
constexpr int kDftCols = 8;  // box columns per workgroup of k_dft_colinv

// y-pass: column c of field f -> out[task][f][y][ostride] for every y < gy.
// ty: [wy][gy] = exp(2 pi i (my0 + r) y / gy).  The box columns of one tile
// go through the same loader as the FFT pass (damping, floor, multiplier,
// scale) into LDS; each output is a wy-term sum.
__global__ __launch_bounds__(NT) void k_dft_colinv(ColInvArgs a, int gy, const cpx<double>* __restrict__ ty) {
  extern __shared__ cpx<double> col[];  // [kDftCols][wy]
  const int f = blockIdx.y;
  const int task = blockIdx.z;
  const int set = task / a.sdiv;
  const int src = a.fsrc[f];
  const int box = a.wx * a.wy;
  ColInvLoad<double> ld;
  ld.spec = a.spec + ((size_t)set * a.nsrc + src) * box;
  ld.damp = a.damp ? a.damp + (size_t)((task % a.dmod) * a.dmul + a.dadd) * box : nullptr;
  ld.level = -1.0;
  if (a.fmax && ((a.floor_src_mask >> src) & 1)) {
    const size_t k = (size_t)set * a.fmax_stride + (a.fmax_stride == 1 ? 0 : src);
    ld.level = a.floor_level * __longlong_as_double((long long)a.fmax[k]);
  }
  ld.scale = a.fscale[f];
  ld.mul = a.tab.mul + (size_t)a.fcode[f] * box;
  ld.col0 = blockIdx.x * kDftCols;
  ld.wx = a.wx;
  ld.wy = a.wy;
  ld.my0 = 0;           // element e of the loader is box row e (the table carries the mode offset)
  ld.gy = a.wy;
  for (int i = threadIdx.x; i < kDftCols * a.wy; i += NT) {
    const int c = i / a.wy, r = i - c * a.wy;
    col[i] = ld(c, r);
  }
  __syncthreads();
  cpx<double>* out = reinterpret_cast<cpx<double>*>(a.out) + ((size_t)task * a.out_nfield + a.out_foff + f) * gy *
                     a.ostride;
  const int ncol = min(kDftCols, a.wx - ld.col0);
  for (int i = threadIdx.x; i < gy * ncol; i += NT) {
    const int y = i / ncol, c = i - y * ncol;
    const cpx<double>* x = col + c * a.wy;
    double re = 0.0, im = 0.0;
    for (int r = 0; r < a.wy; ++r) {
      const cpx<double> w = ty[(size_t)r * gy + y], v = x[r];
      re += v.x * w.x - v.y * w.y;
      im += v.x * w.y + v.y * w.x;
    }
    out[(size_t)y * a.ostride + ld.col0 + c] = {re, im};
  }
}

// x-pass: row y of sequence seq (box columns, istride apart) -> out[seq][y][gx].
// tx: [wx][gx] = exp(2 pi i (mx0 + c) x / gx); one workgroup per row.
__global__ __launch_bounds__(NT) void k_dft_row(SynthArgs a, int gx, const cpx<double>* __restrict__ tx) {
  extern __shared__ cpx<double> row[];  // [wx]
  const int y = blockIdx.x;
  const int seq = blockIdx.y;
  const cpx<double>* in = reinterpret_cast<const cpx<double>*>(a.in) + (size_t)seq * a.in_seq + (size_t)y * a.istride;
  for (int c = threadIdx.x; c < a.wx; c += NT) row[c] = in[c];
  __syncthreads();
  cpx<double>* out = a.out + ((size_t)seq * a.gy + y) * gx;
  for (int x = threadIdx.x; x < gx; x += NT) {
    double re = 0.0, im = 0.0;
    for (int c = 0; c < a.wx; ++c) {
      const cpx<double> w = tx[(size_t)c * gx + x], v = row[c];
      re += v.x * w.x - v.y * w.y;
      im += v.x * w.y + v.y * w.x;
    }
    out[x] = {re, im};
  }
}

__device__ __forceinline__ cpx<double> floored(cpx<double> x, double level) {
  return hypot(x.x, x.y) < level ? cpx<double>{0.0, 0.0} : x;
}
__device__ __forceinline__ double bits_to_d(uint64_t b) { return __longlong_as_double((long long)b); }

// Layout of spectra written by a column-forward pass at a padded row pitch:
// spectrum k, box bin b (row-major over box width w) at k * slab + (b / w) * p + b % w.
struct SpecLayout {
  int w, p;     // box width, row pitch
  size_t slab;  // elements per spectrum (box rows x p)
  __device__ __forceinline__ size_t bin(int b) const {
    const int r = b / w;
    return (size_t)r * p + (b - r * w);
  }
};
inline SpecLayout spec_layout(const glw_box& box) {
  return SpecLayout{box.wx, sw(box.wx), (size_t)box.wy * sw(box.wx)};
}


// Inner GL accumulation of the pair state (_order3.py:240-259).
struct AccInnerArgs {
  const cpx<double>* spec;  // [nto*I][4] spectra in layout lay
  SpecLayout lay;
  const uint64_t* fmax;     // [nto*I][4]
  double floor_level;
  int I, box;
  const double* a;
  const uint8_t* mask;
  const double* ch;         // [I][box]
  const double* sh;
  const double* coef;       // [I]
  cpx<double>* out;         // [nto][4][box]: eta, eta_t, phi, phi_t
  uint64_t* etmax;          // [nto]
};

__global__ __launch_bounds__(NT) void k_acc_inner(AccInnerArgs g) {
  const int b = blockIdx.x * NT + threadIdx.x;
  const int to = blockIdx.y;
  cpx<double> e = {0, 0}, et = {0, 0}, p = {0, 0}, pt = {0, 0};
  const bool on = b < g.box && g.mask[b];
  if (on) {
    const double A = g.a[b];
    const size_t bs = g.lay.bin(b);
    for (int i = 0; i < g.I; ++i) {
      const size_t task = (size_t)to * g.I + i;
      cpx<double> s[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        s[k] = floored(g.spec[(task * 4 + k) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[task * 4 + k]));
      const double c = g.coef[i], chv = g.ch[(size_t)i * g.box + b], shv = g.sh[(size_t)i * g.box + b];
      // e += c (-A sh sd + ch sk)
      const cpx<double> ze = {c * (-A * shv * s[0].x + chv * s[1].x), c * (-A * shv * s[0].y + chv * s[1].y)};
      e = cadd(e, ze);
      // et -= i c (-A sh sdt + ch skt)
      const cpx<double> zt = {c * (-A * shv * s[2].x + chv * s[3].x), c * (-A * shv * s[2].y + chv * s[3].y)};
      et = {et.x + zt.y, et.y - zt.x};
      // p += i c (ch sd - sh sk)
      const cpx<double> zp = {c * (chv * s[0].x - shv * s[1].x), c * (chv * s[0].y - shv * s[1].y)};
      p = {p.x - zp.y, p.y + zp.x};
      // pt += c (ch sdt - sh skt)
      pt = {pt.x + c * (chv * s[2].x - shv * s[3].x), pt.y + c * (chv * s[2].y - shv * s[3].y)};
    }
  }
  if (b < g.box) {
    cpx<double>* o = g.out + (size_t)to * 4 * g.box + b;
    o[0] = e;
    o[g.box] = et;
    o[2 * (size_t)g.box] = p;
    o[3 * (size_t)g.box] = pt;
  }
  block_max_atomic(on ? hypot(et.x, et.y) : 0.0, g.etmax + to);
}

// Floors of the cubic forcing spectra and the maxima of the psi combinations
// (_order3.py:103-122).
struct ComboArgs {
  const cpx<double>* spec;  // [nto][4] spectra in layout lay: F_K, F_D, C, C_t
  SpecLayout lay;
  const uint64_t* fmax;     // [nto][4]
  double floor_level;
  int box;
  const double* a;
  const uint8_t* mask;
  uint64_t* cmax;           // [nto][2]
};

__global__ __launch_bounds__(NT) void k_combo_outer(ComboArgs g) {
  const int b = blockIdx.x * NT + threadIdx.x;
  const int to = blockIdx.y;
  double m0 = 0.0, m1 = 0.0;
  if (b < g.box && g.mask[b]) {
    const size_t bs = g.lay.bin(b);
    cpx<double> s[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      s[k] = floored(g.spec[((size_t)to * 4 + k) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[(size_t)to * 4 + k]));
    const double A = g.a[b];
    const cpx<double> kp = {s[0].x + A * s[2].x, s[0].y + A * s[2].y};
    const cpx<double> dp = csub(s[1], s[3]);
    m0 = hypot(kp.x, kp.y);
    m1 = hypot(dp.x, dp.y);
  }
  block_max_atomic(m0, g.cmax + (size_t)to * 2);
  __syncthreads();
  block_max_atomic(m1, g.cmax + (size_t)to * 2 + 1);
}

// Outer GL accumulation of eta33, psi33 spectra (_order3.py:117-138).
struct AccOuterArgs {
  const cpx<double>* spec;  // [nt*O][4] spectra in layout lay
  SpecLayout lay;
  const uint64_t* fmax;     // [nt*O][4]
  const uint64_t* cmax;     // [nt*O][2]
  double floor_level;
  int O, box;
  const double* a;
  const uint8_t* mask;
  const double* ch;         // [O][box]
  const double* sh;
  const double* coef;       // [O]
  cpx<double>* out;         // [nt][2][box]
};

__global__ __launch_bounds__(NT) void k_acc_outer(AccOuterArgs g) {
  const int b = blockIdx.x * NT + threadIdx.x;
  const int t = blockIdx.y;
  if (b >= g.box) return;
  cpx<double> eta = {0, 0}, psi = {0, 0};
  if (g.mask[b]) {
    const double A = g.a[b];
    const size_t bs = g.lay.bin(b);
    for (int o = 0; o < g.O; ++o) {
      const size_t to = (size_t)t * g.O + o;
      cpx<double> s[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        s[k] = floored(g.spec[(to * 4 + k) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[to * 4 + k]));
      const cpx<double> kp = floored(cpx<double>{s[0].x + A * s[2].x, s[0].y + A * s[2].y},
                                     g.floor_level * bits_to_d(g.cmax[to * 2]));
      const cpx<double> dp = floored(csub(s[1], s[3]), g.floor_level * bits_to_d(g.cmax[to * 2 + 1]));
      const double c = g.coef[o], chv = g.ch[(size_t)o * g.box + b], shv = g.sh[(size_t)o * g.box + b];
      // eta += c (A sh F_D - i ch F_K)
      eta.x += c * (A * shv * s[1].x + chv * s[0].y);
      eta.y += c * (A * shv * s[1].y - chv * s[0].x);
      // psi += c (-i ch F_Dpsi - sh F_Kpsi)
      psi.x += c * (chv * dp.y - shv * kp.x);
      psi.y += c * (-chv * dp.x - shv * kp.y);
    }
  }
  g.out[(size_t)t * 2 * g.box + b] = eta;
  g.out[((size_t)t * 2 + 1) * g.box + b] = psi;
}

// eta22 accumulation (spatial.py:192-221).
struct AccEta22Args {
  const cpx<double>* spec;  // [nt*J][2] spectra in layout lay
  SpecLayout lay;
  const uint64_t* fmax;     // [nt*J][2]
  double floor_level;
  int J, box;
  const double* a;
  const double* ch;         // [J][box]
  const double* sh;
  const double* coef;       // [J]
  cpx<double>* out;         // [nt][2][box] slot 0
};

__global__ __launch_bounds__(NT) void k_acc_eta22(AccEta22Args g) {
  const int b = blockIdx.x * NT + threadIdx.x;
  const int t = blockIdx.y;
  if (b >= g.box) return;
  const double A = g.a[b];
  const size_t bs = g.lay.bin(b);
  cpx<double> acc = {0, 0};
  for (int j = 0; j < g.J; ++j) {
    const size_t task = (size_t)t * g.J + j;
    const cpx<double> sd = floored(g.spec[(task * 2) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[task * 2]));
    const cpx<double> sk =
        floored(g.spec[(task * 2 + 1) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[task * 2 + 1]));
    const double c = g.coef[j], chv = g.ch[(size_t)j * g.box + b], shv = g.sh[(size_t)j * g.box + b];
    acc.x += c * (-A * shv * sd.x + chv * sk.x);
    acc.y += c * (-A * shv * sd.y + chv * sk.y);
  }
  g.out[(size_t)t * 2 * g.box + b] = acc;
}

// psi22 accumulation (spatial.py:276-308).
struct AccPsi22Args {
  const cpx<double>* spec;    // [nt*4][8] spectra in layout lay (branch-node task, pair p: sd = 2p, sk = 2p+1)
  const cpx<double>* contact; // [nt][1] spectra in layout lay
  SpecLayout lay;
  int box;
  const double* inu;
  const uint8_t* mask;
  const double* factor;       // [4][4][box]
  const double* coef;         // [4]
  const double* sign;         // [2]
  cpx<double>* out;           // [nt][2][box] slot 1
};

__global__ __launch_bounds__(NT) void k_acc_psi22(AccPsi22Args g) {
  const int b = blockIdx.x * NT + threadIdx.x;
  const int t = blockIdx.y;
  if (b >= g.box) return;
  cpx<double> res = {0, 0};
  const double inu = g.inu[b];
  const size_t bs = g.lay.bin(b);
  for (int bn = 0; bn < 4; ++bn) {
    const double sgn = g.sign[bn / 2];
    const size_t task = (size_t)t * 4 + bn;
    cpx<double> resp = {0, 0};
    for (int p = 0; p < 4; ++p) {
      const double f = g.factor[((size_t)bn * 4 + p) * g.box + b];
      const cpx<double> sd = g.spec[(task * 8 + 2 * p) * g.lay.slab + bs];
      const cpx<double> sk = g.spec[(task * 8 + 2 * p + 1) * g.lay.slab + bs];
      resp.x += f * (sd.x + sgn * sk.x * inu);
      resp.y += f * (sd.y + sgn * sk.y * inu);
    }
    const double c = g.coef[bn];
    res = {res.x - c * resp.y, res.y + c * resp.x};
  }
  const cpx<double> ct = g.contact[(size_t)t * g.lay.slab + bs];
  cpx<double> v = cadd(res, ct);
  if (!g.mask[b]) v = {0, 0};
  g.out[((size_t)t * 2 + 1) * g.box + b] = v;
}


// ---------------------------------------------------------------------------
// Order-4 accumulations (paper/order4/cpp/gl_eta44_fftw.cpp)
// ---------------------------------------------------------------------------
struct Acc4Args {
  const cpx<double>* spec;  // [nset*J][nsrc] spectra in layout lay
  SpecLayout lay;
  const uint64_t* fmax;     // [nset*J][nsrc]
  double floor_level;
  int J, box, nsrc, nacc;
  const double* a;
  const uint8_t* mask;
  const double* ch;         // [J][box]
  const double* sh;
  const double* coef;       // [J]
  cpx<double>* out;         // [nset][nacc][box]
  uint64_t* omax;           // [nset][nacc]
};

// Pair level (:404-428): 6 sources -> es, ets, ps, pts, ptts, then masked
// (floors applied by the consumer with omax).
__global__ __launch_bounds__(NT) void k_acc_pair44(Acc4Args g) {
  const int b = blockIdx.x * NT + threadIdx.x;
  const int set = blockIdx.y;
  cpx<double> acc[5] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
  const bool on = b < g.box && g.mask[b];
  if (on) {
    const double A = g.a[b];
    const size_t bs = g.lay.bin(b);
    for (int i = 0; i < g.J; ++i) {
      const size_t task = (size_t)set * g.J + i;
      cpx<double> s[6];
#pragma unroll
      for (int k = 0; k < 6; ++k)
        s[k] = floored(g.spec[(task * 6 + k) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[task * 6 + k]));
      const double c = g.coef[i], chv = g.ch[(size_t)i * g.box + b], shv = g.sh[(size_t)i * g.box + b];
      // es += c (-A sh sd + ch sk)
      acc[0].x += c * (-A * shv * s[0].x + chv * s[1].x);
      acc[0].y += c * (-A * shv * s[0].y + chv * s[1].y);
      // ets -= i c (-A sh sds + ch sks)
      const cpx<double> z1 = {c * (-A * shv * s[2].x + chv * s[3].x), c * (-A * shv * s[2].y + chv * s[3].y)};
      acc[1] = {acc[1].x + z1.y, acc[1].y - z1.x};
      // ps += i c (ch sd - sh sk)
      const cpx<double> z2 = {c * (chv * s[0].x - shv * s[1].x), c * (chv * s[0].y - shv * s[1].y)};
      acc[2] = {acc[2].x - z2.y, acc[2].y + z2.x};
      // pts += c (ch sds - sh sks)
      acc[3].x += c * (chv * s[2].x - shv * s[3].x);
      acc[3].y += c * (chv * s[2].y - shv * s[3].y);
      // ptts -= i c (ch sdss - sh skss)
      const cpx<double> z4 = {c * (chv * s[4].x - shv * s[5].x), c * (chv * s[4].y - shv * s[5].y)};
      acc[4] = {acc[4].x + z4.y, acc[4].y - z4.x};
    }
  }
  if (b < g.box)
    for (int k = 0; k < 5; ++k) g.out[((size_t)set * 5 + k) * g.box + b] = acc[k];
  for (int k = 0; k < 5; ++k) {
    block_max_atomic(on ? hypot(acc[k].x, acc[k].y) : 0.0, g.omax + (size_t)set * 5 + k);
    __syncthreads();
  }
}

// Third level (:594-615): 4 sources -> es, ps, pts.
__global__ __launch_bounds__(NT) void k_acc_third44(Acc4Args g) {
  const int b = blockIdx.x * NT + threadIdx.x;
  const int set = blockIdx.y;
  cpx<double> es = {0, 0}, ps = {0, 0}, pts = {0, 0};
  const bool on = b < g.box && g.mask[b];
  if (on) {
    const double A = g.a[b];
    const size_t bs = g.lay.bin(b);
    for (int i = 0; i < g.J; ++i) {
      const size_t task = (size_t)set * g.J + i;
      cpx<double> s[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        s[k] = floored(g.spec[(task * 4 + k) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[task * 4 + k]));
      const double c = g.coef[i], chv = g.ch[(size_t)i * g.box + b], shv = g.sh[(size_t)i * g.box + b];
      // es += c (A sh fd - i ch fk)
      es.x += c * (A * shv * s[1].x + chv * s[0].y);
      es.y += c * (A * shv * s[1].y - chv * s[0].x);
      // ps += c (-i ch fd - sh fk)
      ps.x += c * (chv * s[1].y - shv * s[0].x);
      ps.y += c * (-chv * s[1].x - shv * s[0].y);
      // pts += c (-i ch fdt - sh fkt)
      pts.x += c * (chv * s[3].y - shv * s[2].x);
      pts.y += c * (-chv * s[3].x - shv * s[2].y);
    }
  }
  if (b < g.box) {
    g.out[((size_t)set * 3 + 0) * g.box + b] = es;
    g.out[((size_t)set * 3 + 1) * g.box + b] = ps;
    g.out[((size_t)set * 3 + 2) * g.box + b] = pts;
  }
  block_max_atomic(on ? hypot(es.x, es.y) : 0.0, g.omax + (size_t)set * 3);
  __syncthreads();
  block_max_atomic(on ? hypot(ps.x, ps.y) : 0.0, g.omax + (size_t)set * 3 + 1);
  __syncthreads();
  block_max_atomic(on ? hypot(pts.x, pts.y) : 0.0, g.omax + (size_t)set * 3 + 2);
}

// Quartic level (:736-750): 2 sources -> eta44 spectrum.
__global__ __launch_bounds__(NT) void k_acc_quartic44(Acc4Args g) {
  const int b = blockIdx.x * NT + threadIdx.x;
  const int set = blockIdx.y;
  cpx<double> o = {0, 0};
  const bool on = b < g.box && g.mask[b];
  if (on) {
    const double A = g.a[b];
    const size_t bs = g.lay.bin(b);
    for (int i = 0; i < g.J; ++i) {
      const size_t task = (size_t)set * g.J + i;
      const cpx<double> fk = floored(g.spec[(task * 2) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[task * 2]));
      const cpx<double> fd =
          floored(g.spec[(task * 2 + 1) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[task * 2 + 1]));
      const double c = g.coef[i], chv = g.ch[(size_t)i * g.box + b], shv = g.sh[(size_t)i * g.box + b];
      o.x += c * (A * shv * fd.x + chv * fk.y);
      o.y += c * (A * shv * fd.y - chv * fk.x);
    }
  }
  if (b < g.box) g.out[(size_t)set * g.box + b] = o;
  block_max_atomic(on ? hypot(o.x, o.y) : 0.0, g.omax + set);
}

// ----------------------------------------------------------------------------
// Adjoint kernels (gradients of a real loss of the eta fields; a gradient of
// a complex value z is dL/dRe z + i dL/dIm z).  Supports, floors and masks of
// the forward graph are held fixed (piecewise-constant in the amplitudes).
// ----------------------------------------------------------------------------
__device__ __forceinline__ cpx<double> conj_mul(cpx<double> m, cpx<double> g) {
  return {m.x * g.x + m.y * g.y, m.x * g.y - m.y * g.x};
}
__device__ __forceinline__ bool kept(cpx<double> x, double level) { return !(hypot(x.x, x.y) < level); }

// dst (box d) = / += scale * src (box u), where box u contains box d
__global__ void k_adj_gather(const cpx<double>* src, glw_box u, glw_box d, double scale, int add, cpx<double>* dst) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= d.wx * d.wy) return;
  const int r = b / d.wx, c = b - r * d.wx;
  const cpx<double> v = src[(size_t)(d.my0 + r - u.my0) * u.wx + (d.mx0 + c - u.mx0)];
  const cpx<double> w = {scale * v.x, scale * v.y};
  dst[b] = add ? cadd(dst[b], w) : w;
}

// eta22 accumulation adjoint: gE [box] -> gS [J][2][box] (sd, sk)
__global__ __launch_bounds__(NT) void k_adj_acc_eta22(AccEta22Args g, const cpx<double>* ge, const uint8_t* mask,
                                                      cpx<double>* gs) {
  const int b = blockIdx.x * NT + threadIdx.x;
  const int j = blockIdx.y;
  if (b >= g.box) return;
  const size_t task = j;
  cpx<double> gd = {0, 0}, gk = {0, 0};
  if (mask[b]) {
    const cpx<double> e = ge[b];
    const double A = g.a[b], c = g.coef[j], chv = g.ch[(size_t)j * g.box + b], shv = g.sh[(size_t)j * g.box + b];
    const size_t bs = g.lay.bin(b);
    if (kept(g.spec[(task * 2) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[task * 2])))
      gd = {-c * A * shv * e.x, -c * A * shv * e.y};
    if (kept(g.spec[(task * 2 + 1) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[task * 2 + 1])))
      gk = {c * chv * e.x, c * chv * e.y};
  }
  gs[(task * 2) * g.box + b] = gd;
  gs[(task * 2 + 1) * g.box + b] = gk;
}

// eta33 outer accumulation adjoint: gE [box] -> gS [O][2][box] (forcing_k, forcing_d)
__global__ __launch_bounds__(NT) void k_adj_acc_outer(AccOuterArgs g, const cpx<double>* ge, cpx<double>* gs) {
  const int b = blockIdx.x * NT + threadIdx.x;
  const int o = blockIdx.y;
  if (b >= g.box) return;
  cpx<double> gk = {0, 0}, gd = {0, 0};
  if (g.mask[b]) {
    const cpx<double> e = ge[b];
    const double A = g.a[b], c = g.coef[o], chv = g.ch[(size_t)o * g.box + b], shv = g.sh[(size_t)o * g.box + b];
    // eta += c (A sh F_D - i ch F_K)
    const size_t bs = g.lay.bin(b);
    if (kept(g.spec[((size_t)o * 4) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[(size_t)o * 4])))
      gk = {-c * chv * e.y, c * chv * e.x};  // conj(-i c ch) e = i c ch e
    if (kept(g.spec[((size_t)o * 4 + 1) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[(size_t)o * 4 + 1])))
      gd = {c * A * shv * e.x, c * A * shv * e.y};
  }
  gs[((size_t)o * 2) * g.box + b] = gk;
  gs[((size_t)o * 2 + 1) * g.box + b] = gd;
}

// Inner accumulation adjoint: gpair [O][4][box] (eta, eta_t, phi, phi_t) -> gS [O*I][4][box]
__global__ __launch_bounds__(NT) void k_adj_acc_inner(AccInnerArgs g, const cpx<double>* gp, cpx<double>* gs) {
  const int b = blockIdx.x * NT + threadIdx.x;
  const int task = blockIdx.y;  // o * I + i
  if (b >= g.box) return;
  const int o = task / g.I, i = task - o * g.I;
  cpx<double> r[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  if (g.mask[b]) {
    const cpx<double>* G = gp + (size_t)o * 4 * g.box + b;
    const cpx<double> ge = G[0], get = G[g.box], gph = G[2 * (size_t)g.box], gpt = G[3 * (size_t)g.box];
    const double A = g.a[b], c = g.coef[i], chv = g.ch[(size_t)i * g.box + b], shv = g.sh[(size_t)i * g.box + b];
    // e = c(-A sh s0 + ch s1), et = -i c(-A sh s2 + ch s3), p = i c(ch s0 - sh s1), pt = c(ch s2 - sh s3)
    r[0] = {c * (-A * shv * ge.x + chv * gph.y), c * (-A * shv * ge.y - chv * gph.x)};
    r[1] = {c * (chv * ge.x - shv * gph.y), c * (chv * ge.y + shv * gph.x)};
    r[2] = {c * (A * shv * get.y + chv * gpt.x), c * (-A * shv * get.x + chv * gpt.y)};
    r[3] = {c * (-chv * get.y - shv * gpt.x), c * (chv * get.x - shv * gpt.y)};
    const size_t bs = g.lay.bin(b);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (!kept(g.spec[((size_t)task * 4 + k) * g.lay.slab + bs], g.floor_level * bits_to_d(g.fmax[(size_t)task * 4 + k])))
        r[k] = {0, 0};
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) gs[((size_t)task * 4 + k) * g.box + b] = r[k];
}

// Adjoint of a column-inverse pass (k_colinv) after its y-DFT: x [ntask][nfield][box]
// holds DFT_y of the column gradients; out[set][src] += sum over the set's tasks
// and fields f of conj(multiplier_f) * damping * fscale_f * x (floored bins excluded).
// Workgroup tile of the collection: kCollectBins consecutive bins; the
// (task, field) terms of each bin are split over NT / kCollectBins thread
// groups and summed through LDS in a fixed order (deterministic).
constexpr int kCollectBins = 16;
constexpr int kCollectGroups = NT / kCollectBins;

__global__ __launch_bounds__(NT) void k_adj_collect(ColInvArgs a, const cpx<double>* x, cpx<double>* out) {
  __shared__ cpx<double> part[kCollectGroups][4][kCollectBins];
  const int lb = threadIdx.x % kCollectBins, grp = threadIdx.x / kCollectBins;
  const int b = blockIdx.x * kCollectBins + lb;
  const int set = blockIdx.y;
  const int box = a.wx * a.wy;
  const int nsrc = a.nsrc < 4 ? a.nsrc : 4;
  // source s of field f is summed in acc_s; selects keep the four sums in registers
  cpx<double> acc0 = {0, 0}, acc1 = {0, 0}, acc2 = {0, 0}, acc3 = {0, 0};
  if (b < box) {
    const int nterm = a.sdiv * a.nfield;
    for (int it = grp; it < nterm; it += kCollectGroups) {
      const int u = it / a.nfield, f = it - u * a.nfield;
      const int task = set * a.sdiv + u;
      const double dmp = a.damp ? a.damp[(size_t)((task % a.dmod) * a.dmul + a.dadd) * box + b] : 1.0;
      const cpx<double> m = a.tab.mul[(size_t)a.fcode[f] * box + b];
      const cpx<double> v = x[((size_t)task * a.nfield + f) * box + b];
      const cpx<double> w = conj_mul(m, v);
      const double sc = dmp * a.fscale[f];
      const cpx<double> t = {sc * w.x, sc * w.y};
      const int src = a.fsrc[f];
      if (src == 0) acc0 = cadd(acc0, t);
      else if (src == 1) acc1 = cadd(acc1, t);
      else if (src == 2) acc2 = cadd(acc2, t);
      else acc3 = cadd(acc3, t);
    }
  }
  part[grp][0][lb] = acc0;
  part[grp][1][lb] = acc1;
  part[grp][2][lb] = acc2;
  part[grp][3][lb] = acc3;
  __syncthreads();
  // threads (s, lb) of the first 4 x kCollectBins sum their column over the groups
  if (threadIdx.x < 4 * kCollectBins) {
    const int s = threadIdx.x / kCollectBins, l = threadIdx.x % kCollectBins;
    const int bb = blockIdx.x * kCollectBins + l;
    if (s < nsrc && bb < box) {
      cpx<double> sum = {0, 0};
      for (int g = 0; g < kCollectGroups; ++g) sum = cadd(sum, part[g][s][l]);
      bool keep = true;
      if (a.fmax && ((a.floor_src_mask >> s) & 1)) {
        const size_t k = (size_t)set * a.fmax_stride + (a.fmax_stride == 1 ? 0 : s);
        keep = kept(a.spec[((size_t)set * a.nsrc + s) * box + bb], a.floor_level * bits_to_d(a.fmax[k]));
      }
      cpx<double>* o = out + ((size_t)set * a.nsrc + s) * box + bb;
      if (keep) *o = cadd(*o, sum);
    }
  }
}

// Phase adjoint: g0 = exp(+i omega t) g
__global__ void k_adj_phase(const cpx<double>* g, const double* omega, const double* times, int box, cpx<double>* g0) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= box) return;
  double sn, cs;
  sincos(omega[b] * times[0], &sn, &cs);
  const cpx<double> v = g[b];
  g0[b] = {v.x * cs - v.y * sn, v.y * cs + v.x * sn};
}

__global__ void k_fill_one_u8(uint8_t* p, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 1;
}

__global__ void k_fill_zero_c(cpx<double>* p, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = {0.0, 0.0};
}

// ----------------------------------------------------------------------------
// Plan
// ----------------------------------------------------------------------------
struct DevTables {
  double *qx = nullptr, *qy = nullptr, *nu = nullptr, *inu = nullptr, *q2 = nullptr, *a = nullptr;
  cpx<double>* mul = nullptr;
  BoxTables view() const { return BoxTables{qx, qy, nu, inu, q2, a, mul}; }
};

// mul[code][b] = apply_mul(code, 1) on every bin of a box.
__global__ void k_mul_table(BoxTables t, int n, cpx<double>* mul) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int code = blockIdx.y;
  if (b < n) mul[(size_t)code * n + b] = apply_mul(code, cpx<double>{1.0, 0.0}, t, b);
}

int box_size(const glw_box& b) { return b.wx * b.wy; }

}  // namespace

int glw_report_error(int code, const std::string& msg) { return fail(code, msg); }

struct glw_plan_s {
  glw_plan_desc d{};
  int device = 0;
  std::vector<void*> allocs;
  std::map<int, void*> tw64;
  // direct-DFT synthesis tables of output lengths that are not FFT lengths:
  // (first signed mode, count, length) -> [count][length] exp(2 pi i m n / length)
  std::map<std::tuple<int, int, int>, cpx<double>*> dft;
  DevTables t1, t2, t3;
  double *omega = nullptr, *damp_o = nullptr, *damp_oi = nullptr;
  uint8_t *mask2 = nullptr, *mask3 = nullptr;
  double *in_ch = nullptr, *in_sh = nullptr, *in_coef = nullptr;
  double *out_ch = nullptr, *out_sh = nullptr, *out_coef = nullptr;
  // order 4
  DevTables t4;
  uint8_t* mask4 = nullptr;
  double *o4_dq = nullptr, *o4_dt = nullptr, *o4_dp3 = nullptr, *o4_dp4 = nullptr;
  double *o4_pch = nullptr, *o4_psh = nullptr, *o4_pco = nullptr;
  double *o4_tch = nullptr, *o4_tsh = nullptr, *o4_tco = nullptr;
  double *o4_qch = nullptr, *o4_qsh = nullptr, *o4_qco = nullptr;
  // order 2
  double *e22_damp = nullptr, *e22_ch = nullptr, *e22_sh = nullptr, *e22_coef = nullptr;
  double *p22_damp = nullptr, *p22_fac = nullptr, *p22_coef = nullptr, *p22_sign = nullptr;
  // workspace: name -> (ptr, bytes)
  std::map<std::string, std::pair<void*, size_t>> ws;
  // per-stage schedule (absent: plan 0, tile 0) and the kind of every stage seen
  std::map<std::string, Sched> sched;
  std::map<std::string, int> stage_kind;
  const char* cur = "";
  hipStream_t cur_stream = nullptr;
  // scratch of the split row path, grown on demand outside graph capture;
  // while holds > 0 (captured graphs that replay this plan) it never moves
  void* split_ptr = nullptr;
  size_t split_bytes = 0;
  int holds = 0;
  // stage timing
  std::vector<hipEvent_t> ev;
  std::vector<const char*> ev_name;
  std::vector<double> ev_bytes;  // compulsory HBM bytes of each timed launch
  size_t ev_used = 0;
  size_t ev_dropped = 0;

  ~glw_plan_s() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
    if (split_ptr) (void)hipFree(split_ptr);
  }

  static int scratch(void* ctx, size_t bytes, void** ptr) {
    glw_plan_s* P = static_cast<glw_plan_s*>(ctx);
    if (bytes > P->split_bytes) {
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      GLW_HIP(hipStreamIsCapturing(P->cur_stream, &cs));
      if (cs != hipStreamCaptureStatusNone)
        return fail(GLW_EINVAL, "split row scratch must be sized before graph capture: execute the plan once first");
      if (P->holds > 0)
        return fail(GLW_EINVAL, "split row scratch would move under a captured graph that holds this plan "
                                "(size it before capture, or release the hold first)");
      if (P->split_ptr) GLW_HIP(hipFree(P->split_ptr));  // hipFree waits for work in flight
      P->split_ptr = nullptr;
      P->split_bytes = 0;
      const size_t rounded = (bytes + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
      GLW_HIP(hipMalloc(&P->split_ptr, rounded));
      P->split_bytes = rounded;
    }
    *ptr = P->split_ptr;
    return GLW_OK;
  }

  // Schedule of the stage being launched (named by the last mark_begin).
  Exec exec(int kind) {
    stage_kind[cur] = kind;
    Exec x;
    auto it = sched.find(cur);
    if (it != sched.end()) x.s = it->second;
    x.scratch = &glw_plan_s::scratch;
    x.ctx = this;
    return x;
  }

  // Bracket one launch with two events (no-op unless profiling is enabled).
  void mark_begin(const char* name, hipStream_t st, double bytes) {
    cur = name;
    if (ev.empty()) return;
    if (ev_used + 2 > ev.size()) {
      ++ev_dropped;
      return;
    }
    ev_name[ev_used / 2] = name;
    ev_bytes[ev_used / 2] = bytes;
    (void)hipEventRecord(ev[ev_used], st);
  }
  void mark_end(hipStream_t st) {
    if (ev.empty() || ev_used + 2 > ev.size()) return;
    (void)hipEventRecord(ev[ev_used + 1], st);
    ev_used += 2;
  }

  template <typename X>
  int upload(X** dst, const X* src, size_t n, bool required = true) {
    *dst = nullptr;
    if (!src) {
      if (required) return fail(GLW_EINVAL, "missing required plan table");
      return GLW_OK;
    }
    void* p = nullptr;
    GLW_HIP(hipMalloc(&p, n * sizeof(X)));
    allocs.push_back(p);
    GLW_HIP(hipMemcpy(p, src, n * sizeof(X), hipMemcpyHostToDevice));
    *dst = (X*)p;
    return GLW_OK;
  }

  int upload_tables(DevTables& dt, const glw_box_tables& h, size_t n) {
    int rc;
    if ((rc = upload(&dt.qx, h.qx, n))) return rc;
    if ((rc = upload(&dt.qy, h.qy, n))) return rc;
    if ((rc = upload(&dt.nu, h.nu, n))) return rc;
    if ((rc = upload(&dt.inu, h.inu, n))) return rc;
    if ((rc = upload(&dt.q2, h.q2, n))) return rc;
    if ((rc = upload(&dt.a, h.a, n))) return rc;
    void* p = nullptr;
    GLW_HIP(hipMalloc(&p, n * kNumMulCodes * sizeof(cpx<double>)));
    allocs.push_back(p);
    dt.mul = (cpx<double>*)p;
    hipLaunchKernelGGL(k_mul_table, dim3((unsigned)((n + 255) / 256), kNumMulCodes), dim3(256), 0, 0, dt.view(),
                       (int)n, dt.mul);
    GLW_HIP(hipGetLastError());
    GLW_HIP(hipDeviceSynchronize());
    return GLW_OK;
  }

  int twiddles(int n) {
    if (tw64.count(n)) return GLW_OK;
    std::vector<cpx<double>> w(n);
    for (int m = 0; m < n; ++m) {
      long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
      w[m] = {(double)cosl(ang), (double)sinl(ang)};
    }
    cpx<double>* p = nullptr;
    int rc = upload(&p, w.data(), n);
    if (rc) return rc;
    tw64[n] = p;
    return GLW_OK;
  }

  int reserve(const char* name, size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) bytes = 256;
    GLW_HIP(hipMalloc(&p, bytes));
    allocs.push_back(p);
    ws[name] = {p, bytes};
    return GLW_OK;
  }
  template <typename X>
  X* buf(const char* name) {
    return (X*)ws.at(name).first;
  }
  // FFT twiddles exp(-2 pi i m / n), FP64 (every GL plan is FP64)
  const void* tw(int n, bool /*fp32*/) { return tw64.at(n); }

  // exp(+2 pi i m k / n) for m = m0 .. m0 + count - 1 and k < n, from the exact integer residue
  // (m k) mod n; built once per (box axis, length) at plan creation
  int dft_table(int m0, int count, int n) {
    const auto key = std::make_tuple(m0, count, n);
    if (dft.count(key)) return GLW_OK;
    std::vector<cpx<double>> t((size_t)count * n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int i = 0; i < count; ++i) {
      const long long m = (((long long)(m0 + i) % n) + n) % n;
      for (int k = 0; k < n; ++k) {
        const long double ang = two_pi * (long double)((m * k) % n) / (long double)n;
        t[(size_t)i * n + k] = {(double)cosl(ang), (double)sinl(ang)};
      }
    }
    cpx<double>* p = nullptr;
    int rc = upload(&p, t.data(), t.size());
    if (rc) return rc;
    dft[key] = p;
    return GLW_OK;
  }
  const cpx<double>* dft_at(int m0, int count, int n) const { return dft.at(std::make_tuple(m0, count, n)); }
};

namespace {

int validate(const glw_plan_desc* d) {
  if (!d) return fail(GLW_EINVAL, "null descriptor");
  if (d->nx < 8 || d->ny < 8 || d->nx % 2 || d->ny % 2) return fail(GLW_EINVAL, "grid must be even and >= 8");
  if (d->nx > kMaxOutputLength || d->ny > kMaxOutputLength)
    return fail(GLW_EUNSUPPORTED, "output grid longer than " + std::to_string(kMaxOutputLength) + " per side");
  if (d->order < 1 || d->order > 3) return fail(GLW_EINVAL, "order must be 1, 2 or 3");
  if (d->precision == GLW_PREC_FP32)
    return fail(GLW_EUNSUPPORTED,
                "FP32 transforms are refused for GL plans: FP32 rounding of the intermediate spectra, amplified by "
                "the GL node factors, exceeds relative L2 1e-5 against FP64 (eta22, eta33, psi33, eta44); use FP64");
  if (d->precision != GLW_PREC_FP64) return fail(GLW_EINVAL, "bad precision");
  if (d->max_instants < 1) return fail(GLW_EINVAL, "max_instants must be positive");
  auto box_ok = [&](const glw_box& b, int gx, int gy) {
    return b.wx >= 1 && b.wy >= 1 && b.wx <= gx && b.wy <= gy;
  };
  if (!box_ok(d->p1, d->nx, d->ny)) return fail(GLW_EINVAL, "parent box does not fit the grid");
  if (d->order >= 2) {
    if (!length_supported(d->g2x) || !length_supported(d->g2y))
      return fail(GLW_EUNSUPPORTED, "quadratic working grid length not supported");
    if (2 * d->p1.wx - 1 > d->g2x || 2 * d->p1.wy - 1 > d->g2y)
      return fail(GLW_EINVAL, "quadratic working grid narrower than the pair support");
    if (!box_ok(d->p2, d->g2x, d->g2y) || !box_ok(d->p2, d->nx, d->ny))
      return fail(GLW_EINVAL, "pair box does not fit the working grid");
  }
  if (d->order >= 3) {
    if (!length_supported(d->g3x) || !length_supported(d->g3y))
      return fail(GLW_EUNSUPPORTED, "cubic working grid length not supported");
    if (3 * d->p1.wx - 2 > d->g3x || 3 * d->p1.wy - 2 > d->g3y)
      return fail(GLW_EINVAL, "cubic working grid narrower than the triple support");
    if (!box_ok(d->p3, d->g3x, d->g3y) || !box_ok(d->p3, d->nx, d->ny))
      return fail(GLW_EINVAL, "triple box does not fit the working grid");
    if (d->inner_rank < 1 || d->inner_rank > 16 || d->outer_rank < 1 || d->outer_rank > 16)
      return fail(GLW_EINVAL, "GL ranks must be 1..16");
    if (d->p2.wx > d->g3x || d->p2.wy > d->g3y) return fail(GLW_EINVAL, "pair box wider than the cubic grid");
  }
  if (d->with_order4) {
    const glw_order4_desc& o = d->o4;
    if (o.rank < 1 || o.rank > 12) return fail(GLW_EINVAL, "order-4 GL rank must be 1..12");
    if (o.forcing_terms != 0 && o.forcing_terms != 1) return fail(GLW_EINVAL, "order-4 forcing_terms must be 0 or 1");
    if (!length_supported(d->g2x) || !length_supported(d->g2y) || !length_supported(d->g3x) ||
        !length_supported(d->g3y) || !length_supported(o.g4x) || !length_supported(o.g4y))
      return fail(GLW_EUNSUPPORTED, "order-4 working grid length not supported");
    if (2 * d->p1.wx - 1 > d->g2x || 2 * d->p1.wy - 1 > d->g2y || 3 * d->p1.wx - 2 > d->g3x ||
        3 * d->p1.wy - 2 > d->g3y || 4 * d->p1.wx - 3 > o.g4x || 4 * d->p1.wy - 3 > o.g4y)
      return fail(GLW_EINVAL, "order-4 working grids narrower than the product supports");
    if (!box_ok(d->p2, d->g2x, d->g2y) || !box_ok(d->p3, d->g3x, d->g3y) || !box_ok(o.p4, o.g4x, o.g4y) ||
        !box_ok(o.p4, d->nx, d->ny))
      return fail(GLW_EINVAL, "order-4 boxes do not fit the working grids");
    if (d->p2.wx > o.g4x || d->p2.wy > o.g4y || d->p3.wx > o.g4x || d->p3.wy > o.g4y || d->p2.wx > d->g3x ||
        d->p2.wy > d->g3y)
      return fail(GLW_EINVAL, "lower-order boxes wider than the higher-order grids");
    if ((long long)o.rank * o.rank * o.rank * d->max_instants > 65535)
      return fail(GLW_EINVAL, "order-4 task count exceeds the launch grid (lower max_instants)");
  }
  if (d->order >= 2 && (d->o2.eta_rank < 1 || d->o2.eta_rank > 16 || d->o2.channels != 2))
    return fail(GLW_EINVAL, "order-2 tables: eta rank 1..16 and two psi22 channels required");
  return GLW_OK;
}

int build(glw_plan* P) {
  const glw_plan_desc& d = P->d;
  const size_t b1 = box_size(d.p1), b2 = box_size(d.p2), b3 = box_size(d.p3);
  const size_t csz = sizeof(cpx<double>);
  const size_t zsz = sizeof(cpx<double>);
  const size_t nt = d.max_instants;
  int rc;
  if ((rc = P->upload_tables(P->t1, d.t1, b1))) return rc;
  if ((rc = P->upload(&P->omega, d.omega, b1))) return rc;
  // output synthesis: FFT twiddles, or direct-DFT tables of every box the plan synthesizes
  std::vector<glw_box> out_boxes{d.p1};
  if (d.order >= 2) out_boxes.push_back(d.p2);
  if (d.order >= 3) out_boxes.push_back(d.p3);
  if (d.with_order4) out_boxes.push_back(d.o4.p4);
  for (const glw_box& b : out_boxes) {
    if ((rc = length_supported(d.nx) ? P->twiddles(d.nx) : P->dft_table(b.mx0, b.wx, d.nx))) return rc;
    if ((rc = length_supported(d.ny) ? P->twiddles(d.ny) : P->dft_table(b.my0, b.wy, d.ny))) return rc;
  }
  int wfin = d.p1.wx;
  if ((rc = P->reserve("st", nt * b1 * zsz))) return rc;
  if (d.order >= 2) {
    wfin = std::max(wfin, (int)d.p2.wx);
    if ((rc = P->upload_tables(P->t2, d.t2, b2))) return rc;
    if ((rc = P->upload(&P->mask2, d.mask2, b2))) return rc;
    if ((rc = P->twiddles(d.g2x))) return rc;
    if ((rc = P->twiddles(d.g2y))) return rc;
    const int J = d.o2.eta_rank;
    if ((rc = P->upload(&P->e22_damp, d.o2.eta_damp, (size_t)J * b1))) return rc;
    if ((rc = P->upload(&P->e22_ch, d.o2.eta_cosh, (size_t)J * b2))) return rc;
    if ((rc = P->upload(&P->e22_sh, d.o2.eta_sinhr, (size_t)J * b2))) return rc;
    if ((rc = P->upload(&P->e22_coef, d.o2.eta_coef, (size_t)J))) return rc;
    if ((rc = P->upload(&P->p22_damp, d.o2.psi_damp, (size_t)8 * b1))) return rc;
    if ((rc = P->upload(&P->p22_fac, d.o2.psi_factor, (size_t)16 * b2))) return rc;
    if ((rc = P->upload(&P->p22_coef, d.o2.psi_coef, (size_t)4))) return rc;
    if ((rc = P->upload(&P->p22_sign, d.o2.psi_sign, (size_t)2))) return rc;
    const size_t g2 = (size_t)d.g2y;
    if ((rc = P->reserve("e22_t", nt * J * 8 * g2 * sw(d.p1.wx) * csz))) return rc;
    if ((rc = P->reserve("e22_u", nt * J * 2 * g2 * sw(d.p2.wx) * csz))) return rc;
    if ((rc = P->reserve("e22_s", nt * J * 2 * spec_layout(d.p2).slab * zsz))) return rc;
    if ((rc = P->reserve("e22_max", nt * J * 2 * sizeof(uint64_t)))) return rc;
    if ((rc = P->reserve("p22_t", nt * 4 * 16 * g2 * sw(d.p1.wx) * csz))) return rc;
    if ((rc = P->reserve("p22_u", nt * 4 * 8 * g2 * sw(d.p2.wx) * csz))) return rc;
    if ((rc = P->reserve("p22_s", nt * 4 * 8 * spec_layout(d.p2).slab * zsz))) return rc;
    if ((rc = P->reserve("c22_t", nt * 2 * g2 * sw(d.p1.wx) * csz))) return rc;
    if ((rc = P->reserve("c22_u", nt * 1 * g2 * sw(d.p2.wx) * csz))) return rc;
    if ((rc = P->reserve("c22_s", nt * 1 * spec_layout(d.p2).slab * zsz))) return rc;
    if ((rc = P->reserve("ep22", nt * 2 * b2 * zsz))) return rc;
  }
  if (d.order >= 3) {
    wfin = std::max(wfin, (int)d.p3.wx);
    const size_t I = d.inner_rank, O = d.outer_rank;
    if ((rc = P->upload_tables(P->t3, d.t3, b3))) return rc;
    if ((rc = P->upload(&P->mask3, d.mask3, b3))) return rc;
    if ((rc = P->twiddles(d.g3x))) return rc;
    if ((rc = P->twiddles(d.g3y))) return rc;
    if ((rc = P->upload(&P->damp_o, d.damp_outer, O * b1))) return rc;
    if ((rc = P->upload(&P->damp_oi, d.damp_inner, O * I * b1))) return rc;
    if ((rc = P->upload(&P->in_ch, d.inner_cosh, I * b2))) return rc;
    if ((rc = P->upload(&P->in_sh, d.inner_sinhr, I * b2))) return rc;
    if ((rc = P->upload(&P->in_coef, d.inner_coef, I))) return rc;
    if ((rc = P->upload(&P->out_ch, d.outer_cosh, O * b3))) return rc;
    if ((rc = P->upload(&P->out_sh, d.outer_sinhr, O * b3))) return rc;
    if ((rc = P->upload(&P->out_coef, d.outer_coef, O))) return rc;
    const size_t g2 = d.g2y, g3 = d.g3y;
    const size_t wmax = sw(std::max(d.p1.wx, d.p2.wx));
    if ((rc = P->reserve("tin", nt * O * I * 12 * g2 * sw(d.p1.wx) * csz))) return rc;
    if ((rc = P->reserve("uin", nt * O * I * 4 * g2 * sw(d.p2.wx) * csz))) return rc;
    if ((rc = P->reserve("sin", nt * O * I * 4 * spec_layout(d.p2).slab * zsz))) return rc;
    if ((rc = P->reserve("max_in", nt * O * I * 4 * sizeof(uint64_t)))) return rc;
    if ((rc = P->reserve("pair", nt * O * 4 * b2 * zsz))) return rc;
    if ((rc = P->reserve("max_et", nt * O * sizeof(uint64_t)))) return rc;
    if ((rc = P->reserve("tout", nt * O * 21 * g3 * wmax * csz))) return rc;
    if ((rc = P->reserve("uout", nt * O * 4 * g3 * sw(d.p3.wx) * csz))) return rc;
    if ((rc = P->reserve("sout", nt * O * 4 * spec_layout(d.p3).slab * zsz))) return rc;
    if ((rc = P->reserve("max_out", nt * O * 4 * sizeof(uint64_t)))) return rc;
    if ((rc = P->reserve("max_combo", nt * O * 2 * sizeof(uint64_t)))) return rc;
    if ((rc = P->reserve("ep33", nt * 2 * b3 * zsz))) return rc;
  }
  if (d.with_order4) {
    const glw_order4_desc& o = d.o4;
    const size_t J = o.rank, b4 = box_size(o.p4);
    wfin = std::max({wfin, (int)d.p2.wx, (int)o.p4.wx});
    if (d.order < 2) {  // order-4-only plan: pair tables and grids
      if ((rc = P->upload_tables(P->t2, d.t2, b2))) return rc;
      if ((rc = P->upload(&P->mask2, d.mask2, b2))) return rc;
      if ((rc = P->twiddles(d.g2x))) return rc;
      if ((rc = P->twiddles(d.g2y))) return rc;
    }
    if (d.order < 3) {
      if ((rc = P->upload_tables(P->t3, d.t3, b3))) return rc;
      if ((rc = P->upload(&P->mask3, d.mask3, b3))) return rc;
      if ((rc = P->twiddles(d.g3x))) return rc;
      if ((rc = P->twiddles(d.g3y))) return rc;
    }
    if ((rc = P->upload_tables(P->t4, o.t4, b4))) return rc;
    if ((rc = P->upload(&P->mask4, o.mask4, b4))) return rc;
    if ((rc = P->twiddles(o.g4x))) return rc;
    if ((rc = P->twiddles(o.g4y))) return rc;
    if ((rc = P->upload(&P->o4_dq, o.damp_quartic, J * b1))) return rc;
    if ((rc = P->upload(&P->o4_dt, o.damp_third, J * J * b1))) return rc;
    if ((rc = P->upload(&P->o4_dp3, o.damp_pair3, J * J * J * b1))) return rc;
    if ((rc = P->upload(&P->o4_dp4, o.damp_pair4, J * J * b1))) return rc;
    if ((rc = P->upload(&P->o4_pch, o.pair_cosh, J * b2))) return rc;
    if ((rc = P->upload(&P->o4_psh, o.pair_sinhr, J * b2))) return rc;
    if ((rc = P->upload(&P->o4_pco, o.pair_coef, J))) return rc;
    if ((rc = P->upload(&P->o4_tch, o.third_cosh, J * b3))) return rc;
    if ((rc = P->upload(&P->o4_tsh, o.third_sinhr, J * b3))) return rc;
    if ((rc = P->upload(&P->o4_tco, o.third_coef, J))) return rc;
    if ((rc = P->upload(&P->o4_qch, o.quartic_cosh, J * b4))) return rc;
    if ((rc = P->upload(&P->o4_qsh, o.quartic_sinhr, J * b4))) return rc;
    if ((rc = P->upload(&P->o4_qco, o.quartic_coef, J))) return rc;
    const size_t g2 = d.g2y, g3 = d.g3y, g4 = o.g4y;
    const size_t w12 = sw(std::max(d.p1.wx, d.p2.wx)), w123 = sw(std::max<int>((int)w12, d.p3.wx));
    const size_t np = nt * J * J * J;  // pair tasks of the third level (the larger pair batch)
    if ((rc = P->reserve("o4_pt", np * 16 * g2 * sw(d.p1.wx) * csz))) return rc;
    if ((rc = P->reserve("o4_pu", np * 6 * g2 * sw(d.p2.wx) * csz))) return rc;
    if ((rc = P->reserve("o4_ps", np * 6 * spec_layout(d.p2).slab * zsz))) return rc;
    if ((rc = P->reserve("o4_pm", np * 6 * sizeof(uint64_t)))) return rc;
    if ((rc = P->reserve("o4_set3", nt * J * J * 5 * b2 * zsz))) return rc;
    if ((rc = P->reserve("o4_max3", nt * J * J * 5 * sizeof(uint64_t)))) return rc;
    if ((rc = P->reserve("o4_set4", nt * J * 5 * b2 * zsz))) return rc;
    if ((rc = P->reserve("o4_max4", nt * J * 5 * sizeof(uint64_t)))) return rc;
    if ((rc = P->reserve("o4_tt", nt * J * J * 31 * g3 * w12 * csz))) return rc;
    if ((rc = P->reserve("o4_tu", nt * J * J * 4 * g3 * sw(d.p3.wx) * csz))) return rc;
    if ((rc = P->reserve("o4_ts", nt * J * J * 4 * spec_layout(d.p3).slab * zsz))) return rc;
    if ((rc = P->reserve("o4_tm", nt * J * J * 4 * sizeof(uint64_t)))) return rc;
    if ((rc = P->reserve("o4_th", nt * J * 3 * b3 * zsz))) return rc;
    if ((rc = P->reserve("o4_thm", nt * J * 3 * sizeof(uint64_t)))) return rc;
    if ((rc = P->reserve("o4_qt", nt * J * 45 * g4 * w123 * csz))) return rc;
    if ((rc = P->reserve("o4_qu", nt * J * 2 * g4 * sw(o.p4.wx) * csz))) return rc;
    if ((rc = P->reserve("o4_qs", nt * J * 2 * spec_layout(o.p4).slab * zsz))) return rc;
    if ((rc = P->reserve("o4_qm", nt * J * 2 * sizeof(uint64_t)))) return rc;
    if ((rc = P->reserve("o4_out", nt * b4 * zsz))) return rc;
    if ((rc = P->reserve("o4_outm", nt * sizeof(uint64_t)))) return rc;
  }
  if ((rc = P->reserve("tfin", nt * 2 * (size_t)d.ny * sw(wfin) * zsz))) return rc;
  return GLW_OK;
}

// ----------------------------------------------------------------------------
// Compulsory HBM bytes of one launch on the reduced grids: every input
// element the launch consumes and every output element it produces, once.
// Tables that all tasks share (multipliers, damping rows, node factors, masks,
// twiddles) are excluded, and a spectrum that several tasks read (one
// spectrum set per `sdiv` tasks) counts once per set.  These are the
// algorithmic bytes of the roofline: a stage's bytes over its launch time,
// against the HBM peak.
// ----------------------------------------------------------------------------
constexpr double kZ = sizeof(cpx<double>);  // bytes per complex value

double colinv_bytes(const ColInvArgs& a, int gy, int ntask) {
  bool used[kMaxFields] = {};
  int nused = 0;
  for (int f = 0; f < a.nfield; ++f)
    if (!used[a.fsrc[f]]) {
      used[a.fsrc[f]] = true;
      ++nused;
    }
  const double sets = std::ceil((double)ntask / a.sdiv);
  return kZ * (sets * nused * (double)a.wx * a.wy + (double)ntask * a.nfield * gy * a.wx);
}

double row_bytes(const RowArgs& r, int ntask) {
  double cols = 0.0;
  for (int f = 0; f < r.nin; ++f) cols += r.fwx[f];
  for (int o = 0; o < r.nout; ++o) cols += r.out_per_field ? r.fwx[o] : r.owx;
  return kZ * ntask * (double)r.gy * cols;
}

double colfwd_bytes(const ColFwdArgs& c, int gy, int ntask) {
  return kZ * ntask * (double)c.nsrc * c.owx * ((double)gy + c.owy);
}

double synth_bytes(const SynthArgs& s, int gx, int nseq) { return kZ * nseq * (double)s.gy * (s.wx + gx); }

// split-path forward row pass: full physical rows in, box columns out
double rowfwd_bytes(const RowArgs& r, int gx, int ntask) {
  double cols = 0.0;
  for (int o = 0; o < r.nout; ++o) cols += gx + (r.out_per_field ? r.fwx[o] : r.owx);
  return kZ * ntask * (double)r.gy * cols;
}

// row adjoint: forward inputs and output gradients in, input gradients out
double rowadj_bytes(const RowArgs& fwd, const RowArgs& gin, const RowArgs& gout, int ntask) {
  return row_bytes(fwd, ntask) - kZ * ntask * (double)fwd.gy * fwd.nout * fwd.owx + row_bytes(gin, ntask) +
         row_bytes(gout, ntask);
}

int fill_zero(uint64_t* p, size_t n, hipStream_t st) {
  if (n == 0) return GLW_OK;
  hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  GLW_HIP(hipGetLastError());
  return GLW_OK;
}

// Launch with stage-timing brackets; evaluates to the launch's return code.
#define GLW_STAGE(P, name, st, bytes, expr) \
  ([&]() -> int {                           \
    (P)->mark_begin(name, st, (bytes));     \
    int rc_ = (expr);                       \
    (P)->mark_end(st);                      \
    return rc_;                             \
  }())

int launch_plain(hipError_t e) {
  if (e != hipSuccess) return fail(GLW_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  return GLW_OK;
}

ColInvArgs colinv_base(const cpx<double>* spec, int nsrc, int sdiv, const glw_box& box, const DevTables& t) {
  ColInvArgs a;
  memset(&a, 0, sizeof(a));
  a.spec = spec;
  a.nsrc = nsrc;
  a.sdiv = sdiv;
  a.mx0 = box.mx0;
  a.my0 = box.my0;
  a.wx = box.wx;
  a.wy = box.wy;
  a.tab = t.view();
  a.dmod = 1;
  a.dmul = 1;
  a.fmax_stride = nsrc;
  return a;
}

// Spectra on `box` ([nt][nsrc][box]) -> physical [nt][ny][nx] fields, one per
// (src, scale, multiplier) triple.
int synthesize(glw_plan* P, const cpx<double>* spec, int nsrc, const glw_box& box, const DevTables& tab,
               int nt, const int* srcs, const int* codes, const double* scales, double** outs, int nout,
               hipStream_t st, const uint64_t* fmax = nullptr, int fmax_stride = 1, int floor_mask = 0) {
  const glw_plan_desc& d = P->d;
  ColInvArgs a = colinv_base(spec, nsrc, 1, box, tab);
  a.fmax = fmax;
  a.fmax_stride = fmax_stride;
  a.floor_src_mask = floor_mask;
  a.floor_level = d.floor_level;
  a.nfield = nout;
  for (int f = 0; f < nout; ++f) {
    a.fsrc[f] = srcs[f];
    a.fcode[f] = codes[f];
    a.fscale[f] = scales[f];
  }
  cpx<double>* tfin = P->buf<cpx<double>>("tfin");
  a.out = tfin;
  a.out_nfield = nout;
  a.out_foff = 0;
  a.ostride = sw(box.wx);
  int rc;
  if (length_supported(d.ny)) {
    rc = GLW_STAGE(P, "synth_colinv", st, colinv_bytes(a, d.ny, nt),
                   launch_colinv<double>(d.ny, a, nt, P->tw(d.ny, false), P->exec(STAGE_COL), st));
  } else {  // direct DFT: kDftCols box columns per workgroup, all box rows in LDS
    const size_t shm = (size_t)kDftCols * box.wy * sizeof(cpx<double>);
    if (shm > (size_t)160 * 1024) return fail(GLW_EUNSUPPORTED, "spectral box too tall for the direct-DFT synthesis");
    if (shm > (size_t)64 * 1024)
      GLW_HIP(hipFuncSetAttribute((const void*)k_dft_colinv, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    P->mark_begin("synth_dft_col", st, colinv_bytes(a, d.ny, nt));
    hipLaunchKernelGGL(k_dft_colinv, dim3((box.wx + kDftCols - 1) / kDftCols, nout, nt), dim3(NT), shm, st, a, d.ny,
                       P->dft_at(box.my0, box.wy, d.ny));
    P->mark_end(st);
    rc = launch_plain(hipGetLastError());
  }
  if (rc) return rc;
  for (int f = 0; f < nout; ++f) {
    // rows of field f of every instant: sequence t starts at (t*nout + f)*ny*wx
    SynthArgs s;
    s.in = tfin + (size_t)f * d.ny * sw(box.wx);
    s.in_seq = (size_t)nout * d.ny * sw(box.wx);
    s.istride = sw(box.wx);
    s.mx0 = box.mx0;
    s.wx = box.wx;
    s.gy = d.ny;
    s.out = (cpx<double>*)outs[f];
    if (length_supported(d.nx)) {
      if ((rc = GLW_STAGE(P, "synth_row", st, synth_bytes(s, d.nx, nt), launch_synth(d.nx, s, nt, P->tw(d.nx, false), st))))
        return rc;
    } else {  // direct DFT: one row per workgroup, its box columns in LDS
      const size_t shm = (size_t)box.wx * sizeof(cpx<double>);
      if (shm > (size_t)64 * 1024)
        GLW_HIP(hipFuncSetAttribute((const void*)k_dft_row, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
      P->mark_begin("synth_dft_row", st, synth_bytes(s, d.nx, nt));
      hipLaunchKernelGGL(k_dft_row, dim3(d.ny, nt), dim3(NT), shm, st, s, d.nx, P->dft_at(box.mx0, box.wx, d.nx));
      P->mark_end(st);
      if ((rc = launch_plain(hipGetLastError()))) return rc;
    }
  }
  return GLW_OK;
}

int phase(glw_plan* P, const double* s0, const double* times, int nt, hipStream_t st) {
  const int b1 = box_size(P->d.p1);
  P->mark_begin("phase", st, kZ * (1.0 + nt) * b1);
  hipLaunchKernelGGL(k_phase, dim3((b1 + 255) / 256, nt), dim3(256), 0, st, (const cpx<double>*)s0, P->omega,
                     times, b1, P->buf<cpx<double>>("st"));
  P->mark_end(st);
  GLW_HIP(hipGetLastError());
  return GLW_OK;
}

template <typename T>
int order3_chunk(glw_plan* P, int nt, double* eta, double* psi, hipStream_t st) {
  const glw_plan_desc& d = P->d;
  const bool f32 = std::is_same<T, float>::value;
  const int I = d.inner_rank, O = d.outer_rank;
  const int b2 = box_size(d.p2), b3 = box_size(d.p3);
  const int n_in = nt * O * I, n_to = nt * O;
  int rc;
  const cpx<double>* stp = P->buf<cpx<double>>("st");
  // ---- inner pair stage: tasks (t, o, i)
  {
    ColInvArgs a = colinv_base(stp, 1, O * I, d.p1, P->t1);
    a.damp = P->damp_oi;
    a.dmod = O * I;
    static const int codes[12] = {M_ONE, M_NU, M_NU2, M_NU3, M_QX_NU, M_QY_NU, M_QX, M_QY, M_Q2_NU, M_Q2, M_NU_QX, M_NU_QY};
    const double sc = 1.0 / ((double)d.nx * d.ny);
    a.nfield = 12;
    for (int f = 0; f < 12; ++f) {
      a.fsrc[f] = 0;
      a.fcode[f] = codes[f];
      a.fscale[f] = sc;
    }
    a.out = P->ws.at("tin").first;
    a.out_nfield = 12;
    a.ostride = sw(d.p1.wx);
    if ((rc = GLW_STAGE(P, "o3_inner_colinv", st, colinv_bytes(a, d.g2y, n_in), launch_colinv<T>(d.g2y, a, n_in, P->tw(d.g2y, f32), P->exec(STAGE_COL), st)))) return rc;
  }
  {
    RowArgs r;
    memset(&r, 0, sizeof(r));
    r.in = P->ws.at("tin").first;
    r.nin = 12;
    r.istride = sw(d.p1.wx);
    for (int f = 0; f < 12; ++f) {
      r.fmx0[f] = d.p1.mx0;
      r.fwx[f] = d.p1.wx;
    }
    r.out = P->ws.at("uin").first;
    r.nout = 4;
    r.ostride = sw(d.p2.wx);
    r.omx0 = d.p2.mx0;
    r.owx = d.p2.wx;
    r.gy = d.g2y;
    if ((rc = GLW_STAGE(P, "o3_inner_row", st, row_bytes(r, n_in), (launch_row<T, 12, 4, InnerOp>(d.g2x, r, n_in, P->tw(d.g2x, f32), P->exec(STAGE_ROW), st))))) return rc;
  }
  if ((rc = fill_zero(P->buf<uint64_t>("max_in"), (size_t)n_in * 4, st))) return rc;
  {
    ColFwdArgs c;
    c.in = P->ws.at("uin").first;
    c.nsrc = 4;
    c.istride = sw(d.p2.wx);
    c.omx0 = d.p2.mx0;
    c.owx = d.p2.wx;
    c.omy0 = d.p2.my0;
    c.owy = d.p2.wy;
    c.mask = P->mask2;
    c.scale = (double)d.nx * d.ny / ((double)d.g2x * d.g2y);
    c.out = P->buf<cpx<double>>("sin");
    c.fmax = P->buf<uint64_t>("max_in");
    c.opitch = spec_layout(d.p2).p;
    if ((rc = GLW_STAGE(P, "o3_inner_colfwd", st, colfwd_bytes(c, d.g2y, n_in), launch_colfwd<T>(d.g2y, c, n_in, P->tw(d.g2y, f32), P->exec(STAGE_COL), st)))) return rc;
  }
  if ((rc = fill_zero(P->buf<uint64_t>("max_et"), (size_t)n_to, st))) return rc;
  {
    AccInnerArgs g;
    g.spec = P->buf<cpx<double>>("sin");
    g.lay = spec_layout(d.p2);
    g.fmax = P->buf<uint64_t>("max_in");
    g.floor_level = d.floor_level;
    g.I = I;
    g.box = b2;
    g.a = P->t2.a;
    g.mask = P->mask2;
    g.ch = P->in_ch;
    g.sh = P->in_sh;
    g.coef = P->in_coef;
    g.out = P->buf<cpx<double>>("pair");
    g.etmax = P->buf<uint64_t>("max_et");
    P->mark_begin("o3_inner_acc", st, kZ * n_to * (4.0 * I + 4.0) * b2);
    hipLaunchKernelGGL(k_acc_inner, dim3((b2 + NT - 1) / NT, n_to), dim3(NT), 0, st, g);
    P->mark_end(st);
    GLW_HIP(hipGetLastError());
  }
  // ---- outer stage: tasks (t, o)
  const int wmax = sw(std::max(d.p1.wx, d.p2.wx));
  {
    ColInvArgs a = colinv_base(stp, 1, O, d.p1, P->t1);
    a.damp = P->damp_o;
    a.dmod = O;
    static const int codes[12] = {M_ONE,  M_I_QX, M_I_QY,     M_QX_NU,   M_QY_NU, M_MI_NU,
                                  M_NU_QX, M_NU_QY, M_MI_Q2_NU, M_MI_Q2NU, M_MA,    M_MQ2};
    const double sc = 1.0 / ((double)d.nx * d.ny);
    a.nfield = 12;
    for (int f = 0; f < 12; ++f) {
      a.fsrc[f] = 0;
      a.fcode[f] = codes[f];
      a.fscale[f] = sc;
    }
    a.out = P->ws.at("tout").first;
    a.out_nfield = 21;
    a.out_foff = 0;
    a.ostride = wmax;
    if ((rc = GLW_STAGE(P, "o3_outer_colinv", st, colinv_bytes(a, d.g3y, n_to), launch_colinv<T>(d.g3y, a, n_to, P->tw(d.g3y, f32), P->exec(STAGE_COL), st)))) return rc;
  }
  {
    ColInvArgs a = colinv_base(P->buf<cpx<double>>("pair"), 4, 1, d.p2, P->t2);
    a.fmax = P->buf<uint64_t>("max_et");
    a.fmax_stride = 1;          // one maximum per (t, o)
    a.floor_src_mask = 1 << 1;  // eta_t is floored after accumulation (_order3.py:259)
    a.floor_level = d.floor_level;
    static const int srcs[9] = {0, 0, 0, 2, 2, 2, 2, 3, 1};
    static const int codes[9] = {M_ONE, M_I_QX, M_I_QY, M_I_QX, M_I_QY, M_A, M_Q2, M_A, M_ONE};
    const double sc = 1.0 / ((double)d.nx * d.ny);
    a.nfield = 9;
    for (int f = 0; f < 9; ++f) {
      a.fsrc[f] = srcs[f];
      a.fcode[f] = codes[f];
      a.fscale[f] = sc;
    }
    a.out = P->ws.at("tout").first;
    a.out_nfield = 21;
    a.out_foff = 12;
    a.ostride = wmax;
    if ((rc = GLW_STAGE(P, "o3_outer_colinv", st, colinv_bytes(a, d.g3y, n_to), launch_colinv<T>(d.g3y, a, n_to, P->tw(d.g3y, f32), P->exec(STAGE_COL), st)))) return rc;
  }
  {
    RowArgs r;
    memset(&r, 0, sizeof(r));
    r.in = P->ws.at("tout").first;
    r.nin = 21;
    r.istride = wmax;
    for (int f = 0; f < 21; ++f) {
      r.fmx0[f] = f < 12 ? d.p1.mx0 : d.p2.mx0;
      r.fwx[f] = f < 12 ? d.p1.wx : d.p2.wx;
    }
    r.out = P->ws.at("uout").first;
    r.nout = 4;
    r.ostride = sw(d.p3.wx);
    r.omx0 = d.p3.mx0;
    r.owx = d.p3.wx;
    r.gy = d.g3y;
    if ((rc = GLW_STAGE(P, "o3_outer_row", st, row_bytes(r, n_to), (launch_row<T, 21, 4, OuterOp>(d.g3x, r, n_to, P->tw(d.g3x, f32), P->exec(STAGE_ROW), st))))) return rc;
  }
  if ((rc = fill_zero(P->buf<uint64_t>("max_out"), (size_t)n_to * 4, st))) return rc;
  if ((rc = fill_zero(P->buf<uint64_t>("max_combo"), (size_t)n_to * 2, st))) return rc;
  {
    ColFwdArgs c;
    c.in = P->ws.at("uout").first;
    c.nsrc = 4;
    c.istride = sw(d.p3.wx);
    c.omx0 = d.p3.mx0;
    c.owx = d.p3.wx;
    c.omy0 = d.p3.my0;
    c.owy = d.p3.wy;
    c.mask = P->mask3;
    c.scale = (double)d.nx * d.ny / ((double)d.g3x * d.g3y);
    c.out = P->buf<cpx<double>>("sout");
    c.fmax = P->buf<uint64_t>("max_out");
    c.opitch = spec_layout(d.p3).p;
    if ((rc = GLW_STAGE(P, "o3_outer_colfwd", st, colfwd_bytes(c, d.g3y, n_to), launch_colfwd<T>(d.g3y, c, n_to, P->tw(d.g3y, f32), P->exec(STAGE_COL), st)))) return rc;
  }
  {
    ComboArgs g;
    g.spec = P->buf<cpx<double>>("sout");
    g.lay = spec_layout(d.p3);
    g.fmax = P->buf<uint64_t>("max_out");
    g.floor_level = d.floor_level;
    g.box = b3;
    g.a = P->t3.a;
    g.mask = P->mask3;
    g.cmax = P->buf<uint64_t>("max_combo");
    P->mark_begin("o3_outer_combo", st, kZ * n_to * 4.0 * b3);
    hipLaunchKernelGGL(k_combo_outer, dim3((b3 + NT - 1) / NT, n_to), dim3(NT), 0, st, g);
    P->mark_end(st);
    GLW_HIP(hipGetLastError());
  }
  {
    AccOuterArgs g;
    g.spec = P->buf<cpx<double>>("sout");
    g.lay = spec_layout(d.p3);
    g.fmax = P->buf<uint64_t>("max_out");
    g.cmax = P->buf<uint64_t>("max_combo");
    g.floor_level = d.floor_level;
    g.O = O;
    g.box = b3;
    g.a = P->t3.a;
    g.mask = P->mask3;
    g.ch = P->out_ch;
    g.sh = P->out_sh;
    g.coef = P->out_coef;
    g.out = P->buf<cpx<double>>("ep33");
    P->mark_begin("o3_outer_acc", st, kZ * nt * (4.0 * O + 2.0) * b3);
    hipLaunchKernelGGL(k_acc_outer, dim3((b3 + NT - 1) / NT, nt), dim3(NT), 0, st, g);
    P->mark_end(st);
    GLW_HIP(hipGetLastError());
  }
  const double h = d.depth, pot = d.depth * std::sqrt(d.gravity * d.depth), nn = (double)d.nx * d.ny;
  const int srcs[2] = {0, 1}, codes[2] = {M_ONE, M_ONE};
  const double scales[2] = {h / nn, pot / nn};
  double* outs[2] = {eta, psi};
  return synthesize(P, P->buf<cpx<double>>("ep33"), 2, d.p3, P->t3, nt, srcs, codes, scales, outs, 2, st);
}

template <typename T>
int order2_chunk(glw_plan* P, int nt, double* eta, double* psi, hipStream_t st) {
  const glw_plan_desc& d = P->d;
  const bool f32 = std::is_same<T, float>::value;
  const int J = d.o2.eta_rank;
  const int b2 = box_size(d.p2);
  const double sc = 1.0 / ((double)d.nx * d.ny);
  const double up = (double)d.nx * d.ny / ((double)d.g2x * d.g2y);
  const cpx<double>* stp = P->buf<cpx<double>>("st");
  int rc;
  auto row_args = [&](const char* tin, int nin, const char* uout, int nout) {
    RowArgs r;
    memset(&r, 0, sizeof(r));
    r.in = P->ws.at(tin).first;
    r.nin = nin;
    r.istride = sw(d.p1.wx);
    for (int f = 0; f < nin; ++f) {
      r.fmx0[f] = d.p1.mx0;
      r.fwx[f] = d.p1.wx;
    }
    r.out = P->ws.at(uout).first;
    r.nout = nout;
    r.ostride = sw(d.p2.wx);
    r.omx0 = d.p2.mx0;
    r.owx = d.p2.wx;
    r.gy = d.g2y;
    return r;
  };
  auto fwd_args = [&](const char* uin, int nsrc, const char* sout, uint64_t* fmax, const uint8_t* mask) {
    ColFwdArgs c;
    c.in = P->ws.at(uin).first;
    c.nsrc = nsrc;
    c.istride = sw(d.p2.wx);
    c.omx0 = d.p2.mx0;
    c.owx = d.p2.wx;
    c.omy0 = d.p2.my0;
    c.owy = d.p2.wy;
    c.mask = mask;
    c.scale = up;
    c.out = P->buf<cpx<double>>(sout);
    c.fmax = fmax;
    c.opitch = spec_layout(d.p2).p;
    return c;
  };
  static const int e_codes[8] = {M_ONE, M_NU, M_NU2, M_QX_NU, M_QY_NU, M_Q2_NU, M_QX, M_QY};
  // ---- eta22: tasks (t, j)
  {
    ColInvArgs a = colinv_base(stp, 1, J, d.p1, P->t1);
    a.damp = P->e22_damp;
    a.dmod = J;
    a.nfield = 8;
    for (int f = 0; f < 8; ++f) {
      a.fsrc[f] = 0;
      a.fcode[f] = e_codes[f];
      a.fscale[f] = sc;
    }
    a.out = P->ws.at("e22_t").first;
    a.out_nfield = 8;
    a.ostride = sw(d.p1.wx);
    if ((rc = GLW_STAGE(P, "o2_eta_colinv", st, colinv_bytes(a, d.g2y, nt * J), launch_colinv<T>(d.g2y, a, nt * J, P->tw(d.g2y, f32), P->exec(STAGE_COL), st)))) return rc;
    if ((rc = GLW_STAGE(P, "o2_eta_row", st, row_bytes(row_args("e22_t", 8, "e22_u", 2), nt * J), (launch_row<T, 8, 2, Eta22Op>(d.g2x, row_args("e22_t", 8, "e22_u", 2), nt * J, P->tw(d.g2x, f32), P->exec(STAGE_ROW), st)))
     )) return rc;
    if ((rc = fill_zero(P->buf<uint64_t>("e22_max"), (size_t)nt * J * 2, st))) return rc;
    if ((rc = GLW_STAGE(P, "o2_eta_colfwd", st, colfwd_bytes(fwd_args("e22_u", 2, "e22_s", P->buf<uint64_t>("e22_max"), P->mask2), d.g2y, nt * J), launch_colfwd<T>(d.g2y, fwd_args("e22_u", 2, "e22_s", P->buf<uint64_t>("e22_max"), P->mask2), nt * J,
                               P->tw(d.g2y, f32), P->exec(STAGE_COL), st)))
     ) return rc;
    AccEta22Args g;
    g.spec = P->buf<cpx<double>>("e22_s");
    g.lay = spec_layout(d.p2);
    g.fmax = P->buf<uint64_t>("e22_max");
    g.floor_level = d.floor_level;
    g.J = J;
    g.box = b2;
    g.a = P->t2.a;
    g.ch = P->e22_ch;
    g.sh = P->e22_sh;
    g.coef = P->e22_coef;
    g.out = P->buf<cpx<double>>("ep22");
    P->mark_begin("o2_eta_acc", st, kZ * nt * (2.0 * J + 1.0) * b2);
    hipLaunchKernelGGL(k_acc_eta22, dim3((b2 + NT - 1) / NT, nt), dim3(NT), 0, st, g);
    P->mark_end(st);
    GLW_HIP(hipGetLastError());
  }
  // ---- psi22: tasks (t, branch*2 + node), two channels of 8 fields each
  {
    ColInvArgs a = colinv_base(stp, 1, 4, d.p1, P->t1);
    static const int p_codes[8] = {M_ONE, M_NU, M_NU2, M_QX_NU, M_QY_NU, M_Q2_NU, M_QX, M_QY};
    // damping rows: [branch][node][channel]; a task (t, bn) uses rows 2*bn + ch.
    // colinv picks one damping row per task, so run the two channels as two
    // launches writing field offsets 0 and 8.
    for (int ch = 0; ch < 2; ++ch) {
      a.damp = P->p22_damp;
      a.dmod = 4;
      a.dmul = 2;
      a.dadd = ch;
      a.nfield = 8;
      for (int f = 0; f < 8; ++f) {
        a.fsrc[f] = 0;
        a.fcode[f] = p_codes[f];
        a.fscale[f] = sc;
      }
      a.out = P->ws.at("p22_t").first;
      a.out_nfield = 16;
      a.out_foff = 8 * ch;
      a.ostride = sw(d.p1.wx);
      if ((rc = GLW_STAGE(P, "o2_psi_colinv", st, colinv_bytes(a, d.g2y, nt * 4), launch_colinv<T>(d.g2y, a, nt * 4, P->tw(d.g2y, f32), P->exec(STAGE_COL), st)))) return rc;
    }
    if ((rc = GLW_STAGE(P, "o2_psi_row", st, row_bytes(row_args("p22_t", 16, "p22_u", 8), nt * 4), (launch_row<T, 16, 8, Psi22Op>(d.g2x, row_args("p22_t", 16, "p22_u", 8), nt * 4, P->tw(d.g2x, f32), P->exec(STAGE_ROW), st)))
     )) return rc;
    if ((rc = GLW_STAGE(P, "o2_psi_colfwd", st, colfwd_bytes(fwd_args("p22_u", 8, "p22_s", nullptr, P->mask2), d.g2y, nt * 4), launch_colfwd<T>(d.g2y, fwd_args("p22_u", 8, "p22_s", nullptr, P->mask2), nt * 4, P->tw(d.g2y, f32), P->exec(STAGE_COL), st)))
     ) return rc;
  }
  // ---- psi22 Taylor contact: tasks t
  {
    ColInvArgs a = colinv_base(stp, 1, 1, d.p1, P->t1);
    a.nfield = 2;
    a.fsrc[0] = 0;
    a.fcode[0] = M_ONE;
    a.fscale[0] = sc;
    a.fsrc[1] = 0;
    a.fcode[1] = M_MI_NU;
    a.fscale[1] = sc;
    a.out = P->ws.at("c22_t").first;
    a.out_nfield = 2;
    a.ostride = sw(d.p1.wx);
    if ((rc = GLW_STAGE(P, "o2_contact_colinv", st, colinv_bytes(a, d.g2y, nt), launch_colinv<T>(d.g2y, a, nt, P->tw(d.g2y, f32), P->exec(STAGE_COL), st)))) return rc;
    if ((rc = GLW_STAGE(P, "o2_contact_row", st, row_bytes(row_args("c22_t", 2, "c22_u", 1), nt), (launch_row<T, 2, 1, ContactOp>(d.g2x, row_args("c22_t", 2, "c22_u", 1), nt, P->tw(d.g2x, f32), P->exec(STAGE_ROW), st)))
     )) return rc;
    if ((rc = GLW_STAGE(P, "o2_contact_colfwd", st, colfwd_bytes(fwd_args("c22_u", 1, "c22_s", nullptr, P->mask2), d.g2y, nt), launch_colfwd<T>(d.g2y, fwd_args("c22_u", 1, "c22_s", nullptr, P->mask2), nt, P->tw(d.g2y, f32), P->exec(STAGE_COL), st)))
     ) return rc;
  }
  {
    AccPsi22Args g;
    g.spec = P->buf<cpx<double>>("p22_s");
    g.lay = spec_layout(d.p2);
    g.contact = P->buf<cpx<double>>("c22_s");
    g.box = b2;
    g.inu = P->t2.inu;
    g.mask = P->mask2;
    g.factor = P->p22_fac;
    g.coef = P->p22_coef;
    g.sign = P->p22_sign;
    g.out = P->buf<cpx<double>>("ep22");
    P->mark_begin("o2_psi_acc", st, kZ * nt * (32.0 + 1.0 + 1.0) * b2);
    hipLaunchKernelGGL(k_acc_psi22, dim3((b2 + NT - 1) / NT, nt), dim3(NT), 0, st, g);
    P->mark_end(st);
    GLW_HIP(hipGetLastError());
  }
  const double h = d.depth, pot = d.depth * std::sqrt(d.gravity * d.depth), nn = (double)d.nx * d.ny;
  const int srcs[2] = {0, 1}, codes[2] = {M_ONE, M_ONE};
  const double scales[2] = {h / nn, pot / nn};
  double* outs[2] = {eta, psi};
  return synthesize(P, P->buf<cpx<double>>("ep22"), 2, d.p2, P->t2, nt, srcs, codes, scales, outs, 2, st);
}


template <typename T>
int order4_chunk(glw_plan* P, int nt, double* eta, hipStream_t st) {
  const glw_plan_desc& d = P->d;
  const glw_order4_desc& o = d.o4;
  const bool f32 = std::is_same<T, float>::value;
  const int J = o.rank;
  const int b2 = box_size(d.p2), b3 = box_size(d.p3), b4 = box_size(o.p4);
  const double nn = (double)d.nx * d.ny, sc = 1.0 / nn;
  const cpx<double>* stp = P->buf<cpx<double>>("st");
  int rc;
  auto fields = [](ColInvArgs& a, int n, const int* srcs, const int* codes, double scale) {
    a.nfield = n;
    for (int f = 0; f < n; ++f) {
      a.fsrc[f] = srcs ? srcs[f] : 0;
      a.fcode[f] = codes[f];
      a.fscale[f] = scale;
    }
  };
  // ---- pair level (paper/order4/cpp/gl_eta44_fftw.cpp:299-529)
  static const int pair_codes[16] = {M_ONE,   M_NU,    M_NU2, M_NU3, M_NU4, M_QX_NU,  M_QY_NU,  M_QX,
                                     M_QY,    M_Q2_NU, M_Q2,  M_Q2NU, M_NU_QX, M_NU_QY, M_NU2_QX, M_NU2_QY};
  auto pair_level = [&](int per_instant, const double* damp, const char* set, const char* setmax) -> int {
    const int ntask = nt * per_instant, nset = ntask / J;
    int rc2;
    ColInvArgs a = colinv_base(stp, 1, per_instant, d.p1, P->t1);
    a.damp = damp;
    a.dmod = per_instant;
    fields(a, 16, nullptr, pair_codes, sc);
    a.out = P->ws.at("o4_pt").first;
    a.out_nfield = 16;
    a.ostride = sw(d.p1.wx);
    if ((rc2 = GLW_STAGE(P, "o4_pair_colinv", st, colinv_bytes(a, d.g2y, ntask), launch_colinv<T>(d.g2y, a, ntask, P->tw(d.g2y, f32), P->exec(STAGE_COL), st))))
      return rc2;
    RowArgs r;
    memset(&r, 0, sizeof(r));
    r.in = P->ws.at("o4_pt").first;
    r.nin = 16;
    r.istride = sw(d.p1.wx);
    for (int f = 0; f < 16; ++f) {
      r.fmx0[f] = d.p1.mx0;
      r.fwx[f] = d.p1.wx;
    }
    r.out = P->ws.at("o4_pu").first;
    r.nout = 6;
    r.ostride = sw(d.p2.wx);
    r.omx0 = d.p2.mx0;
    r.owx = d.p2.wx;
    r.gy = d.g2y;
    if ((rc2 = GLW_STAGE(P, "o4_pair_row", st, row_bytes(r, ntask), (launch_row<T, 16, 6, Pair44Op>(d.g2x, r, ntask, P->tw(d.g2x, f32), P->exec(STAGE_ROW), st)))))
      return rc2;
    if ((rc2 = fill_zero(P->buf<uint64_t>("o4_pm"), (size_t)ntask * 6, st))) return rc2;
    ColFwdArgs c;
    c.in = P->ws.at("o4_pu").first;
    c.nsrc = 6;
    c.istride = sw(d.p2.wx);
    c.omx0 = d.p2.mx0;
    c.owx = d.p2.wx;
    c.omy0 = d.p2.my0;
    c.owy = d.p2.wy;
    c.mask = P->mask2;
    c.scale = nn / ((double)d.g2x * d.g2y);
    c.out = P->buf<cpx<double>>("o4_ps");
    c.opitch = spec_layout(d.p2).p;
    c.fmax = P->buf<uint64_t>("o4_pm");
    if ((rc2 = GLW_STAGE(P, "o4_pair_colfwd", st, colfwd_bytes(c, d.g2y, ntask), launch_colfwd<T>(d.g2y, c, ntask, P->tw(d.g2y, f32), P->exec(STAGE_COL), st))))
      return rc2;
    if ((rc2 = fill_zero(P->buf<uint64_t>(setmax), (size_t)nset * 5, st))) return rc2;
    Acc4Args g;
    g.spec = P->buf<cpx<double>>("o4_ps");
    g.lay = spec_layout(d.p2);
    g.fmax = P->buf<uint64_t>("o4_pm");
    g.floor_level = d.floor_level;
    g.J = J;
    g.box = b2;
    g.nsrc = 6;
    g.nacc = 5;
    g.a = P->t2.a;
    g.mask = P->mask2;
    g.ch = P->o4_pch;
    g.sh = P->o4_psh;
    g.coef = P->o4_pco;
    g.out = P->buf<cpx<double>>(set);
    g.omax = P->buf<uint64_t>(setmax);
    P->mark_begin("o4_pair_acc", st, kZ * nset * (6.0 * J + 5.0) * b2);
    hipLaunchKernelGGL(k_acc_pair44, dim3((b2 + NT - 1) / NT, nset), dim3(NT), 0, st, g);
    P->mark_end(st);
    GLW_HIP(hipGetLastError());
    return GLW_OK;
  };
  if ((rc = pair_level(J * J * J, P->o4_dp3, "o4_set3", "o4_max3"))) return rc;  // pairs of the third level
  if ((rc = pair_level(J * J, P->o4_dp4, "o4_set4", "o4_max4"))) return rc;      // pairs of the quartic level
  static const int first_codes[19] = {M_ONE,     M_I_QX,    M_I_QY,     M_QX_NU,     M_QY_NU,     M_MI_NU,  M_NU_QX,
                                      M_NU_QY,   M_MI_Q2_NU, M_MI_Q2NU, M_MA,        M_MQ2,       M_MI_NU2_QX,
                                      M_MI_NU2_QY, M_MQ2A,  M_I_ANU,    M_QXQ2_NU,   M_QYQ2_NU,   M_MI_Q4_NU};
  static const int pair_src[18] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 2, 2, 2};
  static const int pair_out_codes[18] = {M_ONE, M_I_QX, M_I_QY, M_ONE, M_I_QX, M_I_QY, M_I_QX,  M_I_QY,  M_A,
                                         M_Q2,  M_A,    M_I_QX, M_I_QY, M_Q2,  M_A,    M_I_QXA, M_I_QYA, M_Q2A};
  // ---- third level (:537-637): tasks (t, o4, o3)
  const int w12 = sw(std::max(d.p1.wx, d.p2.wx));
  {
    const int ntask = nt * J * J;
    ColInvArgs a = colinv_base(stp, 1, J * J, d.p1, P->t1);
    a.damp = P->o4_dt;
    a.dmod = J * J;
    fields(a, 16, nullptr, first_codes, sc);
    a.out = P->ws.at("o4_tt").first;
    a.out_nfield = 31;
    a.out_foff = 0;
    a.ostride = w12;
    if ((rc = GLW_STAGE(P, "o4_third_colinv", st, colinv_bytes(a, d.g3y, ntask), launch_colinv<T>(d.g3y, a, ntask, P->tw(d.g3y, f32), P->exec(STAGE_COL), st))))
      return rc;
    ColInvArgs b = colinv_base(P->buf<cpx<double>>("o4_set3"), 5, 1, d.p2, P->t2);
    b.fmax = P->buf<uint64_t>("o4_max3");
    b.fmax_stride = 5;
    b.floor_src_mask = 0x1f;
    b.floor_level = d.floor_level;
    fields(b, 15, pair_src, pair_out_codes, sc);
    b.out = P->ws.at("o4_tt").first;
    b.out_nfield = 31;
    b.out_foff = 16;
    b.ostride = w12;
    if ((rc = GLW_STAGE(P, "o4_third_colinv", st, colinv_bytes(b, d.g3y, ntask), launch_colinv<T>(d.g3y, b, ntask, P->tw(d.g3y, f32), P->exec(STAGE_COL), st))))
      return rc;
    RowArgs r;
    memset(&r, 0, sizeof(r));
    r.in = P->ws.at("o4_tt").first;
    r.nin = 31;
    r.istride = w12;
    for (int f = 0; f < 31; ++f) {
      r.fmx0[f] = f < 16 ? d.p1.mx0 : d.p2.mx0;
      r.fwx[f] = f < 16 ? d.p1.wx : d.p2.wx;
    }
    r.out = P->ws.at("o4_tu").first;
    r.nout = 4;
    r.ostride = sw(d.p3.wx);
    r.omx0 = d.p3.mx0;
    r.owx = d.p3.wx;
    r.gy = d.g3y;
    if ((rc = GLW_STAGE(P, "o4_third_row", st, row_bytes(r, ntask), (launch_row<T, 31, 4, Third44Op>(d.g3x, r, ntask, P->tw(d.g3x, f32), P->exec(STAGE_ROW), st)))))
      return rc;
    if ((rc = fill_zero(P->buf<uint64_t>("o4_tm"), (size_t)ntask * 4, st))) return rc;
    ColFwdArgs c;
    c.in = P->ws.at("o4_tu").first;
    c.nsrc = 4;
    c.istride = sw(d.p3.wx);
    c.omx0 = d.p3.mx0;
    c.owx = d.p3.wx;
    c.omy0 = d.p3.my0;
    c.owy = d.p3.wy;
    c.mask = P->mask3;
    c.scale = nn / ((double)d.g3x * d.g3y);
    c.out = P->buf<cpx<double>>("o4_ts");
    c.opitch = spec_layout(d.p3).p;
    c.fmax = P->buf<uint64_t>("o4_tm");
    if ((rc = GLW_STAGE(P, "o4_third_colfwd", st, colfwd_bytes(c, d.g3y, ntask), launch_colfwd<T>(d.g3y, c, ntask, P->tw(d.g3y, f32), P->exec(STAGE_COL), st))))
      return rc;
    if ((rc = fill_zero(P->buf<uint64_t>("o4_thm"), (size_t)nt * J * 3, st))) return rc;
    Acc4Args g;
    g.spec = P->buf<cpx<double>>("o4_ts");
    g.lay = spec_layout(d.p3);
    g.fmax = P->buf<uint64_t>("o4_tm");
    g.floor_level = d.floor_level;
    g.J = J;
    g.box = b3;
    g.nsrc = 4;
    g.nacc = 3;
    g.a = P->t3.a;
    g.mask = P->mask3;
    g.ch = P->o4_tch;
    g.sh = P->o4_tsh;
    g.coef = P->o4_tco;
    g.out = P->buf<cpx<double>>("o4_th");
    g.omax = P->buf<uint64_t>("o4_thm");
    P->mark_begin("o4_third_acc", st, kZ * nt * J * (4.0 * J + 3.0) * b3);
    hipLaunchKernelGGL(k_acc_third44, dim3((b3 + NT - 1) / NT, nt * J), dim3(NT), 0, st, g);
    P->mark_end(st);
    GLW_HIP(hipGetLastError());
  }
  // ---- quartic level (:696-750): tasks (t, o4)
  const int w123 = sw(std::max(w12, (int)d.p3.wx));
  {
    const int ntask = nt * J;
    ColInvArgs a = colinv_base(stp, 1, J, d.p1, P->t1);
    a.damp = P->o4_dq;
    a.dmod = J;
    fields(a, 19, nullptr, first_codes, sc);
    a.out = P->ws.at("o4_qt").first;
    a.out_nfield = 45;
    a.out_foff = 0;
    a.ostride = w123;
    if ((rc = GLW_STAGE(P, "o4_quartic_colinv", st, colinv_bytes(a, o.g4y, ntask), launch_colinv<T>(o.g4y, a, ntask, P->tw(o.g4y, f32), P->exec(STAGE_COL), st))))
      return rc;
    ColInvArgs b = colinv_base(P->buf<cpx<double>>("o4_set4"), 5, 1, d.p2, P->t2);
    b.fmax = P->buf<uint64_t>("o4_max4");
    b.fmax_stride = 5;
    b.floor_src_mask = 0x1f;
    b.floor_level = d.floor_level;
    fields(b, 18, pair_src, pair_out_codes, sc);
    b.out = P->ws.at("o4_qt").first;
    b.out_nfield = 45;
    b.out_foff = 19;
    b.ostride = w123;
    if ((rc = GLW_STAGE(P, "o4_quartic_colinv", st, colinv_bytes(b, o.g4y, ntask), launch_colinv<T>(o.g4y, b, ntask, P->tw(o.g4y, f32), P->exec(STAGE_COL), st))))
      return rc;
    static const int third_src[8] = {0, 0, 0, 1, 1, 1, 1, 2};
    static const int third_codes[8] = {M_ONE, M_I_QX, M_I_QY, M_I_QX, M_I_QY, M_A, M_Q2, M_A};
    ColInvArgs c3 = colinv_base(P->buf<cpx<double>>("o4_th"), 3, 1, d.p3, P->t3);
    c3.fmax = P->buf<uint64_t>("o4_thm");
    c3.fmax_stride = 3;
    c3.floor_src_mask = 0x7;
    c3.floor_level = d.floor_level;
    fields(c3, 8, third_src, third_codes, sc);
    c3.out = P->ws.at("o4_qt").first;
    c3.out_nfield = 45;
    c3.out_foff = 37;
    c3.ostride = w123;
    if ((rc = GLW_STAGE(P, "o4_quartic_colinv", st, colinv_bytes(c3, o.g4y, ntask), launch_colinv<T>(o.g4y, c3, ntask, P->tw(o.g4y, f32), P->exec(STAGE_COL), st))))
      return rc;
    RowArgs r;
    memset(&r, 0, sizeof(r));
    r.in = P->ws.at("o4_qt").first;
    r.nin = 45;
    r.istride = w123;
    for (int f = 0; f < 45; ++f) {
      const glw_box& bx = f < 19 ? d.p1 : (f < 37 ? d.p2 : d.p3);
      r.fmx0[f] = bx.mx0;
      r.fwx[f] = bx.wx;
    }
    r.out = P->ws.at("o4_qu").first;
    r.nout = 2;
    r.ostride = sw(o.p4.wx);
    r.omx0 = o.p4.mx0;
    r.owx = o.p4.wx;
    r.gy = o.g4y;
    if (o.forcing_terms == 1) {  // the 3+1 partition only (split runs)
      if ((rc = GLW_STAGE(P, "o4_quartic_row", st, row_bytes(r, ntask), (launch_row<T, 45, 2, Quartic31Op>(o.g4x, r, ntask, P->tw(o.g4x, f32), P->exec(STAGE_ROW), st)))))
        return rc;
    } else if ((rc = GLW_STAGE(P, "o4_quartic_row", st, row_bytes(r, ntask), (launch_row<T, 45, 2, Quartic44Op>(o.g4x, r, ntask, P->tw(o.g4x, f32), P->exec(STAGE_ROW), st))))) {
      return rc;
    }
    if ((rc = fill_zero(P->buf<uint64_t>("o4_qm"), (size_t)ntask * 2, st))) return rc;
    ColFwdArgs c;
    c.in = P->ws.at("o4_qu").first;
    c.nsrc = 2;
    c.istride = sw(o.p4.wx);
    c.omx0 = o.p4.mx0;
    c.owx = o.p4.wx;
    c.omy0 = o.p4.my0;
    c.owy = o.p4.wy;
    c.mask = P->mask4;
    c.scale = nn / ((double)o.g4x * o.g4y);
    c.out = P->buf<cpx<double>>("o4_qs");
    c.opitch = spec_layout(o.p4).p;
    c.fmax = P->buf<uint64_t>("o4_qm");
    if ((rc = GLW_STAGE(P, "o4_quartic_colfwd", st, colfwd_bytes(c, o.g4y, ntask), launch_colfwd<T>(o.g4y, c, ntask, P->tw(o.g4y, f32), P->exec(STAGE_COL), st))))
      return rc;
    if ((rc = fill_zero(P->buf<uint64_t>("o4_outm"), (size_t)nt, st))) return rc;
    Acc4Args g;
    g.spec = P->buf<cpx<double>>("o4_qs");
    g.lay = spec_layout(o.p4);
    g.fmax = P->buf<uint64_t>("o4_qm");
    g.floor_level = d.floor_level;
    g.J = J;
    g.box = b4;
    g.nsrc = 2;
    g.nacc = 1;
    g.a = P->t4.a;
    g.mask = P->mask4;
    g.ch = P->o4_qch;
    g.sh = P->o4_qsh;
    g.coef = P->o4_qco;
    g.out = P->buf<cpx<double>>("o4_out");
    g.omax = P->buf<uint64_t>("o4_outm");
    P->mark_begin("o4_quartic_acc", st, kZ * nt * (2.0 * J + 1.0) * b4);
    hipLaunchKernelGGL(k_acc_quartic44, dim3((b4 + NT - 1) / NT, nt), dim3(NT), 0, st, g);
    P->mark_end(st);
    GLW_HIP(hipGetLastError());
  }
  const int srcs[1] = {0}, codes[1] = {M_ONE};
  const double scales[1] = {d.depth / nn};
  double* outs[1] = {eta};
  return synthesize(P, P->buf<cpx<double>>("o4_out"), 1, o.p4, P->t4, nt, srcs, codes, scales, outs, 1, st,
                    P->buf<uint64_t>("o4_outm"), 1, 1);
}

int order4_dispatch(glw_plan* P, int nt, double* a, double* /*unused*/, hipStream_t st) {
  return order4_chunk<double>(P, nt, a, st);
}

typedef int (*chunk_fn)(glw_plan*, int, double*, double*, hipStream_t);

int order4_dispatch(glw_plan* P, int nt, double* a, double* b, hipStream_t st);

int run_chunks(glw_plan* P, const double* s0, const double* times, int nt, double* a, double* b, void* stream,
               chunk_fn fn) {
  if (!P) return fail(GLW_EINVAL, "null plan");
  if (nt < 0) return fail(GLW_EINVAL, "negative instant count");
  if (nt == 0) return GLW_OK;
  if (!s0 || !times || !a || (!b && fn != order4_dispatch)) return fail(GLW_EINVAL, "null array argument");
  int dev = -1;
  GLW_HIP(hipGetDevice(&dev));
  if (dev != P->device) GLW_HIP(hipSetDevice(P->device));
  hipStream_t st = (hipStream_t)stream;
  P->cur_stream = st;
  const size_t plane = (size_t)P->d.ny * P->d.nx * 2;  // doubles per output field
  for (int t0 = 0; t0 < nt; t0 += P->d.max_instants) {
    const int n = std::min(P->d.max_instants, nt - t0);
    int rc = phase(P, s0, times + t0, n, st);
    if (rc) return rc;
    rc = fn(P, n, a + (size_t)t0 * plane, b ? b + (size_t)t0 * plane : nullptr, st);
    if (rc) return rc;
  }
  return GLW_OK;
}

int linear_chunk(glw_plan* P, int nt, double* eta, double* psi, hipStream_t st) {
  const glw_plan_desc& d = P->d;
  const double h = d.depth, pot = d.depth * std::sqrt(d.gravity * d.depth), nn = (double)d.nx * d.ny;
  const int srcs[2] = {0, 0}, codes[2] = {M_ONE, M_MI_INU};
  const double scales[2] = {h / nn, pot / nn};
  double* outs[2] = {eta, psi};
  return synthesize(P, P->buf<cpx<double>>("st"), 1, d.p1, P->t1, nt, srcs, codes, scales, outs, 2, st);
}

int order3_dispatch(glw_plan* P, int nt, double* a, double* b, hipStream_t st) {
  return order3_chunk<double>(P, nt, a, b, st);
}
int order2_dispatch(glw_plan* P, int nt, double* a, double* b, hipStream_t st) {
  return order2_chunk<double>(P, nt, a, b, st);
}

// ----------------------------------------------------------------------------
// Adjoint of eta11 + eta22 + eta33 (one instant): reverse pass of the
// forward graph above, stage by stage.  Forward intermediates (column
// passes, forced spectra, maxima) are read from the plan's workspace, so the
// forward stages must have run for the same spectrum and instant (recompute).
// ----------------------------------------------------------------------------
int ensure_ws(glw_plan* P, const char* name, size_t bytes) {
  auto it = P->ws.find(name);
  if (it != P->ws.end())
    return it->second.second >= bytes ? GLW_OK : fail(GLW_EINVAL, std::string("workspace size changed: ") + name);
  return P->reserve(name, bytes);
}

glw_box union_box(const glw_box& a, const glw_box& b) {
  glw_box u;
  u.mx0 = std::min(a.mx0, b.mx0);
  u.my0 = std::min(a.my0, b.my0);
  u.wx = std::max(a.mx0 + a.wx, b.mx0 + b.wx) - u.mx0;
  u.wy = std::max(a.my0 + a.wy, b.my0 + b.wy) - u.my0;
  return u;
}

int gather(const cpx<double>* u, const glw_box& ub, const glw_box& db, double scale, int add, cpx<double>* dst,
           hipStream_t st) {
  const int n = box_size(db);
  hipLaunchKernelGGL(k_adj_gather, dim3((n + 255) / 256), dim3(256), 0, st, u, ub, db, scale, add, dst);
  GLW_HIP(hipGetLastError());
  return GLW_OK;
}

RowArgs row_desc(const void* in, int nin, int istride, const glw_box* boxes, int gy) {
  RowArgs r;
  memset(&r, 0, sizeof(r));
  r.in = in;
  r.nin = nin;
  r.istride = istride;
  for (int f = 0; f < nin; ++f) {
    r.fmx0[f] = boxes[f].mx0;
    r.fwx[f] = boxes[f].wx;
  }
  r.gy = gy;
  return r;
}

RowArgs row_grad_out(void* out, int nout, int ostride, const glw_box* boxes, int gy) {
  RowArgs r;
  memset(&r, 0, sizeof(r));
  r.out = out;
  r.nout = nout;
  r.ostride = ostride;
  r.out_per_field = 1;
  for (int f = 0; f < nout; ++f) {
    r.fmx0[f] = boxes[f].mx0;
    r.fwx[f] = boxes[f].wx;
  }
  r.gy = gy;
  return r;
}

ColFwdArgs col_dft(const void* in, int in_nfield, int in_foff, int nsrc, int istride, const glw_box& box,
                   const uint8_t* ones, cpx<double>* out) {
  ColFwdArgs c;
  c.in = in;
  c.nsrc = nsrc;
  c.in_nfield = in_nfield;
  c.in_foff = in_foff;
  c.istride = istride;
  c.omx0 = box.mx0;
  c.owx = box.wx;
  c.omy0 = box.my0;
  c.owy = box.wy;
  c.mask = ones;
  c.scale = 1.0;
  c.out = out;
  c.fmax = nullptr;
  return c;
}

// Spectra [ntask][nsrc][box] -> unnormalized inverse y-pass columns, all sources with unit multipliers.
ColInvArgs col_idft(const cpx<double>* spec, int nsrc, const glw_box& box, const DevTables& t, double scale,
                    void* out) {
  ColInvArgs a = colinv_base(spec, nsrc, 1, box, t);
  a.nfield = nsrc;
  for (int f = 0; f < nsrc; ++f) {
    a.fsrc[f] = f;
    a.fcode[f] = M_ONE;
    a.fscale[f] = scale;
  }
  a.out = out;
  a.out_nfield = nsrc;
  a.ostride = sw(box.wx);
  return a;
}

int collect(const ColInvArgs& fwd, const cpx<double>* x, int nset, cpx<double>* out, hipStream_t st) {
  const int n = box_size(glw_box{fwd.mx0, fwd.my0, fwd.wx, fwd.wy});
  hipLaunchKernelGGL(k_adj_collect, dim3((n + kCollectBins - 1) / kCollectBins, nset), dim3(NT), 0, st, fwd, x, out);
  GLW_HIP(hipGetLastError());
  return GLW_OK;
}

int fill_ones(uint8_t* p, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_one_u8, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  GLW_HIP(hipGetLastError());
  return GLW_OK;
}

int adjoint_run(glw_plan* P, const double* s0, const double* times, int orders, const cpx<double>* gfield,
                cpx<double>* gs0, int recompute, hipStream_t st) {
  const glw_plan_desc& d = P->d;
  if (d.precision != GLW_PREC_FP64) return fail(GLW_EINVAL, "the adjoint runs on FP64 plans");
  if (!length_supported(d.nx) || !length_supported(d.ny))
    return fail(GLW_EUNSUPPORTED, "the adjoint needs an output grid of FFT lengths (2^a 3^b, 16..2048 per side)");
  if (!orders || (orders & ~7)) return fail(GLW_EINVAL, "orders must be a mask of 1 (eta11), 2 (eta22), 4 (eta33)");
  if ((orders & 2) && d.order < 2) return fail(GLW_EINVAL, "plan built without order-2 tables");
  if ((orders & 4) && d.order < 3) return fail(GLW_EINVAL, "plan built without order-3 tables");
  const size_t z = sizeof(cpx<double>);
  const int b1 = box_size(d.p1), b2 = box_size(d.p2), b3 = box_size(d.p3);
  const double nn = (double)d.nx * d.ny, sc = 1.0 / nn;
  glw_box U = d.p1;  // analysis box: every box of the plan, so the workspace does not depend on `orders`
  if (d.order >= 2) U = union_box(U, d.p2);
  if (d.order >= 3) U = union_box(U, d.p3);
  int rc;
#define GLW_WS(name, bytes) \
  if ((rc = ensure_ws(P, name, bytes))) return rc
  GLW_WS("adj_fwd", 2 * (size_t)d.nx * d.ny * z);
  GLW_WS("adj_rows", (size_t)d.ny * U.wx * z);
  GLW_WS("adj_u", (size_t)box_size(U) * z);
  GLW_WS("adj_onesU", (size_t)box_size(U));
  GLW_WS("adj_ones1", (size_t)b1);
  GLW_WS("adj_gs", (size_t)b1 * z);
  if (d.order >= 2) GLW_WS("adj_ones2", (size_t)b2);
  const int J = d.o2.eta_rank, O = d.outer_rank, I = d.inner_rank;
  const int wmax = sw(std::max(d.p1.wx, d.p2.wx));
  if (orders & 2) {
    GLW_WS("adj_e22", (size_t)b2 * z);
    GLW_WS("adj_s22", (size_t)J * 2 * b2 * z);
    GLW_WS("adj_u22", (size_t)J * 2 * d.g2y * sw(d.p2.wx) * z);
    GLW_WS("adj_t22", (size_t)J * 8 * d.g2y * sw(d.p1.wx) * z);
    GLW_WS("adj_x22", (size_t)J * 8 * b1 * z);
  }
  if (orders & 4) {
    GLW_WS("adj_e33", (size_t)b3 * z);
    GLW_WS("adj_s33", (size_t)O * 2 * b3 * z);
    GLW_WS("adj_u33", (size_t)O * 2 * d.g3y * sw(d.p3.wx) * z);
    GLW_WS("adj_t33", (size_t)O * 21 * d.g3y * wmax * z);
    GLW_WS("adj_x33a", (size_t)O * 12 * b1 * z);
    GLW_WS("adj_x33b", (size_t)O * 9 * b2 * z);
    GLW_WS("adj_pair", (size_t)O * 4 * b2 * z);
    GLW_WS("adj_sin", (size_t)O * I * 4 * b2 * z);
    GLW_WS("adj_uin", (size_t)O * I * 4 * d.g2y * sw(d.p2.wx) * z);
    GLW_WS("adj_tin", (size_t)O * I * 12 * d.g2y * sw(d.p1.wx) * z);
    GLW_WS("adj_xin", (size_t)O * I * 12 * b1 * z);
  }
#undef GLW_WS
  auto B = [&](const char* n) { return P->buf<cpx<double>>(n); };
  if ((rc = fill_ones(P->buf<uint8_t>("adj_onesU"), box_size(U), st))) return rc;
  if ((rc = fill_ones(P->buf<uint8_t>("adj_ones1"), b1, st))) return rc;
  if (d.order >= 2 && (rc = fill_ones(P->buf<uint8_t>("adj_ones2"), b2, st))) return rc;
  const uint8_t* ones1 = P->buf<uint8_t>("adj_ones1");
  const uint8_t* ones2 = d.order >= 2 ? P->buf<uint8_t>("adj_ones2") : nullptr;

  if (recompute) {  // forward stages for this spectrum and instant
    double* fo = P->buf<double>("adj_fwd");
    const size_t plane = (size_t)d.nx * d.ny * 2;
    if ((rc = phase(P, s0, times, 1, st))) return rc;
    if ((orders & 2) && (rc = order2_chunk<double>(P, 1, fo, fo + plane, st))) return rc;
    if ((orders & 4) && (rc = order3_chunk<double>(P, 1, fo, fo + plane, st))) return rc;
  }
  // analysis: (h / N^2) DFT2 of the field gradient on the union box
  {
    RowArgs ra;
    memset(&ra, 0, sizeof(ra));
    ra.gy = d.ny;
    ra.nout = 1;
    ra.out = B("adj_rows");
    ra.ostride = U.wx;
    ra.omx0 = U.mx0;
    ra.owx = U.wx;
    if ((rc = GLW_STAGE(P, "adj_analysis_row", st, rowfwd_bytes(ra, d.nx, 1), launch_rowfwd<double>(d.nx, ra, 1, P->tw(d.nx, false), 0, gfield, st))))
      return rc;
    ColFwdArgs ca = col_dft(B("adj_rows"), 1, 0, 1, U.wx, U, P->buf<uint8_t>("adj_onesU"), B("adj_u"));
    ca.scale = d.depth / nn;
    if ((rc = GLW_STAGE(P, "adj_analysis_col", st, colfwd_bytes(ca, d.ny, 1), launch_colfwd<double>(d.ny, ca, 1, P->tw(d.ny, false), P->exec(STAGE_COL), st))))
      return rc;
  }
  cpx<double>* gs = B("adj_gs");
  if (orders & 1) {
    if ((rc = gather(B("adj_u"), U, d.p1, 1.0, 0, gs, st))) return rc;
  } else {
    hipLaunchKernelGGL(k_fill_zero_c, dim3((b1 + 255) / 256), dim3(256), 0, st, gs, (size_t)b1);
    GLW_HIP(hipGetLastError());
  }
  const cpx<double>* stp = P->buf<cpx<double>>("st");
  if (orders & 2) {
    static const int e_codes[8] = {M_ONE, M_NU, M_NU2, M_QX_NU, M_QY_NU, M_Q2_NU, M_QX, M_QY};
    if ((rc = gather(B("adj_u"), U, d.p2, 1.0, 0, B("adj_e22"), st))) return rc;
    AccEta22Args g;
    g.spec = B("e22_s");
    g.lay = spec_layout(d.p2);
    g.fmax = P->buf<uint64_t>("e22_max");
    g.floor_level = d.floor_level;
    g.J = J;
    g.box = b2;
    g.a = P->t2.a;
    g.ch = P->e22_ch;
    g.sh = P->e22_sh;
    g.coef = P->e22_coef;
    g.out = nullptr;
    P->mark_begin("adj_o2_acc", st, 0.0);
    hipLaunchKernelGGL(k_adj_acc_eta22, dim3((b2 + NT - 1) / NT, J), dim3(NT), 0, st, g, B("adj_e22"), P->mask2,
                       B("adj_s22"));
    P->mark_end(st);
    GLW_HIP(hipGetLastError());
    const double up = nn / ((double)d.g2x * d.g2y);
    ColInvArgs a = col_idft(B("adj_s22"), 2, d.p2, P->t2, up, B("adj_u22"));
    if ((rc = GLW_STAGE(P, "adj_o2_colinv", st, colinv_bytes(a, d.g2y, J), launch_colinv<double>(d.g2y, a, J, P->tw(d.g2y, false), P->exec(STAGE_COL), st))))
      return rc;
    glw_box in_boxes[8], g_boxes[2] = {d.p2, d.p2};
    for (int f = 0; f < 8; ++f) in_boxes[f] = d.p1;
    const RowArgs fwd = row_desc(P->ws.at("e22_t").first, 8, sw(d.p1.wx), in_boxes, d.g2y);
    const RowArgs gin = row_desc(B("adj_u22"), 2, sw(d.p2.wx), g_boxes, d.g2y);
    const RowArgs gout = row_grad_out(B("adj_t22"), 8, sw(d.p1.wx), in_boxes, d.g2y);
    if ((rc = GLW_STAGE(P, "adj_o2_row", st, rowadj_bytes(fwd, gin, gout, J), (launch_row_adjoint<double, 8, 2, Eta22Op>(d.g2x, fwd, gin, gout, J, P->tw(d.g2x, false), P->exec(STAGE_ROW), st)))))
      return rc;
    const ColFwdArgs c = col_dft(B("adj_t22"), 8, 0, 8, sw(d.p1.wx), d.p1, ones1, B("adj_x22"));
    if ((rc = GLW_STAGE(P, "adj_o2_colfwd", st, colfwd_bytes(c, d.g2y, J), launch_colfwd<double>(d.g2y, c, J, P->tw(d.g2y, false), P->exec(STAGE_COL), st))))
      return rc;
    ColInvArgs fa = colinv_base(stp, 1, J, d.p1, P->t1);
    fa.damp = P->e22_damp;
    fa.dmod = J;
    fa.nfield = 8;
    for (int f = 0; f < 8; ++f) {
      fa.fsrc[f] = 0;
      fa.fcode[f] = e_codes[f];
      fa.fscale[f] = sc;
    }
    if ((rc = collect(fa, B("adj_x22"), 1, gs, st))) return rc;
  }
  if (orders & 4) {
    const double up2 = nn / ((double)d.g2x * d.g2y), up3 = nn / ((double)d.g3x * d.g3y);
    if ((rc = gather(B("adj_u"), U, d.p3, 1.0, 0, B("adj_e33"), st))) return rc;
    AccOuterArgs g;
    g.spec = B("sout");
    g.lay = spec_layout(d.p3);
    g.fmax = P->buf<uint64_t>("max_out");
    g.cmax = P->buf<uint64_t>("max_combo");
    g.floor_level = d.floor_level;
    g.O = O;
    g.box = b3;
    g.a = P->t3.a;
    g.mask = P->mask3;
    g.ch = P->out_ch;
    g.sh = P->out_sh;
    g.coef = P->out_coef;
    g.out = nullptr;
    P->mark_begin("adj_o3_outer_acc", st, 0.0);
    hipLaunchKernelGGL(k_adj_acc_outer, dim3((b3 + NT - 1) / NT, O), dim3(NT), 0, st, g, B("adj_e33"), B("adj_s33"));
    P->mark_end(st);
    GLW_HIP(hipGetLastError());
    ColInvArgs a = col_idft(B("adj_s33"), 2, d.p3, P->t3, up3, B("adj_u33"));
    if ((rc = GLW_STAGE(P, "adj_o3_outer_colinv", st, colinv_bytes(a, d.g3y, O), launch_colinv<double>(d.g3y, a, O, P->tw(d.g3y, false), P->exec(STAGE_COL), st))))
      return rc;
    glw_box fb[21], g3b[2] = {d.p3, d.p3};
    for (int f = 0; f < 21; ++f) fb[f] = f < 12 ? d.p1 : d.p2;
    const RowArgs fwd = row_desc(P->ws.at("tout").first, 21, wmax, fb, d.g3y);
    const RowArgs gin = row_desc(B("adj_u33"), 2, sw(d.p3.wx), g3b, d.g3y);
    const RowArgs gout = row_grad_out(B("adj_t33"), 21, wmax, fb, d.g3y);
    if ((rc = GLW_STAGE(P, "adj_o3_outer_row", st, rowadj_bytes(fwd, gin, gout, O), (launch_row_adjoint<double, 21, 2, OuterOp>(d.g3x, fwd, gin, gout, O, P->tw(d.g3x, false), P->exec(STAGE_ROW), st)))))
      return rc;
    const ColFwdArgs ca = col_dft(B("adj_t33"), 21, 0, 12, wmax, d.p1, ones1, B("adj_x33a"));
    if ((rc = GLW_STAGE(P, "adj_o3_outer_colfwd", st, colfwd_bytes(ca, d.g3y, O), launch_colfwd<double>(d.g3y, ca, O, P->tw(d.g3y, false), P->exec(STAGE_COL), st))))
      return rc;
    const ColFwdArgs cb = col_dft(B("adj_t33"), 21, 12, 9, wmax, d.p2, ones2, B("adj_x33b"));
    if ((rc = GLW_STAGE(P, "adj_o3_outer_colfwd", st, colfwd_bytes(cb, d.g3y, O), launch_colfwd<double>(d.g3y, cb, O, P->tw(d.g3y, false), P->exec(STAGE_COL), st))))
      return rc;
    {  // first-order fields of the outer nodes -> parent spectrum
      static const int codes[12] = {M_ONE,  M_I_QX, M_I_QY,     M_QX_NU,   M_QY_NU, M_MI_NU,
                                    M_NU_QX, M_NU_QY, M_MI_Q2_NU, M_MI_Q2NU, M_MA,    M_MQ2};
      ColInvArgs fa = colinv_base(stp, 1, O, d.p1, P->t1);
      fa.damp = P->damp_o;
      fa.dmod = O;
      fa.nfield = 12;
      for (int f = 0; f < 12; ++f) {
        fa.fsrc[f] = 0;
        fa.fcode[f] = codes[f];
        fa.fscale[f] = sc;
      }
      if ((rc = collect(fa, B("adj_x33a"), 1, gs, st))) return rc;
    }
    {  // pair-state fields -> pair spectra (eta, eta_t, phi, phi_t)
      hipLaunchKernelGGL(k_fill_zero_c, dim3((O * 4 * b2 + 255) / 256), dim3(256), 0, st, B("adj_pair"),
                         (size_t)O * 4 * b2);
      GLW_HIP(hipGetLastError());
      static const int srcs[9] = {0, 0, 0, 2, 2, 2, 2, 3, 1};
      static const int codes[9] = {M_ONE, M_I_QX, M_I_QY, M_I_QX, M_I_QY, M_A, M_Q2, M_A, M_ONE};
      ColInvArgs fa = colinv_base(B("pair"), 4, 1, d.p2, P->t2);
      fa.fmax = P->buf<uint64_t>("max_et");
      fa.fmax_stride = 1;
      fa.floor_src_mask = 1 << 1;
      fa.floor_level = d.floor_level;
      fa.nfield = 9;
      for (int f = 0; f < 9; ++f) {
        fa.fsrc[f] = srcs[f];
        fa.fcode[f] = codes[f];
        fa.fscale[f] = sc;
      }
      if ((rc = collect(fa, B("adj_x33b"), O, B("adj_pair"), st))) return rc;
    }
    AccInnerArgs gi;
    gi.spec = B("sin");
    gi.lay = spec_layout(d.p2);
    gi.fmax = P->buf<uint64_t>("max_in");
    gi.floor_level = d.floor_level;
    gi.I = I;
    gi.box = b2;
    gi.a = P->t2.a;
    gi.mask = P->mask2;
    gi.ch = P->in_ch;
    gi.sh = P->in_sh;
    gi.coef = P->in_coef;
    gi.out = nullptr;
    gi.etmax = nullptr;
    P->mark_begin("adj_o3_inner_acc", st, 0.0);
    hipLaunchKernelGGL(k_adj_acc_inner, dim3((b2 + NT - 1) / NT, O * I), dim3(NT), 0, st, gi, B("adj_pair"), B("adj_sin"));
    P->mark_end(st);
    GLW_HIP(hipGetLastError());
    ColInvArgs ai = col_idft(B("adj_sin"), 4, d.p2, P->t2, up2, B("adj_uin"));
    if ((rc = GLW_STAGE(P, "adj_o3_inner_colinv", st, colinv_bytes(ai, d.g2y, O * I), launch_colinv<double>(d.g2y, ai, O * I, P->tw(d.g2y, false), P->exec(STAGE_COL), st))))
      return rc;
    glw_box ib[12], g2b[4] = {d.p2, d.p2, d.p2, d.p2};
    for (int f = 0; f < 12; ++f) ib[f] = d.p1;
    const RowArgs ifwd = row_desc(P->ws.at("tin").first, 12, sw(d.p1.wx), ib, d.g2y);
    const RowArgs igin = row_desc(B("adj_uin"), 4, sw(d.p2.wx), g2b, d.g2y);
    const RowArgs igout = row_grad_out(B("adj_tin"), 12, sw(d.p1.wx), ib, d.g2y);
    if ((rc = GLW_STAGE(P, "adj_o3_inner_row", st, rowadj_bytes(ifwd, igin, igout, O * I), (launch_row_adjoint<double, 12, 4, InnerOp>(d.g2x, ifwd, igin, igout, O * I, P->tw(d.g2x, false), P->exec(STAGE_ROW), st)))))
      return rc;
    const ColFwdArgs ci = col_dft(B("adj_tin"), 12, 0, 12, sw(d.p1.wx), d.p1, ones1, B("adj_xin"));
    if ((rc = GLW_STAGE(P, "adj_o3_inner_colfwd", st, colfwd_bytes(ci, d.g2y, O * I), launch_colfwd<double>(d.g2y, ci, O * I, P->tw(d.g2y, false), P->exec(STAGE_COL), st))))
      return rc;
    static const int icodes[12] = {M_ONE, M_NU, M_NU2, M_NU3, M_QX_NU, M_QY_NU, M_QX, M_QY, M_Q2_NU, M_Q2, M_NU_QX, M_NU_QY};
    ColInvArgs fi = colinv_base(stp, 1, O * I, d.p1, P->t1);
    fi.damp = P->damp_oi;
    fi.dmod = O * I;
    fi.nfield = 12;
    for (int f = 0; f < 12; ++f) {
      fi.fsrc[f] = 0;
      fi.fcode[f] = icodes[f];
      fi.fscale[f] = sc;
    }
    if ((rc = collect(fi, B("adj_xin"), 1, gs, st))) return rc;
  }
  hipLaunchKernelGGL(k_adj_phase, dim3((b1 + 255) / 256), dim3(256), 0, st, gs, P->omega, times, b1, gs0);
  GLW_HIP(hipGetLastError());
  return GLW_OK;
}

// Per-stage time (ms, summed over launches) of the events recorded since the
// pool was last reset; resets the pool.
int collect_stage_ms(glw_plan* P, std::map<std::string, double>& agg) {
  agg.clear();
  if (P->ev_dropped) {
    P->ev_used = P->ev_dropped = 0;
    return fail(GLW_EINVAL, "event pool exhausted while timing stages (fewer instants or repetitions)");
  }
  if (P->ev_used) GLW_HIP(hipEventSynchronize(P->ev[P->ev_used - 1]));
  for (size_t i = 0; i + 1 < P->ev_used; i += 2) {
    float t = 0.f;
    GLW_HIP(hipEventElapsedTime(&t, P->ev[i], P->ev[i + 1]));
    agg[P->ev_name[i / 2]] += t;
  }
  P->ev_used = 0;
  return GLW_OK;
}

int enable_events(glw_plan* P, bool on) {
  for (hipEvent_t e : P->ev) (void)hipEventDestroy(e);
  P->ev.clear();
  P->ev_name.clear();
  P->ev_bytes.clear();
  P->ev_used = P->ev_dropped = 0;
  if (!on) return GLW_OK;
  const size_t n = 8192;
  P->ev.resize(n);
  P->ev_name.resize(n / 2);
  P->ev_bytes.resize(n / 2);
  for (size_t i = 0; i < n; ++i) GLW_HIP(hipEventCreate(&P->ev[i]));
  return GLW_OK;
}

// Candidate schedules, tried all at once: pass k gives every row stage
// kRowSched[k] and every column stage kColSched[k].
constexpr Sched kRowSched[6] = {{0, 0}, {0, 1}, {0, 2}, {1, 0}, {1, 1}, {1, 2}};
constexpr Sched kColSched[6] = {{0, 0}, {0, 1}, {0, 2}, {1, 0}, {1, 1}, {1, 2}};

int autotune(glw_plan* P, int kind, const double* s0, const double* times, int nt, double* a, double* b, int reps,
             hipStream_t st) {
  int rc;
  P->sched.clear();
  P->stage_kind.clear();
  if ((rc = glw_execute_kind(P, kind, s0, times, nt, a, b, st))) return rc;  // discovers the stages
  GLW_HIP(hipStreamSynchronize(st));
  std::vector<std::string> names;
  for (const auto& kv : P->stage_kind)
    if (kv.second == STAGE_ROW || kv.second == STAGE_COL) names.push_back(kv.first);
  std::map<std::string, std::pair<double, Sched>> best;
  std::map<std::string, double> agg;
  for (int k = 0; k < 6; ++k) {
    for (const std::string& nm : names)
      P->sched[nm] = P->stage_kind[nm] == STAGE_ROW ? kRowSched[k] : kColSched[k];
    if ((rc = glw_execute_kind(P, kind, s0, times, nt, a, b, st))) return rc;  // warm-up, sizes scratch
    GLW_HIP(hipStreamSynchronize(st));
    P->ev_used = 0;
    for (int r = 0; r < reps; ++r)
      if ((rc = glw_execute_kind(P, kind, s0, times, nt, a, b, st))) return rc;
    if ((rc = collect_stage_ms(P, agg))) return rc;
    for (const std::string& nm : names) {
      const double t = agg.count(nm) ? agg[nm] / reps : 0.0;
      auto it = best.find(nm);
      if (it == best.end() || t < it->second.first) best[nm] = {t, P->sched[nm]};
    }
  }
  P->sched.clear();
  for (const auto& kv : best) P->sched[kv.first] = kv.second.second;
  return GLW_OK;
}

}  // namespace

int glw_execute_kind(glw_plan* P, int kind, const double* s0, const double* times, int nt, double* a, double* b,
                     hipStream_t st) {
  if (!P) return fail(GLW_EINVAL, "null plan");
  switch (kind) {
    case GLW_KIND_LINEAR:
      return run_chunks(P, s0, times, nt, a, b, st, linear_chunk);
    case GLW_KIND_ORDER2:
      if (P->d.order < 2) return fail(GLW_EINVAL, "plan built without order-2 tables");
      return run_chunks(P, s0, times, nt, a, b, st, order2_dispatch);
    case GLW_KIND_ORDER3:
      if (P->d.order < 3) return fail(GLW_EINVAL, "plan built without order-3 tables");
      return run_chunks(P, s0, times, nt, a, b, st, order3_dispatch);
    case GLW_KIND_ORDER4:
      if (!P->d.with_order4) return fail(GLW_EINVAL, "plan built without order-4 tables");
      return run_chunks(P, s0, times, nt, a, nullptr, st, order4_dispatch);
    default:
      return fail(GLW_EINVAL, "unknown evaluation kind " + std::to_string(kind));
  }
}

bool glw_plan_profiling(const glw_plan* P) { return P && !P->ev.empty(); }

int glw_plan_release_scratch(glw_plan* P) {
  if (P->holds > 0) return fail(GLW_EINVAL, "split row scratch is held by a captured graph");
  if (P->split_ptr) GLW_HIP(hipFree(P->split_ptr));
  P->split_ptr = nullptr;
  P->split_bytes = 0;
  return GLW_OK;
}

double glw_plan_cost(const glw_plan* P, int kind, int nt) {
  const glw_plan_desc& d = P->d;
  const double g2 = (double)d.g2x * d.g2y, g3 = (double)d.g3x * d.g3y, n2 = (double)d.nx * d.ny;
  double c = 4.0 * n2;
  if (kind == GLW_KIND_ORDER2) c += (8.0 * d.o2.eta_rank + 64.0 + 2.0) * g2;
  if (kind == GLW_KIND_ORDER3) c += 16.0 * d.inner_rank * d.outer_rank * g2 + 25.0 * d.outer_rank * g3;
  if (kind == GLW_KIND_ORDER4) c += 64.0 * d.o4.rank * d.o4.rank * d.o4.rank * g2;
  return c * nt;
}

// Streaming kernels of the HBM calibration probe (glw_hbm_probe): one 16-byte
// element per thread, so every launch has many short workgroups in flight.
__global__ __launch_bounds__(256) void k_probe_copy(const double2* __restrict__ src, double2* __restrict__ dst,
                                                    size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
__global__ __launch_bounds__(256) void k_probe_fill(double2* __restrict__ dst, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = make_double2(1.0, (double)i);
}
// read-only: the sum is written only if it equals a value no finite input sums to
__global__ __launch_bounds__(256) void k_probe_read(const double2* __restrict__ src, double2* __restrict__ sink,
                                                    size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  const double v = i < n ? src[i].x + src[i].y : 0.0;
  if (v == __builtin_inf()) sink[0] = make_double2(v, v);
}

extern "C" {

int glw_plan_create(const glw_plan_desc* desc, glw_plan** plan) {
  if (!plan) return fail(GLW_EINVAL, "null plan pointer");
  *plan = nullptr;
  int rc = validate(desc);
  if (rc) return rc;
  glw_plan* P = new glw_plan_s();
  P->d = *desc;
  if (hipGetDevice(&P->device) != hipSuccess) {
    delete P;
    return fail(GLW_EDEVICE, "no HIP device");
  }
  rc = build(P);
  if (rc) {
    delete P;
    return rc;
  }
  // host table pointers are not kept past creation
  P->d.t1 = P->d.t2 = P->d.t3 = glw_box_tables{};
  *plan = P;
  return GLW_OK;
}

int glw_plan_destroy(glw_plan* plan) {
  if (plan && plan->holds > 0) return fail(GLW_EINVAL, "plan is held by a captured graph: destroy the graph first");
  delete plan;
  return GLW_OK;
}

int glw_linear_execute(glw_plan* plan, const double* s0, const double* times, int32_t nt, double* eta11,
                       double* psi11, void* stream) {
  return glw_execute_kind(plan, GLW_KIND_LINEAR, s0, times, nt, eta11, psi11, (hipStream_t)stream);
}

int glw_order2_execute(glw_plan* plan, const double* s0, const double* times, int32_t nt, double* eta22,
                       double* psi22, void* stream) {
  return glw_execute_kind(plan, GLW_KIND_ORDER2, s0, times, nt, eta22, psi22, (hipStream_t)stream);
}

int glw_order3_execute(glw_plan* plan, const double* s0, const double* times, int32_t nt, double* eta33,
                       double* psi33, void* stream) {
  return glw_execute_kind(plan, GLW_KIND_ORDER3, s0, times, nt, eta33, psi33, (hipStream_t)stream);
}

int glw_order4_execute(glw_plan* plan, const double* s0, const double* times, int32_t nt, double* eta44,
                       void* stream) {
  return glw_execute_kind(plan, GLW_KIND_ORDER4, s0, times, nt, eta44, nullptr, (hipStream_t)stream);
}

int glw_plan_autotune(glw_plan* plan, int32_t kind, const double* s0, const double* times, int32_t nt, double* a,
                      double* b, int32_t reps, void* stream) {
  if (!plan) return fail(GLW_EINVAL, "null plan");
  if (reps < 1 || reps > 64) return fail(GLW_EINVAL, "reps must be 1..64");
  if (plan->holds > 0) return fail(GLW_EINVAL, "autotuning would resize the scratch of a plan a captured graph holds");
  if (nt < 1) return fail(GLW_EINVAL, "autotuning needs at least one instant");
  hipStream_t st = (hipStream_t)stream;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  GLW_HIP(hipStreamIsCapturing(st, &cs));
  if (cs != hipStreamCaptureStatusNone) return fail(GLW_EINVAL, "autotuning cannot run under graph capture");
  const bool profiling = !plan->ev.empty();
  int rc = enable_events(plan, true);
  if (!rc) rc = autotune(plan, kind, s0, times, nt, a, b, reps, st);
  // size the split-row scratch for the chosen schedule only
  if (!rc) rc = glw_plan_release_scratch(plan);
  if (!rc) rc = glw_execute_kind(plan, kind, s0, times, nt, a, b, st);
  if (!rc) rc = hipStreamSynchronize(st) == hipSuccess ? GLW_OK : fail(GLW_EDEVICE, "stream synchronize failed");
  const int rc2 = enable_events(plan, profiling);
  return rc ? rc : rc2;
}

int glw_adjoint_execute(glw_plan* plan, const double* s0, const double* times, int32_t orders,
                        const double* grad_field, double* grad_s0, int32_t recompute, void* stream) {
  if (!plan) return fail(GLW_EINVAL, "null plan");
  if (!s0 || !times || !grad_field || !grad_s0) return fail(GLW_EINVAL, "null array argument");
  int dev = -1;
  GLW_HIP(hipGetDevice(&dev));
  if (dev != plan->device) GLW_HIP(hipSetDevice(plan->device));
  hipStream_t st = (hipStream_t)stream;
  plan->cur_stream = st;
  return adjoint_run(plan, s0, times, orders, (const cpx<double>*)grad_field, (cpx<double>*)grad_s0, recompute, st);
}

int glw_plan_hold(glw_plan* plan, int32_t hold) {
  if (!plan) return fail(GLW_EINVAL, "null plan");
  if (hold) {
    ++plan->holds;
    return GLW_OK;
  }
  if (plan->holds < 1) return fail(GLW_EINVAL, "plan hold count is already zero");
  --plan->holds;
  return GLW_OK;
}

int glw_plan_get_schedule(glw_plan* plan, char* text, size_t len) {
  if (!plan || !text || len < 1) return fail(GLW_EINVAL, "null argument");
  std::string out;
  for (const auto& kv : plan->sched)
    out += (out.empty() ? "" : ";") + kv.first + "=" + std::to_string(kv.second.plan) + "/" +
           std::to_string(kv.second.tile);
  if (out.size() + 1 > len) return fail(GLW_EINVAL, "schedule text buffer too small");
  memcpy(text, out.c_str(), out.size() + 1);
  return GLW_OK;
}

int glw_plan_set_schedule(glw_plan* plan, const char* text) {
  if (!plan || !text) return fail(GLW_EINVAL, "null argument");
  std::map<std::string, Sched> sched;
  std::string all(text);
  size_t pos = 0;
  while (pos < all.size()) {
    size_t end = all.find(';', pos);
    if (end == std::string::npos) end = all.size();
    const std::string item = all.substr(pos, end - pos);
    pos = end + 1;
    if (item.empty()) continue;
    const size_t eq = item.find('='), sl = item.find('/');
    if (eq == std::string::npos || sl == std::string::npos || sl < eq || eq == 0)
      return fail(GLW_EINVAL, "bad schedule entry '" + item + "' (want stage=plan/tile)");
    const std::string pl = item.substr(eq + 1, sl - eq - 1), tl = item.substr(sl + 1);
    if (pl.size() != 1 || tl.size() != 1 || pl[0] < '0' || pl[0] > '1' || tl[0] < '0' || tl[0] > '2')
      return fail(GLW_EINVAL, "bad schedule entry '" + item + "' (plan 0..1, tile 0..2)");
    sched[item.substr(0, eq)] = Sched{pl[0] - '0', tl[0] - '0'};
  }
  plan->sched = sched;
  return GLW_OK;
}

int glw_plan_profile(glw_plan* plan, int32_t enable) {
  if (!plan) return fail(GLW_EINVAL, "null plan");
  GLW_HIP(hipSetDevice(plan->device));
  return enable_events(plan, enable != 0);
}

int glw_plan_stage_times(glw_plan* plan, int32_t max, char* names, size_t names_len, double* ms,
                         int32_t* launches, double* bytes, int32_t* count) {
  if (!plan || !names || !ms || !launches || !count || max < 1 || names_len < 1) return fail(GLW_EINVAL, "null argument");
  *count = 0;
  names[0] = 0;
  if (plan->ev.empty()) return fail(GLW_EINVAL, "profiling is not enabled on this plan");
  std::vector<std::string> order;
  std::map<std::string, std::pair<double, int>> agg;
  std::map<std::string, double> moved;
  if (plan->ev_used) GLW_HIP(hipEventSynchronize(plan->ev[plan->ev_used - 1]));
  for (size_t i = 0; i + 1 < plan->ev_used; i += 2) {
    float t = 0.f;
    GLW_HIP(hipEventElapsedTime(&t, plan->ev[i], plan->ev[i + 1]));
    std::string nm = plan->ev_name[i / 2];
    if (!agg.count(nm)) order.push_back(nm);
    agg[nm].first += t;
    agg[nm].second += 1;
    moved[nm] += plan->ev_bytes[i / 2];
  }
  std::string csv;
  int k = 0;
  for (const std::string& nm : order) {
    if (k >= max) break;
    ms[k] = agg[nm].first;
    launches[k] = agg[nm].second;
    if (bytes) bytes[k] = moved[nm];
    csv += (k ? "," : "") + nm;
    ++k;
  }
  if (csv.size() + 1 > names_len) return fail(GLW_EINVAL, "stage name buffer too small");
  memcpy(names, csv.c_str(), csv.size() + 1);
  *count = k;
  plan->ev_used = 0;
  if (plan->ev_dropped) {
    plan->ev_dropped = 0;
    return fail(GLW_EINVAL, "event pool exhausted: some launches were not timed");
  }
  return GLW_OK;
}

int glw_plan_buffer(glw_plan* plan, const char* name, void** ptr, size_t* bytes) {
  if (!plan || !name || !ptr || !bytes) return fail(GLW_EINVAL, "null argument");
  auto it = plan->ws.find(name);
  if (it == plan->ws.end()) return fail(GLW_EINVAL, std::string("no workspace buffer ") + name);
  *ptr = it->second.first;
  *bytes = it->second.second;
  return GLW_OK;
}

int glw_fft_length_supported(int32_t n) { return length_supported(n) ? 1 : 0; }

const char* glw_last_error(void) { return g_err.c_str(); }

const char* glw_version(void) { return "glw 0.1.0 gfx950"; }

int glw_hbm_probe(int32_t kind, const double* src, double* dst, size_t bytes, int32_t reps, double* ms) {
  if (!ms || reps < 1 || bytes < 16 || (kind != 1 && !src) || !dst || kind < 0 || kind > 2)
    return fail(GLW_EINVAL, "glw_hbm_probe: bad arguments");
  const size_t n = bytes / 16;
  const dim3 grid((unsigned)((n + 255) / 256));
  if ((n + 255) / 256 > 0x7fffffffull) return fail(GLW_EINVAL, "glw_hbm_probe: buffer too large");
  auto launch = [&]() {
    if (kind == 0) hipLaunchKernelGGL(k_probe_copy, grid, dim3(256), 0, 0, (const double2*)src, (double2*)dst, n);
    else if (kind == 1) hipLaunchKernelGGL(k_probe_fill, grid, dim3(256), 0, 0, (double2*)dst, n);
    else hipLaunchKernelGGL(k_probe_read, grid, dim3(256), 0, 0, (const double2*)src, (double2*)dst, n);
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  GLW_HIP_CHECK(hipEventCreate(&e0));
  GLW_HIP_CHECK(hipEventCreate(&e1));
  launch();  // warm-up
  GLW_HIP_CHECK(hipGetLastError());
  GLW_HIP_CHECK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; ++r) launch();
  GLW_HIP_CHECK(hipEventRecord(e1, 0));
  GLW_HIP_CHECK(hipEventSynchronize(e1));
  float t = 0.0f;
  GLW_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)t / reps;
  return GLW_OK;
}

size_t glw_abi_sizeof(int32_t which) {
  switch (which) {
    case GLW_ABI_BOX: return sizeof(glw_box);
    case GLW_ABI_BOX_TABLES: return sizeof(glw_box_tables);
    case GLW_ABI_ORDER2_DESC: return sizeof(glw_order2_desc);
    case GLW_ABI_ORDER4_DESC: return sizeof(glw_order4_desc);
    case GLW_ABI_PLAN_DESC: return sizeof(glw_plan_desc);
    case GLW_ABI_WIT_DESC: return sizeof(glw_wit_desc);
    case GLW_ABI_BATCH_ITEM: return sizeof(glw_batch_item);
    case GLW_ABI_MODAL_DESC: return sizeof(glw_modal_desc);
    case GLW_ABI_ETA20_DESC: return sizeof(glw_eta20_desc);
    default: return 0;
  }
}

}  // extern "C"
