// WIT sector-44 quartet sum on gfx950.
//
// The reference (paper/order4/cpp/directional_eta4_openmp.cpp:157-453)
// evaluates, for every unordered quartet, the states of all 15 subsets of its
// four parents in increasing degree.  A subset's state depends only on the
// multiset of its parents, so this build tabulates the states of every single,
// pair (i<=j) and triple (i<=j<=k) once and evaluates only the degree-4 level
// per quartet.
//
// Partition sums.  The reference sums over ordered set partitions with
// labelled blocks: block 0 carries the potential phi, blocks 1..n the
// elevations eta, weighted 1/n!.  For a fixed potential block P the n!
// labellings of one unordered partition of the complement C = S\P cancel the
// 1/n!, so with E_n(C) = sum over unordered partitions of C into n blocks of
// prod eta[block]:
//   taylor    = sum_P phi[P] sum_n zd(K_P, n)   E_n(C)
//   lowerVert = sum_P phi[P] sum_n zd(K_P, n+1) E_n(C)
//   lowerTime = sum_P (-i Omega_P) phi[P] sum_n zd(K_P, n) E_n(C)
//   surfaceH  = sum_{P subset of S} i K_P zd(K_P, n) phi[P] E_n(C), sum over n (n=0 for P=S)
//   surfaceV  = sum_{P subset of S} zd(K_P, n+1) phi[P] E_n(C)
// The ordinary source (:264-282) keeps the reference's ordered two-block sum.
// In sector 44 every parent sign is +1, so charge = degree: the retained-V3
// push (:284-298) and the y31 coordinate (:300-336) vanish; only the
// degree-4 near-resonant branch (:393-411) retains a V3 term.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "glw.h"
#include "glw_fft.hpp"
#include "glw_internal.hpp"

using glw::cpx;

namespace {

int wfail(int code, const std::string& m) { return glw_report_error(code, m); }
#define W_HIP(call)                                                                   \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return wfail(GLW_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int WNT = 256;
// Largest per-workgroup LDS accumulator (eta and psi over the pair-sum window
// of the modes); inputs whose window is larger accumulate with global atomics.
constexpr size_t kWitLdsBytes = 128 * 1024;
constexpr size_t kLdsPerCu = 160 * 1024;

// Subset record: state of one subset (multiset of parents) in arithmetic T.
template <typename T>
struct RecT {
  cpx<T> eta, phi, sh0, sh1, sv;
  T kx, ky, om, q, k2;
  T zero_src;  // strict-zero source magnitude of this subset (0 if not strict zero)
  T zero;      // 1 if this subset is a strict zero, else 0
};
using Rec = RecT<double>;

template <typename T>
__device__ __forceinline__ cpx<T> cm(cpx<T> a, cpx<T> b) { return glw::cmul(a, b); }
template <typename T>
__device__ __forceinline__ cpx<T> ca(cpx<T> a, cpx<T> b) { return glw::cadd(a, b); }
template <typename T>
__device__ __forceinline__ cpx<T> cs(cpx<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T>
__device__ __forceinline__ cpx<T> ci(T s, cpx<T> a) { return {-s * a.y, s * a.x}; }  // (i s) a
template <typename T>
__device__ __forceinline__ T cabs2(cpx<T> a) { return hypot(a.x, a.y); }

__host__ __device__ constexpr int popc(int m) { return (m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1) + ((m >> 3) & 1); }
__host__ __device__ constexpr int lowbit(int m) { return m & (-m); }

template <typename T>
__device__ __forceinline__ T zd(const RecT<T>& r, int n) {
  switch (n) {
    case 0: return T(1);
    case 1: return r.q;
    case 2: return r.k2;
    case 3: return r.k2 * r.q;
    default: return r.k2 * r.k2;
  }
}

struct Model {
  double g, h, thr;
  template <typename T>
  __device__ T q(T kx, T ky, T* k2) const {
    const T k = hypot(kx, ky);
    *k2 = k * k;
    return k * tanh(T(h) * k);
  }
};

template <typename T>
struct TopOut {
  cpx<T> eta, psi, phi, src0, src1, retained;
  T rho;
  bool zero, near;
  T zero_src;
};

template <typename T>
struct Limits;
template <>
struct Limits<double> {
  static constexpr double eps = 2.220446049250313e-16, tiny = 2.2250738585072014e-308;
};
template <>
struct Limits<float> {
  static constexpr float eps = 1.1920928955078125e-7f, tiny = 1.1754943508222875e-38f;
};

// State of the full set FULL = (1<<D)-1 from the records of its proper
// subsets.  `rec(m)` returns the record of local sub-mask m.
template <int D, class Acc, typename T>
__device__ __forceinline__ void subset_state(const Acc& rec, const RecT<T>& S, const Model& md, TopOut<T>& o) {
  constexpr int FULL = (1 << D) - 1;
  const T g = T(md.g);
  cpx<T> horiz = {0, 0}, kin = {0, 0}, taylor = {0, 0}, lv = {0, 0}, lt = {0, 0};
#pragma unroll
  for (int L = 1; L < FULL; ++L) {
    const int R = FULL ^ L;
    const RecT<T>& a = rec(L);
    const RecT<T>& b = rec(R);
    horiz = ca(horiz, cm(ca(ci(b.kx, a.sh0), ci(b.ky, a.sh1)), b.eta));
    kin = ca(kin, cs(ca(ca(cm(a.sh0, b.sh0), cm(a.sh1, b.sh1)), cm(a.sv, b.sv)), T(0.5)));
  }
#pragma unroll
  for (int P = 1; P < FULL; ++P) {
    const int C = FULL ^ P;
    const RecT<T>& p = rec(P);
    cpx<T> t = {0, 0}, v = {0, 0};
    const cpx<T> e1 = rec(C).eta;
    t = ca(t, cs(e1, zd(p, 1)));
    v = ca(v, cs(e1, zd(p, 2)));
    if (popc(C) == 2) {
      const int a = lowbit(C), b = C ^ a;
      const cpx<T> e2 = cm(rec(a).eta, rec(b).eta);
      t = ca(t, cs(e2, zd(p, 2)));
      v = ca(v, cs(e2, zd(p, 3)));
    } else if (popc(C) == 3) {
      const int a = lowbit(C), b = lowbit(C ^ a), c = C ^ a ^ b;
      const cpx<T> e2 = ca(ca(cm(rec(a).eta, rec(b | c).eta), cm(rec(b).eta, rec(a | c).eta)),
                           cm(rec(c).eta, rec(a | b).eta));
      const cpx<T> e3 = cm(cm(rec(a).eta, rec(b).eta), rec(c).eta);
      t = ca(ca(t, cs(e2, zd(p, 2))), cs(e3, zd(p, 3)));
      v = ca(ca(v, cs(e2, zd(p, 3))), cs(e3, zd(p, 4)));
    }
    const cpx<T> pt = cm(p.phi, t);
    taylor = ca(taylor, pt);
    lv = ca(lv, cm(p.phi, v));
    lt = ca(lt, ci(-p.om, pt));
  }
  // source = [-(horiz - lv) - q taylor, -(lt + kin) - i Omega taylor]
  o.src0 = {-(horiz.x - lv.x) - S.q * taylor.x, -(horiz.y - lv.y) - S.q * taylor.y};
  const cpx<T> it = ci(S.om, taylor);
  o.src1 = {-(lt.x + kin.x) - it.x, -(lt.y + kin.y) - it.y};
  const T mag = sqrt(S.k2);
  o.zero = mag <= T(100) * Limits<T>::eps * fmax(mag, T(1));
  o.near = false;
  o.rho = T(0);
  o.retained = {0, 0};
  o.zero_src = T(0);
  if (o.zero) {
    o.eta = o.psi = {0, 0};
    o.zero_src = hypot(cabs2(o.src0), cabs2(o.src1));
  } else {
    bool solve = true;
    if constexpr (D == 4) {
      const T lin = g * S.q, tmp = S.om * S.om;
      o.rho = fabs(lin - tmp) / fmax(lin + tmp, Limits<T>::tiny);
      if (o.rho < T(md.thr)) {
        const T w = sqrt(lin), ms = sqrt(w / (T(2) * g)), ns = sqrt(g / (T(2) * w));
        const T sg = S.om >= T(0) ? T(1) : T(-1);
        const cpx<T> a0 = cs(o.src0, T(1) / (T(2) * ms));
        const cpx<T> a1 = ci(sg, cs(o.src1, T(1) / (T(2) * ns)));
        o.retained = ca(a0, a1);
        const cpx<T> comp = glw::csub(a0, a1);
        const T den = S.om + sg * w;
        o.eta = cs(ci(ms, comp), T(1) / den);
        o.psi = cs(comp, -sg * ns / den);
        o.near = true;
        solve = false;
      }
    }
    if (solve) {
      const T den = S.om * S.om - g * S.q;
      const cpx<T> a = glw::csub(ci(S.om, o.src0), cs(o.src1, S.q));
      const cpx<T> b = ca(cs(o.src0, g), ci(S.om, o.src1));
      o.eta = cs(a, T(1) / den);
      o.psi = cs(b, T(1) / den);
    }
  }
  o.phi = glw::csub(o.psi, taylor);
}

// Surface derivatives of the full set (degree < 4), reference :202-225.
template <int D, class Acc>
__device__ __forceinline__ void surface(const Acc& rec, const Rec& S, cpx<double> phiS, Rec& out) {
  constexpr int FULL = (1 << D) - 1;
  cpx<double> h0 = ci(S.kx, phiS), h1 = ci(S.ky, phiS), sv = cs(phiS, S.q);
#pragma unroll
  for (int P = 1; P < FULL; ++P) {
    const int C = FULL ^ P;
    const Rec& p = rec(P);
    cpx<double> e[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    e[1] = rec(C).eta;
    if (popc(C) == 2) {
      const int a = lowbit(C), b = C ^ a;
      e[2] = cm(rec(a).eta, rec(b).eta);
    } else if (popc(C) == 3) {
      const int a = lowbit(C), b = lowbit(C ^ a), c = C ^ a ^ b;
      e[2] = ca(ca(cm(rec(a).eta, rec(b | c).eta), cm(rec(b).eta, rec(a | c).eta)), cm(rec(c).eta, rec(a | b).eta));
      e[3] = cm(cm(rec(a).eta, rec(b).eta), rec(c).eta);
    }
#pragma unroll
    for (int n = 1; n <= 3; ++n) {
      if (n > popc(C)) break;
      const cpx<double> c = cm(p.phi, e[n]);
      const double z = zd(p, n);
      h0 = ca(h0, ci(p.kx * z, c));
      h1 = ca(h1, ci(p.ky * z, c));
      sv = ca(sv, cs(c, zd(p, n + 1)));
    }
  }
  out.sh0 = h0;
  out.sh1 = h1;
  out.sv = sv;
}

struct WitDev {
  int M, nx, ny;
  Model md;
  const int* modes;      // [M][2]
  Rec* single;           // [M]
  Rec* pair;             // [M(M+1)/2]
  Rec* triple;           // [C(M+2,3)]
  RecT<float>* single32; // FP32 copies of the tables (FP32 quartet evaluation only)
  RecT<float>* pair32;
  RecT<float>* triple32;
  const long long* offP; // [M+1]  pairs with first index < i
  const long long* offT; // [M+1]  triples with first index < i
};

__device__ __forceinline__ long long pidx(const WitDev& w, int i, int j) { return w.offP[i] + (j - i); }
__device__ __forceinline__ long long tidx(const WitDev& w, int i, int j, int k) {
  return w.offT[i] + (w.offP[j] - w.offP[i]) + (k - j);
}

template <typename T>
__device__ void base_record(const WitDev& w, const RecT<T>** sub, int D, RecT<T>& S) {
  T kx = 0, ky = 0, om = 0;
  for (int s = 0; s < D; ++s) {
    kx += sub[s]->kx;
    ky += sub[s]->ky;
    om += sub[s]->om;
  }
  S.kx = kx;
  S.ky = ky;
  S.om = om;
  S.q = w.md.q(kx, ky, &S.k2);
}

__global__ void k_wit_single_fill(WitDev w, double dkx, double dky) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.M) return;
  Rec r;
  r.kx = w.modes[2 * i] * dkx;
  r.ky = w.modes[2 * i + 1] * dky;
  r.q = w.md.q(r.kx, r.ky, &r.k2);
  r.om = sqrt(w.md.g * r.q);
  r.eta = {0.5, 0.0};
  r.phi = {0.0, -r.om / (2.0 * r.q)};
  r.sh0 = ci(r.kx, r.phi);
  r.sh1 = ci(r.ky, r.phi);
  r.sv = cs(r.phi, r.q);
  r.zero = 0.0;
  r.zero_src = 0.0;
  w.single[i] = r;
}

template <typename T, int N>
struct AccT {
  const RecT<T>* r[N];
  __device__ const RecT<T>& operator()(int m) const { return *r[m]; }
};
using Acc2 = AccT<double, 4>;
using Acc3 = AccT<double, 8>;

template <typename T>
struct Tables;
template <>
struct Tables<double> {
  static __device__ const Rec* single(const WitDev& w) { return w.single; }
  static __device__ const Rec* pair(const WitDev& w) { return w.pair; }
  static __device__ const Rec* triple(const WitDev& w) { return w.triple; }
};
template <>
struct Tables<float> {
  static __device__ const RecT<float>* single(const WitDev& w) { return w.single32; }
  static __device__ const RecT<float>* pair(const WitDev& w) { return w.pair32; }
  static __device__ const RecT<float>* triple(const WitDev& w) { return w.triple32; }
};

__device__ __forceinline__ RecT<float> to_f32(const Rec& r) {
  auto c = [](cpx<double> v) { return cpx<float>{(float)v.x, (float)v.y}; };
  return RecT<float>{c(r.eta), c(r.phi), c(r.sh0), c(r.sh1), c(r.sv), (float)r.kx, (float)r.ky, (float)r.om,
                     (float)r.q, (float)r.k2, (float)r.zero_src, (float)r.zero};
}

__global__ void k_wit_to_f32(const Rec* src, RecT<float>* dst, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = to_f32(src[i]);
}

__global__ void k_wit_pairs(WitDev w) {
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long np = (long long)w.M * (w.M + 1) / 2;
  if (p >= np) return;
  // unrank p -> (i, j): largest i with offP[i] <= p
  int lo = 0, hi = w.M - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (w.offP[mid] <= p) lo = mid; else hi = mid - 1;
  }
  const int i = lo, j = i + (int)(p - w.offP[i]);
  Acc2 a;
  a.r[1] = &w.single[i];
  a.r[2] = &w.single[j];
  const Rec* sub[2] = {a.r[1], a.r[2]};
  Rec S;
  base_record(w, sub, 2, S);
  TopOut<double> o;
  subset_state<2>(a, S, w.md, o);
  S.eta = o.eta;
  S.phi = o.phi;
  S.zero = o.zero ? 1.0 : 0.0;
  S.zero_src = o.zero_src;
  surface<2>(a, S, o.phi, S);
  w.pair[p] = S;
}

__global__ void k_wit_triples(WitDev w, long long nt) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int lo = 0, hi = w.M - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (w.offT[mid] <= t) lo = mid; else hi = mid - 1;
  }
  const int i = lo;
  // within first index i: pair rank among (j,k), i<=j<=k
  const long long r = t - w.offT[i] + w.offP[i];
  int lj = i, hj = w.M - 1;
  while (lj < hj) {
    const int mid = (lj + hj + 1) / 2;
    if (w.offP[mid] <= r) lj = mid; else hj = mid - 1;
  }
  const int j = lj, k = j + (int)(r - w.offP[j]);
  Acc3 a;
  a.r[1] = &w.single[i];
  a.r[2] = &w.single[j];
  a.r[4] = &w.single[k];
  a.r[3] = &w.pair[pidx(w, i, j)];
  a.r[5] = &w.pair[pidx(w, i, k)];
  a.r[6] = &w.pair[pidx(w, j, k)];
  const Rec* sub[3] = {a.r[1], a.r[2], a.r[4]};
  Rec S;
  base_record(w, sub, 3, S);
  TopOut<double> o;
  subset_state<3>(a, S, w.md, o);
  S.eta = o.eta;
  S.phi = o.phi;
  S.zero = o.zero ? 1.0 : 0.0;
  S.zero_src = o.zero_src;
  surface<3>(a, S, o.phi, S);
  w.triple[t] = S;
}

struct QuartetArgs {
  WitDev w;
  const cpx<double>* amps;  // [cases][M]
  int cases;
  double* out;              // [5][cases][ny][nx] complex
  unsigned long long* cnt;  // kept, near, zero_out, zero_int
  unsigned long long* dmax; // bits: max residual, max zero source, min detuning
  // Workgroup accumulator in LDS (0 bins: off).  The quartets of one leading
  // pair (i, j) land on mode(i) + mode(j) + the pair sums of the other two,
  // all inside the pw x ph box starting at (pxmin, pymin) = 2 min(mode).
  int lds_bins, pw, ph, pxmin, pymin;
  long long pair0;          // leading pair of workgroup 0 (first-index ranges)
};

__device__ __forceinline__ void atomic_addc(double* p, cpx<double> v) {
  unsafeAtomicAdd(p, v.x);
  unsafeAtomicAdd(p + 1, v.y);
}

__device__ __forceinline__ unsigned long long wsum_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wmax_d(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wmin_d(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

// One workgroup per leading pair (i, j); threads sweep the (k, l) triangle.
// T = arithmetic of the degree-4 evaluation (the subset tables are built in
// FP64; the FP32 path reads FP32 copies).  Weights and accumulation are FP64
// in both: every FP32 term is exact in FP64 and the FP64 sums bound the
// summation error far below a compensated FP32 sum.
// OCC: minimum waves per SIMD asked of the register allocator (2 fits the FP32
// evaluation with a small spill; used when the LDS window leaves room for two
// workgroups per CU).
template <typename T, int OCC = 1>
__global__ __launch_bounds__(WNT, OCC) void k_wit_quartets(QuartetArgs g) {
  const WitDev& w = g.w;
  const long long p = g.pair0 + blockIdx.x;
  int lo = 0, hi = w.M - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (w.offP[mid] <= p) lo = mid; else hi = mid - 1;
  }
  const int i = lo, j = i + (int)(p - w.offP[i]);
  const int n = w.M - j;  // k, l in [j, M)
  const long long tri = (long long)n * (n + 1) / 2;
  unsigned long long kept = 0, nearc = 0, zo = 0, zi = 0;
  double maxres = 0.0, maxz = 0.0, mindet = INFINITY;
  const size_t plane = (size_t)w.nx * w.ny * g.cases;
  // LDS accumulator [bin][case][eta, psi] (complex), cleared before use
  extern __shared__ double wacc[];
  const int nacc = g.lds_bins * g.cases * 4;
  const int sx = w.modes[2 * i] + w.modes[2 * j], sy = w.modes[2 * i + 1] + w.modes[2 * j + 1];
  if (g.lds_bins) {
    for (int e = threadIdx.x; e < nacc; e += WNT) wacc[e] = 0.0;
    __syncthreads();
  }
  for (long long f = threadIdx.x; f < tri; f += WNT) {
    // unrank f -> (k', l'), 0 <= k' <= l' < n: row k' starts at k' n - k'(k'-1)/2
    double disc = (2.0 * n + 1.0) * (2.0 * n + 1.0) - 8.0 * (double)f;
    int kk = (int)floor(((2.0 * n + 1.0) - sqrt(fmax(disc, 0.0))) / 2.0);
    if (kk < 0) kk = 0;
    if (kk > n - 1) kk = n - 1;
    while (kk > 0 && (long long)kk * n - (long long)kk * (kk - 1) / 2 > f) --kk;
    while (kk + 1 < n && (long long)(kk + 1) * n - (long long)(kk + 1) * kk / 2 <= f) ++kk;
    const long long start = (long long)kk * n - (long long)kk * (kk - 1) / 2;
    const int k = j + kk, l = k + (int)(f - start);
    const int tx = w.modes[2 * i] + w.modes[2 * j] + w.modes[2 * k] + w.modes[2 * l];
    const int ty = w.modes[2 * i + 1] + w.modes[2 * j + 1] + w.modes[2 * k + 1] + w.modes[2 * l + 1];
    if (tx == 0 && ty == 0) {
      ++zo;
      continue;
    }
    if (abs(tx) >= w.nx / 2 || abs(ty) >= w.ny / 2) continue;
    const RecT<T>* single = Tables<T>::single(w);
    const RecT<T>* pair = Tables<T>::pair(w);
    const RecT<T>* triple = Tables<T>::triple(w);
    AccT<T, 16> a;
    a.r[1] = &single[i];
    a.r[2] = &single[j];
    a.r[4] = &single[k];
    a.r[8] = &single[l];
    a.r[3] = &pair[pidx(w, i, j)];
    a.r[5] = &pair[pidx(w, i, k)];
    a.r[9] = &pair[pidx(w, i, l)];
    a.r[6] = &pair[pidx(w, j, k)];
    a.r[10] = &pair[pidx(w, j, l)];
    a.r[12] = &pair[pidx(w, k, l)];
    a.r[7] = &triple[tidx(w, i, j, k)];
    a.r[11] = &triple[tidx(w, i, j, l)];
    a.r[13] = &triple[tidx(w, i, k, l)];
    a.r[14] = &triple[tidx(w, j, k, l)];
    const RecT<T>* sub[4] = {a.r[1], a.r[2], a.r[4], a.r[8]};
    RecT<T> S;
    base_record(w, sub, 4, S);
    TopOut<T> o;
    subset_state<4>(a, S, w.md, o);
    // strict-zero intermediates of the ten proper subsets of degree >= 2
    double zcount = 0.0, zsrc = 0.0;
    constexpr int kInner[10] = {3, 5, 6, 9, 10, 12, 7, 11, 13, 14};
#pragma unroll
    for (int q = 0; q < 10; ++q) {
      zcount += (double)a.r[kInner[q]]->zero;
      zsrc = fmax(zsrc, (double)a.r[kInner[q]]->zero_src);
    }
    // paired residual (:413-429)
    const T gT = T(w.md.g);
    cpx<T> er = glw::csub(glw::csub(ci(-S.om, o.eta), cs(o.psi, S.q)), o.src0);
    cpx<T> pr = glw::csub(ca(ci(-S.om, o.psi), cs(o.eta, gT)), o.src1);
    if (o.near) {
      const T ww = sqrt(gT * S.q), ms = sqrt(ww / (T(2) * gT)), ns = sqrt(gT / (T(2) * ww));
      const T sg = S.om >= T(0) ? T(1) : T(-1);
      er = ca(er, cs(o.retained, ms));
      pr = ca(pr, ci(-sg * ns, o.retained));
    }
    const double res = (double)hypot(cabs2(er), cabs2(pr));
    // repeated-index factor (:486-499)
    const int id[4] = {i, j, k, l};
    int fac = 1, run = 1;
#pragma unroll
    for (int s = 1; s < 4; ++s) {
      if (id[s] == id[s - 1]) {
        ++run;
      } else {
        for (int v = 2; v <= run; ++v) fac *= v;
        run = 1;
      }
    }
    for (int v = 2; v <= run; ++v) fac *= v;
    const size_t ix = (size_t)(((tx % w.nx) + w.nx) % w.nx), iy = (size_t)(((ty % w.ny) + w.ny) % w.ny);
    const cpx<double> eta = {(double)o.eta.x, (double)o.eta.y}, psi = {(double)o.psi.x, (double)o.psi.y};
    const cpx<double> ret = {(double)o.retained.x, (double)o.retained.y};
    for (int c = 0; c < g.cases; ++c) {
      const cpx<double>* ac = g.amps + (size_t)c * w.M;
      cpx<double> wt = cm(cm(cm(cs(ac[i], 2.0), ac[j]), ac[k]), ac[l]);
      wt = cs(wt, 1.0 / (double)fac);
      const size_t loc = (size_t)c * w.nx * w.ny + ix + (size_t)w.nx * iy;
      if (g.lds_bins) {
        const int lb = (ty - sy - g.pymin) * g.pw + (tx - sx - g.pxmin);
        double* a4 = wacc + ((size_t)lb * g.cases + c) * 4;
        const cpx<double> ve = cm(wt, eta), vp = cm(wt, psi);
        atomicAdd(a4 + 0, ve.x);
        atomicAdd(a4 + 1, ve.y);
        atomicAdd(a4 + 2, vp.x);
        atomicAdd(a4 + 3, vp.y);
      } else {
        atomic_addc(g.out + 2 * loc, cm(wt, eta));
        atomic_addc(g.out + 2 * (plane + loc), cm(wt, psi));
      }
      if (o.near) atomic_addc(g.out + 2 * (4 * plane + loc), cm(wt, ret));
    }
    ++kept;
    nearc += o.near ? 1 : 0;
    zi += (unsigned long long)zcount;
    maxres = fmax(maxres, res);
    mindet = fmin(mindet, (double)o.rho);
    maxz = fmax(maxz, zsrc);
  }
  if (g.lds_bins) {  // flush: one global add per touched bin
    __syncthreads();
    for (int e = threadIdx.x; e < g.lds_bins * g.cases; e += WNT) {
      const double* a4 = wacc + (size_t)e * 4;
      if (a4[0] == 0.0 && a4[1] == 0.0 && a4[2] == 0.0 && a4[3] == 0.0) continue;
      const int lb = e / g.cases, c = e - lb * g.cases;
      const int tx = sx + g.pxmin + lb % g.pw, ty = sy + g.pymin + lb / g.pw;
      const size_t ix = (size_t)(((tx % w.nx) + w.nx) % w.nx), iy = (size_t)(((ty % w.ny) + w.ny) % w.ny);
      const size_t loc = (size_t)c * w.nx * w.ny + ix + (size_t)w.nx * iy;
      atomic_addc(g.out + 2 * loc, cpx<double>{a4[0], a4[1]});
      atomic_addc(g.out + 2 * (plane + loc), cpx<double>{a4[2], a4[3]});
    }
  }
  kept = wsum_u64(kept);
  nearc = wsum_u64(nearc);
  zo = wsum_u64(zo);
  zi = wsum_u64(zi);
  maxres = wmax_d(maxres);
  maxz = wmax_d(maxz);
  mindet = wmin_d(mindet);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(g.cnt + 0, kept);
    atomicAdd(g.cnt + 1, nearc);
    atomicAdd(g.cnt + 2, zo);
    atomicAdd(g.cnt + 3, zi);
    atomicMax(g.dmax + 0, (unsigned long long)__double_as_longlong(maxres));
    atomicMax(g.dmax + 1, (unsigned long long)__double_as_longlong(maxz));
    if (mindet < INFINITY) atomicMin(g.dmax + 2, (unsigned long long)__double_as_longlong(mindet));
  }
}

// ----------------------------------------------------------------------------
// Sectors 40 and 42: quartets with signed parents (directional_eta4_openmp.cpp
// sector loops :603-688).  A subset's state now depends on its members' signs,
// so nothing is tabulated: every candidate evaluates all 15 subset states
// (reference evaluate, :157-453) in registers and private memory, including
// the two terms that vanish in sector 44: the retained-V3 state of a cubic
// subset whose charge is not +-3 (and its push on the quartic source,
// :282-298) and the y31 coordinate of the quartic output (:300-336).
// ----------------------------------------------------------------------------
struct Sgn {
  double g, h, thr;
  __device__ double q(double kx, double ky) const {
    const double r = hypot(kx, ky);
    return r * tanh(h * r);
  }
  __device__ double om(double kx, double ky) const { return sqrt(g * q(kx, ky)); }
  // vertical-derivative factor |k|^n -> q, k^2, k^2 q, k^4 (n = 0..4)
  __device__ double zd(double kx, double ky, int n) const {
    const double k2 = kx * kx + ky * ky;
    switch (n) {
      case 0: return 1.0;
      case 1: return q(kx, ky);
      case 2: return k2;
      case 3: return k2 * q(kx, ky);
      default: return k2 * k2;
    }
  }
  // 2x2 forced linear response (:111-121)
  __device__ void solve(cpx<double> s0, cpx<double> s1, double kx, double ky, double w, cpx<double>& eta,
                        cpx<double>& psi) const {
    const double qq = q(kx, ky), den = w * w - g * qq;
    eta = {(-w * s0.y - qq * s1.x) / den, (w * s0.x - qq * s1.y) / den};
    psi = {(g * s0.x - w * s1.y) / den, (g * s0.y + w * s1.x) / den};
  }
  __device__ double g1(double ax, double ay, double bx, double by) const {
    return -q(ax + bx, ay + by) * q(bx, by) + (bx * bx + by * by) + (ax * bx + ay * by);
  }
  // second-order response of two physical states (:136-146)
  __device__ void cross(double ax, double ay, double wa, cpx<double> ea, cpx<double> pa, double bx, double by,
                        double wb, cpx<double> eb, cpx<double> pb, cpx<double>& eta, cpx<double>& psi) const {
    eta = psi = {0.0, 0.0};
    if (hypot(ax + bx, ay + by) == 0.0) return;
    const double gab = g1(ax, ay, bx, by), gba = g1(bx, by, ax, ay);
    const cpx<double> s0 = glw::cadd(cs(glw::cmul(ea, pb), gab), cs(glw::cmul(eb, pa), gba));
    const cpx<double> s1 = cs(glw::cmul(pa, pb), ax * bx + ay * by + q(ax, ay) * q(bx, by));
    solve(s0, s1, ax + bx, ay + by, wa + wb, eta, psi);
  }
  // normal-form map of a cubic mode (unit normal amplitude) and a parent (:122-135)
  __device__ void normal(double ax, double ay, int sa, double bx, double by, int sb, cpx<double>& eta,
                         cpx<double>& psi) const {
    const double wa = om(ax, ay), wb = om(bx, by);
    const double ma = sqrt(wa / (2.0 * g)), na = sqrt(g / (2.0 * wa));
    const cpx<double> ea = {ma, 0.0}, pa = {0.0, -sa * na};
    const cpx<double> eb = {0.5, 0.0}, pb = {0.0, -sb * wb / (2.0 * q(bx, by))};
    const double gab = g1(ax, ay, bx, by), gba = g1(bx, by, ax, ay);
    const cpx<double> s0 = glw::cadd(cs(glw::cmul(ea, pb), gab), cs(glw::cmul(eb, pa), gba));
    const cpx<double> s1 = cs(glw::cmul(pa, pb), ax * bx + ay * by + q(ax, ay) * q(bx, by));
    solve(s0, s1, ax + bx, ay + by, sa * wa + sb * wb, eta, psi);
  }
};

struct SignedOut {
  cpx<double> eta, psi, ecoord, pcoord, retained;
  double rho, residual, zsrc;
  int zcount;
  bool near;
};

__device__ __forceinline__ cpx<double> mul_i(double s, cpx<double> a) { return {-s * a.y, s * a.x}; }  // (i s) a

// State of the four-parent set for parents `pk` (physical wavevectors) with signs `sg`.
__device__ SignedOut evaluate_signed(const Sgn& m, const double (&pkx)[4], const double (&pky)[4], const int (&sg)[4]) {
  const double inv_fact[4] = {1.0, 1.0, 0.5, 1.0 / 6.0};
  double kx[16], ky[16], w[16];
  int deg[16], chg[16];
  double skx[4], sky[4], sw[4];
  for (int s = 0; s < 4; ++s) {
    skx[s] = sg[s] * pkx[s];
    sky[s] = sg[s] * pky[s];
    sw[s] = sg[s] * m.om(pkx[s], pky[s]);
  }
  for (int mask = 0; mask < 16; ++mask) {
    kx[mask] = ky[mask] = w[mask] = 0.0;
    deg[mask] = chg[mask] = 0;
    for (int s = 0; s < 4; ++s)
      if (mask >> s & 1) {
        kx[mask] += skx[s];
        ky[mask] += sky[s];
        w[mask] += sw[s];
        deg[mask] += 1;
        chg[mask] += sg[s];
      }
  }
  cpx<double> eta[16], phi[16], psi[16], ret[16], shx[16], shy[16], sv[16];
  for (int mask = 0; mask < 16; ++mask) eta[mask] = phi[mask] = psi[mask] = ret[mask] = shx[mask] = shy[mask] = sv[mask] = {0.0, 0.0};
  SignedOut o{};
  cpx<double> eff0 = {0.0, 0.0}, eff1 = {0.0, 0.0};
  // Partitions of `mask` into nb labelled non-empty blocks: the members are
  // assigned base-nb digits of `code`; block 0 carries the potential, blocks
  // 1.. the elevations (product `ep`).  f(block0, order, ep) is applied to each.
  auto partitions = [&](int mask, int nb, auto&& f) {
    int slots[4], n = 0;
    for (int s = 0; s < 4; ++s)
      if (mask >> s & 1) slots[n++] = s;
    int count = 1;
    for (int i = 0; i < n; ++i) count *= nb;
    for (int code = 0; code < count; ++code) {
      int blocks[4] = {0, 0, 0, 0}, seen = 0, c = code;
      for (int i = 0; i < n; ++i) {
        const int a = c % nb;
        c /= nb;
        seen |= 1 << a;
        blocks[a] |= 1 << slots[i];
      }
      if (seen != (1 << nb) - 1) continue;
      cpx<double> ep = {1.0, 0.0};
      for (int b = 1; b < nb; ++b) ep = glw::cmul(ep, eta[blocks[b]]);
      f(blocks, ep);
    }
  };
  for (int d = 1; d <= 4; ++d) {
    for (int mask = 1; mask < 16; ++mask) {
      if (deg[mask] != d) continue;
      cpx<double> taylor = {0.0, 0.0};
      if (d == 1) {
        eta[mask] = {0.5, 0.0};
        phi[mask] = {0.0, -w[mask] / (2.0 * m.q(kx[mask], ky[mask]))};
        psi[mask] = phi[mask];
      } else {
        // ordinary two-block source (:264-282) and the Taylor terms of orders 1..d-1
        cpx<double> hor = {0.0, 0.0}, kin = {0.0, 0.0}, lv = {0.0, 0.0}, lt = {0.0, 0.0};
        partitions(mask, 2, [&](const int* b, cpx<double>) {
          const int l = b[0], r = b[1];
          const cpx<double> hk = glw::cadd(glw::cmul(shx[l], {0.0, kx[r]}), glw::cmul(shy[l], {0.0, ky[r]}));
          hor = glw::cadd(hor, glw::cmul(hk, eta[r]));
          const cpx<double> kk = glw::cadd(glw::cadd(glw::cmul(shx[l], shx[r]), glw::cmul(shy[l], shy[r])),
                                           glw::cmul(sv[l], sv[r]));
          kin = glw::cadd(kin, cs(kk, 0.5));
        });
        for (int order = 1; order < d; ++order)
          partitions(mask, order + 1, [&](const int* b, cpx<double> ep) {
            const int pm = b[0];
            const cpx<double> base = cs(glw::cmul(phi[pm], ep), inv_fact[order]);
            const double z0 = m.zd(kx[pm], ky[pm], order), z1 = m.zd(kx[pm], ky[pm], order + 1);
            taylor = glw::cadd(taylor, cs(base, z0));
            lv = glw::cadd(lv, cs(base, z1));
            lt = glw::cadd(lt, mul_i(-w[pm] * z0, base));
          });
        const double qm = m.q(kx[mask], ky[mask]);
        const cpx<double> src0 = {-(hor.x - lv.x) - qm * taylor.x, -(hor.y - lv.y) - qm * taylor.y};
        const cpx<double> it = mul_i(w[mask], taylor);
        const cpx<double> src1 = {-(lt.x + kin.x) - it.x, -(lt.y + kin.y) - it.y};
        const double mag = hypot(kx[mask], ky[mask]);
        if (mag <= 100.0 * 2.220446049250313e-16 * fmax(mag, 1.0)) {  // strict zero (:356-364)
          eta[mask] = psi[mask] = {0.0, 0.0};
          o.zcount += 1;
          o.zsrc = fmax(o.zsrc, hypot(hypot(src0.x, src0.y), hypot(src1.x, src1.y)));
        } else if (d == 2 || (d == 3 && abs(chg[mask]) == 3)) {
          m.solve(src0, src1, kx[mask], ky[mask], w[mask], eta[mask], psi[mask]);
        } else {
          cpx<double> f0 = src0, f1 = src1;
          bool resonant = d == 3;  // a cubic subset of charge +-1 keeps its free-wave part (:370-386)
          int ts = chg[mask] > 0 ? 1 : -1;
          if (d == 4) {
            // push of the retained cubic states on the quartic source (:282-298)
            for (int s = 0; s < 4; ++s) {
              const int cm = mask & ~(1 << s);
              if (ret[cm].x == 0.0 && ret[cm].y == 0.0) continue;
              cpx<double> pe, pp;
              m.normal(kx[cm], ky[cm], chg[cm] > 0 ? 1 : -1, skx[s], sky[s], sg[s], pe, pp);
              f0 = glw::csub(f0, glw::cmul(ret[cm], pe));
              f1 = glw::csub(f1, glw::cmul(ret[cm], pp));
            }
            eff0 = f0;
            eff1 = f1;
            // y31 coordinate (:300-336)
            for (int s = 0; s < 4; ++s) {
              const int cm = mask & ~(1 << s), cc = chg[cm];
              if (abs(cc) != 1) continue;
              if (hypot(eta[cm].x, eta[cm].y) + hypot(psi[cm].x, psi[cm].y) == 0.0) continue;
              const double cf = m.om(kx[cm], ky[cm]);
              const cpx<double> ps = mul_i(cc * cf / m.g, psi[cm]);
              const cpx<double> ae = cs(glw::cadd(eta[cm], ps), 0.5), oe = cs(glw::csub(eta[cm], ps), 0.5);
              const cpx<double> ap = mul_i(-m.g * cc / cf, ae), op = mul_i(m.g * cc / cf, oe);
              const double sq = m.q(skx[s], sky[s]);
              const cpx<double> se = {0.5, 0.0}, spsi = {0.0, -sg[s] * m.om(pkx[s], pky[s]) / (2.0 * sq)};
              cpx<double> e1, p1, e2, p2;
              m.cross(kx[cm], ky[cm], cc * cf, ae, ap, skx[s], sky[s], sw[s], se, spsi, e1, p1);
              m.cross(kx[cm], ky[cm], -cc * cf, oe, op, skx[s], sky[s], sw[s], se, spsi, e2, p2);
              o.ecoord = glw::cadd(o.ecoord, glw::cadd(e1, e2));
              o.pcoord = glw::cadd(o.pcoord, glw::cadd(p1, p2));
            }
            const double qo = m.q(kx[mask], ky[mask]), lin = m.g * qo, tmp = w[mask] * w[mask];
            const double rho = fabs(lin - tmp) / fmax(lin + tmp, 2.2250738585072014e-308);
            resonant = rho < m.thr;
            ts = w[mask] >= 0.0 ? 1 : -1;
          }
          if (resonant) {
            const double wo = m.om(kx[mask], ky[mask]);
            const double ms = sqrt(wo / (2.0 * m.g)), ns = sqrt(m.g / (2.0 * wo));
            const cpx<double> a0 = cs(f0, 1.0 / (2.0 * ms)), a1 = mul_i(ts, cs(f1, 1.0 / (2.0 * ns)));
            ret[mask] = glw::cadd(a0, a1);
            const cpx<double> comp = glw::csub(a0, a1);
            const double den = w[mask] + ts * wo;
            eta[mask] = cs(mul_i(ms, comp), 1.0 / den);
            psi[mask] = cs(comp, -ts * ns / den);
          } else {
            m.solve(f0, f1, kx[mask], ky[mask], w[mask], eta[mask], psi[mask]);
          }
        }
        phi[mask] = glw::csub(psi[mask], taylor);
      }
      if (d < 4) {  // surface derivatives of this subset, orders 0..d-1 (:185-225)
        cpx<double> hx = {0.0, 0.0}, hy = {0.0, 0.0}, v = {0.0, 0.0};
        for (int order = 0; order < d; ++order)
          partitions(mask, order + 1, [&](const int* b, cpx<double> ep) {
            const int pm = b[0];
            const cpx<double> base = cs(glw::cmul(phi[pm], ep), inv_fact[order]);
            const double z0 = m.zd(kx[pm], ky[pm], order), z1 = m.zd(kx[pm], ky[pm], order + 1);
            hx = glw::cadd(hx, mul_i(kx[pm] * z0, base));
            hy = glw::cadd(hy, mul_i(ky[pm] * z0, base));
            v = glw::cadd(v, cs(base, z1));
          });
        shx[mask] = hx;
        shy[mask] = hy;
        sv[mask] = v;
      }
    }
  }
  // residual of the quartic output against its effective source (:413-433)
  const double qo = m.q(kx[15], ky[15]), lin = m.g * qo, tmp = w[15] * w[15];
  o.rho = fabs(lin - tmp) / fmax(lin + tmp, 2.2250738585072014e-308);
  cpx<double> er = glw::csub(glw::csub(mul_i(-w[15], eta[15]), cs(psi[15], qo)), eff0);
  cpx<double> pr = glw::csub(glw::cadd(mul_i(-w[15], psi[15]), cs(eta[15], m.g)), eff1);
  o.near = o.rho < m.thr;
  if (o.near) {
    const double wo = m.om(kx[15], ky[15]);
    const double ms = sqrt(wo / (2.0 * m.g)), ns = sqrt(m.g / (2.0 * wo));
    const int ts = w[15] >= 0.0 ? 1 : -1;
    er = glw::cadd(er, cs(ret[15], ms));
    pr = glw::cadd(pr, mul_i(-ts * ns, ret[15]));
  }
  o.eta = eta[15];
  o.psi = psi[15];
  o.retained = ret[15];
  o.residual = hypot(hypot(er.x, er.y), hypot(pr.x, pr.y));
  return o;
}

struct SignedArgs {
  WitDev w;
  double dkx, dky;
  const cpx<double>* amps;  // [cases][M]
  int cases;
  double* out;              // [5][cases][ny][nx] complex
  unsigned long long* cnt;
  unsigned long long* dmax;
  long long total;          // candidates
};

// pair index p (i <= j, row i starting at offP[i]) -> (i, j)
__device__ __forceinline__ void unrank_pair(const WitDev& w, long long p, int& i, int& j) {
  int lo = 0, hi = w.M - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (w.offP[mid] <= p) lo = mid; else hi = mid - 1;
  }
  i = lo;
  j = i + (int)(p - w.offP[i]);
}

// triple index t (i <= j <= k, tidx order) -> (i, j, k)
__device__ __forceinline__ void unrank_triple(const WitDev& w, long long t, int& i, int& j, int& k) {
  int lo = 0, hi = w.M - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (w.offT[mid] <= t) lo = mid; else hi = mid - 1;
  }
  i = lo;
  int a, b;
  unrank_pair(w, w.offP[i] + (t - w.offT[i]), a, b);
  j = a;
  k = b;
}

// SECTOR 40: candidate c = minus pair * P + plus pair (signs + + - -);
// SECTOR 42: c = minus mode * T + plus triple (signs + + + -).
// Four waves per SIMD (128 VGPRs, 2.8 KB of private memory per lane): the per-candidate
// evaluation is a long dependent FP64 chain, so waves in flight matter more than spills
// (M=200 on MI355X: 1.52x (sector 42) and 1.63x (sector 40) over one wave with 348 registers;
// each step 1 -> 2 -> 3 -> 4 waves was faster).
template <int SECTOR>
__global__ __launch_bounds__(WNT, 4) void k_wit_signed(SignedArgs g) {
  const WitDev& w = g.w;
  const Sgn m{w.md.g, w.md.h, w.md.thr};
  unsigned long long kept = 0, nearc = 0, zo = 0, zi = 0;
  double maxres = 0.0, maxz = 0.0, mindet = INFINITY;
  const size_t plane = (size_t)w.nx * w.ny * g.cases;
  for (long long c = blockIdx.x * (long long)WNT + threadIdx.x; c < g.total; c += (long long)gridDim.x * WNT) {
    int id[4], sg[4];
    int repeat;
    if constexpr (SECTOR == 40) {
      const long long P = w.offP[w.M];
      unrank_pair(w, c % P, id[0], id[1]);
      unrank_pair(w, c / P, id[2], id[3]);
      sg[0] = sg[1] = 1;
      sg[2] = sg[3] = -1;
      repeat = (1 + (id[0] == id[1])) * (1 + (id[2] == id[3]));
    } else {
      const long long T = w.offT[w.M];
      unrank_triple(w, c % T, id[0], id[1], id[2]);
      id[3] = (int)(c / T);
      sg[0] = sg[1] = sg[2] = 1;
      sg[3] = -1;
      repeat = id[0] == id[2] ? 6 : ((id[0] == id[1] || id[1] == id[2]) ? 2 : 1);
    }
    int tx = 0, ty = 0;
    for (int s = 0; s < 4; ++s) {
      tx += sg[s] * w.modes[2 * id[s]];
      ty += sg[s] * w.modes[2 * id[s] + 1];
    }
    if (tx == 0 && ty == 0) {
      ++zo;
      continue;
    }
    if (SECTOR == 40 && !(tx > 0 || (tx == 0 && ty > 0))) continue;  // analytic half plane (:566-568)
    if (abs(tx) >= w.nx / 2 || abs(ty) >= w.ny / 2) continue;
    double pkx[4], pky[4];
    for (int s = 0; s < 4; ++s) {
      pkx[s] = w.modes[2 * id[s]] * g.dkx;
      pky[s] = w.modes[2 * id[s] + 1] * g.dky;
    }
    const SignedOut o = evaluate_signed(m, pkx, pky, sg);
    const size_t ix = (size_t)(((tx % w.nx) + w.nx) % w.nx), iy = (size_t)(((ty % w.ny) + w.ny) % w.ny);
    for (int cs_ = 0; cs_ < g.cases; ++cs_) {
      const cpx<double>* ac = g.amps + (size_t)cs_ * w.M;
      cpx<double> wt = {2.0, 0.0};
      for (int s = 0; s < 4; ++s) {
        const cpx<double> a = ac[id[s]];
        wt = glw::cmul(wt, sg[s] > 0 ? a : cpx<double>{a.x, -a.y});
      }
      wt = {wt.x / repeat, wt.y / repeat};
      const size_t loc = (size_t)cs_ * w.nx * w.ny + ix + (size_t)w.nx * iy;
      atomic_addc(g.out + 2 * loc, glw::cmul(wt, o.eta));
      atomic_addc(g.out + 2 * (plane + loc), glw::cmul(wt, o.psi));
      atomic_addc(g.out + 2 * (2 * plane + loc), glw::cmul(wt, o.ecoord));
      atomic_addc(g.out + 2 * (3 * plane + loc), glw::cmul(wt, o.pcoord));
      if (o.near) atomic_addc(g.out + 2 * (4 * plane + loc), glw::cmul(wt, o.retained));
    }
    ++kept;
    nearc += o.near ? 1 : 0;
    zi += (unsigned long long)o.zcount;
    maxres = fmax(maxres, o.residual);
    mindet = fmin(mindet, o.rho);
    maxz = fmax(maxz, o.zsrc);
  }
  kept = wsum_u64(kept);
  nearc = wsum_u64(nearc);
  zo = wsum_u64(zo);
  zi = wsum_u64(zi);
  maxres = wmax_d(maxres);
  maxz = wmax_d(maxz);
  mindet = wmin_d(mindet);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(g.cnt + 0, kept);
    atomicAdd(g.cnt + 1, nearc);
    atomicAdd(g.cnt + 2, zo);
    atomicAdd(g.cnt + 3, zi);
    atomicMax(g.dmax + 0, (unsigned long long)__double_as_longlong(maxres));
    atomicMax(g.dmax + 1, (unsigned long long)__double_as_longlong(maxz));
    if (mindet < INFINITY) atomicMin(g.dmax + 2, (unsigned long long)__double_as_longlong(mindet));
  }
}

__global__ void k_wit_clear(double* out, size_t n, unsigned long long* cnt, unsigned long long* dmax) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = 0.0;
  if (i < 4) cnt[i] = 0;
  if (i < 2) dmax[i] = 0;
  if (i == 2) dmax[2] = (unsigned long long)__double_as_longlong(INFINITY);
}

__global__ void k_wit_counters(const unsigned long long* cnt, const unsigned long long* dmax, double total,
                               double* counters) {
  if (threadIdx.x != 0) return;
  counters[0] = total;
  counters[1] = (double)cnt[0];
  counters[2] = (double)cnt[1];
  counters[3] = (double)cnt[2];
  counters[4] = (double)cnt[3];
  counters[5] = __longlong_as_double((long long)dmax[0]);
  counters[6] = __longlong_as_double((long long)dmax[2]);
  counters[7] = __longlong_as_double((long long)dmax[1]);
}

}  // namespace

struct glw_wit_s {
  glw_wit_desc d{};  // d.modes is the caller's array: read during create only
  int device = 0;
  int xmin = 0, xmax = 0, ymin = 0, ymax = 0;  // mode box (LDS accumulator window)
  WitDev w{};
  std::vector<void*> allocs;
  unsigned long long* cnt = nullptr;
  unsigned long long* dmax = nullptr;
  ~glw_wit_s() {
    for (void* p : allocs) (void)hipFree(p);
  }
  int alloc(void** p, size_t bytes) {
    W_HIP(hipMalloc(p, bytes ? bytes : 64));
    allocs.push_back(*p);
    return GLW_OK;
  }
};

extern "C" {

int glw_wit_create(const glw_wit_desc* d, glw_wit** out) {
  if (!out || !d) return wfail(GLW_EINVAL, "null argument");
  *out = nullptr;
  if (d->modes_count < 1 || !d->modes || d->nx < 2 || d->ny < 2 || d->gravity <= 0 || d->depth <= 0)
    return wfail(GLW_EINVAL, "invalid WIT descriptor");
  if (d->precision != GLW_PREC_FP64 && d->precision != GLW_PREC_FP32) return wfail(GLW_EINVAL, "bad WIT precision");
  glw_wit* W = new glw_wit_s();
  W->d = *d;
  int rc = GLW_OK;
  const int M = d->modes_count;
  W->xmin = W->xmax = d->modes[0];
  W->ymin = W->ymax = d->modes[1];
  for (int m = 1; m < M; ++m) {
    W->xmin = std::min(W->xmin, (int)d->modes[2 * m]);
    W->xmax = std::max(W->xmax, (int)d->modes[2 * m]);
    W->ymin = std::min(W->ymin, (int)d->modes[2 * m + 1]);
    W->ymax = std::max(W->ymax, (int)d->modes[2 * m + 1]);
  }
  if (hipGetDevice(&W->device) != hipSuccess) {
    delete W;
    return wfail(GLW_EDEVICE, "no HIP device");
  }
  const long long np = (long long)M * (M + 1) / 2;
  std::vector<long long> offP(M + 1), offT(M + 1);
  offP[0] = offT[0] = 0;
  for (int i = 0; i < M; ++i) {
    offP[i + 1] = offP[i] + (M - i);
    const long long n = M - i;
    offT[i + 1] = offT[i] + n * (n + 1) / 2;
  }
  const long long nt = offT[M];
  void *modes, *single, *pair, *triple, *op, *ot, *cnt, *dmax;
  rc = W->alloc(&modes, sizeof(int) * 2 * M);
  if (!rc) rc = W->alloc(&single, sizeof(Rec) * M);
  if (!rc) rc = W->alloc(&pair, sizeof(Rec) * np);
  if (!rc) rc = W->alloc(&triple, sizeof(Rec) * nt);
  if (!rc) rc = W->alloc(&op, sizeof(long long) * (M + 1));
  if (!rc) rc = W->alloc(&ot, sizeof(long long) * (M + 1));
  if (!rc) rc = W->alloc(&cnt, sizeof(unsigned long long) * 4);
  if (!rc) rc = W->alloc(&dmax, sizeof(unsigned long long) * 4);
  if (rc) {
    delete W;
    return rc;
  }
  auto copy = [&](void* dst, const void* src, size_t n) -> int {
    W_HIP(hipMemcpy(dst, src, n, hipMemcpyHostToDevice));
    return GLW_OK;
  };
  if ((rc = copy(modes, d->modes, sizeof(int) * 2 * M)) || (rc = copy(op, offP.data(), sizeof(long long) * (M + 1))) ||
      (rc = copy(ot, offT.data(), sizeof(long long) * (M + 1)))) {
    delete W;
    return rc;
  }
  WitDev& w = W->w;
  w.M = M;
  w.nx = d->nx;
  w.ny = d->ny;
  w.md = Model{d->gravity, d->depth, d->threshold};
  w.modes = (const int*)modes;
  w.single = (Rec*)single;
  w.pair = (Rec*)pair;
  w.triple = (Rec*)triple;
  w.offP = (const long long*)op;
  w.offT = (const long long*)ot;
  W->cnt = (unsigned long long*)cnt;
  W->dmax = (unsigned long long*)dmax;
  hipLaunchKernelGGL(k_wit_single_fill, dim3((M + 255) / 256), dim3(256), 0, 0, w, d->dkx, d->dky);
  hipLaunchKernelGGL(k_wit_pairs, dim3((np + 255) / 256), dim3(256), 0, 0, w);
  hipLaunchKernelGGL(k_wit_triples, dim3((nt + 255) / 256), dim3(256), 0, 0, w, nt);
  w.single32 = w.pair32 = w.triple32 = nullptr;
  if (d->precision == GLW_PREC_FP32) {  // FP32 copies of the FP64 tables
    void *s32, *p32, *t32;
    rc = W->alloc(&s32, sizeof(RecT<float>) * M);
    if (!rc) rc = W->alloc(&p32, sizeof(RecT<float>) * np);
    if (!rc) rc = W->alloc(&t32, sizeof(RecT<float>) * nt);
    if (rc) {
      delete W;
      return rc;
    }
    w.single32 = (RecT<float>*)s32;
    w.pair32 = (RecT<float>*)p32;
    w.triple32 = (RecT<float>*)t32;
    hipLaunchKernelGGL(k_wit_to_f32, dim3((M + 255) / 256), dim3(256), 0, 0, w.single, w.single32, (long long)M);
    hipLaunchKernelGGL(k_wit_to_f32, dim3((np + 255) / 256), dim3(256), 0, 0, w.pair, w.pair32, np);
    hipLaunchKernelGGL(k_wit_to_f32, dim3((nt + 255) / 256), dim3(256), 0, 0, w.triple, w.triple32, nt);
  }
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    delete W;
    return wfail(GLW_EDEVICE, "WIT subset tables failed");
  }
  *out = W;
  return GLW_OK;
}

int glw_wit_destroy(glw_wit* w) {
  delete w;
  return GLW_OK;
}

int glw_wit_sector44(glw_wit* W, const double* amps, int32_t cases, double* out, double* counters, void* stream) {
  if (!W) return wfail(GLW_EINVAL, "null argument");
  return glw_wit_sector44_range(W, 0, W->d.modes_count, 0, amps, cases, out, counters, stream);
}

int glw_wit_sector44_range(glw_wit* W, int32_t first_begin, int32_t first_end, int32_t flags, const double* amps,
                           int32_t cases, double* out, double* counters, void* stream) {
  if (!W || !amps || !out || !counters || cases < 1) return wfail(GLW_EINVAL, "null argument");
  if (first_begin < 0 || first_end > W->d.modes_count || first_begin >= first_end)
    return wfail(GLW_EINVAL, "first-index range must satisfy 0 <= begin < end <= M");
  if (flags & ~GLW_WIT_GLOBAL_ATOMICS) return wfail(GLW_EINVAL, "unknown WIT flag");
  W_HIP(hipSetDevice(W->device));
  hipStream_t st = (hipStream_t)stream;
  const size_t n = 5ull * cases * W->d.nx * W->d.ny * 2;
  hipLaunchKernelGGL(k_wit_clear, dim3((n + 255) / 256), dim3(256), 0, st, out, n, W->cnt, W->dmax);
  QuartetArgs g;
  g.w = W->w;
  g.amps = (const cpx<double>*)amps;
  g.cases = cases;
  g.out = out;
  g.cnt = W->cnt;
  g.dmax = W->dmax;
  const long long M = W->d.modes_count;
  // LDS accumulator window: pair sums of the modes, (2 W - 1) x (2 H - 1) bins
  g.pw = 2 * (W->xmax - W->xmin) + 1;
  g.ph = 2 * (W->ymax - W->ymin) + 1;
  g.pxmin = 2 * W->xmin;
  g.pymin = 2 * W->ymin;
  const size_t lds = (size_t)g.pw * g.ph * cases * 4 * sizeof(double);
  g.lds_bins = (lds <= kWitLdsBytes && !(flags & GLW_WIT_GLOBAL_ATOMICS)) ? g.pw * g.ph : 0;
  const size_t shm = g.lds_bins ? lds : 0;
  // leading pairs (i, j) with i in [first_begin, first_end): offP[i] = i M - i (i - 1) / 2
  auto off_p = [M](long long i) { return i * M - i * (i - 1) / 2; };
  g.pair0 = off_p(first_begin);
  const unsigned nblk = (unsigned)(off_p(first_end) - g.pair0);
  if (W->d.precision == GLW_PREC_FP32 && shm <= kLdsPerCu / 2) {  // two workgroups per CU fit: occupancy 2
    if (shm) W_HIP(hipFuncSetAttribute((const void*)k_wit_quartets<float, 2>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL((k_wit_quartets<float, 2>), dim3(nblk), dim3(WNT), shm, st, g);
  } else if (W->d.precision == GLW_PREC_FP32) {
    if (shm) W_HIP(hipFuncSetAttribute((const void*)k_wit_quartets<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)shm));
    hipLaunchKernelGGL(k_wit_quartets<float>, dim3(nblk), dim3(WNT), shm, st, g);
  } else {
    if (shm) W_HIP(hipFuncSetAttribute((const void*)k_wit_quartets<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)shm));
    hipLaunchKernelGGL(k_wit_quartets<double>, dim3(nblk), dim3(WNT), shm, st, g);
  }
  // quartets i <= j <= k <= l < M with i in the range: sum over i of C(M - i + 2, 3)
  double total = 0.0;
  for (long long i = first_begin; i < first_end; ++i) total += (double)(M - i + 2) * (M - i + 1) * (M - i) / 6.0;
  hipLaunchKernelGGL(k_wit_counters, dim3(1), dim3(64), 0, st, W->cnt, W->dmax, total, counters);
  W_HIP(hipGetLastError());
  return GLW_OK;
}

int glw_wit_sector(glw_wit* W, int32_t sector, const double* amps, int32_t cases, double* out, double* counters,
                   void* stream) {
  if (sector == 44) return glw_wit_sector44(W, amps, cases, out, counters, stream);
  if (sector != 40 && sector != 42) return wfail(GLW_EINVAL, "sector must be 40, 42, or 44");
  if (!W || !amps || !out || !counters || cases < 1) return wfail(GLW_EINVAL, "null argument");
  if (W->d.precision != GLW_PREC_FP64) return wfail(GLW_EUNSUPPORTED, "sectors 40 and 42 evaluate in FP64 only");
  W_HIP(hipSetDevice(W->device));
  hipStream_t st = (hipStream_t)stream;
  const size_t n = 5ull * cases * W->d.nx * W->d.ny * 2;
  hipLaunchKernelGGL(k_wit_clear, dim3((n + 255) / 256), dim3(256), 0, st, out, n, W->cnt, W->dmax);
  SignedArgs g;
  g.w = W->w;
  g.dkx = W->d.dkx;
  g.dky = W->d.dky;
  g.amps = (const cpx<double>*)amps;
  g.cases = cases;
  g.out = out;
  g.cnt = W->cnt;
  g.dmax = W->dmax;
  const long long M = W->d.modes_count, P = M * (M + 1) / 2, T = M * (M + 1) * (M + 2) / 6;
  g.total = sector == 40 ? P * P : T * M;
  const long long blocks = std::min<long long>((g.total + WNT - 1) / WNT, 1 << 20);
  if (sector == 40) hipLaunchKernelGGL(k_wit_signed<40>, dim3((unsigned)blocks), dim3(WNT), 0, st, g);
  else hipLaunchKernelGGL(k_wit_signed<42>, dim3((unsigned)blocks), dim3(WNT), 0, st, g);
  hipLaunchKernelGGL(k_wit_counters, dim3(1), dim3(64), 0, st, W->cnt, W->dmax, (double)g.total, counters);
  W_HIP(hipGetLastError());
  return GLW_OK;
}

}  // extern "C"
