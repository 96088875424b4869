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
// Sectors 40 and 42 have signed parents.  Flipping every sign of a subset
// conjugates its state, so two more tables -- pairs {a+, b-} and triples
// {a+, b+, c-} -- give every signed subset, and those sectors too evaluate
// only the degree-4 level per candidate, plus the push and y31 terms, which
// pair each cubic subset with the parent it lacks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
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

// Source of the full set FULL = (1<<D)-1 from the records of its proper
// subsets (`rec(m)` returns the record of local sub-mask m): the Taylor sum
// and the two source components (reference :264-282 and :337-350).
template <int D, class Acc, typename T>
__device__ __forceinline__ void subset_sources(const Acc& rec, const RecT<T>& S, cpx<T>& taylor, cpx<T>& src0,
                                               cpx<T>& src1) {
  constexpr int FULL = (1 << D) - 1;
  cpx<T> horiz = {0, 0}, kin = {0, 0}, lv = {0, 0}, lt = {0, 0};
  taylor = {0, 0};
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
  src0 = {-(horiz.x - lv.x) - S.q * taylor.x, -(horiz.y - lv.y) - S.q * taylor.y};
  const cpx<T> it = ci(S.om, taylor);
  src1 = {-(lt.x + kin.x) - it.x, -(lt.y + kin.y) - it.y};
}

// Degree-4 source by complementary pairs: the same sums as subset_sources<4>,
// with the etas of the fourteen proper subsets read first and then one pair of
// records (L, 15 ^ L) at a time.  Each pair contributes both orders of the
// two-block source and the Taylor terms whose potential block is L or 15 ^ L.
template <class Acc, typename T>
__device__ __forceinline__ void quartet_sources(const Acc& rec, const RecT<T>& S, cpx<T>& taylor, cpx<T>& src0,
                                                cpx<T>& src1) {
  cpx<T> eta[16];
#pragma unroll
  for (int m = 1; m < 15; ++m) eta[m] = rec(m).eta;
  // E_n(C), n = 1..3: sums over unordered partitions of C into n blocks of prod eta[block]
  auto en = [&](int C, cpx<T>& e1, cpx<T>& e2, cpx<T>& e3) {
    e1 = eta[C];
    e2 = e3 = {0, 0};
    if (popc(C) == 2) {
      const int a = lowbit(C), b = C ^ a;
      e2 = cm(eta[a], eta[b]);
    } else if (popc(C) == 3) {
      const int a = lowbit(C), b = lowbit(C ^ a), c = C ^ a ^ b;
      e2 = ca(ca(cm(eta[a], eta[b | c]), cm(eta[b], eta[a | c])), cm(eta[c], eta[a | b]));
      e3 = cm(cm(eta[a], eta[b]), eta[c]);
    }
  };
  cpx<T> horiz = {0, 0}, kin = {0, 0}, lv = {0, 0}, lt = {0, 0};
  taylor = {0, 0};
  auto taylor_term = [&](const RecT<T>& p, int C) {
    cpx<T> e1, e2, e3;
    en(C, e1, e2, e3);
    const cpx<T> t = ca(ca(cs(e1, zd(p, 1)), cs(e2, zd(p, 2))), cs(e3, zd(p, 3)));
    const cpx<T> v = ca(ca(cs(e1, zd(p, 2)), cs(e2, zd(p, 3))), cs(e3, zd(p, 4)));
    const cpx<T> pt = cm(p.phi, t);
    taylor = ca(taylor, pt);
    lv = ca(lv, cm(p.phi, v));
    lt = ca(lt, ci(-p.om, pt));
  };
  constexpr int kLeft[7] = {1, 2, 4, 8, 3, 5, 6};
#pragma unroll
  for (int u = 0; u < 7; ++u) {
    const int L = kLeft[u], R = 15 ^ L;
    const RecT<T> a = rec(L), b = rec(R);
    horiz = ca(horiz, cm(ca(ci(b.kx, a.sh0), ci(b.ky, a.sh1)), eta[R]));
    horiz = ca(horiz, cm(ca(ci(a.kx, b.sh0), ci(a.ky, b.sh1)), eta[L]));
    kin = ca(kin, ca(ca(cm(a.sh0, b.sh0), cm(a.sh1, b.sh1)), cm(a.sv, b.sv)));
    taylor_term(a, R);
    taylor_term(b, L);
  }
  src0 = {-(horiz.x - lv.x) - S.q * taylor.x, -(horiz.y - lv.y) - S.q * taylor.y};
  const cpx<T> it = ci(S.om, taylor);
  src1 = {-(lt.x + kin.x) - it.x, -(lt.y + kin.y) - it.y};
}

// Strict zero of a subset's wavevector (reference :356-364).
template <typename T>
__device__ __forceinline__ bool strict_zero(const RecT<T>& S) {
  const T mag = sqrt(S.k2);
  return mag <= T(100) * Limits<T>::eps * fmax(mag, T(1));
}

// Free-wave split of a forced response whose frequency is `ts` times the
// dispersion frequency of its wavevector (reference :370-411): the retained
// part `ret` and the bound part (eta, psi).
template <typename T>
__device__ __forceinline__ void retained_split(cpx<T> f0, cpx<T> f1, T q, T om, T ts, T g, cpx<T>& ret, cpx<T>& eta,
                                               cpx<T>& psi) {
  const T w = sqrt(g * q), ms = sqrt(w / (T(2) * g)), ns = sqrt(g / (T(2) * w));
  const cpx<T> a0 = cs(f0, T(1) / (T(2) * ms));
  const cpx<T> a1 = ci(ts, cs(f1, T(1) / (T(2) * ns)));
  ret = ca(a0, a1);
  const cpx<T> comp = glw::csub(a0, a1);
  const T den = om + ts * w;
  eta = cs(ci(ms, comp), T(1) / den);
  psi = cs(comp, -ts * ns / den);
}

// 2x2 forced linear response at wavevector K (q = q(K)) and frequency om
// (reference :111-121).
template <typename T>
__device__ __forceinline__ void forced_solve(cpx<T> f0, cpx<T> f1, T q, T om, T g, cpx<T>& eta, cpx<T>& psi) {
  const T den = om * om - g * q;
  const cpx<T> a = glw::csub(ci(om, f0), cs(f1, q));
  const cpx<T> b = ca(cs(f0, g), ci(om, f1));
  eta = cs(a, T(1) / den);
  psi = cs(b, T(1) / den);
}

// State of the full set FULL = (1<<D)-1 from the records of its proper
// subsets.  split != 0 (a cubic subset of charge +-1, sign `split`) keeps the
// free-wave part (:370-386); degree 4 keeps it when near resonant.
template <int D, class Acc, typename T>
__device__ __forceinline__ void subset_state(const Acc& rec, const RecT<T>& S, const Model& md, TopOut<T>& o,
                                             T split = T(0)) {
  const T g = T(md.g);
  cpx<T> taylor;
  // degree 4 in FP32 by complementary pairs (M=400 on MI355X: 92 ms against 118 ms);
  // in FP64 the pair order costs registers at one wave per SIMD (298 ms against 282 ms)
  if constexpr (D == 4 && std::is_same<T, float>::value) quartet_sources(rec, S, taylor, o.src0, o.src1);
  else subset_sources<D>(rec, S, taylor, o.src0, o.src1);
  o.zero = strict_zero(S);
  o.near = false;
  o.rho = T(0);
  o.retained = {0, 0};
  o.zero_src = T(0);
  if (o.zero) {
    o.eta = o.psi = {0, 0};
    o.zero_src = hypot(cabs2(o.src0), cabs2(o.src1));
  } else {
    if constexpr (D == 4) {
      const T lin = g * S.q, tmp = S.om * S.om;
      o.rho = fabs(lin - tmp) / fmax(lin + tmp, Limits<T>::tiny);
      o.near = o.rho < T(md.thr);
      split = o.near ? (S.om >= T(0) ? T(1) : T(-1)) : T(0);
    }
    if (split != T(0)) retained_split(o.src0, o.src1, S.q, S.om, split, g, o.retained, o.eta, o.psi);
    else forced_solve(o.src0, o.src1, S.q, S.om, g, o.eta, o.psi);
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
  // Signed subsets (sectors 40 and 42; built by their first call).  Flipping
  // the sign of every parent conjugates a state (wavevector and frequency
  // negate), so the sign patterns (+-) and (++-) give every signed subset.
  Rec* mpair;            // [M][M]  {a+, b-} at a M + b
  Rec* mtrip;            // [P][M]  {a+, b+, c-} (a <= b) at pidx(a, b) M + c
  cpx<double>* mtrip_x;  // [P][M][2] psi and retained V3 of those triples
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

// f -> (k, l), 0 <= k <= l < n, in row-major order of the upper triangle:
// row k starts at k n - k (k - 1) / 2.
__device__ __forceinline__ void unrank_upper(long long f, int n, int& k, int& l) {
  const double disc = (2.0 * n + 1.0) * (2.0 * n + 1.0) - 8.0 * (double)f;
  int kk = (int)floor(((2.0 * n + 1.0) - sqrt(fmax(disc, 0.0))) / 2.0);
  if (kk < 0) kk = 0;
  if (kk > n - 1) kk = n - 1;
  while (kk > 0 && (long long)kk * n - (long long)kk * (kk - 1) / 2 > f) --kk;
  while (kk + 1 < n && (long long)(kk + 1) * n - (long long)(kk + 1) * kk / 2 <= f) ++kk;
  k = kk;
  l = kk + (int)(f - ((long long)kk * n - (long long)kk * (kk - 1) / 2));
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
    int kk, ll;
    unrank_upper(f, n, kk, ll);
    const int k = j + kk, l = j + ll;
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

// ----------------------------------------------------------------------------
// Tabulated signed sectors.
// ----------------------------------------------------------------------------
// State of the same subset with every parent sign flipped: the conjugate
// state at the negated wavevector and frequency.
__device__ __forceinline__ Rec flip(Rec r) {
  r.eta.y = -r.eta.y;
  r.phi.y = -r.phi.y;
  r.sh0.y = -r.sh0.y;
  r.sh1.y = -r.sh1.y;
  r.sv.y = -r.sv.y;
  r.kx = -r.kx;
  r.ky = -r.ky;
  r.om = -r.om;
  return r;
}
__device__ __forceinline__ cpx<double> conj(cpx<double> a) { return {a.x, -a.y}; }

struct LocalAcc {
  const Rec* r;
  __device__ const Rec& operator()(int m) const { return r[m]; }
};

// {a+, b-} for a, b in [0, M): ordinary pair solve (:357-369).
__global__ void k_wit_mixed_pairs(WitDev w) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= (long long)w.M * w.M) return;
  const int a = (int)(t / w.M), b = (int)(t % w.M);
  Rec loc[4];
  loc[1] = w.single[a];
  loc[2] = flip(w.single[b]);
  const Rec* sub[2] = {&loc[1], &loc[2]};
  Rec S;
  base_record(w, sub, 2, S);
  TopOut<double> o;
  const LocalAcc acc{loc};
  subset_state<2>(acc, S, w.md, o);
  S.eta = o.eta;
  S.phi = o.phi;
  S.zero = o.zero ? 1.0 : 0.0;
  S.zero_src = o.zero_src;
  surface<2>(acc, S, o.phi, S);
  w.mpair[t] = S;
}

// {a+, b+, c-} (a <= b): a cubic subset of charge +1 keeps its free-wave
// part (:370-386).
__global__ void k_wit_mixed_triples(WitDev w, long long n) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long long p = t / w.M;
  const int c = (int)(t % w.M);
  int a, b;
  unrank_pair(w, p, a, b);
  Rec loc[8];
  loc[1] = w.single[a];
  loc[2] = w.single[b];
  loc[4] = flip(w.single[c]);
  loc[3] = w.pair[p];
  loc[5] = w.mpair[(long long)a * w.M + c];
  loc[6] = w.mpair[(long long)b * w.M + c];
  const Rec* sub[3] = {&loc[1], &loc[2], &loc[4]};
  Rec S;
  base_record(w, sub, 3, S);
  TopOut<double> o;
  const LocalAcc acc{loc};
  subset_state<3>(acc, S, w.md, o, 1.0);
  S.eta = o.eta;
  S.phi = o.phi;
  S.zero = o.zero ? 1.0 : 0.0;
  S.zero_src = o.zero_src;
  surface<3>(acc, S, o.phi, S);
  w.mtrip[t] = S;
  w.mtrip_x[2 * t] = o.psi;
  w.mtrip_x[2 * t + 1] = o.retained;
}

// Record accessor of a signed quartet: sub-mask m reads *r[m], flipped when
// bit m of FLIPS is set.
template <int FLIPS>
struct SignedAcc {
  const Rec* r[16];
  __device__ Rec operator()(int m) const { return (FLIPS >> m & 1) ? flip(*r[m]) : *r[m]; }
};

struct SignedTabArgs {
  WitDev w;
  const cpx<double>* amps;  // [cases][M]
  int cases;
  double* out;              // [5][cases][ny][nx] complex
  unsigned long long* cnt;
  unsigned long long* dmax;
  // LDS window of a workgroup (0 bins: global atomics): bin
  // (ty - sy + offy) pw + (tx - sx + offx), (sx, sy) = mode(i) + mode(j).
  int lds_bins, pw, ph, offx, offy;
};

// Sector 40 (signs + + - -; plus pair (i, j), minus pair (k, l)) and sector 42
// (signs + + + -; plus triple (i, j, k), minus mode l), reference sector loops
// :603-688 and evaluate :157-453.  One workgroup per plus pair (i, j) =
// leading pair; its threads sweep the minus pairs (k <= l) (40) or the
// (k >= j, l) rectangle (42).  Subset states come from the tables; per
// candidate only the degree-4 level is evaluated: the partition-sum source,
// the push of the four cubic subsets' retained V3 through the normal map
// (:282-298), the y31 coordinate (:300-336), the near-resonance split and the
// residual (:393-433).
// One wave per SIMD (256 VGPRs + ~160 AGPRs, no spill): held to two waves the
// kernel spills 600-800 B per lane and runs 1.2-1.6x slower (M=200, MI355X).
template <int SECTOR>
__global__ __launch_bounds__(WNT, 1) void k_wit_signed_tab(SignedTabArgs g) {
  const WitDev& w = g.w;
  const int M = w.M;
  int i, j;
  unrank_pair(w, blockIdx.x, i, j);
  const long long pij = blockIdx.x;
  const long long nf = SECTOR == 40 ? w.offP[M] : (long long)(M - j) * M;
  constexpr int FLIPS = SECTOR == 40 ? (1 << 4 | 1 << 8 | 1 << 12 | 1 << 13 | 1 << 14) : (1 << 8);
  // charge of the cubic sub-mask without parent s (cm = 15 ^ (1 << s))
  constexpr int kCharge[4] = {SECTOR == 40 ? -1 : 1, SECTOR == 40 ? -1 : 1, SECTOR == 40 ? 1 : 1,
                              SECTOR == 40 ? 1 : 3};
  constexpr int kSign[4] = {1, 1, SECTOR == 40 ? -1 : 1, -1};
  const double gv = w.md.g;
  unsigned long long kept = 0, nearc = 0, zo = 0, zi = 0;
  double maxres = 0.0, maxz = 0.0, mindet = INFINITY;
  const size_t plane = (size_t)w.nx * w.ny * g.cases;
  extern __shared__ double wacc[];
  const int sx = w.modes[2 * i] + w.modes[2 * j], sy = w.modes[2 * i + 1] + w.modes[2 * j + 1];
  if (g.lds_bins) {
    for (int e = threadIdx.x; e < g.lds_bins * g.cases * 8; e += WNT) wacc[e] = 0.0;
    __syncthreads();
  }
  for (long long f = threadIdx.x; f < nf; f += WNT) {
    int k, l;
    if constexpr (SECTOR == 40) {
      unrank_upper(f, M, k, l);
    } else {
      k = j + (int)(f / M);
      l = (int)(f % M);
    }
    const int id[4] = {i, j, k, l};
    int tx = 0, ty = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      tx += kSign[s] * w.modes[2 * id[s]];
      ty += kSign[s] * w.modes[2 * id[s] + 1];
    }
    if (tx == 0 && ty == 0) {
      ++zo;
      continue;
    }
    if (SECTOR == 40 && !(tx > 0 || (tx == 0 && ty > 0))) continue;  // analytic half plane (:566-568)
    if (abs(tx) >= w.nx / 2 || abs(ty) >= w.ny / 2) continue;
    SignedAcc<FLIPS> a;
    const cpx<double>* xs[4];  // psi, retained of the cubic sub-mask without parent s
    a.r[1] = &w.single[i];
    a.r[2] = &w.single[j];
    a.r[4] = &w.single[k];
    a.r[8] = &w.single[l];
    if constexpr (SECTOR == 40) {
      const long long pkl = pidx(w, k, l);
      a.r[3] = &w.pair[pij];
      a.r[12] = &w.pair[pkl];
      a.r[5] = &w.mpair[(long long)i * M + k];
      a.r[9] = &w.mpair[(long long)i * M + l];
      a.r[6] = &w.mpair[(long long)j * M + k];
      a.r[10] = &w.mpair[(long long)j * M + l];
      const long long t7 = pij * M + k, t11 = pij * M + l, t13 = pkl * M + i, t14 = pkl * M + j;
      a.r[7] = &w.mtrip[t7];
      a.r[11] = &w.mtrip[t11];
      a.r[13] = &w.mtrip[t13];
      a.r[14] = &w.mtrip[t14];
      xs[0] = w.mtrip_x + 2 * t14;
      xs[1] = w.mtrip_x + 2 * t13;
      xs[2] = w.mtrip_x + 2 * t11;
      xs[3] = w.mtrip_x + 2 * t7;
    } else {
      const long long pik = pidx(w, i, k), pjk = pidx(w, j, k);
      a.r[3] = &w.pair[pij];
      a.r[5] = &w.pair[pik];
      a.r[6] = &w.pair[pjk];
      a.r[9] = &w.mpair[(long long)i * M + l];
      a.r[10] = &w.mpair[(long long)j * M + l];
      a.r[12] = &w.mpair[(long long)k * M + l];
      a.r[7] = &w.triple[tidx(w, i, j, k)];
      const long long t11 = pij * M + l, t13 = pik * M + l, t14 = pjk * M + l;
      a.r[11] = &w.mtrip[t11];
      a.r[13] = &w.mtrip[t13];
      a.r[14] = &w.mtrip[t14];
      xs[0] = w.mtrip_x + 2 * t14;
      xs[1] = w.mtrip_x + 2 * t13;
      xs[2] = w.mtrip_x + 2 * t11;
      xs[3] = nullptr;
    }
    Rec S;
    {
      double kx = 0, ky = 0, om = 0;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const Rec b = a(1 << s);
        kx += b.kx;
        ky += b.ky;
        om += b.om;
      }
      S.kx = kx;
      S.ky = ky;
      S.om = om;
      S.q = w.md.q(kx, ky, &S.k2);
    }
    cpx<double> taylor, f0, f1;
    subset_sources<4>(a, S, taylor, f0, f1);
    cpx<double> ecoord = {0, 0}, pcoord = {0, 0};
    const double qK = S.q;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int cc = kCharge[s];
      if (cc != 1 && cc != -1) continue;  // a cubic subset of charge +-3 has no free-wave part
      const int cmask = 15 ^ (1 << s);
      const Rec C = a(cmask);
      const Rec b = a(1 << s);
      const bool fl = (FLIPS >> cmask) & 1;
      const cpx<double> psic = fl ? conj(xs[s][0]) : xs[s][0];
      const cpx<double> retc = fl ? conj(xs[s][1]) : xs[s][1];
      const double adb = C.kx * b.kx + C.ky * b.ky;
      const double gab = -qK * b.q + b.k2 + adb, gba = -qK * C.q + C.k2 + adb, kk = adb + C.q * b.q;
      const double cf = sqrt(gv * C.q);
      // b.phi = -i (signed frequency) / (2 q): the parent's potential, pb of the maps below
      if (retc.x != 0.0 || retc.y != 0.0) {  // push through the normal map (unit normal amplitude)
        const double ma = sqrt(cf / (2.0 * gv)), na = sqrt(gv / (2.0 * cf));
        const cpx<double> pa = {0.0, -cc * na};
        const cpx<double> s0 = ca(cs(b.phi, ma * gab), cs(pa, 0.5 * gba));
        const cpx<double> s1 = cs(cm(pa, b.phi), kk);
        cpx<double> pe, pp;
        forced_solve(s0, s1, qK, cc * cf + b.om, gv, pe, pp);
        f0 = glw::csub(f0, cm(retc, pe));
        f1 = glw::csub(f1, cm(retc, pp));
      }
      if (C.zero == 0.0) {  // y31 coordinate: the cubic state split into its +- frequency parts
        const cpx<double> ps = ci(cc * cf / gv, psic);
        const cpx<double> ae = cs(ca(C.eta, ps), 0.5), oe = cs(glw::csub(C.eta, ps), 0.5);
        const cpx<double> ap = ci(-gv * cc / cf, ae), op = ci(gv * cc / cf, oe);
        cpx<double> e1, p1, e2, p2;
        forced_solve(ca(cs(cm(ae, b.phi), gab), cs(ap, 0.5 * gba)), cs(cm(ap, b.phi), kk), qK, cc * cf + b.om, gv,
                     e1, p1);
        forced_solve(ca(cs(cm(oe, b.phi), gab), cs(op, 0.5 * gba)), cs(cm(op, b.phi), kk), qK, -cc * cf + b.om, gv,
                     e2, p2);
        ecoord = ca(ecoord, ca(e1, e2));
        pcoord = ca(pcoord, ca(p1, p2));
      }
    }
    const double lin = gv * qK, tmp = S.om * S.om;
    const double rho = fabs(lin - tmp) / fmax(lin + tmp, Limits<double>::tiny);
    const bool near = rho < w.md.thr;
    const double ts = S.om >= 0.0 ? 1.0 : -1.0;
    cpx<double> eta, psi, ret = {0.0, 0.0};
    if (near) retained_split(f0, f1, qK, S.om, ts, gv, ret, eta, psi);
    else forced_solve(f0, f1, qK, S.om, gv, eta, psi);
    // residual against the pushed source (:413-433)
    cpx<double> er = glw::csub(glw::csub(ci(-S.om, eta), cs(psi, qK)), f0);
    cpx<double> pr = glw::csub(ca(ci(-S.om, psi), cs(eta, gv)), f1);
    if (near) {
      const double wo = sqrt(lin), ms = sqrt(wo / (2.0 * gv)), ns = sqrt(gv / (2.0 * wo));
      er = ca(er, cs(ret, ms));
      pr = ca(pr, ci(-ts * ns, ret));
    }
    const double res = hypot(cabs2(er), cabs2(pr));
    double zcount = 0.0, zsrc = 0.0;
    constexpr int kInner[10] = {3, 5, 6, 9, 10, 12, 7, 11, 13, 14};
#pragma unroll
    for (int q = 0; q < 10; ++q) {
      zcount += a.r[kInner[q]]->zero;
      zsrc = fmax(zsrc, a.r[kInner[q]]->zero_src);
    }
    const int repeat = SECTOR == 40 ? (1 + (i == j)) * (1 + (k == l))
                                    : (i == k ? 6 : ((i == j || j == k) ? 2 : 1));
    const size_t ix = (size_t)(((tx % w.nx) + w.nx) % w.nx), iy = (size_t)(((ty % w.ny) + w.ny) % w.ny);
    for (int c = 0; c < g.cases; ++c) {
      const cpx<double>* ac = g.amps + (size_t)c * M;
      cpx<double> wt = {2.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4; ++s) wt = cm(wt, kSign[s] > 0 ? ac[id[s]] : conj(ac[id[s]]));
      wt = {wt.x / repeat, wt.y / repeat};
      const cpx<double> ve = cm(wt, eta), vp = cm(wt, psi), vc = cm(wt, ecoord), vq = cm(wt, pcoord);
      const size_t loc = (size_t)c * w.nx * w.ny + ix + (size_t)w.nx * iy;
      if (g.lds_bins) {
        const int lb = (ty - sy + g.offy) * g.pw + (tx - sx + g.offx);
        double* a8 = wacc + ((size_t)lb * g.cases + c) * 8;
        atomicAdd(a8 + 0, ve.x);
        atomicAdd(a8 + 1, ve.y);
        atomicAdd(a8 + 2, vp.x);
        atomicAdd(a8 + 3, vp.y);
        atomicAdd(a8 + 4, vc.x);
        atomicAdd(a8 + 5, vc.y);
        atomicAdd(a8 + 6, vq.x);
        atomicAdd(a8 + 7, vq.y);
      } else {
        atomic_addc(g.out + 2 * loc, ve);
        atomic_addc(g.out + 2 * (plane + loc), vp);
        atomic_addc(g.out + 2 * (2 * plane + loc), vc);
        atomic_addc(g.out + 2 * (3 * plane + loc), vq);
      }
      if (near) atomic_addc(g.out + 2 * (4 * plane + loc), cm(wt, ret));
    }
    ++kept;
    nearc += near ? 1 : 0;
    zi += (unsigned long long)zcount;
    maxres = fmax(maxres, res);
    mindet = fmin(mindet, rho);
    maxz = fmax(maxz, zsrc);
  }
  if (g.lds_bins) {  // flush: one global add per touched bin and array
    __syncthreads();
    for (int e = threadIdx.x; e < g.lds_bins * g.cases; e += WNT) {
      const double* a8 = wacc + (size_t)e * 8;
      const int lb = e / g.cases, c = e - lb * g.cases;
      const int tx = sx - g.offx + lb % g.pw, ty = sy - g.offy + lb / g.pw;
      const size_t ix = (size_t)(((tx % w.nx) + w.nx) % w.nx), iy = (size_t)(((ty % w.ny) + w.ny) % w.ny);
      const size_t loc = (size_t)c * w.nx * w.ny + ix + (size_t)w.nx * iy;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (a8[2 * q] != 0.0 || a8[2 * q + 1] != 0.0)
          atomic_addc(g.out + 2 * (q * plane + loc), cpx<double>{a8[2 * q], a8[2 * q + 1]});
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
  const long long M = W->d.modes_count, P = M * (M + 1) / 2, T = M * (M + 1) * (M + 2) / 6;
  if (!W->w.mpair) {  // signed subset tables, built once on this stream
    const long long nm = M * M, n3 = P * M;
    void *mp, *mt, *mx;
    int rc = W->alloc(&mp, sizeof(Rec) * nm);
    if (!rc) rc = W->alloc(&mt, sizeof(Rec) * n3);
    if (!rc) rc = W->alloc(&mx, sizeof(cpx<double>) * 2 * n3);
    if (rc) return rc;
    WitDev w = W->w;
    w.mpair = (Rec*)mp;
    w.mtrip = (Rec*)mt;
    w.mtrip_x = (cpx<double>*)mx;
    hipLaunchKernelGGL(k_wit_mixed_pairs, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_wit_mixed_triples, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, w, n3);
    W_HIP(hipGetLastError());
    W->w = w;
  }
  SignedTabArgs g;
  g.w = W->w;
  g.amps = (const cpx<double>*)amps;
  g.cases = cases;
  g.out = out;
  g.cnt = W->cnt;
  g.dmax = W->dmax;
  // LDS window: (tx - sx, ty - sy) is minus a pair sum (40) or a mode difference (42)
  g.pw = 2 * (W->xmax - W->xmin) + 1;
  g.ph = 2 * (W->ymax - W->ymin) + 1;
  g.offx = sector == 40 ? 2 * W->xmax : W->xmax - W->xmin;
  g.offy = sector == 40 ? 2 * W->ymax : W->ymax - W->ymin;
  const size_t lds = (size_t)g.pw * g.ph * cases * 8 * sizeof(double);
  g.lds_bins = lds <= kWitLdsBytes ? g.pw * g.ph : 0;
  const size_t shm = g.lds_bins ? lds : 0;
  const double total = (double)(sector == 40 ? P * P : T * M);
  if (sector == 40) {
    if (shm) W_HIP(hipFuncSetAttribute((const void*)k_wit_signed_tab<40>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)shm));
    hipLaunchKernelGGL(k_wit_signed_tab<40>, dim3((unsigned)P), dim3(WNT), shm, st, g);
  } else {
    if (shm) W_HIP(hipFuncSetAttribute((const void*)k_wit_signed_tab<42>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)shm));
    hipLaunchKernelGGL(k_wit_signed_tab<42>, dim3((unsigned)P), dim3(WNT), shm, st, g);
  }
  hipLaunchKernelGGL(k_wit_counters, dim3(1), dim3(64), 0, st, W->cnt, W->dmax, total, counters);
  W_HIP(hipGetLastError());
  return GLW_OK;
}

}  // extern "C"
