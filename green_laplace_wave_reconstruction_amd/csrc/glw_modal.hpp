// Single-direction modal time path: kernels of the one-dimensional modal GL
// resolvent (python/src/green_laplace/modal_time.py:187-341) and of the
// difference-frequency eta20 (python/src/green_laplace/_eta20.py:13-132).
//
// Grids.  The reference works on [P][W] arrays: P modal wavenumbers
// q_k = k dq by W = order*span + 1 work frequencies nu_w = (base + w) fstep.
// The input spectrum is a deposit: every input bin w puts its amplitude on the
// two modal nodes k_lo(w), k_lo(w)+1 that bracket its dispersion wavenumber.
// All parents therefore lie in the modal band [k0, k0 + K], the n-th order
// products in [n k0, n k0 + n K], and all frequencies in [n base, n (base +
// span)].  The kernels keep every field band-shifted: modal index m = k - n k0
// on an x grid of Q >= order K + 1 points, frequency index w = b - n base on a
// t grid of T >= order span + 1 points.  Products of band-shifted fields are
// band-shifted products, and with those lengths no product wraps, so every
// spectrum the reference reads is reproduced exactly (up to rounding) on grids
// about half the reference's size.  eta20 multiplies a field by a conjugate,
// which cancels the shifts: its spectra are signed, m in [-K, K] and
// w in [-span, span].
//
// Layout in HBM.  Every array is written with whole 128-byte lines: the time
// pass writes its products row-contiguous, [field][Q][ldw] (ldw = kept
// frequencies rounded up to 8), and the modal pass writes the pair fields
// column-contiguous, [field][w][Q].  Each pass reads the other's layout
// across the grain; workgroups that share lines are scheduled on one XCD
// (row_block, col_block) so the lines are fetched into its L2 once.
//
// Kernels.
//   k_modal_t<MODE>  one workgroup per RB x rows: builds each first-order
//                    field's row from the deposit (two nodes per frequency,
//                    exp(2 pi i m j / Q) from the Q-point table), or loads a
//                    pair field's row, transforms it over t in LDS (two
//                    fields per transform on power-of-two T), folds it
//                    into the pointwise products held in registers, and
//                    finally inverse-transforms the products and stores the
//                    kept frequencies.
//   k_modal_q<MODE>  one workgroup per frequency column: forward modal FFT
//                    of the two forcing columns, GL node response on the
//                    valid set s > sqrt(|q| tanh |q|), workgroup sum into
//                    the output coefficient of that frequency (and the
//                    source-energy audit); for order 3 it also forms the
//                    pair spectra e2, p2 and inverse-transforms the five
//                    modal multiples the time pass needs.
//   k_modal_pre/post modal axes longer than one workgroup's LDS (Q = R0 Q',
//                    R0 = 2, 4, 8): radix-R0 fold of the forcing columns
//                    before k_modal_q runs its R0 sub-transforms of Q'
//                    points per column, and the matching unfold of the pair
//                    fields after it.
#pragma once

#include <hip/hip_runtime.h>

#include "glw_fft.hpp"

namespace glw {
namespace modal {

using C = cpx<double>;
constexpr int kMaxRank = 16;
constexpr int kPairFields = 5;  // e2, e2 q, p2 q, p2 q tanh q, p2 q^2

enum Mode : int { MODE_O2 = 2, MODE_O3 = 3, MODE_E20 = 20 };

// Deposit of one input bin: amplitude, bracket weights, lower node index.
struct Dep {
  double re, im, wlo, whi, klo;
};

struct Args {
  int q_len, t_len;      // reduced grid lengths Q, T
  int count;             // samples of the record
  int base, span;        // input bins base .. base + span (work index w = bin - base)
  int k0, kspan;         // deposit band k0 .. k0 + kspan on the modal grid
  int order;             // 2 or 3 (modal path)
  double dq, fstep, depth;
  const Dep* dep;        // [span + 1]
  const C* tw_t;         // exp(-2 pi i m / T), m < T
  const C* tw_q;         // exp(-2 pi i m / Q), m < Q
  // Modal lengths beyond one workgroup's LDS run as R0 sub-transforms of
  // length Q / R0 (k_modal_pre, k_modal_q over blockIdx.y, k_modal_post):
  const C* tw_qs;        // exp(-2 pi i m / (Q / R0)), m < Q / R0 (tw_q when R0 = 1)
  double* partial;       // [ncol][R0][4] per-sub-transform sums (R0 > 1)
  C* src;                // [2][Q][ldw2]  order-2 sources (d, k); eta20: forcing (d, k) at signed columns
  C* pair;               // [5][2 span + 1][Q]  pair fields in x
  C* pair_scratch;       // [2][2 span + 1][Q]  e2, p2 spectra between the eta22 pass and k_modal_pairs
  C* forcing;            // [2][Q][ldw3]  order-3 forcing (d, k)
  C* coef;               // [nw] output coefficient per kept frequency of the modal pass
  double* energy;        // [nw][2] source energy (total, invalid) per kept frequency
  int rank;
  double lam2, lam_in, lam3;  // GL scales: eta22, inner pair, eta33 (eta20: delta q)
  double nodes[kMaxRank], weights[kMaxRank];
};

__device__ __forceinline__ C conj_(C a) { return {a.x, -a.y}; }
__device__ __forceinline__ C cfma(C acc, C a, C b) {  // acc + a b
  return {acc.x + a.x * b.x - a.y * b.y, acc.y + a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ C cmulr(C a, double s) { return {a.x * s, a.y * s}; }

// Frequency multipliers of the time-pass fields (functions of nu_w).
enum WMul : int { W_ONE, W_NU, W_NU2, W_INU, W_I, W_MI_NU, W_MI_INU, W_MNU2, W_MONE };

__device__ __forceinline__ C wmul(int wm, C v, double nu) {
  switch (wm) {
    case W_NU: return cmulr(v, nu);
    case W_NU2: return cmulr(v, nu * nu);
    case W_INU: return cmulr(v, 1.0 / fmax(nu, 2.2250738585072014e-308));
    case W_I: return {-v.y, v.x};
    case W_MI_NU: return {v.y * nu, -v.x * nu};
    case W_MI_INU: {
      const double inu = 1.0 / fmax(nu, 2.2250738585072014e-308);
      return {v.y * inu, -v.x * inu};
    }
    case W_MNU2: return cmulr(v, -(nu * nu));
    case W_MONE: return {-v.x, -v.y};
    default: return v;
  }
}

// One field of the time pass: a first-order field (modal power q^idx of the
// deposit) or a pair field (x-space array idx of the modal pass), with a
// frequency multiplier.  Pair multipliers use nu_in = (2 base + w) fstep.
struct Recipe {
  bool pair;
  int idx, wm;
};

// Field sequences.  The order keeps at most five per-point values live:
//   MODE_O2  v, vn, vn2, hx, jx, radial                 (modal_time.py:220-227)
//   MODE_O3  eta, phizzz, phitzz, Gphizz, Gphitz, etax, phix, phixz, Getax,
//            Gphix, phiz, phizz, Gphiz, Geta, phitz     (modal_time.py:273-310)
//   MODE_E20 one, dno, nu, bx, qx, radial               (_eta20.py:64-86)
template <int MODE>
struct Steps;
template <>
struct Steps<MODE_O2> {
  static constexpr int N = 6;
};
template <>
struct Steps<MODE_O3> {
  static constexpr int N = 15;
};
template <>
struct Steps<MODE_E20> {
  static constexpr int N = 6;
};

__device__ __forceinline__ Recipe recipe(int mode, int step) {
  if (mode == MODE_O3) {
    switch (step) {
      case 0: return {false, 0, W_ONE};     // eta
      case 1: return {false, 2, W_MI_NU};   // phizzz = -i q^2 nu
      case 2: return {false, 2, W_MONE};    // phitzz = -q^2
      case 3: return {true, 4, W_ONE};      // Gphizz = p2 q^2
      case 4: return {true, 3, W_MI_NU};    // Gphitz = -i nu_in p2 A
      case 5: return {false, 1, W_I};       // etax = i q
      case 6: return {false, 1, W_INU};     // phix = q / nu
      case 7: return {false, 1, W_NU};      // phixz = nu q
      case 8: return {true, 1, W_I};        // Getax = i e2 q
      case 9: return {true, 2, W_I};        // Gphix = i p2 q
      case 10: return {false, 0, W_MI_NU};  // phiz = -i nu
      case 11: return {false, 2, W_MI_INU}; // phizz = -i q^2 / nu
      case 12: return {true, 3, W_ONE};     // Gphiz = p2 A
      case 13: return {true, 0, W_ONE};     // Geta = e2
      default: return {false, 0, W_MNU2};   // phitz = -nu^2
    }
  }
  switch (step) {  // MODE_O2 and MODE_E20 share the six first-order fields
    case 0: return {false, 0, W_ONE};
    case 1: return {false, 0, mode == MODE_O2 ? W_NU : W_NU2};
    case 2: return {false, 0, mode == MODE_O2 ? W_NU2 : W_NU};
    case 3: return {false, 1, W_INU};
    case 4: return {false, 1, W_ONE};
    default: return {false, 2, W_INU};
  }
}

// First-order row element (x row j, work frequency w < span + 1) for the
// modal powers q^0, q^1, q^2: sum over the two deposit nodes k of
// A_w wt_k (k dq)^p exp(2 pi i (k - k0) j / Q), the phase read from the
// Q-point table ((k - k0) j < Q^2 fits 32 bits for Q <= 8192).
struct Powers {
  C p0, p1, p2;
};
__device__ __forceinline__ Powers first_values(const Dep* __restrict__ dep, const C* __restrict__ twq, int k0, int q_len,
                                               double dq, int j, int w) {
  const Dep d = dep[w];
  const int klo = (int)d.klo;
  const int i0 = ((klo - k0) * j) % q_len;
  int i1 = i0 + j;
  if (i1 >= q_len) i1 -= q_len;
  const C e0 = conj_(twq[i0]), e1 = conj_(twq[i1]);
  const C amp = {d.re, d.im};
  const C u0 = cmulr(cmul(amp, e0), d.wlo), u1 = cmulr(cmul(amp, e1), d.whi);
  const double q0 = (double)klo * dq, q1 = (double)(klo + 1) * dq;
  Powers b;
  b.p0 = cadd(u0, u1);
  b.p1 = {u0.x * q0 + u1.x * q1, u0.y * q0 + u1.y * q1};
  b.p2 = {u0.x * (q0 * q0) + u1.x * (q1 * q1), u0.y * (q0 * q0) + u1.y * (q1 * q1)};
  return b;
}

// Row stride of a time-pass output with nw kept frequencies (whole lines).
__host__ __device__ constexpr int out_ld(int nw) { return (nw + 7) & ~7; }

// Store of the kept frequencies of an inverse t transform, scaled by 1/T,
// into rows of `out` (row j0 + r at r * ld): w < nw (shifted spectra) or
// signed w in [-span, span] at column w + span.
template <bool SIGNED>
struct TStore {
  static constexpr bool SEQ_FAST = false;  // lanes walk the contiguous frequency axis
  C* out;  // out + field * Q * ld + j0 * ld
  int ld, t_len, nw, span;
  double scale;
  int rb;              // rows per field: sequence r >= rb is row r - rb of the next field
  size_t field_step;   // elements between fields (Q ld)
  __device__ void operator()(int seq, int w, C v) const {
    const int fld = seq >= rb ? 1 : 0;
    const int r = seq - fld * rb;
    int c = w;
    if constexpr (SIGNED) {
      if (w <= span) c = w + span;
      else if (w >= t_len - span) c = w - t_len + span;
      else return;
    } else {
      if (w >= nw) return;
    }
    out[fld * field_step + (size_t)r * ld + c] = cmulr(v, scale);
  }
};

// Column loader of the modal pass: element m of a column at stride ld.
struct ColLoad {
  static constexpr bool SEQ_FAST = false;
  const C* col;
  int ld;
  __device__ C operator()(int, int m) const { return col[(size_t)m * ld]; }
};
struct ColStore {
  static constexpr bool SEQ_FAST = false;
  C* col;
  __device__ void operator()(int, int m, C v) const { col[m] = v; }
};

// Launch shape of the time pass at length T: RB rows per workgroup (RB T <=
// kRowWork), NT threads, about kPoints points per thread.
#ifndef GLW_MODAL_ROW_WORK
#define GLW_MODAL_ROW_WORK 2048
#endif
#ifndef GLW_MODAL_POINTS
#define GLW_MODAL_POINTS 4
#endif
constexpr int kRowWork = GLW_MODAL_ROW_WORK;
constexpr int kPoints = GLW_MODAL_POINTS;
// The order-2 and order-3 passes at T = 1024 take one row per workgroup (256
// threads) held to 256 registers (two waves per SIMD, two workgroups per CU):
// their transforms are barrier-bound, and a second workgroup per CU fills the
// barrier waits (order 3 on MI355X: 0.586 ms against 0.631 ms for two rows,
// 512 threads and one workgroup per CU; one row with 512 registers: 0.997 ms).
__host__ __device__ constexpr bool t_one_row(int T, int MODE) { return (MODE == MODE_O3 || MODE == MODE_O2) && T == 1024; }
__host__ __device__ constexpr int t_row_work(int T, int MODE) { return t_one_row(T, MODE) ? 1024 : kRowWork; }
__host__ __device__ constexpr int t_rows(int T, int MODE) {
  return T >= t_row_work(T, MODE) ? 1 : (t_row_work(T, MODE) / T > 4 ? 4 : t_row_work(T, MODE) / T);
}
__host__ __device__ constexpr int t_threads(int T, int MODE) {
  // past 2048 points twice the points per thread: 512 threads keep the 256-VGPR budget
  const int n = (t_rows(T, MODE) * T / (T > 2048 ? 2 * kPoints : kPoints)) / 64 * 64;
  return n < 64 ? 64 : (n > 1024 ? 1024 : n);
}
__host__ __device__ constexpr int t_min_waves(int T, int MODE) { return t_one_row(T, MODE) ? 2 : 1; }
// Radix plan of the time-pass transforms (1: greedy radix 8/4/2/3, fewer
// registers per butterfly than the composite radix-16 plan 0).
#ifndef GLW_MODAL_TPLAN
#define GLW_MODAL_TPLAN 1
#endif
constexpr int kTPlan = GLW_MODAL_TPLAN;
__host__ __device__ constexpr int q_threads(int Q) { return Q <= 512 ? 64 : (Q <= 2048 ? 256 : 512); }
// The modal passes at Q = 4096 are held to 128 registers (four waves per SIMD,
// two workgroups per CU; eta33 on MI355X: 0.186 ms against 0.213 ms at 168
// registers and one workgroup per CU).
template <int Q, int MODE>
__host__ __device__ constexpr int q_min_waves() { return Q == 4096 ? 4 : 1; }

// Row block of workgroup b.  A 128-byte line of a [w][Q] array holds 8
// consecutive x rows; the G = 8 / RB workgroups that share those lines are
// given consecutive dispatch slots of one XCD (workgroups are dealt to the
// 8 XCDs round robin), so that XCD's L2 merges their partial-line writes and
// serves their shared line reads.  Identity when the grid is not a multiple
// of 8 G.
__device__ __forceinline__ int row_block(int b, int nblocks, int rb) {
  const int g = 8 / rb;
  if (g <= 1 || b >= nblocks / (8 * g) * (8 * g)) return b;
  return g * (8 * (b / (8 * g)) + b % 8) + (b / 8) % g;
}

// Column of modal-pass workgroup b: the 8 columns that share the lines of a
// row-contiguous time-pass output go to consecutive slots of one XCD (the
// tail beyond a multiple of 64 keeps its own index).
__device__ __forceinline__ int col_block(int b, int nblocks) {
  if (b >= nblocks / 64 * 64) return b;
  return 8 * (8 * (b / 64) + b % 8) + (b / 8) % 8;
}

// Time pass.  MODE_O2: sources d = 2 v vn2 + vn^2 - hx^2, k = 2 v radial +
// 2 hx jx (modal_time.py:220-227) into a.src.  MODE_O3: forcing
// (modal_time.py:273-310) into a.forcing.  MODE_E20: difference forcing
// (_eta20.py:64-86, the qy = 0 plane) into a.src at signed columns.
// Every field runs through the same staged LDS transform (one FFT instance).
// Fields transformed per LDS FFT call in the time pass (1 or 2): two fields
// give the radix stages twice the butterflies (all threads busy at RB = 2)
// and half the barriers (order-3 pass at T = 1024: 0.63 ms instead of
// 0.75 ms).  Lengths with a radix-3 stage keep one field: their butterflies
// need the registers (two fields spill there).
#ifndef GLW_MODAL_FIELDS_PER_FFT
#define GLW_MODAL_FIELDS_PER_FFT 2
#endif
__host__ __device__ constexpr int fields_per_fft(int T) { return T % 3 == 0 ? 1 : GLW_MODAL_FIELDS_PER_FFT; }

// LDS of the time pass beyond its transform buffer.  The twiddles of T
// points are copied to LDS when the workgroup's LDS stays within 80 KB (two
// workgroups per CU): every radix stage of the fifteen transforms reads them.
// The order-3 pass also builds its row's three first-order base rows (q^0,
// q^1, q^2 at every input bin, order * span + 1 <= T, so at most T / 3 + 1
// bins) once, in dynamic LDS, where nine of its fields read them; each
// element is two dependent global gathers (deposit, then phase table) when
// built per field.
template <int T, int MODE>
__host__ __device__ constexpr int t_buf_bytes() {
  return fields_per_fft(T) * t_rows(T, MODE) * seq_stride<T>() * 16;
}
template <int T, int MODE>
__host__ __device__ constexpr bool t_twiddles_lds() {
  return t_buf_bytes<T, MODE>() + T * 16 <= 80 * 1024;
}
template <int T, int MODE>
__host__ __device__ constexpr bool t_rows_lds() {
  return MODE == MODE_O3 && t_buf_bytes<T, MODE>() + (t_twiddles_lds<T, MODE>() ? T * 16 : 0) +
                                    3 * t_rows(T, MODE) * (T / 3 + 1) * 16 <=
                                96 * 1024;
}
// dynamic LDS bytes of a launch with nbin input bins
template <int T, int MODE>
__host__ __device__ constexpr int t_dyn_bytes(int nbin) {
  return t_rows_lds<T, MODE>() ? 3 * t_rows(T, MODE) * nbin * 16 : 0;
}

template <int T, int MODE>
__global__ __launch_bounds__(t_threads(T, MODE), t_min_waves(T, MODE)) void k_modal_t(Args a) {
  constexpr int RB = t_rows(T, MODE), NT = t_threads(T, MODE);
  constexpr int SS = seq_stride<T>();
  constexpr int PT = (RB * T + NT - 1) / NT;
  constexpr int FB = fields_per_fft(T);
  __shared__ C buf[FB * RB * SS];
  constexpr bool TWL = t_twiddles_lds<T, MODE>(), ROWL = t_rows_lds<T, MODE>();
  __shared__ C twl[TWL ? T : 1];
  extern __shared__ C rowl[];  // ROWL: [3][RB][nbin] first-order base rows q^0, q^1, q^2
  const int j0 = row_block(blockIdx.x, gridDim.x, RB) * RB;
  const int nbin = a.span + 1;
  const int nw2 = 2 * a.span + 1;
  const size_t fs2 = (size_t)nw2 * a.q_len;
  const C* tw = a.tw_t;
  if constexpr (TWL) {  // visible after the first barrier of the field loop
    for (int i = threadIdx.x; i < T; i += NT) twl[i] = a.tw_t[i];
    tw = twl;
  }
  if constexpr (ROWL) {
    for (int e = threadIdx.x; e < RB * nbin; e += NT) {
      const int r = e % RB, w = e / RB;
      const Powers b = first_values(a.dep, a.tw_q, a.k0, a.q_len, a.dq, j0 + r, w);
      rowl[(0 * RB + r) * nbin + w] = b.p0;
      rowl[(1 * RB + r) * nbin + w] = b.p1;
      rowl[(2 * RB + r) * nbin + w] = b.p2;
    }
  }
  C f[PT], e[PT], x[PT], h[PT], acc_d[PT], acc_k[PT];
#pragma unroll
  for (int i = 0; i < PT; ++i) acc_d[i] = acc_k[i] = x[i] = h[i] = e[i] = {0.0, 0.0};
  // the first-order rows (q^0, q^1, q^2) of this thread's staging elements,
  // element el = tid + k NT being (row el % RB, w = el / RB): built once and
  // rescaled per field where the registers allow (order 2, eta20), rebuilt
  // per field for order 3 (whose fifteen-field products need the registers)
  constexpr bool kCache = MODE != MODE_O3;
  constexpr int PC = kCache ? PT : 1;
  C b0[PC], b1[PC], b2[PC];
  if constexpr (kCache) {
#pragma unroll
    for (int k = 0; k < PT; ++k) {
      const int el = threadIdx.x + k * NT;
      const int r = el % RB, w = el / RB;
      Powers b = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
      if (el < RB * T && w < nbin) b = first_values(a.dep, a.tw_q, a.k0, a.q_len, a.dq, j0 + r, w);
      b0[k] = b.p0;
      b1[k] = b.p1;
      b2[k] = b.p2;
    }
  }
  // stage field `step` of the sequence into LDS sequences slot * RB + r
  auto stage = [&](int step, int slot) {
    const Recipe rc = recipe(MODE, step < Steps<MODE>::N ? step : 0);
    const bool live = step < Steps<MODE>::N;
#pragma unroll
    for (int k = 0; k < PT; ++k) {
      const int el = threadIdx.x + k * NT;
      if (el >= RB * T) continue;
      const int r = el % RB, w = el / RB;  // rows fastest: pair loads coalesce over x
      C v = {0.0, 0.0};
      if (!live) {
      } else if (!rc.pair) {
        C u = {0.0, 0.0};
        if constexpr (ROWL) {
          if (w < nbin) u = rowl[(rc.idx * RB + r) * nbin + w];
        } else {
          Powers b;
          if constexpr (kCache) {
            b = {b0[k], b1[k], b2[k]};
          } else {
            b = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
            if (w < nbin) b = first_values(a.dep, a.tw_q, a.k0, a.q_len, a.dq, j0 + r, w);
          }
          // arithmetic select: a branch or ?: on the runtime power is lowered to an indexed private array
          const double s0 = rc.idx == 0, s1 = rc.idx == 1, s2 = rc.idx == 2;
          u = {s0 * b.p0.x + s1 * b.p1.x + s2 * b.p2.x, s0 * b.p0.y + s1 * b.p1.y + s2 * b.p2.y};
        }
        v = wmul(rc.wm, u, (double)(a.base + w) * a.fstep);
      } else if (w < nw2) {
        v = wmul(rc.wm, a.pair[rc.idx * fs2 + (size_t)w * a.q_len + j0 + r], (double)(2 * a.base + w) * a.fstep);
      }
      buf[lds_at<T>(slot * RB + r, w)] = v;
    }
  };
  // fold the transformed field `step` (in f) into the per-point products
  auto update = [&](int step) {
      if constexpr (MODE == MODE_O2) {
        // e = v, h = hx
  #pragma unroll
        for (int i = 0; i < PT; ++i) {
          switch (step) {
            case 0: e[i] = f[i]; break;
            case 1: acc_d[i] = cmul(f[i], f[i]); break;
            case 2: acc_d[i] = cfma(acc_d[i], cmulr(e[i], 2.0), f[i]); break;
            case 3: h[i] = f[i]; acc_d[i] = csub(acc_d[i], cmul(f[i], f[i])); break;
            case 4: acc_k[i] = cmul(cmulr(h[i], 2.0), f[i]); break;
            default: acc_k[i] = cfma(acc_k[i], cmulr(e[i], 2.0), f[i]); break;
          }
        }
      } else if constexpr (MODE == MODE_O3) {
        // fk = (eta phixz etax - eta^2 phizzz / 2) / 4 + (phix Getax + Gphix etax - eta Gphizz - Geta phizz) / 2
        // fd = (eta^2 phitzz / 2 + eta (phix phixz + phiz phizz)) / 4
        //      + (eta Gphitz + Geta phitz + phix Gphix + phiz Gphiz) / 2
        // live values: e = eta; x = etax, then phiz, then Geta; h = phix, then phizz
  #pragma unroll
        for (int i = 0; i < PT; ++i) {
          switch (step) {
            case 0: e[i] = f[i]; break;
            case 1: acc_k[i] = cmul(cmulr(cmul(e[i], e[i]), -0.125), f[i]); break;
            case 2: acc_d[i] = cmul(cmulr(cmul(e[i], e[i]), 0.125), f[i]); break;
            case 3: acc_k[i] = csub(acc_k[i], cmul(cmulr(e[i], 0.5), f[i])); break;
            case 4: acc_d[i] = cfma(acc_d[i], cmulr(e[i], 0.5), f[i]); break;
            case 5: x[i] = f[i]; break;
            case 6: h[i] = f[i]; break;
            case 7: {
              const C ef = cmulr(cmul(e[i], f[i]), 0.25);
              acc_k[i] = cfma(acc_k[i], ef, x[i]);
              acc_d[i] = cfma(acc_d[i], ef, h[i]);
              break;
            }
            case 8: acc_k[i] = cfma(acc_k[i], cmulr(h[i], 0.5), f[i]); break;
            case 9: {
              const C g = cmulr(f[i], 0.5);
              acc_k[i] = cfma(acc_k[i], g, x[i]);
              acc_d[i] = cfma(acc_d[i], g, h[i]);
              break;
            }
            case 10: x[i] = f[i]; break;
            case 11: h[i] = f[i]; acc_d[i] = cfma(acc_d[i], cmulr(cmul(e[i], x[i]), 0.25), f[i]); break;
            case 12: acc_d[i] = cfma(acc_d[i], cmulr(x[i], 0.5), f[i]); break;
            case 13: x[i] = f[i]; acc_k[i] = csub(acc_k[i], cmul(cmulr(f[i], 0.5), h[i])); break;
            default: acc_d[i] = cfma(acc_d[i], cmulr(x[i], 0.5), f[i]); break;
          }
        }
      } else {
        // fd = -dno one* - one dno* + nu nu* + bx bx*
        // fk = i (radial one* - bx qx* + qx bx* - one radial*);   e = one, h = bx
  #pragma unroll
        for (int i = 0; i < PT; ++i) {
          switch (step) {
            case 0: e[i] = f[i]; break;
            case 1: acc_d[i] = csub(cmul(cmulr(f[i], -1.0), conj_(e[i])), cmul(e[i], conj_(f[i]))); break;
            case 2: acc_d[i] = cfma(acc_d[i], f[i], conj_(f[i])); break;
            case 3: h[i] = f[i]; acc_d[i] = cfma(acc_d[i], f[i], conj_(f[i])); break;
            case 4: acc_k[i] = csub(cmul(f[i], conj_(h[i])), cmul(h[i], conj_(f[i]))); break;
            default: {
              const C t = csub(cadd(acc_k[i], cmul(f[i], conj_(e[i]))), cmul(e[i], conj_(f[i])));
              acc_k[i] = {-t.y, t.x};
              break;
            }
          }
        }
      }

  };
#pragma unroll 1
  for (int s0 = 0; s0 < Steps<MODE>::N; s0 += FB) {
    __syncthreads();  // previous transform fully read
#pragma unroll
    for (int slot = 0; slot < FB; ++slot) stage(s0 + slot, slot);
    __syncthreads();
    fft_lds<double, T, FB * RB, SS, NT, -1, kTPlan>(buf, tw);
#pragma unroll
    for (int slot = 0; slot < FB; ++slot) {
      if (s0 + slot >= Steps<MODE>::N) break;
#pragma unroll
      for (int i = 0; i < PT; ++i) {
        const int p = threadIdx.x + i * NT;
        f[i] = p < RB * T ? buf[lds_at<T>(slot * RB + p / T, p % T)] : C{0.0, 0.0};
      }
      update(s0 + slot);
    }
  }
  // inverse t transforms of (d, k), kept frequencies stored at 1/T
  C* out = MODE == MODE_O3 ? a.forcing : a.src;
  const int nw = MODE == MODE_O3 ? 3 * a.span + 1 : nw2;
  const int ld = out_ld(nw);
  NoIO none;
  if constexpr (FB == 2) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PT; ++i) {
      const int p = threadIdx.x + i * NT;
      if (p < RB * T) {
        buf[lds_at<T>(p / T, p % T)] = acc_d[i];
        buf[lds_at<T>(RB + p / T, p % T)] = acc_k[i];
      }
    }
    __syncthreads();
    // sequences 0..RB-1: d rows, RB..2RB-1: k rows (field stride Q ld)
    TStore<MODE == MODE_E20> st{out + (size_t)j0 * ld, ld, T, nw, a.span, 1.0 / T, RB, (size_t)a.q_len * ld};
    fft_io<double, T, FB * RB, SS, NT, +1, kTPlan>(buf, tw, none, st);
  } else {
#pragma unroll 1
    for (int o = 0; o < 2; ++o) {
      const double od = o == 0, ok = o == 1;  // arithmetic select of the output (see above)
      __syncthreads();
#pragma unroll
      for (int i = 0; i < PT; ++i) {
        const int p = threadIdx.x + i * NT;
        if (p < RB * T) buf[lds_at<T>(p / T, p % T)] = C{od * acc_d[i].x + ok * acc_k[i].x, od * acc_d[i].y + ok * acc_k[i].y};
      }
      __syncthreads();
      TStore<MODE == MODE_E20> st{out + ((size_t)o * a.q_len + j0) * ld, ld, T, nw, a.span, 1.0 / T, RB, 0};
      fft_io<double, T, RB, SS, NT, +1, kTPlan>(buf, tw, none, st);
    }
  }
}

// GL node sums on the valid set s > av (modal_time.py:383-399):
// S = sum (E+ - E-) / (2 av), C = sum (E+ + E-) / 2 with
// E+- = (w / lam) exp(x - (s -+ av) x / lam); at av == 0 the limits
// S = sum (w / lam) exp(x - s tau) tau, C = sum (w / lam) exp(x - s tau).
__device__ __forceinline__ void node_sums(const Args& a, double lam, double s, double av, double& S, double& Cs) {
  S = 0.0;
  Cs = 0.0;
  for (int n = 0; n < a.rank; ++n) {
    const double x = a.nodes[n];
    const double tau = x / lam;
    const double wl = a.weights[n] / lam;
    if (av > 0.0) {
      const double plus = wl * exp(x - (s - av) * tau);
      const double minus = wl * exp(x - (s + av) * tau);
      S += (plus - minus) / (2.0 * av);
      Cs += (plus + minus) / 2.0;
    } else {
      const double b = wl * exp(x - s * tau);
      S += b * tau;
      Cs += b;
    }
  }
}

template <int NT>
__device__ __forceinline__ void block_sum4(double (&v)[4], double* red) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    v[k] = x;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) red[wv * 4 + k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < 4; ++k) {
      double s = 0.0;
      for (int i = 0; i < NT / 64; ++i) s += red[i * 4 + k];
      v[k] = s;
    }
  }
}

// Modal pass over frequency column w = blockIdx.x.
//   MODE_O2: eta22 coefficient of bin 2 base + w from the sources, and (order
//            3) the pair spectra e2, p2 (modal_time.py:240-271) transformed
//            to x as e2, e2 q, p2 q, p2 A, p2 q^2.
//   MODE_O3: eta33 coefficient of bin 3 base + w (modal_time.py:316-327).
//   MODE_E20: eta20 coefficient of signed bin w - span (_eta20.py:103-123).
// Q here is the length of the transform in LDS.  With R0 = gridDim.y > 1 the
// modal length is a.q_len = R0 Q: workgroup r = blockIdx.y transforms the
// r-th sequence k_modal_pre left in rows r Q .. r Q + Q - 1 (the modes
// m = R0 m' + r), writes its sums to a.partial, and leaves the inverse
// transforms of its pair spectra for k_modal_post to combine.
template <int Q, int MODE>
__global__ __launch_bounds__(q_threads(Q), (q_min_waves<Q, MODE>())) void k_modal_q(Args a) {
  constexpr int NT = q_threads(Q);
  constexpr int SS = seq_stride<Q>();
  constexpr int PQ = (Q + NT - 1) / NT;
  __shared__ C buf[SS];
  __shared__ double red[NT / 64 * 4];
  const int w = col_block(blockIdx.x, gridDim.x);
  const int r0 = gridDim.y, rs = blockIdx.y;  // R0 sub-transforms; this one's index
  const int nw = MODE == MODE_O3 ? 3 * a.span + 1 : 2 * a.span + 1;
  const C* in = MODE == MODE_O3 ? a.forcing : a.src;
  const double inv_q = 1.0 / a.q_len;
  NoIO none;
  // forward modal FFT of the two forcing columns: d kept in registers, k left in LDS
  C fd[PQ];
#pragma unroll 1
  for (int o = 0; o < 2; ++o) {
    const double od = o == 0, ok = o == 1;  // arithmetic select of the output (see above)
    __syncthreads();
    ColLoad ld{in + ((size_t)o * a.q_len + (size_t)rs * Q) * out_ld(nw) + w, out_ld(nw)};
    fft_io<double, Q, 1, SS, NT, -1, Q >= 3072 ? 1 : 0>(buf, a.tw_qs, ld, none);
    if (o == 0) {
#pragma unroll
      for (int i = 0; i < PQ; ++i) {
        const int m = threadIdx.x + i * NT;
        fd[i] = m < Q ? cmulr(buf[lds_pad(m)], inv_q) : C{0.0, 0.0};
      }
    }
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};  // coefficient re, im; energy total, invalid
  const int n_ord = MODE == MODE_O3 ? 3 : 2;
  const int mmax = n_ord * a.kspan;  // kept modal band 0 .. n K (shifted) or -K .. K (signed)
  // modal_time.py:241-252 (eta22), 316-327 (eta33); _eta20.py:105-123 (eta20)
  const double s = MODE == MODE_E20 ? (double)(w - a.span) * a.fstep : (double)(n_ord * a.base + w) * a.fstep;
  const double lam = MODE == MODE_O2 ? a.lam2 : a.lam3;
  const bool pairs = MODE == MODE_O2 && a.order == 3;
#pragma unroll
  for (int i = 0; i < PQ; ++i) {
    const int ml = threadIdx.x + i * NT;  // element of this transform
    if (ml >= Q) continue;
    const int m = ml * r0 + rs;           // mode index on the a.q_len grid
    const C sd = fd[i], sk = cmulr(buf[lds_pad(ml)], inv_q);
    if (pairs) {  // fd[i] becomes e2, the LDS element p2 (zero off the valid band)
      fd[i] = C{0.0, 0.0};
      buf[lds_pad(ml)] = C{0.0, 0.0};
    }
    int k;
    if constexpr (MODE == MODE_E20) {
      if (m <= a.kspan) k = m;
      else if (m >= a.q_len - a.kspan) k = m - a.q_len;
      else continue;
    } else {
      if (m > mmax) continue;
      k = n_ord * a.k0 + m;
    }
    const double q = fabs((double)k * a.dq);
    const double aq = q * tanh(q);
    const double av = sqrt(aq);
    if constexpr (MODE != MODE_E20) {
      const double energy = sd.x * sd.x + sd.y * sd.y + sk.x * sk.x + sk.y * sk.y;
      acc[2] += energy;
      if (!(s > av)) {
        acc[3] += energy;
        continue;
      }
      double S, Cs;
      node_sums(a, lam, s, av, S, Cs);
      if constexpr (MODE == MODE_O2) {
        // (-A S sd + C sk) / (4 h)
        acc[0] += (-aq * S * sd.x + Cs * sk.x) / (4.0 * a.depth);
        acc[1] += (-aq * S * sd.y + Cs * sk.y) / (4.0 * a.depth);
        if (pairs) {
          double Si, Ci;
          node_sums(a, a.lam_in, s, av, Si, Ci);
          // e2 = (-A S sd + C sk) / 4, p2 = i (C sd - S sk) / 4
          fd[i] = {(-aq * Si * sd.x + Ci * sk.x) / 4.0, (-aq * Si * sd.y + Ci * sk.y) / 4.0};
          const C t = {Ci * sd.x - Si * sk.x, Ci * sd.y - Si * sk.y};
          buf[lds_pad(ml)] = {-t.y / 4.0, t.x / 4.0};
        }
      } else {
        // (A S fd - i C fk) / h^2
        const double h2 = a.depth * a.depth;
        acc[0] += (aq * S * sd.x + Cs * sk.y) / h2;
        acc[1] += (aq * S * sd.y - Cs * sk.x) / h2;
      }
    } else {
      if (!(q > 0.0 && av > fabs(s))) continue;
      double md = 0.0, mki = 0.0;  // mk = i mki
      for (int n = 0; n < a.rank; ++n) {
        const double x = a.nodes[n];
        const double tau = x / lam;
        const double wl = a.weights[n] / lam;
        const double plus = wl * exp(x - (av - s) * tau);
        const double minus = wl * exp(x - (av + s) * tau);
        md -= av * (plus + minus) / 4.0;
        mki += (plus - minus) / 4.0;
      }
      // (md fd + i mki fk) / (2 h)
      acc[0] += (md * sd.x - mki * sk.y) / (2.0 * a.depth);
      acc[1] += (md * sd.y + mki * sk.x) / (2.0 * a.depth);
    }
  }
  block_sum4<NT>(acc, red);
  if (threadIdx.x == 0) {
    if (r0 == 1) {
      a.coef[w] = {acc[0], acc[1]};
      a.energy[2 * w] = acc[2];
      a.energy[2 * w + 1] = acc[3];
    } else {
      double* pt = a.partial + ((size_t)w * r0 + rs) * 4;
      for (int k = 0; k < 4; ++k) pt[k] = acc[k];
    }
  }
  if constexpr (MODE == MODE_O2) {
    if (!pairs) return;
    // e2 and p2 spectra of this column to the pair scratch [2][ncol][q_len]
    // (sub-transform rs at rs Q); k_modal_pairs transforms them
    C* e2 = a.pair_scratch + (size_t)w * a.q_len + (size_t)rs * Q;
    C* p2 = e2 + (size_t)gridDim.x * a.q_len;
#pragma unroll
    for (int i = 0; i < PQ; ++i) {
      const int ml = threadIdx.x + i * NT;
      if (ml < Q) {
        e2[ml] = fd[i];
        p2[ml] = buf[lds_pad(ml)];  // own elements (block_sum4 synchronised)
      }
    }
  }
}

// Pair fields in x (modal_time.py:285-292, _spectral_field = fft_t(ifft_q(.) P)):
// e2, e2 q, p2 q, p2 A, p2 q^2, each inverse-transformed over the modal axis,
// one workgroup per (column, sub-transform, field): the 5 ncol transforms run
// in parallel instead of five in a row per column.
template <int Q>
__global__ __launch_bounds__(q_threads(Q)) void k_modal_pairs(Args a) {
  constexpr int NT = q_threads(Q);
  constexpr int SS = seq_stride<Q>();
  __shared__ C buf[SS];
  const int w = blockIdx.x, rs = blockIdx.y, f = blockIdx.z, r0 = gridDim.y;
  const C* src = a.pair_scratch + ((size_t)(f < 2 ? 0 : 1) * gridDim.x + w) * a.q_len + (size_t)rs * Q;
  for (int ml = threadIdx.x; ml < Q; ml += NT) {
    const double q = (double)(2 * a.k0 + ml * r0 + rs) * a.dq;
    double mul = 1.0;
    if (f == 1 || f == 2) mul = q;
    else if (f == 3) mul = q * tanh(q);
    else if (f == 4) mul = q * q;
    buf[lds_pad(ml)] = cmulr(src[ml], mul);
  }
  __syncthreads();
  const size_t fs = (size_t)gridDim.x * a.q_len;
  ColStore st{a.pair + f * fs + (size_t)w * a.q_len + (size_t)rs * Q};
  NoIO none;
  fft_io<double, Q, 1, SS, NT, +1, Q >= 3072 ? 1 : 0>(buf, a.tw_qs, none, st);
}

// Split modal transform, forward half (R0 sub-transforms of length QS = Q / R0):
// in place on the two columns [2][Q][ld] of `x`, row n' + p QS becomes
//   y_r[n'] = exp(-2 pi i n' r / Q) sum_p x[n' + p QS] exp(-2 pi i p r / R0)
// at row r QS + n', so that the QS-point transform of y_r holds modes R0 m' + r.
// Threads run along the columns (whole lines); each owns the R0 rows it reads.
template <int R0>
__global__ __launch_bounds__(64) void k_modal_pre(C* x, int q_len, int ld, int ncol, const C* tw_q) {
  const int w = blockIdx.x * 64 + threadIdx.x;
  const int np = blockIdx.y;  // n'
  if (w >= ncol) return;
  const int qs = q_len / R0;
  C* col = x + (size_t)blockIdx.z * q_len * ld + w;
  C v[R0], y[R0];
#pragma unroll
  for (int p = 0; p < R0; ++p) v[p] = col[(size_t)(np + p * qs) * ld];
#pragma unroll
  for (int r = 0; r < R0; ++r) {
    C s = v[0];
#pragma unroll
    for (int p = 1; p < R0; ++p) s = cfma(s, v[p], tw_q[((p * r) % R0) * qs]);
    y[r] = r == 0 ? s : glw::cmul(s, tw_q[np * r]);
  }
#pragma unroll
  for (int r = 0; r < R0; ++r) col[(size_t)(r * qs + np) * ld] = y[r];
}

// Split modal transform, inverse half: in place on each pair column
// [5][ncol][Q] (k_modal_q left z_r = the QS-point inverse transform of modes
// R0 m' + r at r QS .. r QS + QS - 1):
//   x[n' + p QS] = sum_r exp(2 pi i p r / R0) exp(2 pi i n' r / Q) z_r[n'].
template <int R0>
__global__ __launch_bounds__(256) void k_modal_post(C* pair, int q_len, const C* tw_q) {
  const int qs = q_len / R0;
  const int np = blockIdx.x * 256 + threadIdx.x;
  if (np >= qs) return;
  C* col = pair + ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * q_len;
  C z[R0], x[R0];
#pragma unroll
  for (int r = 0; r < R0; ++r) z[r] = r == 0 ? col[np] : glw::cmul(col[r * qs + np], conj_(tw_q[np * r]));
#pragma unroll
  for (int p = 0; p < R0; ++p) {
    C s = z[0];
#pragma unroll
    for (int r = 1; r < R0; ++r) s = cfma(s, z[r], conj_(tw_q[((p * r) % R0) * qs]));
    x[p] = s;
  }
#pragma unroll
  for (int p = 0; p < R0; ++p) col[p * qs + np] = x[p];
}

}  // namespace modal
}  // namespace glw
