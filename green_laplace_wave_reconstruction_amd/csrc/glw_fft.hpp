// Workgroup-cooperative mixed-radix Stockham FFT in LDS for gfx950 (CDNA4).
//
// A workgroup transforms NSEQ sequences of length L that sit in LDS at a
// stride SSTR (complex elements).  Stages use radices 8, 4, 3 and 2 chosen at
// compile time; each stage reads every butterfly's R inputs into registers,
// synchronises, and writes the R outputs back in place (autosort, natural
// order out).  Twiddles come from a per-length table w[m] = exp(-2 pi i m / L)
// built on the host in FP64; the inverse transform conjugates them.  The
// transforms are unnormalised in both directions.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>

namespace glw {

template <typename T>
struct cpx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by i*SIGN  (SIGN = +1: i, SIGN = -1: -i)
template <int SIGN, typename T>
__device__ __forceinline__ cpx<T> cmul_i(cpx<T> a) {
  if constexpr (SIGN > 0) return {-a.y, a.x};
  else return {a.y, -a.x};
}

// LDS layout of a batch of sequences: element e of sequence s lives at
// s * seq_stride<L>() + e + e / 16.  One pad slot every 16 elements breaks the
// power-of-two strides of the Stockham writes, and the odd sequence stride
// keeps lanes that walk across sequences (column tiles) on distinct banks.
__host__ __device__ constexpr int lds_pad(int e) { return e + (e >> 4); }
// Element offset within a sequence: padded (the layout above) or dense.
template <bool PAD>
__host__ __device__ constexpr int elem_at(int e) { return PAD ? lds_pad(e) : e; }
template <int L>
__host__ __device__ constexpr int seq_stride() {
  return (L + (L >> 4)) | 1;
}
template <int L>
__device__ __forceinline__ int lds_at(int s, int e) {
  return s * seq_stride<L>() + lds_pad(e);
}

// Stage radices: two passes where an in-register radix <= 18 allows it
// (fewer LDS round trips and barriers), else greedy 16/8/4/2/3 factors.
__host__ __device__ constexpr int radix_greedy(int L, int s) {
  int rem = L;
  for (int i = 0; i <= s; ++i) {
    int r = 0;
    if (rem <= 1) return 0;
    if (rem % 8 == 0) r = 8;
    else if (rem % 4 == 0) r = 4;
    else if (rem % 2 == 0) r = 2;
    else if (rem % 3 == 0) r = 3;
    else return -1;
    if (i == s) return r;
    rem /= r;
  }
  return 0;
}

// Schedule PLAN 0: two passes of composite radices; PLAN 1: greedy 8/4/2/3.
template <int PLAN>
__host__ __device__ constexpr int radix_plan(int L, int s);

__host__ __device__ constexpr int radix_of(int L, int s) {
  switch (L) {  // explicit plans for the lengths the kernels use most
    case 16: return s == 0 ? 16 : 0;
    case 24: return s == 0 ? 6 : (s == 1 ? 4 : 0);
    case 32: return s == 0 ? 8 : (s == 1 ? 4 : 0);
    case 48: return s == 0 ? 12 : (s == 1 ? 4 : 0);
    case 64: return s == 0 ? 8 : (s == 1 ? 8 : 0);
    case 96: return s == 0 ? 12 : (s == 1 ? 8 : 0);
    case 128: return s == 0 ? 16 : (s == 1 ? 8 : 0);
    case 144: return s == 0 ? 12 : (s == 1 ? 12 : 0);
    case 192: return s == 0 ? 16 : (s == 1 ? 12 : 0);
    case 256: return s == 0 ? 16 : (s == 1 ? 16 : 0);
    case 288: return s == 0 ? 18 : (s == 1 ? 16 : 0);
    case 384: return s == 0 ? 16 : (s == 1 ? 12 : (s == 2 ? 2 : 0));
    case 512: return s == 0 ? 8 : (s == 1 ? 8 : (s == 2 ? 8 : 0));
    case 576: return s == 0 ? 16 : (s == 1 ? 12 : (s == 2 ? 3 : 0));
    case 768: return s == 0 ? 16 : (s == 1 ? 16 : (s == 2 ? 3 : 0));
    case 1024: return s == 0 ? 16 : (s == 1 ? 8 : (s == 2 ? 8 : 0));
    default: break;
  }
  int rem = L;
  for (int i = 0; i <= s; ++i) {
    int r = 0;
    if (rem <= 1) return 0;
    if (rem % 16 == 0) r = 16;
    else if (rem % 8 == 0) r = 8;
    else if (rem % 4 == 0) r = 4;
    else if (rem % 2 == 0) r = 2;
    else if (rem % 3 == 0) r = 3;
    else return -1;  // unsupported factor
    if (i == s) return r;
    rem /= r;
  }
  return 0;
}

template <>
__host__ __device__ constexpr int radix_plan<0>(int L, int s) { return radix_of(L, s); }
template <>
__host__ __device__ constexpr int radix_plan<1>(int L, int s) { return radix_greedy(L, s); }

// Schedule actually compiled for PLAN at length L: PLAN 1 folds onto PLAN 0
// where both give the same radix sequence, so no kernel is built twice.
template <int L, int PLAN>
constexpr int canonical_plan() {
  if (PLAN == 0) return 0;
  for (int s = 0; s < 16; ++s)
    if (radix_plan<0>(L, s) != radix_plan<1>(L, s)) return PLAN;
  return 0;
}

// Compile-time cos/sin of 2 pi m / n (Taylor series after reduction to
// |x| <= pi/4), accurate to a few ulp of double.
__host__ __device__ constexpr long double taylor_sin(long double x) {
  long double term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2 * k) * (2 * k + 1));
    sum += term;
  }
  return sum;
}
__host__ __device__ constexpr long double taylor_cos(long double x) {
  long double term = 1, sum = 1;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2 * k - 1) * (2 * k));
    sum += term;
  }
  return sum;
}
// returns cos (which=0) or sin (which=1) of 2 pi m / n
__host__ __device__ constexpr double unit_root(int m, int n, int which) {
  const long double pi4 = 0.785398163397448309615660845819875721L;
  m = ((m % n) + n) % n;
  // angle = 2 pi m / n = (8 m / n) * pi/4 ; octant o and remainder r in [0, 1)
  const long double t = 8.0L * m / n;
  int o = (int)t;
  long double r = t - o;
  if (r > 0.5L) {
    r -= 1.0L;
    o += 1;
  }
  const long double c = taylor_cos(r * pi4), s = taylor_sin(r * pi4);
  // rotate (c, s) by o * pi/4
  long double cc = c, ss = s;
  const long double h = 0.707106781186547524400844362104849039L;
  switch (o & 7) {
    case 0: cc = c; ss = s; break;
    case 1: cc = h * (c - s); ss = h * (c + s); break;
    case 2: cc = -s; ss = c; break;
    case 3: cc = -h * (c + s); ss = h * (c - s); break;
    case 4: cc = -c; ss = -s; break;
    case 5: cc = h * (s - c); ss = -h * (c + s); break;
    case 6: cc = s; ss = -c; break;
    default: cc = h * (c + s); ss = h * (s - c); break;
  }
  return (double)(which ? ss : cc);
}

__host__ __device__ constexpr int small_factor(int R) {
  return R % 4 == 0 ? 4 : (R % 3 == 0 ? 3 : 2);
}

__host__ __device__ constexpr bool fft_length_ok(int L) {
  int rem = L;
  while (rem % 2 == 0) rem /= 2;
  while (rem % 3 == 0) rem /= 3;
  return rem == 1 && L >= 8;
}

template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft_reg(cpx<T>* v);

// Composite in-register DFT (Cooley-Tukey, R = R1 R2, n = R2 n1 + n2,
// k = k1 + R1 k2) with compile-time twiddles exp(SIGN 2 pi i n2 k1 / R).
template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft_composite(cpx<T>* v) {
  constexpr int R1 = small_factor(R), R2 = R / R1;
  cpx<T> y[R2][R1];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    cpx<T> a[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) a[n1] = v[R2 * n1 + n2];
    dft_reg<R1, SIGN>(a);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      if (n2 * k1 == 0) {
        y[n2][k1] = a[k1];
      } else {
        const T c = T(unit_root(n2 * k1, R, 0)), sn = T(SIGN * unit_root(n2 * k1, R, 1));
        y[n2][k1] = {a[k1].x * c - a[k1].y * sn, a[k1].x * sn + a[k1].y * c};
      }
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    cpx<T> b[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) b[n2] = y[n2][k1];
    dft_reg<R2, SIGN>(b);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = b[k2];
  }
}

// In-register R-point DFT with exponent sign SIGN (forward = -1).
template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft_reg(cpx<T>* v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    // c = cos(2pi/3) = -1/2, s = SIGN*sin(2pi/3)
    const T half = T(0.5);
    const T s3 = T(0.86602540378443864676372317075294);
    cpx<T> a = v[0], b = v[1], c = v[2];
    cpx<T> sum = cadd(b, c);
    cpx<T> dif = csub(b, c);
    cpx<T> m = {a.x - half * sum.x, a.y - half * sum.y};
    // SIGN * i * s3 * dif
    cpx<T> r = cmul_i<SIGN>(cscale(dif, s3));
    v[0] = cadd(a, sum);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  } else if constexpr (R == 4) {
    cpx<T> a0 = cadd(v[0], v[2]);
    cpx<T> a1 = csub(v[0], v[2]);
    cpx<T> b0 = cadd(v[1], v[3]);
    cpx<T> b1 = cmul_i<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[3] = csub(a1, b1);
  } else if constexpr (R == 8) {
    const T h = T(0.70710678118654752440084436210485);
    // stage 1: radix-2 across stride 4
    cpx<T> a0 = cadd(v[0], v[4]), a4 = csub(v[0], v[4]);
    cpx<T> a1 = cadd(v[1], v[5]), a5 = csub(v[1], v[5]);
    cpx<T> a2 = cadd(v[2], v[6]), a6 = csub(v[2], v[6]);
    cpx<T> a3 = cadd(v[3], v[7]), a7 = csub(v[3], v[7]);
    // twiddles w8^k for the odd half: w8 = exp(SIGN i pi/4)
    a5 = {h * (a5.x - SIGN * a5.y), h * (a5.y + SIGN * a5.x)};
    a6 = cmul_i<SIGN>(a6);
    a7 = {h * (-a7.x - SIGN * a7.y), h * (-a7.y + SIGN * a7.x)};
    // two radix-4 transforms on (a0,a1,a2,a3) and (a4,a5,a6,a7)
    cpx<T> e0 = cadd(a0, a2), e1 = csub(a0, a2);
    cpx<T> e2 = cadd(a1, a3), e3 = cmul_i<SIGN>(csub(a1, a3));
    cpx<T> o0 = cadd(a4, a6), o1 = csub(a4, a6);
    cpx<T> o2 = cadd(a5, a7), o3 = cmul_i<SIGN>(csub(a5, a7));
    v[0] = cadd(e0, e2);
    v[4] = csub(e0, e2);
    v[2] = cadd(e1, e3);
    v[6] = csub(e1, e3);
    v[1] = cadd(o0, o2);
    v[5] = csub(o0, o2);
    v[3] = cadd(o1, o3);
    v[7] = csub(o1, o3);
  } else {
    dft_composite<R, SIGN>(v);
  }
}

// Optional global-memory ends of a transform: the first stage can read its
// inputs through a loader `T ld(seq, e)` (element e of sequence seq) instead
// of LDS, and the last stage can hand its outputs to a storer
// `st(seq, e, v)` instead of LDS.  That removes the LDS staging pass (and its
// zero fill) before the transform and the read-back pass after it.  Each
// functor declares SEQ_FAST: whether consecutive sequences are contiguous in
// its global layout, which picks the lane mapping of that stage.
struct NoIO {
  static constexpr bool SEQ_FAST = false;
};

// A functor may also gather or scatter all R elements of one butterfly at
// once -- elements j, j + STEP, ..., j + (R - 1) STEP of sequence seq -- by
// declaring BATCH = true and providing
//   template <int R, int STEP> void load_batch(int seq, int j, cpx<T>* v) const;
//   template <int R, int STEP> void store_batch(int seq, int j, const cpx<T>* v);
// so that index wrapping and row addressing happen once per butterfly instead
// of once per element.  The values are the same as R calls of operator().
template <class F, class = void>
struct io_batch : std::false_type {};
template <class F>
struct io_batch<F, std::void_t<decltype(F::BATCH)>> : std::integral_constant<bool, F::BATCH> {};

// Compile-time shape of stage S: radix, butterflies per sequence and per
// thread, twiddle step, which ends touch global memory and the lane mapping.
template <int L, int NSEQ, int NT, int S, int NS, int PLAN, class LD, class ST>
struct StageShape {
  static constexpr int R = radix_plan<PLAN>(L, S);
  static constexpr int NB = R > 0 ? L / R : 1;    // butterflies per sequence
  static constexpr int TOTAL = NSEQ * NB;         // butterflies in the workgroup
  static constexpr int PER = (TOTAL + NT - 1) / NT;
  static constexpr int TWS = R > 0 ? L / (NS * R) : 1;  // twiddle index step
  static constexpr bool GLOAD = S == 0 && !std::is_same<LD, NoIO>::value;
  static constexpr bool GSTORE = R > 0 && radix_plan<PLAN>(L, S + 1) == 0 && !std::is_same<ST, NoIO>::value;
  // Butterfly b -> (sequence, index).  With 8+ sequences in LDS the sequence
  // runs fastest across lanes: neighbouring lanes then touch the same element
  // of sequences an odd stride apart (no LDS bank conflicts) and read the
  // same twiddle (broadcast).  A global end follows its functor's layout.
  static constexpr bool SEQ_FAST = GLOAD ? LD::SEQ_FAST : (GSTORE ? ST::SEQ_FAST : NSEQ >= 8);
  static constexpr int NTW = (R > 1 && NS > 1) ? R - 1 : 0;  // twiddles per butterfly
  static constexpr int NTW1 = NTW > 0 ? NTW : 1;
};

// Twiddles of stage SH for every butterfly this thread owns (unconjugated).
template <class SH, typename T, int NT, int NSEQ, int NS>
__device__ __forceinline__ void stage_twiddle_regs(const cpx<T>* tw, cpx<T> (&w)[SH::PER][SH::NTW1]) {
  if constexpr (SH::NTW > 0) {
#pragma unroll
    for (int p = 0; p < SH::PER; ++p) {
      const int b = threadIdx.x + p * NT;
      if (SH::TOTAL % NT == 0 || b < SH::TOTAL) {
        const int j = SH::SEQ_FAST ? b / NSEQ : b % SH::NB;
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < SH::R; ++r) w[p][r - 1] = tw[r * k * SH::TWS];
      }
    }
  }
}

// One Stockham stage (radix R, NS = product of earlier radices), then the
// remaining stages by compile-time recursion.  PREFETCH: the twiddles of stage
// S arrive in `wcur` (loaded during stage S - 1), and this stage issues the
// global loads of stage S + 1's twiddles before its own work, so they are in
// flight across its LDS barriers instead of exposed between a stage's LDS
// reads and its twiddle multiply.  Only for twiddle tables in global memory
// (an LDS table may not be staged yet when the next stage's loads issue).
template <typename T, int L, int NSEQ, int SSTR, int NT, int SIGN, int S, int NS, int PLAN, bool PREFETCH, bool FINAL,
          bool PAD, class LD, class ST, class WCUR>
__device__ __forceinline__ void fft_stages(cpx<T>* buf, const cpx<T>* tw, const LD& ld, ST& st, const WCUR& wcur) {
  using SH = StageShape<L, NSEQ, NT, S, NS, PLAN, LD, ST>;
  constexpr int R = SH::R;
  if constexpr (R > 0) {
    constexpr int NB = SH::NB, TOTAL = SH::TOTAL, PER = SH::PER, TWS = SH::TWS;
    constexpr bool GLOAD = SH::GLOAD, GSTORE = SH::GSTORE, SEQ_FAST = SH::SEQ_FAST;
    using SN = StageShape<L, NSEQ, NT, S + 1, NS * R, PLAN, LD, ST>;
    cpx<T> v[PER][R];
    cpx<T> wnext[SN::PER][SN::NTW1];
    const int tid = threadIdx.x;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int b = tid + p * NT;
      if (TOTAL % NT == 0 || b < TOTAL) {
        const int seq = SEQ_FAST ? b % NSEQ : b / NB;
        const int j = SEQ_FAST ? b / NSEQ : b - seq * NB;
        if constexpr (GLOAD && io_batch<LD>::value) {
          ld.template load_batch<R, NB>(seq, j, v[p]);
        } else if constexpr (GLOAD) {
#pragma unroll
          for (int r = 0; r < R; ++r) v[p][r] = ld(seq, j + r * NB);
        } else {
          const cpx<T>* src = buf + seq * SSTR;
#pragma unroll
          for (int r = 0; r < R; ++r) v[p][r] = src[elem_at<PAD>(j + r * NB)];
        }
      }
    }
    if constexpr (PREFETCH && SN::R > 0) stage_twiddle_regs<SN, T, NT, NSEQ, NS * R>(tw, wnext);
    if constexpr (NS > 1) {
#pragma unroll
      for (int p = 0; p < PER; ++p) {
        const int b = tid + p * NT;
        if (TOTAL % NT == 0 || b < TOTAL) {
          const int j = SEQ_FAST ? b / NSEQ : b % NB;
          const int k = j % NS;
#pragma unroll
          for (int r = 1; r < R; ++r) {
            cpx<T> w;
            if constexpr (PREFETCH) w = wcur[p][r - 1];
            else w = tw[r * k * TWS];
            if constexpr (SIGN > 0) w.y = -w.y;
            v[p][r] = cmul(v[p][r], w);
          }
        }
      }
    }
    // in place: every read before any write (a last stage that stores to global
    // memory writes no LDS; FINAL says the caller does not reuse the buffer)
    if constexpr (!GLOAD && !(FINAL && GSTORE)) __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int b = tid + p * NT;
      if (TOTAL % NT == 0 || b < TOTAL) {
        const int seq = SEQ_FAST ? b % NSEQ : b / NB;
        const int j = SEQ_FAST ? b / NSEQ : b - seq * NB;
        const int k = j % NS;
        dft_reg<R, SIGN>(v[p]);
        const int base = (j / NS) * NS * R + k;
        if constexpr (GSTORE && io_batch<ST>::value) {
          st.template store_batch<R, NS>(seq, base, v[p]);
        } else if constexpr (GSTORE) {
#pragma unroll
          for (int r = 0; r < R; ++r) st(seq, base + r * NS, v[p][r]);
        } else {
          cpx<T>* dst = buf + seq * SSTR;
#pragma unroll
          for (int r = 0; r < R; ++r) dst[elem_at<PAD>(base + r * NS)] = v[p][r];
        }
      }
    }
    if constexpr (!GSTORE) __syncthreads();
    fft_stages<T, L, NSEQ, SSTR, NT, SIGN, S + 1, NS * R, PLAN, PREFETCH, FINAL, PAD>(buf, tw, ld, st, wnext);
  }
}

// Transform NSEQ sequences in LDS, laid out as lds_at<L>(s, e) with
// SSTR = seq_stride<L>().  `tw` should point to an LDS copy of the twiddle
// table (stage_twiddles).  Callers must __syncthreads() before the call (data
// and twiddles staged) and may read results right after it returns.
template <typename T, int L, int NT>
__device__ __forceinline__ void stage_twiddles(cpx<T>* dst, const cpx<T>* __restrict__ src) {
  for (int i = threadIdx.x; i < L; i += NT) dst[i] = src[i];
}

template <typename T, int L, int NSEQ, int SSTR, int NT, int SIGN, int PLAN = 0>
__device__ __forceinline__ void fft_lds(cpx<T>* buf, const cpx<T>* tw) {
  static_assert(fft_length_ok(L), "FFT length must be 2^a 3^b >= 8");
  NoIO none;
  const cpx<T> w0[1][1] = {};
  fft_stages<T, L, NSEQ, SSTR, NT, SIGN, 0, 1, PLAN, false, false, true>(buf, tw, none, none, w0);
}

// Same transform with a global loader for the first stage and/or a storer
// for the last (NoIO for an LDS end).  With a loader the caller stages
// nothing (the first stage fills the LDS buffer); with a storer the results
// are not left in LDS and the call ends without a barrier.  PREFETCH (global
// twiddle tables only): each stage's twiddles load one stage ahead.  FINAL
// (with a storer): the caller never touches `buf` again, so the last stage
// skips its read-before-write barrier.
template <typename T, int L, int NSEQ, int SSTR, int NT, int SIGN, int PLAN, bool PREFETCH, bool FINAL = false,
          bool PAD = true, class LD, class ST>
__device__ __forceinline__ void fft_io_pf(cpx<T>* buf, const cpx<T>* tw, const LD& ld, ST& st) {
  static_assert(fft_length_ok(L), "FFT length must be 2^a 3^b >= 8");
  const cpx<T> w0[1][1] = {};  // stage 0 has no twiddles (NS = 1)
  fft_stages<T, L, NSEQ, SSTR, NT, SIGN, 0, 1, PLAN, PREFETCH, FINAL, PAD>(buf, tw, ld, st, w0);
}
template <typename T, int L, int NSEQ, int SSTR, int NT, int SIGN, int PLAN, class LD, class ST>
__device__ __forceinline__ void fft_io(cpx<T>* buf, const cpx<T>* tw, const LD& ld, ST& st) {
  fft_io_pf<T, L, NSEQ, SSTR, NT, SIGN, PLAN, false>(buf, tw, ld, st);
}

}  // namespace glw
