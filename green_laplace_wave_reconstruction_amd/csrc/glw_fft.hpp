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
template <int L>
__host__ __device__ constexpr int seq_stride() {
  return (L + (L >> 4)) | 1;
}
template <int L>
__device__ __forceinline__ int lds_at(int s, int e) {
  return s * seq_stride<L>() + lds_pad(e);
}

// Greedy radix decomposition: 8s, then a 4, then a 2, then 3s.
__host__ __device__ constexpr int radix_of(int L, int s) {
  int rem = L;
  for (int i = 0; i <= s; ++i) {
    int r = 0;
    if (rem <= 1) return 0;
    if (rem % 8 == 0) r = 8;
    else if (rem % 4 == 0) r = 4;
    else if (rem % 2 == 0) r = 2;
    else if (rem % 3 == 0) r = 3;
    else return -1;  // unsupported factor
    if (i == s) return r;
    rem /= r;
  }
  return 0;
}

__host__ __device__ constexpr bool fft_length_ok(int L) {
  int rem = L;
  while (rem % 2 == 0) rem /= 2;
  while (rem % 3 == 0) rem /= 3;
  return rem == 1 && L >= 8;
}

// In-register R-point DFT with exponent sign SIGN (forward = -1).
template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft_reg(cpx<T>* v) {
  if constexpr (R == 2) {
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
  }
}

// One Stockham stage (radix R, NS = product of earlier radices), then the
// remaining stages by compile-time recursion.
template <typename T, int L, int NSEQ, int SSTR, int NT, int SIGN, int S, int NS>
__device__ __forceinline__ void fft_stages(cpx<T>* buf, const cpx<T>* __restrict__ tw) {
  constexpr int R = radix_of(L, S);
  if constexpr (R > 0) {
    constexpr int NB = L / R;           // butterflies per sequence
    constexpr int TOTAL = NSEQ * NB;    // butterflies in the workgroup
    constexpr int PER = (TOTAL + NT - 1) / NT;
    constexpr int TWS = L / (NS * R);   // twiddle index step
    cpx<T> v[PER][R];
    const int tid = threadIdx.x;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int b = tid + p * NT;
      if (TOTAL % NT == 0 || b < TOTAL) {
        const int seq = b / NB;
        const int j = b - seq * NB;
        const cpx<T>* src = buf + seq * SSTR;
#pragma unroll
        for (int r = 0; r < R; ++r) v[p][r] = src[lds_pad(j + r * NB)];
        if constexpr (NS > 1) {
          const int k = j % NS;
#pragma unroll
          for (int r = 1; r < R; ++r) {
            cpx<T> w = tw[r * k * TWS];
            if constexpr (SIGN > 0) w.y = -w.y;
            v[p][r] = cmul(v[p][r], w);
          }
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int b = tid + p * NT;
      if (TOTAL % NT == 0 || b < TOTAL) {
        const int seq = b / NB;
        const int j = b - seq * NB;
        const int k = j % NS;
        dft_reg<R, SIGN>(v[p]);
        cpx<T>* dst = buf + seq * SSTR;
        const int base = (j / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[lds_pad(base + r * NS)] = v[p][r];
      }
    }
    __syncthreads();
    fft_stages<T, L, NSEQ, SSTR, NT, SIGN, S + 1, NS * R>(buf, tw);
  }
}

// Transform NSEQ sequences in LDS, laid out as lds_at<L>(s, e) with
// SSTR = seq_stride<L>().  Callers must __syncthreads() before the call (data
// staged) and may read results right after it returns.
template <typename T, int L, int NSEQ, int SSTR, int NT, int SIGN>
__device__ __forceinline__ void fft_lds(cpx<T>* buf, const cpx<T>* __restrict__ tw) {
  static_assert(fft_length_ok(L), "FFT length must be 2^a 3^b >= 8");
  fft_stages<T, L, NSEQ, SSTR, NT, SIGN, 0, 1>(buf, tw);
}

}  // namespace glw
