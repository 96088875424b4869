// Pointwise product operators of the GL graphs (physical space, one point).
#pragma once

#include "glw_fft.hpp"

namespace glw {

template <typename T>
__device__ __forceinline__ cpx<T> mul(cpx<T> a, cpx<T> b) { return cmul(a, b); }
template <typename T>
__device__ __forceinline__ cpx<T> add(cpx<T> a, cpx<T> b) { return cadd(a, b); }
template <typename T>
__device__ __forceinline__ cpx<T> sub(cpx<T> a, cpx<T> b) { return csub(a, b); }
template <typename T>
__device__ __forceinline__ cpx<T> scl(cpx<T> a, T s) { return cscale(a, s); }

// Order-3 inner pair sources (python/src/green_laplace/_order3.py:216-233).
// in: v, v_nu, v_nu2, v_nu3, hx, hy, jx, jy, radial, q2_field, nu_jx, nu_jy
// out: source_d, source_k, source_d_t, source_k_t
struct InnerOp {
  template <typename T>
  __device__ __forceinline__ static void apply(const cpx<T>* f, cpx<T>* o) {
    const cpx<T> v = f[0], vn = f[1], vn2 = f[2], vn3 = f[3], hx = f[4], hy = f[5];
    const cpx<T> jx = f[6], jy = f[7], rad = f[8], q2f = f[9], njx = f[10], njy = f[11];
    const cpx<T> hxjx = mul(hx, jx), hyjy = mul(hy, jy);
    o[0] = sub(sub(add(scl(mul(v, vn2), T(2)), mul(vn, vn)), mul(hx, hx)), mul(hy, hy));
    o[1] = add(scl(mul(v, rad), T(2)), scl(add(hxjx, hyjy), T(2)));
    o[2] = sub(sub(add(scl(mul(v, vn3), T(2)), scl(mul(vn, vn2), T(4))), scl(hxjx, T(2))), scl(hyjy, T(2)));
    o[3] = add(scl(add(mul(vn, rad), mul(v, q2f)), T(2)),
               scl(add(add(mul(jx, jx), mul(hx, njx)), add(mul(jy, jy), mul(hy, njy))), T(2)));
  }
};

// Order-3 forcing and surface Taylor contact (_order3.py:76-102, 275-287).
// in (first-order): 0 eta 1 etax 2 etay 3 phix 4 phiy 5 phiz 6 phixz 7 phiyz
//                   8 phizz 9 phizzz 10 phitz 11 phitzz
//    (pair):        12 eta 13 etax 14 etay 15 phix 16 phiy 17 phiz 18 phizz
//                   19 phitz 20 eta_t
// out: forcing_k, forcing_d, contact, contact_t
struct OuterOp {
  template <typename T>
  __device__ __forceinline__ static void apply(const cpx<T>* f, cpx<T>* o) {
    const cpx<T> fe = f[0], fex = f[1], fey = f[2], fpx = f[3], fpy = f[4], fpz = f[5];
    const cpx<T> fpxz = f[6], fpyz = f[7], fpzz = f[8], fpzzz = f[9], fptz = f[10], fptzz = f[11];
    const cpx<T> pe = f[12], pex = f[13], pey = f[14], ppx = f[15], ppy = f[16], ppz = f[17];
    const cpx<T> ppzz = f[18], pptz = f[19], pet = f[20];
    const cpx<T> fet = fpz, fpzt = fptz, fpzzt = fptzz, ppzt = pptz;
    const cpx<T> fe2 = mul(fe, fe);
    const cpx<T> k_pair = sub(sub(add(add(mul(fpx, pex), mul(fpy, pey)), add(mul(ppx, fex), mul(ppy, fey))),
                                  mul(fe, ppzz)),
                              mul(pe, fpzz));
    const cpx<T> k_direct = sub(mul(fe, add(mul(fpxz, fex), mul(fpyz, fey))), scl(mul(fe2, fpzzz), T(0.5)));
    const cpx<T> d_pair = add(add(add(mul(fe, pptz), mul(pe, fptz)), add(mul(fpx, ppx), mul(fpy, ppy))),
                              mul(fpz, ppz));
    const cpx<T> d_direct = add(scl(mul(fe2, fptzz), T(0.5)),
                                mul(fe, add(add(mul(fpx, fpxz), mul(fpy, fpyz)), mul(fpz, fpzz))));
    o[0] = add(scl(k_direct, T(0.25)), scl(k_pair, T(0.5)));
    o[1] = add(scl(d_direct, T(0.25)), scl(d_pair, T(0.5)));
    o[2] = add(add(scl(mul(fe, ppz), T(0.5)), scl(mul(pe, fpz), T(0.5))), scl(mul(fe2, fpzz), T(0.125)));
    o[3] = add(add(add(scl(add(mul(fet, ppz), mul(fe, ppzt)), T(0.5)), scl(add(mul(pet, fpz), mul(pe, fpzt)), T(0.5))),
                   scl(mul(mul(fe, fet), fpzz), T(0.25))),
               scl(mul(fe2, fpzzt), T(0.125)));
  }
};

// eta22 sources (spatial.py:196-205).
// in: v, v_nu, v_nu2, hx, hy, radial_over_nu, jx, jy ; out: source_d, source_k
struct Eta22Op {
  template <typename T>
  __device__ __forceinline__ static void apply(const cpx<T>* f, cpx<T>* o) {
    const cpx<T> v = f[0], vn = f[1], vn2 = f[2], hx = f[3], hy = f[4], rad = f[5], jx = f[6], jy = f[7];
    o[0] = sub(sub(add(scl(mul(v, vn2), T(2)), mul(vn, vn)), mul(hx, hx)), mul(hy, hy));
    o[1] = add(scl(mul(v, rad), T(2)), scl(add(mul(hx, jx), mul(hy, jy)), T(2)));
  }
};

// psi22 channel-pair sources (spatial.py:370-380), two radial channels.
// in: channel 0 fields 0..7 then channel 1 fields 8..15, each
//     v, v_nu, v_nu2, hx, hy, radial, jx, jy
// out: for pair p = 2*first + second: o[2p] = source_d, o[2p+1] = source_k
struct Psi22Op {
  template <typename T>
  __device__ __forceinline__ static void apply(const cpx<T>* f, cpx<T>* o) {
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const cpx<T>* L = f + 8 * a;
        const cpx<T>* R = f + 8 * b;
        const int p = 2 * a + b;
        o[2 * p] = sub(sub(add(scl(mul(L[0], R[2]), T(2)), mul(L[1], R[1])), mul(L[3], R[3])), mul(L[4], R[4]));
        o[2 * p + 1] = add(scl(mul(L[0], R[5]), T(2)), scl(add(mul(L[3], R[6]), mul(L[4], R[7])), T(2)));
      }
    }
  }
};

// psi22 surface Taylor contact 0.5 eta11 phi1z (spatial.py:303-306).
// in: eta11 (unit multiplier), phi1z (-i nu multiplier) ; out: contact
struct ContactOp {
  template <typename T>
  __device__ __forceinline__ static void apply(const cpx<T>* f, cpx<T>* o) {
    o[0] = scl(mul(f[0], f[1]), T(0.5));
  }
};

}  // namespace glw
