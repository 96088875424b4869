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
// conj(d) * g: one term of a holomorphic Jacobian-transpose product
template <typename T>
__device__ __forceinline__ cpx<T> cjm(cpx<T> d, cpx<T> g) { return {d.x * g.x + d.y * g.y, d.x * g.y - d.y * g.x}; }

// Adjoints: for outputs o_s = Op_s(f) (holomorphic polynomials) and output
// gradients g_s (d/dRe + i d/dIm of a real loss), gf_i = sum_s conj(do_s/df_i) g_s.
// NADJ is the number of outputs whose gradients the adjoint takes.

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
  static constexpr int NADJ = 4;
  template <typename T>
  __device__ __forceinline__ static void adjoint(const cpx<T>* f, const cpx<T>* g, cpx<T>* gf) {
    const cpx<T> v = f[0], vn = f[1], vn2 = f[2], vn3 = f[3], hx = f[4], hy = f[5];
    const cpx<T> jx = f[6], jy = f[7], rad = f[8], q2f = f[9], njx = f[10], njy = f[11];
    const cpx<T> G0 = scl(g[0], T(2)), G1 = scl(g[1], T(2)), G2 = scl(g[2], T(2)), G3 = scl(g[3], T(2));
    const cpx<T> G12 = sub(G1, G2);  // hx, hy, jx, jy enter o1 and o2 with opposite signs
    gf[0] = add(add(cjm(vn2, G0), cjm(rad, G1)), add(cjm(vn3, G2), cjm(q2f, G3)));
    gf[1] = add(add(cjm(vn, G0), cjm(scl(vn2, T(2)), G2)), cjm(rad, G3));
    gf[2] = add(cjm(v, G0), cjm(scl(vn, T(2)), G2));
    gf[3] = cjm(v, G2);
    gf[4] = add(sub(cjm(jx, G12), cjm(hx, G0)), cjm(njx, G3));
    gf[5] = add(sub(cjm(jy, G12), cjm(hy, G0)), cjm(njy, G3));
    gf[6] = add(cjm(hx, G12), cjm(scl(jx, T(2)), G3));
    gf[7] = add(cjm(hy, G12), cjm(scl(jy, T(2)), G3));
    gf[8] = add(cjm(v, G1), cjm(vn, G3));
    gf[9] = cjm(v, G3);
    gf[10] = cjm(hx, G3);
    gf[11] = cjm(hy, G3);
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
  // Gradients of forcing_k (g[0]) and forcing_d (g[1]) only: eta33 is built
  // from those two; the contacts feed psi33 alone.
  static constexpr int NADJ = 2;
  template <typename T>
  __device__ __forceinline__ static void adjoint(const cpx<T>* f, const cpx<T>* g, cpx<T>* gf) {
    const cpx<T> fe = f[0], fex = f[1], fey = f[2], fpx = f[3], fpy = f[4], fpz = f[5];
    const cpx<T> fpxz = f[6], fpyz = f[7], fpzz = f[8], fpzzz = f[9], fptz = f[10], fptzz = f[11];
    const cpx<T> pe = f[12], pex = f[13], pey = f[14], ppx = f[15], ppy = f[16], ppz = f[17];
    const cpx<T> ppzz = f[18], pptz = f[19];
    const cpx<T> gk = g[0], gd = g[1];
    const cpx<T> hk = scl(gk, T(0.5)), hd = scl(gd, T(0.5));      // pair terms carry 1/2
    const cpx<T> qk = scl(gk, T(0.25)), qd = scl(gd, T(0.25));    // direct terms carry 1/4
    const cpx<T> fe2 = mul(fe, fe);
    const cpx<T> dfe_k = sub(add(mul(fpxz, fex), mul(fpyz, fey)), mul(fe, fpzzz));
    const cpx<T> dfe_d = add(mul(fe, fptzz), add(add(mul(fpx, fpxz), mul(fpy, fpyz)), mul(fpz, fpzz)));
    gf[0] = add(add(cjm(dfe_k, qk), cjm(dfe_d, qd)), sub(cjm(pptz, hd), cjm(ppzz, hk)));
    gf[1] = add(cjm(mul(fe, fpxz), qk), cjm(ppx, hk));
    gf[2] = add(cjm(mul(fe, fpyz), qk), cjm(ppy, hk));
    gf[3] = add(cjm(pex, hk), add(cjm(mul(fe, fpxz), qd), cjm(ppx, hd)));
    gf[4] = add(cjm(pey, hk), add(cjm(mul(fe, fpyz), qd), cjm(ppy, hd)));
    gf[5] = add(cjm(mul(fe, fpzz), qd), cjm(ppz, hd));
    gf[6] = add(cjm(mul(fe, fex), qk), cjm(mul(fe, fpx), qd));
    gf[7] = add(cjm(mul(fe, fey), qk), cjm(mul(fe, fpy), qd));
    gf[8] = sub(cjm(mul(fe, fpz), qd), cjm(pe, hk));
    gf[9] = scl(cjm(fe2, gk), T(-0.125));
    gf[10] = cjm(pe, hd);
    gf[11] = scl(cjm(fe2, gd), T(0.125));
    gf[12] = sub(cjm(fptz, hd), cjm(fpzz, hk));
    gf[13] = cjm(fpx, hk);
    gf[14] = cjm(fpy, hk);
    gf[15] = add(cjm(fex, hk), cjm(fpx, hd));
    gf[16] = add(cjm(fey, hk), cjm(fpy, hd));
    gf[17] = cjm(fpz, hd);
    gf[18] = scl(cjm(fe, gk), T(-0.5));
    gf[19] = cjm(fe, hd);
    gf[20] = {T(0), T(0)};
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
  static constexpr int NADJ = 2;
  template <typename T>
  __device__ __forceinline__ static void adjoint(const cpx<T>* f, const cpx<T>* g, cpx<T>* gf) {
    const cpx<T> v = f[0], vn = f[1], vn2 = f[2], hx = f[3], hy = f[4], rad = f[5], jx = f[6], jy = f[7];
    const cpx<T> G0 = scl(g[0], T(2)), G1 = scl(g[1], T(2));
    gf[0] = add(cjm(vn2, G0), cjm(rad, G1));
    gf[1] = cjm(vn, G0);
    gf[2] = cjm(v, G0);
    gf[3] = sub(cjm(jx, G1), cjm(hx, G0));
    gf[4] = sub(cjm(jy, G1), cjm(hy, G0));
    gf[5] = cjm(v, G1);
    gf[6] = cjm(hx, G1);
    gf[7] = cjm(hy, G1);
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

namespace glw {

// ---------------------------------------------------------------------------
// Order-4 graph (paper/order4/cpp/gl_eta44_fftw.cpp).  Products are evaluated
// in the reference's association order where it is explicit.
// ---------------------------------------------------------------------------

// Pair state sources with two time derivatives (gl_eta44_fftw.cpp:394-403).
// in: 0 v 1 vn 2 vn2 3 vn3 4 vn4 5 hx 6 hy 7 jx 8 jy 9 ro 10 q2 11 q2n
//     12 njx 13 njy 14 n2jx 15 n2jy
// out: sd, sk, sds, sks, sdss, skss
struct Pair44Op {
  template <typename T>
  __device__ __forceinline__ static void apply(const cpx<T>* f, cpx<T>* o) {
    const cpx<T> v = f[0], vn = f[1], vn2 = f[2], vn3 = f[3], vn4 = f[4], hx = f[5], hy = f[6];
    const cpx<T> jx = f[7], jy = f[8], ro = f[9], q2 = f[10], q2n = f[11], njx = f[12], njy = f[13];
    const cpx<T> n2jx = f[14], n2jy = f[15];
    const cpx<T> hxjx = mul(hx, jx), hyjy = mul(hy, jy);
    const cpx<T> grad = add(add(mul(jx, jx), mul(hx, njx)), add(mul(jy, jy), mul(hy, njy)));
    o[0] = sub(sub(add(scl(mul(v, vn2), T(2)), mul(vn, vn)), mul(hx, hx)), mul(hy, hy));
    o[1] = add(scl(mul(v, ro), T(2)), scl(add(hxjx, hyjy), T(2)));
    o[2] = sub(sub(add(scl(mul(v, vn3), T(2)), scl(mul(vn, vn2), T(4))), scl(hxjx, T(2))), scl(hyjy, T(2)));
    o[3] = add(scl(add(mul(vn, ro), mul(v, q2)), T(2)), scl(grad, T(2)));
    o[4] = sub(add(add(scl(mul(v, vn4), T(2)), scl(mul(vn, vn3), T(6))), scl(mul(vn2, vn2), T(4))), scl(grad, T(2)));
    o[5] = add(add(scl(add(add(mul(vn2, ro), scl(mul(vn, q2), T(2))), mul(v, q2n)), T(2)),
                   scl(add(mul(jx, njx), mul(jy, njy)), T(6))),
               scl(add(mul(hx, n2jx), mul(hy, n2jy)), T(2)));
  }
};

// Third-order forcing with time derivatives (gl_eta44_fftw.cpp:569-593).
// in: first 0 eta 1 ex 2 ey 3 px 4 py 5 pz 6 pxz 7 pyz 8 pzz 9 pzzz 10 ptz 11 ptzz
//           12 pxz_t 13 pyz_t 14 pzzz_t 15 ptz_t
//     pair  16 eta 17 ex 18 ey 19 eta_t 20 ex_t 21 ey_t 22 px 23 py 24 pz 25 pzz
//           26 ptz 27 px_t 28 py_t 29 pzz_t 30 ptz_t
// out: fk, fd, fkt, fdt
struct Third44Op {
  static constexpr int kRowLdsKb = 96;  // two 64-point rows of 31 fields per fused row tile
  template <typename T>
  __device__ __forceinline__ static void apply(const cpx<T>* f, cpx<T>* o) {
    const cpx<T> e = f[0], ex = f[1], ey = f[2], px = f[3], py = f[4], pz = f[5], pxz = f[6], pyz = f[7];
    const cpx<T> pzz = f[8], pzzz = f[9], ptz = f[10], ptzz = f[11], pxzt = f[12], pyzt = f[13], pzzzt = f[14];
    const cpx<T> ptzt = f[15];
    const cpx<T> Pe = f[16], Pex = f[17], Pey = f[18], Pet = f[19], Pext = f[20], Peyt = f[21], Ppx = f[22];
    const cpx<T> Ppy = f[23], Ppz = f[24], Ppzz = f[25], Pptz = f[26], Ppxt = f[27], Ppyt = f[28], Ppzzt = f[29];
    const cpx<T> Pptzt = f[30];
    const cpx<T> e2 = mul(e, e);
    const cpx<T> kp = sub(sub(add(add(add(mul(px, Pex), mul(py, Pey)), mul(Ppx, ex)), mul(Ppy, ey)), mul(e, Ppzz)),
                          mul(Pe, pzz));
    const cpx<T> kd = sub(mul(e, add(mul(pxz, ex), mul(pyz, ey))), scl(mul(e2, pzzz), T(0.5)));
    const cpx<T> dp = add(add(add(add(mul(e, Pptz), mul(Pe, ptz)), mul(px, Ppx)), mul(py, Ppy)), mul(pz, Ppz));
    const cpx<T> dd = add(scl(mul(e2, ptzz), T(0.5)), mul(e, add(add(mul(px, pxz), mul(py, pyz)), mul(pz, pzz))));
    cpx<T> kpt = sub(mul(px, Pext), mul(ex, Pex));
    kpt = add(kpt, sub(mul(py, Peyt), mul(ey, Pey)));
    kpt = add(add(kpt, mul(Ppxt, ex)), mul(Ppx, pxz));
    kpt = add(add(kpt, mul(Ppyt, ey)), mul(Ppy, pyz));
    kpt = sub(sub(kpt, mul(pz, Ppzz)), mul(e, Ppzzt));
    kpt = sub(sub(kpt, mul(Pet, pzz)), mul(Pe, ptzz));
    const cpx<T> kdt = sub(sub(add(mul(pz, add(mul(pxz, ex), mul(pyz, ey))),
                                   mul(e, add(add(add(mul(pxzt, ex), mul(pxz, pxz)), mul(pyzt, ey)), mul(pyz, pyz)))),
                               mul(mul(e, pz), pzzz)),
                           scl(mul(e2, pzzzt), T(0.5)));
    cpx<T> dpt = add(add(add(mul(pz, Pptz), mul(e, Pptzt)), mul(Pet, ptz)), mul(Pe, ptzt));
    dpt = add(sub(dpt, mul(ex, Ppx)), mul(px, Ppxt));
    dpt = add(sub(dpt, mul(ey, Ppy)), mul(py, Ppyt));
    dpt = add(add(dpt, mul(ptz, Ppz)), mul(pz, Pptz));
    cpx<T> inner = add(sub(mul(px, pxzt), mul(ex, pxz)), sub(mul(py, pyzt), mul(ey, pyz)));
    inner = add(add(inner, mul(ptz, pzz)), mul(pz, ptzz));
    const cpx<T> ddt = add(add(sub(mul(mul(e, pz), ptzz), scl(mul(e2, pzzz), T(0.5))),
                               mul(pz, add(add(mul(px, pxz), mul(py, pyz)), mul(pz, pzz)))),
                           mul(e, inner));
    o[0] = add(scl(kd, T(0.25)), scl(kp, T(0.5)));
    o[1] = add(scl(dd, T(0.25)), scl(dp, T(0.5)));
    o[2] = add(scl(kdt, T(0.25)), scl(kpt, T(0.5)));
    o[3] = add(scl(ddt, T(0.25)), scl(dpt, T(0.5)));
  }
};

// Order-4 forcing partitions 3+1, 2+2, 2+1+1, 1+1+1+1 (gl_eta44_fftw.cpp:711-735).
// in: first 0..15 as Third44Op, 16 pxzz 17 pyzz 18 pzzzz
//     pair  19 eta 20 ex 21 ey 22 eta_t 23 ex_t 24 ey_t 25 px 26 py 27 pz 28 pzz
//           29 ptz 30 px_t 31 py_t 32 pzz_t 33 ptz_t 34 pxz 35 pyz 36 pzzz
//     third 37 eta 38 ex 39 ey 40 px 41 py 42 pz 43 pzz 44 pzt
// out: fk, fd
struct Quartic44Op {
  template <typename T>
  __device__ __forceinline__ static void apply(const cpx<T>* f, cpx<T>* o) {
    const cpx<T> e = f[0], ex = f[1], ey = f[2], px = f[3], py = f[4], pz = f[5], pxz = f[6], pyz = f[7];
    const cpx<T> pzz = f[8], pzzz = f[9], ptz = f[10], ptzz = f[11], pzzzt = f[14];
    const cpx<T> pxzz = f[16], pyzz = f[17], pzzzz = f[18];
    const cpx<T> Pe = f[19], Pex = f[20], Pey = f[21], Ppx = f[25], Ppy = f[26], Ppz = f[27], Ppzz = f[28];
    const cpx<T> Pptz = f[29], Ppzzt = f[32], Ppxz = f[34], Ppyz = f[35], Ppzzz = f[36];
    const cpx<T> Te = f[37], Tex = f[38], Tey = f[39], Tpx = f[40], Tpy = f[41], Tpz = f[42], Tpzz = f[43];
    const cpx<T> Tpzt = f[44];
    const cpx<T> e2 = mul(e, e), e3 = mul(e2, e);
    const cpx<T> k31 = sub(sub(add(add(add(mul(px, Tex), mul(py, Tey)), mul(Tpx, ex)), mul(Tpy, ey)), mul(e, Tpzz)),
                           mul(Te, pzz));
    const cpx<T> d31 = add(add(add(add(mul(e, Tpzt), mul(Te, ptz)), mul(px, Tpx)), mul(py, Tpy)), mul(pz, Tpz));
    const cpx<T> k22 = sub(add(mul(Ppx, Pex), mul(Ppy, Pey)), mul(Pe, Ppzz));
    const cpx<T> d22 = add(mul(Pe, Pptz), scl(add(add(mul(Ppx, Ppx), mul(Ppy, Ppy)), mul(Ppz, Ppz)), T(0.5)));
    const cpx<T> ePxz = add(mul(e, Ppxz), mul(Pe, pxz));
    const cpx<T> ePyz = add(mul(e, Ppyz), mul(Pe, pyz));
    cpx<T> k211 = mul(e, add(mul(pxz, Pex), mul(pyz, Pey)));
    k211 = add(add(k211, mul(ePxz, ex)), mul(ePyz, ey));
    k211 = sub(sub(k211, scl(mul(e2, Ppzzz), T(0.5))), mul(mul(e, Pe), pzzz));
    cpx<T> d211 = add(scl(mul(e2, Ppzzt), T(0.5)), mul(mul(e, Pe), ptzz));
    d211 = add(add(add(d211, mul(px, ePxz)), mul(py, ePyz)), mul(pz, add(mul(e, Ppzz), mul(Pe, pzz))));
    d211 = add(add(add(d211, mul(Ppx, mul(e, pxz))), mul(Ppy, mul(e, pyz))), mul(Ppz, mul(e, pzz)));
    const cpx<T> k1111 = sub(scl(mul(e2, add(mul(pxzz, ex), mul(pyzz, ey))), T(0.5)),
                             scl(mul(e3, pzzzz), T(1.0 / 6.0)));
    const cpx<T> d1111 = add(add(scl(mul(e3, pzzzt), T(1.0 / 6.0)),
                                 scl(mul(e2, add(add(mul(px, pxzz), mul(py, pyzz)), mul(pz, pzzz))), T(0.5))),
                             scl(mul(e2, add(add(mul(pxz, pxz), mul(pyz, pyz)), mul(pzz, pzz))), T(0.5)));
    o[0] = add(scl(k31, T(0.5)), add(add(scl(k1111, T(0.125)), scl(k211, T(0.25))), scl(k22, T(0.5))));
    o[1] = add(scl(d31, T(0.5)), add(add(scl(d1111, T(0.125)), scl(d211, T(0.25))), scl(d22, T(0.5))));
  }
};

// The 3+1 partition of Quartic44Op alone (same inputs): the forcing of one
// third-order node in the reference's split runs
// (paper/order4/cpp/gl_eta44_fftw.cpp:704-717, PART_INDEX >= 0).
struct Quartic31Op {
  template <typename T>
  __device__ __forceinline__ static void apply(const cpx<T>* f, cpx<T>* o) {
    const cpx<T> e = f[0], ex = f[1], ey = f[2], px = f[3], py = f[4], pz = f[5];
    const cpx<T> pzz = f[8], ptz = f[10];
    const cpx<T> Te = f[37], Tex = f[38], Tey = f[39], Tpx = f[40], Tpy = f[41], Tpz = f[42], Tpzz = f[43];
    const cpx<T> Tpzt = f[44];
    const cpx<T> k31 = sub(sub(add(add(add(mul(px, Tex), mul(py, Tey)), mul(Tpx, ex)), mul(Tpy, ey)), mul(e, Tpzz)),
                           mul(Te, pzz));
    const cpx<T> d31 = add(add(add(add(mul(e, Tpzt), mul(Te, ptz)), mul(px, Tpx)), mul(py, Tpy)), mul(pz, Tpz));
    o[0] = scl(k31, T(0.5));
    o[1] = scl(d31, T(0.5));
  }
};

}  // namespace glw
