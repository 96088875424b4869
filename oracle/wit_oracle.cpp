// CPU restatement of the reference WIT sector-44 quartet sum.  TEST
// INFRASTRUCTURE ONLY: the checker and the CPU baseline, never the product.
//
// Follows paper/order4/cpp/directional_eta4_openmp.cpp:
//   partition tables            :60-89
//   q, omega, zd, solve, g1     :92-134
//   evaluate (15 subset states) :157-453
//   sector-44 loop + accumulate :595-602, 690-734
//   repeated-index factor       :486-499
//   signed sectors 40 and 42    :282-336 (retained-V3 push, y31 coordinate),
//                                :603-689 (candidate loops), :478-482
// Every subset state of a quartet is evaluated per quartet, in increasing
// degree, from ordered set partitions with labelled blocks (block 0 carries
// the potential, the others the elevations), exactly as the reference does.
// Parents carry signs (sector 44: all +1; 40: + + - -; 42: + + + -); a cubic
// subset of charge +-1 is split into its free-wave part (retained V3, pushed
// onto the quartic source through the quadratic normal map) and the rest,
// and the quartic state gains the y31 coordinate from the aligned and
// opposite parts of those cubic states.  With all signs +1 every cubic
// charge is 3 and both terms vanish.
#include <omp.h>

#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <limits>
#include <vector>

using cd = std::complex<double>;

namespace {

struct Part {
  int nb;
  int blk[4];
};

struct Tables {
  std::vector<Part> p[16][5];  // [mask][block count]
  Tables() {
    for (int mask = 1; mask < 16; ++mask) {
      int slots[4], ns = 0;
      for (int s = 0; s < 4; ++s)
        if (mask >> s & 1) slots[ns++] = s;
      for (int nb = 1; nb <= 4; ++nb) {
        int codes = 1;
        for (int i = 0; i < ns; ++i) codes *= nb;
        for (int code = 0; code < codes; ++code) {
          Part pt{nb, {0, 0, 0, 0}};
          int c = code, seen = 0;
          for (int i = 0; i < ns; ++i) {
            int b = c % nb;
            c /= nb;
            seen |= 1 << b;
            pt.blk[b] |= 1 << slots[i];
          }
          if (seen == (1 << nb) - 1) p[mask][nb].push_back(pt);
        }
      }
    }
  }
};

struct Model {
  double g, h;
  double q(double kx, double ky) const {
    double k = std::hypot(kx, ky);
    return k * std::tanh(h * k);
  }
  double om(double kx, double ky) const { return std::sqrt(g * q(kx, ky)); }
  // quadratic interaction coefficient of an elevation at (ax, ay) with a potential at (bx, by)
  double g1(double ax, double ay, double bx, double by) const {
    return -q(ax + bx, ay + by) * q(bx, by) + bx * bx + by * by + ax * bx + ay * by;
  }
  // linear response (eta, psi) of the surface to source (s0, s1) at wavenumber (kx, ky), frequency w
  void respond(cd s0, cd s1, double kx, double ky, double w, cd& eta, cd& psi) const {
    const double qk = q(kx, ky), den = w * w - g * qk;
    eta = (cd(0.0, w) * s0 - qk * s1) / den;
    psi = (g * s0 + cd(0.0, w) * s1) / den;
  }
  // second-order state of two free waves A (eta_a, psi_a) and B (eta_b, psi_b)
  void quad(double ax, double ay, double wa, cd ea, cd pa, double bx, double by, double wb, cd eb, cd pb, cd& eta,
            cd& psi) const {
    const cd s0 = g1(ax, ay, bx, by) * ea * pb + g1(bx, by, ax, ay) * eb * pa;
    const cd s1 = (ax * bx + ay * by + q(ax, ay) * q(bx, by)) * pa * pb;
    respond(s0, s1, ax + bx, ay + by, wa + wb, eta, psi);
  }
  double zd(double kx, double ky, int order) const {
    double k = std::hypot(kx, ky), k2 = k * k;
    switch (order) {
      case 0: return 1.0;
      case 1: return q(kx, ky);
      case 2: return k2;
      case 3: return k2 * q(kx, ky);
      default: return k2 * k2;
    }
  }
};

struct Result {
  cd eta, psi, retained, ceta, cpsi;
  double rho, residual, zero_src;
  uint64_t zero_count;
  bool near;
};

// kx4, ky4: the parents' wavenumbers; sg: their signs.  Signed wavenumbers
// and frequencies sg k, sg omega(k) build every subset's state.
Result evaluate(const Tables& T, const Model& M, const double kx4[4], const double ky4[4], const int sg[4],
                double thr) {
  static const double fact[5] = {1, 1, 2, 6, 24};
  double Kx[16] = {}, Ky[16] = {}, Om[16] = {}, kx[4], ky[4], om[4];
  int deg[16] = {}, chg[16] = {};
  for (int s = 0; s < 4; ++s) {
    kx[s] = sg[s] * kx4[s];
    ky[s] = sg[s] * ky4[s];
    om[s] = sg[s] * M.om(kx4[s], ky4[s]);
  }
  for (int m = 1; m < 16; ++m)
    for (int s = 0; s < 4; ++s)
      if (m >> s & 1) {
        Kx[m] += kx[s];
        Ky[m] += ky[s];
        Om[m] += om[s];
        deg[m] += 1;
        chg[m] += sg[s];
      }
  cd eta[16], phi[16], psi[16], sh0[16], sh1[16], sv[16], v3[16];
  Result r{};
  auto eprod = [&](const Part& p) {
    cd v = 1.0;
    for (int b = 1; b < p.nb; ++b) v *= eta[p.blk[b]];
    return v;
  };
  cd src0, src1;
  const cd I(0.0, 1.0);
  for (int d = 1; d <= 4; ++d) {
    for (int m = 1; m < 16; ++m) {
      if (deg[m] != d) continue;
      cd taylor = 0.0;
      if (d == 1) {
        eta[m] = 0.5;
        phi[m] = cd(0.0, -Om[m] / (2.0 * M.q(Kx[m], Ky[m])));
        psi[m] = phi[m];
      } else {
        cd horiz = 0.0, kin = 0.0, lv = 0.0, lt = 0.0;
        for (const Part& p : T.p[m][2]) {
          int L = p.blk[0], R = p.blk[1];
          horiz += (sh0[L] * cd(0.0, Kx[R]) + sh1[L] * cd(0.0, Ky[R])) * eta[R];
          kin += 0.5 * (sh0[L] * sh0[R] + sh1[L] * sh1[R] + sv[L] * sv[R]);
        }
        for (int o = 1; o < d; ++o)
          for (const Part& p : T.p[m][o + 1]) {
            int P = p.blk[0];
            cd c = phi[P] * eprod(p) / fact[o];
            taylor += M.zd(Kx[P], Ky[P], o) * c;
            lv += M.zd(Kx[P], Ky[P], o + 1) * c;
            lt += cd(0.0, -Om[P]) * M.zd(Kx[P], Ky[P], o) * c;
          }
        const double qK = M.q(Kx[m], Ky[m]);
        src0 = -(horiz - lv) - qK * taylor;
        src1 = -(lt + kin) - cd(0.0, Om[m]) * taylor;
        const double mag = std::hypot(Kx[m], Ky[m]);
        const bool zero = mag <= 100.0 * std::numeric_limits<double>::epsilon() * std::max(mag, 1.0);
        auto solve = [&]() {
          double den = Om[m] * Om[m] - M.g * qK;
          eta[m] = (cd(0.0, Om[m]) * src0 - qK * src1) / den;
          psi[m] = (M.g * src0 + cd(0.0, Om[m]) * src1) / den;
        };
        if (zero) {
          eta[m] = psi[m] = 0.0;
          r.zero_count += 1;
          r.zero_src = std::max(r.zero_src, std::hypot(std::abs(src0), std::abs(src1)));
        } else if (d == 2 || (d == 3 && std::abs(chg[m]) == 3)) {
          solve();
        } else if (d == 3) {
          // charge +-1: the part on the free-wave branch of sign chg is kept aside (retained V3)
          const double w = M.om(Kx[m], Ky[m]), ms = std::sqrt(w / (2.0 * M.g)), ns = std::sqrt(M.g / (2.0 * w));
          const int ts = chg[m] > 0 ? 1 : -1;
          v3[m] = src0 / (2.0 * ms) + cd(0.0, ts) * src1 / (2.0 * ns);
          const cd comp = src0 / (2.0 * ms) - cd(0.0, ts) * src1 / (2.0 * ns);
          const double den = Om[m] + ts * w;
          eta[m] = cd(0.0, ms) * comp / den;
          psi[m] = -double(ts) * ns * comp / den;
        } else {
          // push of the retained cubic parts through the quadratic normal map, and the y31 coordinate
          cd push0 = 0.0, push1 = 0.0;
          for (int s = 0; s < 4; ++s) {
            const int c = m & ~(1 << s);
            if (std::norm(v3[c]) != 0.0) {
              const double wa = M.om(Kx[c], Ky[c]), ma = std::sqrt(wa / (2.0 * M.g));
              const double na = std::sqrt(M.g / (2.0 * wa)), qb = M.q(kx[s], ky[s]);
              const int sa = chg[c] > 0 ? 1 : -1;
              cd pe, pp;
              M.quad(Kx[c], Ky[c], sa * wa, ma, cd(0.0, -sa * na), kx[s], ky[s], om[s], 0.5,
                     cd(0.0, -om[s] / (2.0 * qb)), pe, pp);
              push0 += v3[c] * pe;
              push1 += v3[c] * pp;
            }
            if (std::abs(chg[c]) == 1 && std::abs(eta[c]) + std::abs(psi[c]) != 0.0 &&
                std::hypot(Kx[c] + kx[s], Ky[c] + ky[s]) != 0.0) {
              const double wc = M.om(Kx[c], Ky[c]);
              const cd rot = cd(0.0, chg[c] * wc / M.g) * psi[c];
              const cd ea = 0.5 * (eta[c] + rot), eo = 0.5 * (eta[c] - rot);
              const cd pa = cd(0.0, -M.g * chg[c] / wc) * ea, po = cd(0.0, M.g * chg[c] / wc) * eo;
              const cd pb = cd(0.0, -om[s] / (2.0 * M.q(kx[s], ky[s])));
              cd e1, p1, e2, p2;
              M.quad(Kx[c], Ky[c], chg[c] * wc, ea, pa, kx[s], ky[s], om[s], 0.5, pb, e1, p1);
              M.quad(Kx[c], Ky[c], -chg[c] * wc, eo, po, kx[s], ky[s], om[s], 0.5, pb, e2, p2);
              r.ceta += e1 + e2;
              r.cpsi += p1 + p2;
            }
          }
          src0 = src0 - push0;
          src1 = src1 - push1;
          const double lin = M.g * qK, tmp = Om[m] * Om[m];
          r.rho = std::abs(lin - tmp) / std::max(lin + tmp, std::numeric_limits<double>::min());
          if (r.rho < thr) {
            const double w = std::sqrt(lin), ms = std::sqrt(w / (2.0 * M.g)), ns = std::sqrt(M.g / (2.0 * w));
            const int sg = Om[m] >= 0.0 ? 1 : -1;
            r.retained = src0 / (2.0 * ms) + cd(0.0, sg) * src1 / (2.0 * ns);
            cd comp = src0 / (2.0 * ms) - cd(0.0, sg) * src1 / (2.0 * ns);
            double den = Om[m] + sg * w;
            eta[m] = cd(0.0, ms) * comp / den;
            psi[m] = -double(sg) * ns * comp / den;
            r.near = true;
          } else {
            solve();
          }
        }
        phi[m] = psi[m] - taylor;
      }
      if (d < 4) {
        sh0[m] = sh1[m] = sv[m] = 0.0;
        for (int o = 0; o < d; ++o)
          for (const Part& p : T.p[m][o + 1]) {
            int P = p.blk[0];
            cd c = phi[P] * eprod(p) / fact[o];
            double z = M.zd(Kx[P], Ky[P], o);
            sh0[m] += cd(0.0, Kx[P] * z) * c;
            sh1[m] += cd(0.0, Ky[P] * z) * c;
            sv[m] += M.zd(Kx[P], Ky[P], o + 1) * c;
          }
      }
    }
  }
  const double qK = M.q(Kx[15], Ky[15]);
  cd er = cd(0.0, -Om[15]) * eta[15] - qK * psi[15] - src0;
  cd pr = cd(0.0, -Om[15]) * psi[15] + M.g * eta[15] - src1;
  if (r.near) {
    const double w = std::sqrt(M.g * qK), ms = std::sqrt(w / (2.0 * M.g)), ns = std::sqrt(M.g / (2.0 * w));
    const int sg = Om[15] >= 0.0 ? 1 : -1;
    er += ms * r.retained;
    pr += cd(0.0, -sg * ns) * r.retained;
  }
  r.residual = std::hypot(std::abs(er), std::abs(pr));
  r.eta = eta[15];
  r.psi = psi[15];
  return r;
}

int repeat3(const int* id) {
  if (id[0] == id[2]) return 6;
  return (id[0] == id[1] || id[1] == id[2]) ? 2 : 1;
}

int repeat4(const int* id) {
  int f = 1, run = 1;
  for (int i = 1; i < 4; ++i) {
    if (id[i] == id[i - 1]) {
      ++run;
    } else {
      for (int v = 2; v <= run; ++v) f *= v;
      run = 1;
    }
  }
  for (int v = 2; v <= run; ++v) f *= v;
  return f;
}

}  // namespace

extern "C" {

// Sector-44 sum over unordered quartets i<=j<=k<=l whose first index lies in
// [first_begin, first_end).  out: 5 arrays complex128 [cases][ny][nx] (x
// fastest; eta, psi, eta coordinate, psi coordinate, retained) accumulated
// into.  counters (doubles): total, kept, near, strict-zero outputs,
// strict-zero intermediates, max residual, min detuning, max strict-zero source.
int wit_oracle_sector44(int64_t M, const int32_t* modes, const double* amps, int64_t cases, double g, double h,
                        double dkx, double dky, int64_t nx, int64_t ny, double thr, int64_t first_begin,
                        int64_t first_end, double* out, double* counters) {
  static const Tables T;
  const Model mdl{g, h};
  const size_t plane = (size_t)nx * ny * cases;
  const cd* a = reinterpret_cast<const cd*>(amps);
  cd* o = reinterpret_cast<cd*>(out);
  uint64_t kept = 0, nearc = 0, zo = 0, zi = 0;
  double maxres = 0.0, mindet = std::numeric_limits<double>::infinity(), maxz = 0.0;
  const int nth = omp_get_max_threads();
  std::vector<std::vector<cd>> acc(nth, std::vector<cd>(3 * plane));
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : kept, nearc, zo, zi) reduction(max : maxres, maxz) \
    reduction(min : mindet)
  for (int64_t i = first_begin; i < first_end; ++i) {
    std::vector<cd>& A = acc[omp_get_thread_num()];
    for (int64_t j = i; j < M; ++j)
      for (int64_t k = j; k < M; ++k)
        for (int64_t l = k; l < M; ++l) {
          const int id[4] = {(int)i, (int)j, (int)k, (int)l};
          int tx = 0, ty = 0;
          double kx4[4], ky4[4];
          for (int s = 0; s < 4; ++s) {
            tx += modes[2 * id[s]];
            ty += modes[2 * id[s] + 1];
            kx4[s] = modes[2 * id[s]] * dkx;
            ky4[s] = modes[2 * id[s] + 1] * dky;
          }
          if (tx == 0 && ty == 0) {
            ++zo;
            continue;
          }
          if (std::abs(tx) >= nx / 2 || std::abs(ty) >= ny / 2) continue;
          static const int plus[4] = {1, 1, 1, 1};
          Result r = evaluate(T, mdl, kx4, ky4, plus, thr);
          const double rep = repeat4(id);
          const size_t ix = (size_t)(((tx % nx) + nx) % nx), iy = (size_t)(((ty % ny) + ny) % ny);
          for (int64_t c = 0; c < cases; ++c) {
            const cd* ac = a + c * M;
            const cd w = 2.0 * ac[i] * ac[j] * ac[k] * ac[l] / rep;
            const size_t loc = (size_t)c * nx * ny + ix + nx * iy;
            A[loc] += w * r.eta;
            A[plane + loc] += w * r.psi;
            if (r.near) A[2 * plane + loc] += w * r.retained;
          }
          ++kept;
          nearc += r.near ? 1 : 0;
          zi += r.zero_count;
          maxres = std::max(maxres, r.residual);
          mindet = std::min(mindet, r.rho);
          maxz = std::max(maxz, r.zero_src);
        }
  }
  for (int t = 0; t < nth; ++t)
    for (size_t x = 0; x < plane; ++x) {
      o[x] += acc[t][x];
      o[plane + x] += acc[t][plane + x];
      o[4 * plane + x] += acc[t][2 * plane + x];
    }
  const double total = (double)M * (M + 1) * (M + 2) * (M + 3) / 24.0;
  counters[0] = total;
  counters[1] += kept;
  counters[2] += nearc;
  counters[3] += zo;
  counters[4] += zi;
  counters[5] = std::max(counters[5], maxres);
  counters[6] = std::min(counters[6] == 0.0 ? mindet : counters[6], mindet);
  counters[7] = std::max(counters[7], maxz);
  return 0;
}

// Signed sectors over the outer candidate loop indices [outer_begin, outer_end):
// sector 40 -- minus pairs (pair index, i <= j) x every plus pair, targets in
// the analytic half plane (mx > 0, or mx = 0 and my > 0); sector 42 -- minus
// modes x every plus triple (i <= j <= k).  Same outputs and counters as
// wit_oracle_sector44 (all five arrays; counters[0] is the sector's whole
// candidate count).
int wit_oracle_signed(int32_t sector, int64_t M, const int32_t* modes, const double* amps, int64_t cases, double g,
                      double h, double dkx, double dky, int64_t nx, int64_t ny, double thr, int64_t outer_begin,
                      int64_t outer_end, double* out, double* counters) {
  if (sector != 40 && sector != 42) return 1;
  static const Tables T;
  const Model mdl{g, h};
  const size_t plane = (size_t)nx * ny * cases;
  const cd* a = reinterpret_cast<const cd*>(amps);
  cd* o = reinterpret_cast<cd*>(out);
  std::vector<std::array<int, 3>> plus;  // sector 40: pairs (third unused); 42: triples
  for (int i = 0; i < M; ++i)
    for (int j = i; j < M; ++j) {
      if (sector == 40) {
        plus.push_back({i, j, -1});
      } else {
        for (int k = j; k < M; ++k) plus.push_back({i, j, k});
      }
    }
  uint64_t kept = 0, nearc = 0, zo = 0, zi = 0;
  double maxres = 0.0, mindet = std::numeric_limits<double>::infinity(), maxz = 0.0;
  const int nth = omp_get_max_threads();
  std::vector<std::vector<cd>> acc(nth, std::vector<cd>(5 * plane));
  const int sg40[4] = {1, 1, -1, -1}, sg42[4] = {1, 1, 1, -1};
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : kept, nearc, zo, zi) reduction(max : maxres, maxz) \
    reduction(min : mindet)
  for (int64_t outer = outer_begin; outer < outer_end; ++outer) {
    std::vector<cd>& A = acc[omp_get_thread_num()];
    for (const auto& pl : plus) {
      int id[4];
      if (sector == 40) {
        const auto& mi = plus[outer];
        id[0] = pl[0], id[1] = pl[1], id[2] = mi[0], id[3] = mi[1];
      } else {
        id[0] = pl[0], id[1] = pl[1], id[2] = pl[2], id[3] = (int)outer;
      }
      const int* sg = sector == 40 ? sg40 : sg42;
      int tx = 0, ty = 0;
      double kx4[4], ky4[4];
      for (int s = 0; s < 4; ++s) {
        tx += sg[s] * modes[2 * id[s]];
        ty += sg[s] * modes[2 * id[s] + 1];
        kx4[s] = modes[2 * id[s]] * dkx;
        ky4[s] = modes[2 * id[s] + 1] * dky;
      }
      if (tx == 0 && ty == 0) {
        ++zo;
        continue;
      }
      if (sector == 40 && !(tx > 0 || (tx == 0 && ty > 0))) continue;
      if (std::abs(tx) >= nx / 2 || std::abs(ty) >= ny / 2) continue;
      Result r = evaluate(T, mdl, kx4, ky4, sg, thr);
      const double rep = sector == 40 ? (1.0 + (id[0] == id[1])) * (1.0 + (id[2] == id[3])) : repeat3(id);
      const size_t ix = (size_t)(((tx % nx) + nx) % nx), iy = (size_t)(((ty % ny) + ny) % ny);
      for (int64_t c = 0; c < cases; ++c) {
        const cd* ac = a + c * M;
        const cd w = sector == 40 ? 2.0 * ac[id[0]] * ac[id[1]] * std::conj(ac[id[2]]) * std::conj(ac[id[3]]) / rep
                                  : 2.0 * ac[id[0]] * ac[id[1]] * ac[id[2]] * std::conj(ac[id[3]]) / rep;
        const size_t loc = (size_t)c * nx * ny + ix + nx * iy;
        A[loc] += w * r.eta;
        A[plane + loc] += w * r.psi;
        A[2 * plane + loc] += w * r.ceta;
        A[3 * plane + loc] += w * r.cpsi;
        if (r.near) A[4 * plane + loc] += w * r.retained;
      }
      ++kept;
      nearc += r.near ? 1 : 0;
      zi += r.zero_count;
      maxres = std::max(maxres, r.residual);
      mindet = std::min(mindet, r.rho);
      maxz = std::max(maxz, r.zero_src);
    }
  }
  for (int t = 0; t < nth; ++t)
    for (size_t x = 0; x < 5 * plane; ++x) o[x] += acc[t][x];
  const double np = (double)M * (M + 1) / 2.0;
  counters[0] = sector == 40 ? np * np : (double)plus.size() * M;
  counters[1] += kept;
  counters[2] += nearc;
  counters[3] += zo;
  counters[4] += zi;
  counters[5] = std::max(counters[5], maxres);
  counters[6] = std::min(counters[6] == 0.0 ? mindet : counters[6], mindet);
  counters[7] = std::max(counters[7], maxz);
  return 0;
}

}  // extern "C"
