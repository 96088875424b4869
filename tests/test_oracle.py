"""Pins the FP64 oracle (CPU): reference fixtures, closed forms, brute force, homogeneity."""

import json
import math

import numpy as np
import pytest

from cases import GOLDEN, STOKES_WAVES, fixture, rel, stokes_wave
from oracle import gl_oracle as O


def _run_fixture(name, order):
    d = fixture(name)
    t = float(d["time"]) if "time" in d else 0.0
    rank = int(d["rank"]) if "rank" in d else 6
    r = O.reconstruct_directional_field(d["eta1"], lx=float(d["lx"]), ly=float(d["ly"]), depth=float(d["depth"]),
                                        gravity=float(d["gravity"]), time=t, order=order, eta22_rank=rank)
    return d, r["components"]


def test_fixture_eta22():
    # python/tests/test_spatial.py:16-34
    d, c = _run_fixture("eta22", 2)
    assert rel(c["eta11"], d["eta1"]) < 2e-13
    assert rel(c["eta22"], d["eta22"]) < 2e-12
    assert rel(c["psi11"], d["psi11"]) < 2e-13
    assert rel(c["psi22"], d["psi22"]) < 3e-12


def test_fixture_order3():
    # python/tests/test_spatial.py:37-49
    d, c = _run_fixture("order3", 3)
    assert rel(c["eta33"], d["eta33"]) < 5e-11
    assert rel(c["psi33"], d["psi33"]) < 5e-11


def test_fixture_nonzero_time():
    # python/tests/test_spatial.py:52-65
    d, c = _run_fixture("nonzero_time", 3)
    for k in ("eta11", "psi11", "eta22", "psi22", "eta33", "psi33"):
        assert rel(c[k], d[k]) < 8e-11, k


@pytest.mark.parametrize("rank", list(range(1, 17)))
def test_gauss_laguerre_factorial_moments(rank):
    # python/tests/test_quadrature.py:8-14 (all ranks the graphs use)
    x, w = O.gauss_laguerre_rule(rank)
    assert np.all(x > 0) and np.all(w > 0)
    for d in range(2 * rank):
        assert np.isclose(np.sum(w * x**d), math.factorial(d), rtol=3e-11)


def test_gl8_matches_frozen_interface():
    frozen = json.loads((GOLDEN / "frozen_gl8.json").read_text())
    x, w = O.gauss_laguerre_rule(8)
    np.testing.assert_allclose(x, frozen["nodes"], rtol=1e-13)
    np.testing.assert_allclose(w, frozen["weights"], rtol=1e-11)


def test_psi22_two_node_rule_is_gl2():
    # closed form nodes 2 -/+ sqrt(2) are the Laguerre L2 roots; weights (2 +/- sqrt 2)/4
    x, w = O.psi22_two_node_rule()
    xg, wg = O.gauss_laguerre_rule(2)
    np.testing.assert_allclose(x, xg, rtol=1e-14)
    np.testing.assert_allclose(w, wg, rtol=1e-14)
    np.testing.assert_allclose(w, [(2 + math.sqrt(2)) / 4, (2 - math.sqrt(2)) / 4], rtol=1e-14)


@pytest.mark.parametrize("mx,my,t", STOKES_WAVES)
def test_single_wave_second_order_is_finite_depth_stokes(mx, my, t):
    w = stokes_wave(mx, my, t)
    errs = []
    for rank in (4, 8, 12, 16):
        c = O.reconstruct_directional_field(w["eta1"], lx=w["lx"], ly=w["lx"], depth=w["depth"],
                                            gravity=w["gravity"], time=t, order=2, eta22_rank=rank)["components"]
        assert rel(c["psi22"], w["closed"]["psi22"]) < 1e-13, rank
        errs.append(rel(c["eta22"], w["closed"]["eta22"]))
    # geometric convergence of the GL eta22 rule: 2.4e-2 .. 6e-8 at kh 0.67, 4e-3 .. 2e-11 at kh 1.62
    assert all(b < 0.1 * a for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < (1e-7 if w["kh"] < 1 else 1e-9), errs


@pytest.mark.parametrize("mx,my,t", STOKES_WAVES[1:])
def test_single_wave_third_harmonic_converges_to_stokes(mx, my, t):
    """Order-3 eta33+ and psi33+ of the nested GL graph (analytic fields in units of h and
    h sqrt(g h)) at inner x outer ranks 4x4, 8x8, 12x10 against the Stokes third harmonics
    (time 0: the order-3 graph is stationary)."""
    w = stokes_wave(mx, my, 0.0)
    h, g = w["depth"], w["gravity"]
    kx, ky = O.wavenumber_grid(w["lx"], w["lx"], w["n"], w["n"])
    spec, _ = O.analytic_spectrum(w["eta1"], h, kx)   # dimensionless (eta / h)
    errs = {"eta33": [], "psi33": []}
    for ri, ro in ((4, 4), (8, 8), (12, 10)):
        e, p, _ = O.order3(spec, h * kx, h * ky, inner_rank=ri, outer_rank=ro)
        errs["eta33"].append(rel(np.real(e) * h, w["closed"]["eta33"]))
        errs["psi33"].append(rel(np.real(p) * h * math.sqrt(g * h), w["closed"]["psi33"]))
    # eta33: 7.9e-2, 5.8e-3, 1.1e-3 at kh 1.08; 6.7e-2, 3.8e-3, 4.4e-4 at kh 1.62
    # psi33: 2.6e-2, 2.2e-3, 4.0e-4 at kh 1.08; 3.3e-2, 2.6e-3, 2.8e-4 at kh 1.62
    for name, e in errs.items():
        assert e[1] < 0.1 * e[0] and e[2] < 0.3 * e[1], (name, e)
        assert e[2] < 2e-3, (name, e)


def test_linear_closed_form():
    # tests/run_release_tests.m:31-41
    r = O.gl_spectral_surface(1, 9.81, 1.0, [0.01], [0.0], [2.0], [0.0], 2 * np.pi, 2 * np.pi, 32, 32, 0.0)
    x = np.arange(32) * (2 * np.pi / 32)
    assert rel(r["eta"], np.tile(0.01 * np.cos(2 * x), (32, 1))) < 5e-14
    om = math.sqrt(9.81 * 2 * math.tanh(2.0))
    assert rel(r["psi"], np.tile(np.real(-1j * 9.81 / om * 0.01 * np.exp(2j * x)), (32, 1))) < 5e-14


def _ordered_pair_eta22(spec, qx, qy, peak_q, nodes, weights):
    """Direct O(M^2) evaluation of the eta22 GL kernel per ordered pair of parents.

    Independent restatement of tests/finite_depth_directional_pure_gl8_ordered_pair.m:20-59:
    no FFT, no balance shift, scalar products of the two parents.
    """
    ny, nx = spec.shape
    nn = nx * ny
    active = np.flatnonzero(np.abs(spec) > 1e-13 * max(1.0, np.max(np.abs(spec))))
    rows, cols = np.unravel_index(active, spec.shape)
    amp = spec.flat[active] / nn
    px, py = qx.flat[active], qy.flat[active]
    q = np.hypot(px, py)
    nu = np.sqrt(q * np.tanh(q))
    lam = math.sqrt((2 * math.sqrt(peak_q * math.tanh(peak_q))) ** 2 - 2 * peak_q * math.tanh(2 * peak_q))
    out = np.zeros_like(spec)
    for i in range(active.size):
        for j in range(active.size):
            dot = px[i] * px[j] + py[i] * py[j]
            Q = math.hypot(px[i] + px[j], py[i] + py[j])
            A = Q * math.tanh(Q)
            a = math.sqrt(A)
            s = nu[i] + nu[j]
            sd = nu[i] ** 2 + nu[i] * nu[j] + nu[j] ** 2 - dot / (nu[i] * nu[j])
            sk = (q[i] ** 2 + dot) / nu[i] + (q[j] ** 2 + dot) / nu[j]
            ker = 0.0
            for x, w in zip(nodes, weights):
                t = x / lam
                sh = t if a == 0 else math.sinh(a * t) / a
                ker += w * math.exp(x) / lam * math.exp(-s * t) / 4 * (-A * sd * sh + sk * math.cosh(a * t))
            out[(rows[i] + rows[j]) % ny, (cols[i] + cols[j]) % nx] += nn * amp[i] * amp[j] * ker
    return np.fft.ifft2(out)


def _release_pair_case():
    # tests/run_release_tests.m:43-62
    n = 64
    mode = np.fft.fftfreq(n) * n
    mx, my = np.meshgrid(mode, mode)
    qx, qy = 0.25 * mx, 0.25 * my
    spec = np.zeros((n, n), dtype=complex)
    modes = [(3, 0), (4, 1), (5, -1), (6, 2)]
    amps = [0.017 * np.exp(0.13j), 0.012 * np.exp(-0.41j), 0.009 * np.exp(0.77j), 0.006 * np.exp(-1.04j)]
    for (a, b), v in zip(modes, amps):
        spec[b % n, a % n] = n * n * v
    return spec, qx, qy


@pytest.mark.parametrize("rank", [6, 8])
def test_eta22_ordered_pair_bruteforce(rank):
    spec, qx, qy = _release_pair_case()
    fast, _ = O.eta22(spec, qx, qy, rank, peak_q=1.0)
    x, w = O.gauss_laguerre_rule(rank)
    direct = _ordered_pair_eta22(spec, qx, qy, 1.0, x, w)
    assert rel(fast, direct) < 3e-11


def test_eta22_ordered_pair_sensitivity():
    """The brute-force check fails for a plausible mistake: a perturbed node scale (a
    wrong lambda, 1e-7 relative) or a wrong sign of the dynamic term."""
    spec, qx, qy = _release_pair_case()
    fast, _ = O.eta22(spec, qx, qy, 8, peak_q=1.0)
    x, w = O.gauss_laguerre_rule(8)
    assert rel(fast, _ordered_pair_eta22(spec, qx, qy, 1.0 + 1e-7, x, w)) > 3e-11
    assert rel(fast, _ordered_pair_eta22(spec, qx, qy, 1.0, x, -w)) > 1.0


@pytest.mark.parametrize("order,power", [(2, 2), (3, 3)])
def test_homogeneity(order, power):
    d = fixture("nonzero_time")
    kw = dict(lx=float(d["lx"]), ly=float(d["ly"]), depth=float(d["depth"]), gravity=float(d["gravity"]),
              time=0.3, order=order, eta22_rank=8)
    # scale by a power of two so the check is exact in FP64 (support max stays >= 1)
    a = O.reconstruct_directional_field(d["eta1"] * 4.0, **kw)["components"]
    b = O.reconstruct_directional_field(d["eta1"] * 8.0, **kw)["components"]
    keys = ("eta22", "psi22") if order == 2 else ("eta33", "psi33")
    for k in keys:
        assert rel(b[k], 2.0**power * a[k]) < 1e-14, k


def test_eta44_gl6_reproduces_reference_gl_wit_metric():
    """Oracle eta44 (GL6) on the reference's n0382 input vs the reference WIT field:
    rel-L2 and rel-Linf equal the recorded metrics to 1e-9 (the reference's own
    C++-vs-archived-MATLAB parity is 1.8e-10, paper/order4/validation/cpp_paper_field.json)."""
    from order4_cases import N0382_METRICS, metrics, n0382

    s, qx, qy, h, ref = n0382()
    l2, linf = metrics(h * np.real(O.eta44(s, qx, qy, 1.0, 6)), ref)
    assert abs(l2 - N0382_METRICS[6][0]) < 1e-9 * N0382_METRICS[6][0]
    assert abs(linf - N0382_METRICS[6][1]) < 1e-8 * N0382_METRICS[6][1]


def test_eta44_homogeneity_degree_four():
    from order4_cases import small_case

    s, qx, qy = small_case()
    a = O.eta44(s * 4.0, qx, qy, 1.0, 4)
    b = O.eta44(s * 8.0, qx, qy, 1.0, 4)
    assert rel(b, 16.0 * a) < 1e-14


@pytest.mark.parametrize("rank", [6, 8, 10])
def test_eta44_and_wit_oracles_reproduce_reference_smoke(rank):
    """GL-vs-WIT metric of the reference's 32^2 three-mode smoke (cpp_matlab_smoke.json), from the
    eta44 restatement and the WIT restatement together (measured agreement ~1e-13 relative)."""
    from oracle import wit as ow
    from order4_cases import SMOKE32_GL_WIT, smoke32, smoke_metric

    s, qx, qy, wit = smoke32()
    r = ow.sector44(wit)
    assert r["kept"] == 15 and r["near"] == 0 and r["max_residual"] < 1e-10
    e = O.eta44(s, qx, qy, 1.0, rank)
    e = e[0] if isinstance(e, tuple) else e
    m = smoke_metric(e, r["eta"][0])
    assert abs(m - SMOKE32_GL_WIT[rank]) < 1e-9 * SMOKE32_GL_WIT[rank]


BROADBAND = ("config1_128", "order3_128_t0", "order3_128_t20")


@pytest.mark.parametrize("name", BROADBAND)
def test_oracle_matches_reference_broadband_vectors(name):
    """Seeded broadband inputs through the reference's own reconstruct_directional_field
    (tests/golden/make_broadband_golden.py): the psi22 radial channels, balance shifts, peak_q and the
    roundoff floors on 378-1078 parents, which the 2-3-mode fixtures do not exercise.  Measured: 0 to
    6e-10 (bit-identical where the arithmetic order is); bounds from tests/test_conditioning.py."""
    z = np.load(GOLDEN / f"broadband_{name}.npz")
    order = int(z["order"])
    r = O.reconstruct_directional_field(z["eta1"], lx=float(z["lx"]), ly=float(z["ly"]), depth=float(z["depth"]),
                                        gravity=float(z["gravity"]), time=float(z["time"]), order=order)
    comps = r["components"] if isinstance(r, dict) else r.components
    assert rel(comps["eta22"], z["eta22"]) < 1e-9
    assert rel(comps["psi22"], z["psi22"]) < 1e-13
    if order == 3:
        assert rel(comps["eta33"], z["eta33"]) < 1e-8 and rel(comps["psi33"], z["psi33"]) < 1e-8
