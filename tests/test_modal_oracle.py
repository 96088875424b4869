"""Pins the single-direction modal time oracle (CPU).

Reference checks: the two MATLAB-generated records at the reference's own
tolerances (python/tests/test_temporal.py:13-57), the on-grid identity with
the spatial GL graph and the off-grid ladder (python/tests/test_modal_time.py).
Closed forms: the GL node sums against their a -> 0 limit and the Laplace
integral they approximate, the Hermitian projection, and the dispersion solve.
"""

import numpy as np
import pytest

from cases import rel
from modal_cases import (eta20_fixture_kwargs, off_grid_case, on_grid_case, order3_fixture_kwargs, stokes_record,
                         unidirectional)
from oracle import gl_oracle as O
from oracle import modal_oracle as M


def test_fixture_order3_eta22_eta33():
    d = unidirectional("modal_order3")
    r = M.reconstruct_unidirectional_modal_timeseries(d["eta1"], d["time"], **order3_fixture_kwargs(d))
    assert rel(r.components["eta22"], d["eta22"]) < 3e-13
    assert rel(r.components["eta33"], d["eta33"]) < 3e-13
    assert not r.audit["auxiliary_space_domain"]


def test_fixture_eta20():
    d = unidirectional("time_eta20")
    r = M.reconstruct_unidirectional_modal_timeseries(d["eta1"], d["time"], **eta20_fixture_kwargs(d))
    assert rel(r.components["eta20"], d["eta20"]) < 8e-11
    assert r.audit["eta20"]["delta_q_rms_pair"] == pytest.approx(float(d["delta_q_rms_pair"]), rel=1e-13)


def test_on_grid_modal_matches_spatial_oracle():
    eta1_t, t, tk, eta1_x, sk = on_grid_case()
    spatial = O.reconstruct_directional_field(eta1_x, **sk)["components"]
    modal = M.reconstruct_unidirectional_modal_timeseries(eta1_t, t, **tk)
    for name in ("eta11", "eta22", "eta33"):
        assert np.isclose(modal.components[name][0], spatial[name][0, 0], rtol=3e-14, atol=1e-16), name
    assert modal.audit["wavevector_projection_rms_relative"] < 1e-14
    assert modal.audit["converged"] and modal.audit["modal_points"] == 256


def test_off_grid_ladder_converges():
    eta1, t, kw = off_grid_case()
    r = M.reconstruct_unidirectional_modal_timeseries(eta1, t, **kw)
    assert r.audit["converged"] and r.audit["modal_points"] == 1024
    assert r.audit["levels"][-1]["eta22_change"] < 1e-2
    assert r.audit["levels"][-1]["eta33_change"] < 1e-2
    proj = [lv["wavevector_projection_rms_relative"] for lv in r.audit["levels"]]
    assert all(b < a for a, b in zip(proj, proj[1:]))


def test_node_sums_limits():
    nodes, weights = O.gauss_laguerre_rule(8)
    s, scale = 1.3, 0.7
    # a -> 0: S -> sum (w/scale) e^{x - s tau} tau, C -> the same without tau (the a == 0 branch)
    S0, C0, _ = M.node_sums(s, 0.0, scale, nodes, weights)
    S1, C1, _ = M.node_sums(s, 1e-7, scale, nodes, weights)
    assert S1 == pytest.approx(float(S0), rel=1e-8) and C1 == pytest.approx(float(C0), rel=1e-12)
    # GL approximates the Laplace integrals int_0^inf e^{-s t} sinh(a t)/a dt = 1/(s^2 - a^2),
    # int_0^inf e^{-s t} cosh(a t) dt = s/(s^2 - a^2) (substitution t = x/scale)
    a = 0.4
    nodes, weights = O.gauss_laguerre_rule(16)
    S, C, ok = M.node_sums(np.array([3.0, 0.3]), a, 2.5, nodes, weights)
    assert ok.tolist() == [True, False] and S[1] == 0 and C[1] == 0
    assert S[0] == pytest.approx(1 / (9 - a * a), rel=1e-9)
    assert C[0] == pytest.approx(3 / (9 - a * a), rel=1e-9)


def test_hermitian_projection_and_dispersion():
    rng = np.random.default_rng(5)
    for n in (9, 10):
        c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        h = M.hermitian_project(c)
        assert np.allclose(np.fft.ifft(h).imag, 0, atol=1e-15)
        assert np.allclose(M.hermitian_project(h), h, atol=0)  # idempotent
    omega = np.array([0.05, 0.7, 3.0, 9.0])
    q = M.wavenumber_from_frequency(omega, 12.0, 9.81)
    assert np.allclose(q * np.tanh(q), omega**2 * 12.0 / 9.81, rtol=1e-14)
    assert q[-1] == pytest.approx(omega[-1] ** 2 * 12.0 / 9.81, rel=1e-12)  # deep water: q = omega^2 h / g


def test_homogeneity():
    """eta22 and eta33 are homogeneous of degree 2 and 3 in the record (band and peak are scale-free)."""
    d = unidirectional("modal_order3")
    kw = order3_fixture_kwargs(d)
    r1 = M.reconstruct_unidirectional_modal_timeseries(d["eta1"], d["time"], **kw)
    r2 = M.reconstruct_unidirectional_modal_timeseries(2.5 * d["eta1"], d["time"], **kw)
    assert rel(r2.components["eta22_analytic"], 2.5**2 * r1.components["eta22_analytic"]) < 1e-14
    assert rel(r2.components["eta33_analytic"], 2.5**3 * r1.components["eta33_analytic"]) < 1e-14


@pytest.mark.parametrize("period", [6.0, 10.0])
def test_single_frequency_record_converges_to_stokes(period):
    """A one-frequency record's eta22 and eta33 approach the finite-depth Stokes harmonics as the
    GL rank grows (kh 1.30 and 0.68; the ladder stops at 4096 modal points)."""
    eta1, t, w = stokes_record(period)
    errs = {"eta22": [], "eta33": []}
    for rank in (4, 8, 10):
        r = M.reconstruct_unidirectional_modal_timeseries(eta1, t, depth=w["depth"], gravity=w["gravity"],
                                                          quadrature_rank=rank, order=3, include_eta20=False)
        for k in errs:
            errs[k].append(rel(r.components[k], w["closed"][k]))
    # eta22: 5.6e-3, 1.3e-5, 5.8e-7 at kh 1.30; 2.3e-2, 3.3e-4, 3.4e-5 at kh 0.68
    # eta33: 7.4e-2, 4.7e-3, 1.4e-3 at kh 1.30; 8.3e-2, 1.1e-2, 3.5e-3 at kh 0.68
    e22, e33 = errs["eta22"], errs["eta33"]
    assert e22[1] < 0.05 * e22[0] and e22[2] < 0.2 * e22[1] and e22[2] < 1e-4, e22
    assert e33[1] < 0.2 * e33[0] and e33[2] < 0.5 * e33[1] and e33[2] < 5e-3, e33
