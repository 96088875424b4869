"""GPU eta44 (pure-sum GL, no Stokes) vs the reference's recorded GL-vs-WIT metrics and the oracle.

The n0382 check reproduces paper/order4/validation/recomputed_field_metrics.csv
(GL6/8/10 against the reference WIT field): 1e-9, 1e-8 and 1e-7 relative for
GL6, GL8 and GL10.  GL10's largest node (x = 29.9) amplifies roundoff most: a
1e-16 relative perturbation of the oracle's own input moves its GL10 n0382
field by 1.6e-10, and the GPU field (reduced alias-free grids, different FFT
sizes) differs from the oracle field by 2.3e-9 (GL8: 2.5e-10), which moves the
recorded metric by 4e-8.  The small case compares fields with the FP64 oracle
at 1e-10 for every rank (measured ~1e-12).
"""

import numpy as np
import pytest

from cases import rel
from oracle import gl_oracle as O
from order4_cases import N0382_METRICS, metrics, n0382, small_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def glw():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import green_laplace_wave_reconstruction_amd as G

    return G


@pytest.mark.parametrize("rank", [6, 8, 10])
def test_eta44_n0382_reproduces_reference_metrics(glw, rank):
    s, qx, qy, h, ref = n0382()
    print(f"eta44 n0382 GL{rank}", flush=True)
    eta, psi, audit = glw.gl_pure_sum_order4(s, qx, qy, 1.0, rank)
    l2, linf = metrics(h * np.real(eta), ref)
    tol = {6: 1e-9, 8: 1e-8, 10: 1e-7}[rank]
    assert abs(l2 - N0382_METRICS[rank][0]) < tol * N0382_METRICS[rank][0]
    assert abs(linf - N0382_METRICS[rank][1]) < 10 * tol * N0382_METRICS[rank][1]
    assert psi is None and audit["quadrature_rank"] == rank


@pytest.mark.parametrize("rank", [4, 6, 10])
def test_eta44_small_case_vs_oracle(glw, rank):
    s, qx, qy = small_case()
    eta, _, _ = glw.gl_pure_sum_order4(s, qx, qy, 1.0, rank)
    assert rel(eta, O.eta44(s, qx, qy, 1.0, rank)) < 1e-10


def test_glrti_gle44_drop_in(glw, tmp_path):
    """GLRTI001 -> GPU eta44 -> GLE44O01, like ``gl_eta44 INPUT OUTPUT 8 THREADS``."""
    from green_laplace_wave_reconstruction_amd import api, formats, plan as PL

    s, qx, qy, h, ref = n0382()
    rules = {r: PL.gauss_laguerre_rule(r) for r in (6, 8, 10)}
    formats.write_glrti(tmp_path / "gl_input.bin", s, dk=0.2, depth=h, peak_kh=1.0, rules=rules)
    api.run_glrti(tmp_path / "gl_input.bin", tmp_path / "gl_out.bin", 8)
    out = formats.read_gle44(tmp_path / "gl_out.bin")
    l2, _ = metrics(h * np.real(out["field"]), ref)
    assert out["rank"] == 8 and abs(l2 - N0382_METRICS[8][0]) < 1e-8 * N0382_METRICS[8][0]


def test_glrti_split_runs_sum_to_the_single_run(glw, tmp_path):
    """The reference's split mode (gl_eta44 ... OUTER4_INDEX [PART_INDEX], merged by
    merge_eta44_outer.cpp) on its 32^2 smoke input at GL6: the outer-node partials, and the
    base + per-third-node partials of every outer node, sum to the single run (the reference
    records 5.97e-16, paper/order4/validation/cpp_matlab_smoke.json)."""
    from green_laplace_wave_reconstruction_amd import api, formats, plan as PL
    from order4_cases import smoke32

    s, qx, qy, _ = smoke32()
    rank = 6
    formats.write_glrti(tmp_path / "in.bin", s, dk=0.5, depth=1.0, peak_kh=1.0, rules={rank: PL.gauss_laguerre_rule(rank)})
    full = api.run_glrti(tmp_path / "in.bin", None, rank)["field"]
    outer = []
    for o in range(rank):
        outer.append(tmp_path / f"o{o}.bin")
        api.run_glrti(tmp_path / "in.bin", outer[-1], rank, outer=o)
    by_outer = formats.merge_gle44(tmp_path / "merged_outer.bin", outer)["field"]
    parts = []
    for o in range(rank):
        for p in range(-1, rank):
            parts.append(tmp_path / f"o{o}p{p}.bin")
            api.run_glrti(tmp_path / "in.bin", parts[-1], rank, outer=o, part=p)
    by_part = formats.merge_gle44(tmp_path / "merged_parts.bin", parts)["field"]
    e1, e2 = rel(by_outer, full), rel(by_part, full)
    print(f"split GL6: outer {e1:.2e}, outer x part {e2:.2e}", flush=True)
    assert e1 < 1e-13 and e2 < 1e-13
    assert rel(formats.read_gle44(tmp_path / "o0p-1.bin")["field"], full) > 1e-3  # a partial is not the whole


@pytest.mark.parametrize("rank", [6, 8, 10])
def test_eta44_and_wit_reproduce_reference_smoke(glw, rank):
    """GPU eta44 and GPU WIT reproduce the GL-vs-WIT metric of the reference's 32^2 three-mode smoke
    (paper/order4/validation/cpp_matlab_smoke.json) to 1e-9 relative."""
    from green_laplace_wave_reconstruction_amd import api, wit as W
    from order4_cases import SMOKE32_GL_WIT, smoke32, smoke_metric

    s, qx, qy, w = smoke32()
    plan = W.WitPlan(w["modes"], nx=32, ny=32, gravity=1.0, depth=1.0, dkx=0.5, dky=0.5, threshold=1e-6)
    r = plan.sector44(w["amplitudes"])
    plan.close()
    eta, _, _ = api.gl_pure_sum_order4(s, qx, qy, 1.0, rank)
    m = smoke_metric(eta, r["eta"][0])
    print(f"smoke GL{rank}: {m:.15f} vs {SMOKE32_GL_WIT[rank]:.15f}", flush=True)
    assert abs(m - SMOKE32_GL_WIT[rank]) < 1e-9 * SMOKE32_GL_WIT[rank]


def test_eta44_seeded_config_vs_oracle(glw):
    """The seeded order-4 sea state (256^2, g = h = 1, kp h = 1.5, kp Hs / 2 = 0.12, JONSWAP over 5 peak
    wavelengths, seastate.order4_dimensionless_256) at GL6 against the FP64 oracle.  Bound 1e-7: the
    roundoff floor makes the GL graphs discontinuous at the 1e-16 rounding level (tests/test_conditioning.py);
    n0382's GL6/GL8 fields sit 1e-10 from the oracle."""
    from green_laplace_wave_reconstruction_amd import api, seastate

    st = seastate.order4_dimensionless_256()
    c = api.gl_spectral_coefficients(1, st.gravity, st.depth, st.a, st.b, st.kx, st.ky, 0, 0)
    kxg, kyg = api._wavenumber_grid(st.lx, st.ly, st.nx, st.ny)
    spec, _ = api.deposit_analytic_input(c, st.lx, st.ly, st.nx, st.ny)
    qx, qy = st.depth * kxg, st.depth * kyg
    print(f"eta44 seeded 256^2 GL6: {st.count} parents", flush=True)
    eta, _, _ = glw.gl_pure_sum_order4(spec, qx, qy, c["peak_depth"], 6)
    ref = O.eta44(spec, qx, qy, c["peak_depth"], 6)
    err = rel(eta, ref)
    print(f"eta44 seeded 256^2 GL6 vs oracle: {err:.2e}", flush=True)
    assert err < 1e-7 and np.linalg.norm(ref) > 0
