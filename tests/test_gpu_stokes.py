"""GPU fields of one oblique wave against the finite-depth Stokes harmonics (closed forms in
cases.stokes_wave) at the highest GL ranks the plan takes: eta22 rank 16, order 3 inner 12 x outer 10.

Tolerances: the GL quadrature error at those ranks (psi22's two-node rule is exact for one wave;
eta22 < 1e-9 for kh > 1; eta33 and psi33 < 2e-3) and, at the same ranks, the FP64 oracle at 1e-11
(a narrow spectrum: the reference's fixture-level bound).
"""

import math

import numpy as np
import pytest

from cases import STOKES_WAVES, rel, stokes_wave
from oracle import gl_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from green_laplace_wave_reconstruction_amd import plan as PL
    from green_laplace_wave_reconstruction_amd._native import DevicePlan

    return PL, DevicePlan


@pytest.mark.parametrize("mx,my,t", STOKES_WAVES[1:])
def test_single_wave_matches_stokes_harmonics(dev, mx, my, t):
    import torch

    PL, DevicePlan = dev
    w = stokes_wave(mx, my, t)
    h, g = w["depth"], w["gravity"]
    kx, ky = O.wavenumber_grid(w["lx"], w["lx"], w["n"], w["n"])
    spec, _ = O.analytic_spectrum(w["eta1"], h, kx)
    host = PL.build_plan(spec, h * kx, h * ky, depth=h, gravity=g, order=3, eta22_rank=16, inner_rank=12,
                         outer_rank=10)
    print(f"stokes wave ({mx},{my}) t={t} kh={w['kh']:.2f}", flush=True)
    dp = DevicePlan(host, host.box1.take(spec), max_instants=1)
    try:
        e2, p2 = dp.run("order2", [t])
        e3, p3 = dp.run("order3", [0.0])
        torch.cuda.synchronize()
        out = {k: np.real(v.cpu().numpy()[0]) for k, v in
               (("eta22", e2), ("psi22", p2), ("eta33", e3), ("psi33", p3))}
    finally:
        dp.close()
    ref3 = stokes_wave(mx, my, 0.0)["closed"]
    assert rel(out["psi22"], w["closed"]["psi22"]) < 1e-12
    assert rel(out["eta22"], w["closed"]["eta22"]) < 1e-9
    assert rel(out["eta33"], ref3["eta33"]) < 2e-3
    assert rel(out["psi33"], ref3["psi33"]) < 2e-3
    # the same ranks through the oracle
    r2 = O.reconstruct_directional_field(w["eta1"], lx=w["lx"], ly=w["lx"], depth=h, gravity=g, time=t, order=2,
                                         eta22_rank=16)["components"]
    assert rel(out["eta22"], r2["eta22"]) < 1e-11
    assert rel(out["psi22"], r2["psi22"]) < 1e-11
    eo, po, _ = O.order3(spec, h * kx, h * ky, inner_rank=12, outer_rank=10)
    assert rel(out["eta33"], np.real(eo) * h) < 1e-11
    assert rel(out["psi33"], np.real(po) * h * math.sqrt(g * h)) < 1e-11
