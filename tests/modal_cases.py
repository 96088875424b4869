"""Single-direction modal time cases shared by the CPU and GPU tests.

The cases follow python/tests/test_modal_time.py:9-117 and
python/tests/test_temporal.py:13-63 of the reference (same records, options
and tolerances); the two MATLAB-generated records come from
tests/golden/reference_unidirectional_*.npz.
"""

from __future__ import annotations

import numpy as np

from cases import GOLDEN


def unidirectional(name: str) -> dict:
    d = np.load(GOLDEN / f"reference_unidirectional_{name}.npz")
    return {k: d[k] for k in d.files}


def order3_fixture_kwargs(d: dict) -> dict:
    """Options of test_temporal.py:13-34 (order 3, GL rank from the record, no eta20)."""
    return dict(depth=float(d["depth"]), gravity=float(d["gravity"]), omega_max=float(d["omega_max"]),
                energy_fraction=1.0, quadrature_rank=int(d["rank"]), order=3, modal_points=int(d["modal_points"]),
                modal_limit=float(d["modal_limit"]), include_eta20=False)


def eta20_fixture_kwargs(d: dict) -> dict:
    """Options of test_temporal.py:37-57 (order 2 at 512 modal points, eta20 on the record's grid)."""
    t = d["time"]
    period = t.size * (t[1] - t[0])
    return dict(depth=float(d["depth"]), gravity=float(d["gravity"]), omega_max=2 * np.pi * 8 / period + 1e-12,
                energy_fraction=1.0, order=2, modal_points=512, include_eta20=True, eta20_rank=int(d["rank"]),
                eta20_grid=tuple(int(v) for v in d["grid"]))


def nu(q):
    return np.sqrt(q * np.tanh(q))


def commensurate_q() -> float:
    """q1 with nu(2 q1) = 1.5 nu(q1) by bisection (test_modal_time.py:109-117)."""
    lo, hi = 1e-6, 20.0
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        if nu(2 * mid) / nu(mid) > 1.5:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def on_grid_case():
    """Two commensurate components on both a time record and a 128x16 spatial grid (test_modal_time.py:9-56)."""
    q1 = commensurate_q()
    q2 = 2 * q1
    gravity, depth = 9.81, 1.0
    w1, w2 = np.sqrt(gravity / depth) * nu(q1), np.sqrt(gravity / depth) * nu(q2)
    n = 256
    time = np.arange(n) * (4 * np.pi / w1) / n
    amps = np.array([0.01 * np.exp(0.2j), 0.006 * np.exp(-0.4j)])
    eta1_time = np.real(amps[0] * np.exp(-1j * w1 * time) + amps[1] * np.exp(-1j * w2 * time))
    nx, ny = 128, 16
    lx, ly = 2 * np.pi / q1, 2 * np.pi
    X, _ = np.meshgrid(np.arange(nx) * lx / nx, np.arange(ny) * ly / ny)
    eta1_space = np.real(amps[0] * np.exp(1j * q1 * X) + amps[1] * np.exp(1j * q2 * X))
    modal_points = 256
    time_kwargs = dict(depth=depth, gravity=gravity, omega_max=w2 + 1e-12, energy_fraction=1.0, quadrature_rank=4,
                       order=3, modal_points=(128, modal_points), modal_limit=modal_points * (q1 / 4) / 2,
                       include_eta20=False)
    space_kwargs = dict(lx=lx, ly=ly, depth=depth, gravity=gravity, order=3, eta22_rank=4)
    return eta1_time, time, time_kwargs, eta1_space, space_kwargs


def off_grid_case():
    """Three off-grid components and the ladder 256..2048 (test_modal_time.py:70-102)."""
    n, dt = 512, 0.2
    time = np.arange(n) * dt
    bins = np.array([20, 22, 24])
    amps = np.array([0.01, 0.006, 0.004])
    phases = np.array([0.2, -0.7, 1.1])
    k = np.arange(n)
    eta1 = sum(a * np.cos(2 * np.pi * b * k / n + p) for a, b, p in zip(amps, bins, phases))
    kwargs = dict(depth=1.0, omega_max=2 * np.pi * bins[-1] / (n * dt) + 1e-12, energy_fraction=1.0,
                  quadrature_rank=4, order=3, modal_points=(256, 512, 1024, 2048), relative_tolerance=1e-2,
                  include_eta20=False)
    return eta1, time, kwargs


def broadband_record(count: int = 2048, seed: int = 11):
    """k_p h = 1 JONSWAP record over 50 peak periods (green_laplace_wave_reconstruction_amd.seastate)."""
    from green_laplace_wave_reconstruction_amd import seastate

    return seastate.unidirectional_record(count, seed)


def short_record(count: int = 128, seed: int = 5):
    """A 128-sample k_p h = 1 record over 12 peak periods: its whole 3 x band fits the record and
    the oracle's P-point grids stay cheap up to P = 131072 (modal lengths up to 65536)."""
    from green_laplace_wave_reconstruction_amd import seastate

    return seastate.unidirectional_record(count, seed, periods=12.0)


def stokes_record(period: float, depth: float = 10.0, gravity: float = 9.81, amp: float = 0.05,
                  n: int = 1024, cycles: int = 16):
    """One-frequency record a cos(w t) over a whole number of periods and the finite-depth Stokes
    harmonics it must reduce to (Fenton 1985, J. Waterway Port Coastal Ocean Eng. 111(2), table 1,
    with S = sinh kh, C = cosh kh): eta22 = a^2 k B22 cos 2wt, B22 = C (2 C^2 + 1) / (4 S^3);
    eta33 = a^3 k^2 B33 cos 3wt, B33 = 3 (8 C^6 + 1) / (64 S^6)."""
    om = 2 * np.pi / period
    t = np.arange(n) * (cycles * period / n)
    target = om * om * depth / gravity   # q tanh q = w^2 h / g by bisection
    lo, hi = 1e-9, max(1.0, 2 * target)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if mid * np.tanh(mid) < target else (lo, mid)
    kh = 0.5 * (lo + hi)
    k = kh / depth
    s, c = np.sinh(kh), np.cosh(kh)
    b22, b33 = c * (2 * c * c + 1) / (4 * s**3), 3 * (8 * c**6 + 1) / (64 * s**6)
    closed = {"eta22": amp**2 * k * b22 * np.cos(2 * om * t), "eta33": amp**3 * k * k * b33 * np.cos(3 * om * t)}
    return amp * np.cos(om * t), t, dict(depth=depth, gravity=gravity, kh=kh, closed=closed)
