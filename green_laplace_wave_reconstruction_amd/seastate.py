"""Seeded synthetic directional sea states on an FFT grid.

The components follow the MF12-style input of the reference API
(src/gl_spectral_coefficients.m:1): amplitudes ``a, b`` with
``a + i b = sqrt(2 S(kx, ky) dkx dky) e^{i theta}`` on strict-forward grid
cells, JONSWAP frequency spectrum times a ``cos^{2s}((theta - theta0)/2)``
spreading, truncated at ``|k| < kmax_factor * kp`` and scaled so that the
discrete variance equals ``(Hs/4)^2``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def peak_wavenumber(tp: float, depth: float, gravity: float = 9.81) -> float:
    """Solve ``(2 pi / Tp)^2 = g k tanh(k h)`` by Newton iteration."""
    w2h_g = (2.0 * math.pi / tp) ** 2 * depth / gravity
    q = max(math.sqrt(w2h_g), w2h_g)
    for _ in range(100):
        t = math.tanh(q)
        f = q * t - w2h_g
        q_new = q - f / (t + q * (1.0 - t * t))
        if abs(q_new - q) < 1e-15 * max(1.0, q):
            q = q_new
            break
        q = q_new
    return q / depth


@dataclass(frozen=True)
class SeaState:
    """Components plus the grid they were drawn on."""

    a: np.ndarray
    b: np.ndarray
    kx: np.ndarray
    ky: np.ndarray
    gravity: float
    depth: float
    lx: float
    ly: float
    nx: int
    ny: int
    kp: float

    @property
    def count(self) -> int:
        return int(self.a.size)


def jonswap_sea_state(*, nx: int, ny: int, lx: float, ly: float, depth: float, hs: float, tp: float,
                      gamma: float = 3.3, spread_s: float = 10.0, theta0: float = 0.0, gravity: float = 9.81,
                      kmax_factor: float = 4.0, min_kh: float = 0.0, seed: int = 0) -> SeaState:
    """Draw a seeded strict-forward JONSWAP sea state on the ``[ny, nx]`` FFT grid.

    Cells kept: ``mx >= 1`` (kx > 0), ``|my| < ny/2``, ``|k| < kmax_factor*kp`` and
    ``k h > min_kh``.  Phases are ``U[0, 2 pi)`` from ``numpy.random.default_rng(seed)``
    drawn in row-major cell order.
    """
    kp = peak_wavenumber(tp, depth, gravity)
    dkx, dky = 2.0 * math.pi / lx, 2.0 * math.pi / ly
    mx = np.arange(1, nx // 2)
    my = np.arange(-(ny // 2) + 1, ny // 2)
    MX, MY = np.meshgrid(mx, my)
    KX, KY = MX * dkx, MY * dky
    K = np.hypot(KX, KY)
    keep = (K < kmax_factor * kp) & (K * depth > min_kh)
    kx, ky, k = KX[keep], KY[keep], K[keep]
    omega = np.sqrt(gravity * k * np.tanh(k * depth))
    wp = 2.0 * math.pi / tp
    sigma = np.where(omega <= wp, 0.07, 0.09)
    peak = gamma ** np.exp(-((omega - wp) ** 2) / (2.0 * sigma**2 * wp**2))
    s_omega = omega**-5.0 * np.exp(-1.25 * (wp / omega) ** 4) * peak
    # d omega / d k for the finite-depth dispersion relation
    th = np.tanh(k * depth)
    domega_dk = gravity * (th + k * depth * (1.0 - th**2)) / (2.0 * omega)
    theta = np.arctan2(ky, kx)
    spread = np.cos(0.5 * (theta - theta0)) ** (2.0 * spread_s)
    density = s_omega * domega_dk / k * spread  # S(kx, ky) up to a constant
    variance = float(np.sum(density) * dkx * dky)
    if variance <= 0:
        raise ValueError("empty sea state: no energetic strict-forward cells")
    density *= (hs / 4.0) ** 2 / variance
    amp = np.sqrt(2.0 * density * dkx * dky)
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0.0, 2.0 * math.pi, size=amp.size)
    return SeaState(a=amp * np.cos(phase), b=amp * np.sin(phase), kx=kx, ky=ky, gravity=gravity, depth=depth,
                    lx=lx, ly=ly, nx=nx, ny=ny, kp=kp)


def order2_plain_128(seed: int = 1) -> SeaState:
    """Plain GL order 2 at 128^2: L=1000 m, h=50 m, Hs=4 m, Tp=10 s, gamma=3.3, s=10."""
    return jonswap_sea_state(nx=128, ny=128, lx=1000.0, ly=1000.0, depth=50.0, hs=4.0, tp=10.0, gamma=3.3,
                             spread_s=10.0, seed=seed)


def order3_headline_512(seed: int = 2) -> SeaState:
    """GL order 3 at 512^2: L=2000 m, h=30 m, Hs=6 m, Tp=12 s, gamma=3.3, s=15, parents kh > 0.5."""
    return jonswap_sea_state(nx=512, ny=512, lx=2000.0, ly=2000.0, depth=30.0, hs=6.0, tp=12.0, gamma=3.3,
                             spread_s=15.0, min_kh=0.5, seed=seed)


def headline_instants(tp: float = 12.0, count: int = 64) -> np.ndarray:
    """The 64 instants uniform over [0, 10 Tp]."""
    return np.linspace(0.0, 10.0 * tp, count)


def order4_dimensionless_256(seed: int = 5) -> SeaState:
    """Order-4 sea state: g = h = 1, kp h = 1.5, kp Hs / 2 = 0.12, 256^2 over 5 peak wavelengths
    (dk = 0.3), JONSWAP gamma = 3.3, s = 10, |k| < 4 kp (quartic-alias safe)."""
    kp = 1.5
    tp = 2.0 * math.pi / math.sqrt(kp * math.tanh(kp))
    length = 5.0 * 2.0 * math.pi / kp
    return jonswap_sea_state(nx=256, ny=256, lx=length, ly=length, depth=1.0, hs=0.24 / kp, tp=tp, gamma=3.3,
                             spread_s=10.0, gravity=1.0, seed=seed)


BATCH_GRIDS = (64, 128, 192, 256, 384, 512, 1024)


@dataclass(frozen=True)
class BatchState:
    """One draw of the uneven batch: the sea state, its GL order and the draw itself."""

    state: SeaState
    order: int
    kp_h: float
    hs: float
    tp: float
    gamma: float
    spread_s: float
    theta0_drawn: float
    theta0: float
    seed: int


def fold_direction(theta0: float) -> float:
    """Map a mean direction into (-pi/4, pi/4] by the symmetries of a square grid.

    Rotations by quarter turns and the flips x -> -x, y -> -y, x <-> y map the
    square FFT grid onto itself; the GL kernels depend only on |k| and on dot
    products of wavenumbers, so a sea state and its image give image fields.
    """
    q = 0.5 * math.pi
    return 0.25 * math.pi - math.fmod(math.fmod(0.25 * math.pi - theta0, q) + q, q)


def batch_sea_states(count: int = 256, seed: int = 3, *, gravity: float = 9.81, kmax_factor: float = 4.0,
                     fill: float = 0.9) -> list[BatchState]:
    """Seeded uneven batch of directional sea states for GL orders 2 and 3.

    Draws (``numpy.random.default_rng(seed)``, in this order per state): grid
    N from ``BATCH_GRIDS``, order in {2, 3}, kp h ~ U[0.5, 5], Hs ~ U[1, 8] m,
    Tp ~ U[6, 16] s, gamma ~ U[1, 5], s ~ U[2, 30], theta0 ~ U[0, 2 pi).
    Then: depth h from the dispersion relation at (Tp, kp h); theta0 folded
    into (-45, 45] degrees (``fold_direction``) so that the strict-forward
    half plane kx > 0 holds the spectrum; square domain
    L = fill * N pi / (order kmax) with kmax = kmax_factor kp, so that
    order kmax / dk = fill N / 2 < N / 2 (no aliasing of the order-n
    products onto the grid); parents with kh <= 0.5 (order 3) or kh < 0.3
    (order 2) dropped; component phases from seed ``seed * 1000 + i``.
    """
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(BATCH_GRIDS))
        order = int(rng.integers(2, 4))
        kp_h = float(rng.uniform(0.5, 5.0))
        hs = float(rng.uniform(1.0, 8.0))
        tp = float(rng.uniform(6.0, 16.0))
        gamma = float(rng.uniform(1.0, 5.0))
        spread = float(rng.uniform(2.0, 30.0))
        theta_drawn = float(rng.uniform(0.0, 2.0 * math.pi))
        omega_p = 2.0 * math.pi / tp
        depth = gravity * kp_h * math.tanh(kp_h) / omega_p**2
        kp = kp_h / depth
        length = fill * n * math.pi / (order * kmax_factor * kp)
        theta = fold_direction(theta_drawn)
        # order 3 keeps kh > 0.5; order 2 keeps kh >= 0.3 (strict test with the bound nudged down)
        min_kh = 0.5 if order == 3 else 0.3 * (1.0 - 1e-12)
        st = jonswap_sea_state(nx=n, ny=n, lx=length, ly=length, depth=depth, hs=hs, tp=tp, gamma=gamma,
                               spread_s=spread, theta0=theta, gravity=gravity, kmax_factor=kmax_factor,
                               min_kh=min_kh, seed=seed * 1000 + i)
        out.append(BatchState(state=st, order=order, kp_h=kp_h, hs=hs, tp=tp, gamma=gamma, spread_s=spread,
                              theta0_drawn=theta_drawn, theta0=theta, seed=seed * 1000 + i))
    return out


def inverse_256(seed: int = 4) -> SeaState:
    """Hidden sea state of the inverse reconstruction: 256^2, h=40 m, Hs=5 m, Tp=11 s, gamma=3.3, s=10,
    L=1700 m (|k| < 4 kp stays cubic-alias safe below L=1787 m), parents kh > 0.5."""
    return jonswap_sea_state(nx=256, ny=256, lx=1700.0, ly=1700.0, depth=40.0, hs=5.0, tp=11.0, gamma=3.3,
                             spread_s=10.0, min_kh=0.5, seed=seed)


def unidirectional_record(count: int = 2048, seed: int = 11, *, periods: float = 50.0, depth: float = 20.0,
                          kp_h: float = 1.0, hs: float = 1.5, gamma: float = 3.3, gravity: float = 9.81):
    """Single-direction JONSWAP elevation record eta1(t) at k_p h = kp_h over `periods` peak periods.

    Random phases (numpy default_rng(seed)) on the record's own FFT bins, peak
    enhancement gamma with sigma 0.07/0.09, scaled to significant height hs.
    Returns (eta1, time, depth).  The reference's modal timing runs use
    k_p h = 1 and 50 T_p records (python/docs/random_hos_modal_comparison.md).
    """
    kp = kp_h / depth
    tp = 2 * math.pi / math.sqrt(gravity * kp * math.tanh(kp * depth))
    dt = periods * tp / count
    t = np.arange(count) * dt
    f = np.fft.rfftfreq(count, dt)
    fp = 1.0 / tp
    s = np.zeros_like(f)
    fk = f[1:]
    sig = np.where(fk <= fp, 0.07, 0.09)
    s[1:] = fk**-5.0 * np.exp(-1.25 * (fp / fk) ** 4) * gamma ** np.exp(-((fk - fp) ** 2) / (2 * sig**2 * fp**2))
    df = f[1] - f[0]
    s *= (hs / 4) ** 2 / np.sum(s * df)
    rng = np.random.default_rng(seed)
    half = 0.5 * count * np.sqrt(2 * s * df) * np.exp(2j * math.pi * rng.random(f.size))
    half[0] = 0.0
    if count % 2 == 0:
        half[-1] = 0.0
    return np.fft.irfft(half, count), t, depth
