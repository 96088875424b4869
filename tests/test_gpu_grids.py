"""Output grids of any even size (GPU): lengths that are not FFT lengths synthesize through direct
DFT passes (the working grids stay FFT lengths).  The reference accepts any even Nx, Ny
(src/gl_spectral_surface.m:7-31; python/src/green_laplace/_spectral.py:9-16).

Seeded JONSWAP states on 100^2, 160^2 and 320^2 (5-smooth, not 2^a 3^b) through orders 1-3
against the FP64 oracle, and the linear fields of a 1000^2 grid against numpy's inverse FFT.  The
synthesis is exact to rounding (linear fields and psi22 ~1e-15); the order-3 bound 1e-6 is the
conditioning of these broadband states: on the 320^2 state (3522 parents) the CPU emulator of the
same reduced-grid decomposition is itself 2.6e-7 from the oracle, as the GPU is, and rounding noise
of 1e-16 of each transform's maximum moves it by 3-7e-7 (floor-threshold flips,
tests/test_conditioning.py).
"""

import math

import numpy as np
import pytest

from cases import rel
from oracle import gl_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def glw():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import green_laplace_wave_reconstruction_amd as G

    return G


def _state(n, seed):
    from green_laplace_wave_reconstruction_amd import seastate

    depth, tp, order = 20.0, 9.0, 3
    kp = seastate.peak_wavenumber(tp, depth)
    length = 0.9 * n * math.pi / (order * 4.0 * kp)  # cubic-alias safe by 10 %
    return seastate.jonswap_sea_state(nx=n, ny=n, lx=length, ly=length, depth=depth, hs=3.0, tp=tp, gamma=3.3,
                                      spread_s=10.0, min_kh=0.5, seed=seed)


@pytest.mark.parametrize("n", [100, 160, 320])
def test_non_fft_grid_orders_1_to_3_vs_oracle(glw, n):
    from green_laplace_wave_reconstruction_amd._native import load_library

    assert not load_library().glw_fft_length_supported(n)
    st = _state(n, 30 + n)
    c = glw.gl_spectral_coefficients(3, st.gravity, st.depth, st.a, st.b, st.kx, st.ky, 0, 0)
    print(f"grid {n}^2: {st.count} parents", flush=True)
    eta, psi, X, Y, comps, audit = glw.gl_spectral_surface(c, st.lx, st.ly, n, n, 3.0)
    ref = O.gl_spectral_surface(3, st.gravity, st.depth, st.a, st.b, st.kx, st.ky, st.lx, st.ly, n, n, 3.0)
    bounds = {"eta11": 1e-12, "psi11": 1e-12, "eta22": 1e-8, "psi22": 1e-12, "eta33": 1e-6, "psi33": 1e-6}
    for k, b in bounds.items():
        err = rel(comps[k], ref["components"][k])
        print(f"  {k}: {err:.2e}", flush=True)
        assert err < b, k
    assert eta.shape == (n, n) and X.shape == (n, n)


def test_linear_fields_on_a_1000_grid(glw):
    n = 1000
    st = _state(n, 7)
    c = glw.gl_spectral_coefficients(1, st.gravity, st.depth, st.a, st.b, st.kx, st.ky, 0, 0)
    eta, psi, X, Y, comps, audit = glw.gl_spectral_surface(c, st.lx, st.ly, n, n, 0.0)
    spec = np.zeros((n, n), complex)
    mx = np.rint(st.kx * st.lx / (2 * np.pi)).astype(int)
    my = np.rint(st.ky * st.ly / (2 * np.pi)).astype(int)
    spec[my % n, mx % n] = st.a + 1j * st.b
    ref = np.fft.ifft2(spec) * n * n
    err = rel(comps["eta11_plus"], ref)
    print(f"1000^2 eta11: {err:.2e}", flush=True)
    assert err < 1e-13
