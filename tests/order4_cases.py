"""Order-4 inputs shared by the CPU and GPU tests."""

import numpy as np

from cases import GOLDEN
from green_laplace_wave_reconstruction_amd import formats

# GL-vs-WIT metrics recorded by the reference for the n0382 field
# (paper/order4/validation/recomputed_field_metrics.csv:2-4): rank -> (rel L2, rel Linf)
N0382_METRICS = {6: (0.0916566749057289, 0.155820536315641), 8: (0.0420839343830129, 0.0683858418291898),
                 10: (0.0203486173251526, 0.0319305990944608)}


def n0382():
    """Spectrum, q grids and the WIT reference field of the n0382 case
    (paper/order4/prepare_order4_inputs.m:7-17; reproduce_order4_fields.m:8)."""
    inp = formats.read_dir4in(GOLDEN / "wit_n0382_input.bin")
    n, h = int(inp["nx"]), inp["depth"]
    s = np.zeros((n, n), complex)
    m = inp["modes"]
    s[m[:, 1] % n, m[:, 0] % n] = n * n * inp["amplitudes"][0] / h
    mode = np.fft.fftfreq(n) * n
    mx, my = np.meshgrid(mode, mode)
    g = np.load(GOLDEN / "wit_n0382_sector44.npz")
    eta = np.zeros(n * n, complex)
    eta[g["index"]] = g["eta"]
    ref = np.real(np.fft.ifft2(eta.reshape(n, n) * n * n))
    return s, h * inp["dkx"] * mx, h * inp["dky"] * my, h, ref


def metrics(field, ref):
    d = field - ref
    return np.linalg.norm(d) / np.linalg.norm(ref), np.abs(d).max() / np.abs(ref).max()


def small_case():
    """Four strict-forward modes on 64^2, quartic-alias safe (4|m| < 32)."""
    n = 64
    mode = np.fft.fftfreq(n) * n
    mx, my = np.meshgrid(mode, mode)
    qx, qy = 0.25 * mx, 0.25 * my
    s = np.zeros((n, n), complex)
    for (a, b), v in zip([(3, 0), (4, 1), (5, -1), (6, 2)],
                         [0.017 * np.exp(0.13j), 0.012 * np.exp(-0.41j), 0.009 * np.exp(0.77j), 0.006 * np.exp(-1.04j)]):
        s[b % n, a % n] = n * n * v
    return s, qx, qy


# The reference's order-4 C++ smoke (paper/order4/prepare_order4_cpp_smoke.m:4-10): 32^2, three
# modes, g = h = 1, dk = 0.5, peak kh = 1; GL-vs-WIT relative L2 of the physical eta44 recorded in
# paper/order4/validation/cpp_matlab_smoke.json ("gl_wit_raw_relative_l2", GL6/8/10).
SMOKE32_GL_WIT = {6: 0.028693176699293679, 8: 0.009226906849614044, 10: 0.0034847786460708808}


def smoke32():
    """(spectrum [32,32], qx, qy, WIT DIR4IN01-style dict) of the reference smoke case."""
    n = 32
    modes = np.array([[2, 0], [3, 1], [3, -1]], dtype=np.int32)
    c = np.array([0.012, 0.007 * np.exp(0.31j), 0.005 * np.exp(-0.47j)])
    s = np.zeros((n, n), dtype=complex)
    for (mx, my), cc in zip(modes, c):
        s[my % n, mx] = n * n * cc
    ax = np.rint(np.fft.fftfreq(n) * n)
    mx, my = np.meshgrid(ax, ax)
    wit = dict(modes=modes, amplitudes=c[None, :], gravity=1.0, depth=1.0, dkx=0.5, dky=0.5, nx=n, ny=n,
               threshold=1e-6)
    return s, 0.5 * mx, 0.5 * my, wit


def smoke_metric(gl_field, wit_eta):
    """validate_order4_cpp_smoke.m:14,27: ||Re GL - real(ifft2(n^2 WIT))|| / ||real(ifft2(n^2 WIT))||."""
    n = wit_eta.shape[0]
    ref = np.real(np.fft.ifft2(wit_eta * n * n))
    return float(np.linalg.norm(np.real(gl_field) - ref) / np.linalg.norm(ref))
