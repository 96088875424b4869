"""Broadband golden vectors from the reference Python package (run in the build container only).

The reference's three MATLAB-generated fixtures hold 2-3 parent modes, and
homogeneity and the Stokes limits pin the GL graphs for one wave.  Neither
exercises what only a broadband spectrum does: the data-dependent psi22
radial-channel edges and balance shifts, peak_q, and which bins the roundoff
floor zeroes.  This script imports the reference's own
``green_laplace.reconstruct_directional_field`` (python/src/green_laplace/
spatial.py:33-153, from its source tree under the reference root, never
copied) and records its outputs for seeded broadband inputs of this
project's generator:

* config1_128: order 2 (eta22 rank 6), 128^2, L = 1000 m, h = 50 m,
  JONSWAP Hs = 4 m, Tp = 10 s, gamma = 3.3, s = 10 (seed 1), t = 0;
* order3_128_t0 / order3_128_t20: order 3, 128^2, L = 700 m, h = 30 m,
  JONSWAP Hs = 6 m, Tp = 12 s, gamma = 3.3, s = 15, parents kh > 0.5
  (seed 21), t = 0 and t = 20 s.

Stored (float64): the input eta1 and the real output components eta22,
psi22 (and eta33, psi33 for order 3).  A strict-forward
component's real part determines its analytic spectrum, so the real fields
pin every bin.

Usage: python tests/golden/make_broadband_golden.py [REFERENCE_ROOT]
"""

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
sys.path.insert(0, str(REF / "python" / "src"))

from green_laplace import reconstruct_directional_field  # noqa: E402  (the reference package)

from green_laplace_wave_reconstruction_amd import seastate  # noqa: E402


def snapshot(st) -> np.ndarray:
    """eta1(x, y) = Re sum (a + i b) exp(i k.x) of a strict-forward sea state on its grid."""
    spec = np.zeros((st.ny, st.nx), complex)
    mx = np.rint(st.kx * st.lx / (2 * np.pi)).astype(int)
    my = np.rint(st.ky * st.ly / (2 * np.pi)).astype(int)
    spec[my % st.ny, mx % st.nx] = st.a + 1j * st.b
    return np.real(np.fft.ifft2(spec) * st.nx * st.ny)


CASES = {
    "config1_128": (seastate.order2_plain_128(), 2, 0.0),
    "order3_128_t0": (seastate.jonswap_sea_state(nx=128, ny=128, lx=700.0, ly=700.0, depth=30.0, hs=6.0, tp=12.0,
                                                 gamma=3.3, spread_s=15.0, min_kh=0.5, seed=21), 3, 0.0),
}
CASES["order3_128_t20"] = (CASES["order3_128_t0"][0], 3, 20.0)

for name, (st, order, t) in CASES.items():
    eta1 = snapshot(st)
    r = reconstruct_directional_field(eta1, lx=st.lx, ly=st.ly, depth=st.depth, gravity=st.gravity, time=t,
                                      order=order)
    keys = ["eta22", "psi22"] + (["eta33", "psi33"] if order == 3 else [])
    np.savez_compressed(HERE / f"broadband_{name}.npz", eta1=eta1, lx=st.lx, ly=st.ly, depth=st.depth,
                        gravity=st.gravity, time=t, order=order, parents=st.count,
                        **{k: np.asarray(r.components[k], dtype=np.float64) for k in keys})
    print(name, "parents", st.count, "order", order, "t", t,
          {k: float(np.linalg.norm(r.components[k])) for k in keys})
