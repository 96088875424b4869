"""Why the GL graphs refuse FP32 transforms (CPU).

The FP32-transform / FP64-accumulate emulation (``fp32_emulation``) of the
512^2 order-3 sea state (inner GL4 x outer GL4) and of the 128^2 order-2 sea
state, against the same graph in FP64.  Each check targets one way an FP32
path could be built; none meets relative L2 1e-5, so ``precision="fp32"``
raises (``api._precision``, and ``glw_plan_create`` in the C ABI).
"""

import numpy as np
import pytest

import fp32_emulation as F
import plan_emulator as E
from cases import rel
from green_laplace_wave_reconstruction_amd import api, plan as PL, seastate


def _host(state, order):
    c = api.gl_spectral_coefficients(order, state.gravity, state.depth, state.a, state.b, state.kx, state.ky, 0, 0)
    kxg, kyg = api._wavenumber_grid(state.lx, state.ly, state.nx, state.ny)
    spec, omega = api.deposit_analytic_input(c, state.lx, state.ly, state.nx, state.ny)
    h = state.depth
    host = PL.build_plan(spec, h * kxg, h * kyg, depth=h, gravity=state.gravity, order=order, peak_q=c["peak_depth"])
    host.omega = host.box1.take(omega)
    return host, host.box1.take(spec)


@pytest.fixture(scope="module")
def headline():
    host, s0 = _host(seastate.order3_headline_512(), 3)
    s = E.phase(host, s0, 0.0)
    no = [[False] * 4 for _ in range(4)]
    e, p = F.order3(host, s, no, [False] * 4, None)
    return host, s, (F.physical(host, e, host.box3), F.physical(host, p, host.box3))


def _order3_err(host, s, ref, inner, outer, level32, level64=None):
    e, p = F.order3(host, s, inner, outer, level32, level64)
    return rel(F.physical(host, e, host.box3), ref[0]), rel(F.physical(host, p, host.box3), ref[1])


def test_emulation_fp64_is_the_plan_emulator(headline):
    """Control: with every task in FP64 the emulation is the kernel-sequence emulator itself."""
    host, s, ref = headline
    e, p = E.order3(host, s)
    assert rel(F.physical(host, e, host.box3), ref[0]) < 1e-13
    assert rel(F.physical(host, p, host.box3), ref[1]) < 1e-13


def test_fp32_with_the_reference_floor_is_garbage(headline):
    """FP32 transforms under the reference's FP64 floor (16 eps64 log2 N^2): FP32 rounding noise passes
    the floor and the GL node factors (up to e^49 at the largest inner node) amplify it."""
    host, s, ref = headline
    yes = [[True] * 4 for _ in range(4)]
    ee, ep = _order3_err(host, s, ref, yes, [True] * 4, host.floor_level)
    assert ee > 1e6 and ep > 1e6


def test_fp32_floor_drops_bins_the_fp64_result_needs(headline):
    """An FP32-scaled floor keeps the noise out but drops real bins: all-FP32 misses 1e-5 by >100x, and so
    does FP64 arithmetic whose floors are merely raised to 1e-7 (the finest level FP32 can resolve)."""
    host, s, ref = headline
    yes = [[True] * 4 for _ in range(4)]
    ee, ep = _order3_err(host, s, ref, yes, [True] * 4, F.floor32(host))
    assert min(ee, ep) > 1e-3
    no = [[False] * 4 for _ in range(4)]
    ee, ep = _order3_err(host, s, ref, no, [False] * 4, None, level64=1e-7)
    assert min(ee, ep) > 1e-4


def test_single_inner_node_in_fp32_misses_the_bound(headline):
    """Even one FP32 inner task (outer node 0, inner node 1) is beyond 1e-5 at the FP32 floor."""
    host, s, ref = headline
    one = [[False] * 4 for _ in range(4)]
    one[0][1] = True
    ee, ep = _order3_err(host, s, ref, one, [False] * 4, F.floor32(host))
    assert max(ee, ep) > 1e-4


def test_eta22_in_fp32_misses_the_bound():
    """Order 2 (the 128^2 sea state, rank 6): eta22 with FP32 transforms, at either floor."""
    host, s0 = _host(seastate.order2_plain_128(), 2)
    ref = F.physical(host, F.eta22(host, s0, False, host.floor_level), host.box2)
    for level in (host.floor_level, 1e-7, F.floor32(host)):
        assert rel(F.physical(host, F.eta22(host, s0, True, level), host.box2), ref) > 1e-5


def test_api_refuses_fp32():
    st = seastate.order2_plain_128()
    c = api.gl_spectral_coefficients(2, st.gravity, st.depth, st.a, st.b, st.kx, st.ky, 0, 0)
    with pytest.raises(ValueError, match="refused"):
        api.gl_spectral_surface(c, st.lx, st.ly, st.nx, st.ny, 0.0, precision="fp32")
    with pytest.raises(ValueError, match="fp64"):
        api.gl_spectral_surface(c, st.lx, st.ly, st.nx, st.ny, 0.0, precision="fp16")
