"""Host side of the GPU modal time path (CPU): band, deposit and reduced-grid planning.

The level plan must reproduce the oracle's band and modal bracket exactly, and
its reduced grids must hold every product without wrap (q_len >= order K + 1,
t_len >= order span + 1), which the kernels rely on.
"""

import ctypes

import numpy as np
import pytest

from green_laplace_wave_reconstruction_amd import _native
from green_laplace_wave_reconstruction_amd import modal as MD
from modal_cases import broadband_record, off_grid_case, order3_fixture_kwargs, unidirectional
from oracle import modal_oracle as M


def _lib_or_skip():
    if not _native.LIB_PATH.exists():
        pytest.skip("libglw.so not built")


def test_band_and_deposit_match_oracle():
    _lib_or_skip()
    d = unidirectional("modal_order3")
    kw = order3_fixture_kwargs(d)
    band = MD._select_band(d["eta1"], d["time"], depth=kw["depth"], gravity=kw["gravity"], order=3,
                           omega_max=kw["omega_max"], energy_fraction=1.0)
    ob = M.select_band(d["eta1"], d["time"], depth=kw["depth"], gravity=kw["gravity"], order=3,
                       omega_max=kw["omega_max"], energy_fraction=1.0)
    assert np.array_equal(band.bins, ob.bins)
    assert np.allclose(band.q, ob.q, rtol=1e-15, atol=0)
    lp = MD.modal_level_plan(band, depth=kw["depth"], gravity=kw["gravity"], rank=kw["quadrature_rank"], order=3,
                             points=kw["modal_points"], modal_limit=kw["modal_limit"])
    axis, dq = M.modal_axis(kw["modal_points"], float(np.max(ob.q)), 3, kw["modal_limit"])
    joint, proj = M.deposit_1d(ob.amplitudes, ob.bins - ob.bins.min(), ob.q, axis, dq, 3 * (ob.bins.size - 1) + 1)
    assert lp.projection == proj and lp.desc.dq == dq
    # rebuild the joint spectrum from the device deposit table
    P = kw["modal_points"]
    re = np.zeros_like(joint)
    for w, (ar, ai, wl, wu, kl) in enumerate(lp.deposit):
        re[int(kl) + P // 2, w] += (ar + 1j * ai) * wl
        re[int(kl) + 1 + P // 2, w] += (ar + 1j * ai) * wu
    assert np.array_equal(re, joint)


@pytest.mark.parametrize("points", [256, 1024, 4096, 16384, 65536])
def test_reduced_grids_hold_every_product(points):
    _lib_or_skip()
    eta, t, h = broadband_record()
    band = MD._select_band(eta, t, depth=h, gravity=9.81, order=3, omega_max=None, energy_fraction=0.999)
    lp = MD.modal_level_plan(band, depth=h, gravity=9.81, rank=8, order=3, points=points, modal_limit=None)
    d = lp.desc
    assert d.q_len >= 3 * d.kspan + 1 and d.t_len >= 3 * d.span + 1
    assert d.q_len <= points // 2 + 16  # the band-shifted axis is at most about half the reference's
    assert d.k0 >= 0 and 3 * (d.base_bin + d.span) < d.count
    assert _native.load_library().glw_modal_length_supported(d.q_len) & 2
    assert _native.load_library().glw_modal_length_supported(d.t_len) & 1


def test_eta20_plan_and_validation():
    _lib_or_skip()
    d = unidirectional("time_eta20")
    t = d["time"]
    band = MD._select_band(d["eta1"], t, depth=1.0, gravity=9.81, order=2,
                           omega_max=2 * np.pi * 8 / (t.size * (t[1] - t[0])) + 1e-12, energy_fraction=1.0)
    desc, dep, audit = MD.eta20_plan(band.amplitudes, band.omega, band.q, band.count, band.dt, depth=1.0,
                                     gravity=9.81, rank=16, nx=128, ny=16)
    assert audit["delta_q_rms_pair"] == pytest.approx(float(d["delta_q_rms_pair"]), rel=1e-13)
    assert desc.q_len >= 2 * desc.kspan + 1 and desc.t_len >= 2 * desc.span + 1
    with pytest.raises(ValueError, match="rank"):
        MD.eta20_plan(band.amplitudes, band.omega, band.q, band.count, band.dt, depth=1.0, gravity=9.81, rank=8,
                      nx=128, ny=16)
    with pytest.raises(TypeError, match="unknown time-reconstruction options"):
        MD.reconstruct_unidirectional_timeseries(d["eta1"], t, depth=1.0, bogus=1)
    eta1, tt, kw = off_grid_case()
    with pytest.raises(ValueError, match="increasing sequence"):
        MD.reconstruct_unidirectional_modal_timeseries(eta1, tt, **dict(kw, modal_points=(512, 256)))
    with pytest.raises(ValueError, match="rank 4, 6, 8, or 10"):
        MD.reconstruct_unidirectional_modal_timeseries(eta1, tt, **dict(kw, quadrature_rank=5))


def test_capi_rejects_bad_descriptors():
    """glw_modal_workspace_size / glw_eta20_workspace_size validate without a device."""
    import ctypes

    _lib_or_skip()
    lib = MD._bind(_native.load_library())
    eta, t, h = broadband_record()
    band = MD._select_band(eta, t, depth=h, gravity=9.81, order=3, omega_max=None, energy_fraction=0.999)
    good = MD.modal_level_plan(band, depth=h, gravity=9.81, rank=8, order=3, points=2048, modal_limit=None).desc
    size = ctypes.c_size_t()
    assert lib.glw_modal_workspace_size(ctypes.byref(good), ctypes.byref(size)) == 0 and size.value > 0
    for field, value, code in (("order", 4, 1), ("q_len", 1000, 3), ("t_len", 48, 1), ("rank", 17, 1),
                               ("q_len", 16, 1), ("dq", 0.0, 1), ("count", 300, 1)):
        bad = MD.CModalDesc.from_buffer_copy(good)
        setattr(bad, field, value)
        assert lib.glw_modal_workspace_size(ctypes.byref(bad), ctypes.byref(size)) == code, field
        assert lib.glw_last_error()
    d20 = MD.CEta20Desc(count=256, base_bin=1, span=10, q_len=64, t_len=32, k0=3, kspan=5, rank=16, dq=0.1,
                        fstep=0.05, depth=1.0, delta_q=0.07)
    assert lib.glw_eta20_workspace_size(ctypes.byref(d20), ctypes.byref(size)) == 0
    d20.t_len = 16  # below 2 span + 1
    assert lib.glw_eta20_workspace_size(ctypes.byref(d20), ctypes.byref(size)) != 0


def test_degenerate_eta20_and_grid_limit():
    _lib_or_skip()
    # a single input bin has no difference band: eta20 is exactly zero (_eta20.py:94-100)
    n = 256
    t = np.arange(n) * 0.25
    one = 0.01 * np.cos(2 * np.pi * 5 * np.arange(n) / n)
    band = MD._select_band(one, t, depth=1.0, gravity=9.81, order=2, omega_max=2 * np.pi * 5 / (n * 0.25) + 1e-12,
                           energy_fraction=1.0)
    ob = M.select_band(one, t, depth=1.0, gravity=9.81, order=2, omega_max=2 * np.pi * 5 / (n * 0.25) + 1e-12,
                       energy_fraction=1.0)
    desc, dep, audit = MD.eta20_plan(band.amplitudes, band.omega, band.q, band.count, band.dt, depth=1.0,
                                     gravity=9.81, rank=16, nx=128, ny=16)
    e20, a20 = M.eta20(ob.amplitudes, ob.omega, ob.q, t, depth=1.0, gravity=9.81)
    assert desc is None and audit["status"] == a20["status"] == "degenerate-zero-difference-output"
    assert not np.any(e20)
    # a level whose reduced modal grid exceeds the largest supported length raises a clear error
    eta, tt, h = broadband_record()
    band = MD._select_band(eta, tt, depth=h, gravity=9.81, order=3, omega_max=None, energy_fraction=0.999)
    with pytest.raises(_native.GlwError, match="no supported modal q grid length"):
        MD.modal_level_plan(band, depth=h, gravity=9.81, rank=8, order=3, points=262144, modal_limit=None)


def test_split_modal_lengths():
    """Modal lengths past one workgroup's LDS (8192) run as 2, 4 or 8 sub-transforms: the
    reference's default ladder (up to 65536 modal points) plans on supported grids."""
    _lib_or_skip()
    lib = _native.load_library()
    for n in (16, 6144, 8192):
        assert lib.glw_modal_length_supported(n) & 6 == 6
    for n in (12288, 16384, 24576, 32768, 49152, 65536):
        assert lib.glw_modal_length_supported(n) & 6 == 2  # modal pass only, not eta20
    for n in (10240, 20480, 98304):
        assert not lib.glw_modal_length_supported(n) & 2
    assert MD.grid_length(8193, "q") == 12288 and MD.grid_length(8192, "q20") == 8192
    with pytest.raises(_native.GlwError):
        MD.grid_length(8193, "q20")
    eta, tt, h = broadband_record()
    band = MD._select_band(eta, tt, depth=h, gravity=9.81, order=3, omega_max=None, energy_fraction=0.999)
    for points in (32768, 65536):
        d = MD.modal_level_plan(band, depth=h, gravity=9.81, rank=8, order=3, points=points, modal_limit=None).desc
        assert d.q_len > 8192 and d.q_len >= 3 * d.kspan + 1
        size = ctypes.c_size_t()
        assert lib.glw_modal_workspace_size(ctypes.byref(d), ctypes.byref(size)) == 0
