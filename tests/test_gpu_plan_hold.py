"""Plan precision and graph holds through the C ABI (GPU).

* glw_plan_create refuses FP32 GL plans (GLW_EUNSUPPORTED): FP32 transforms
  miss relative L2 1e-5 against FP64 (tests/test_precision.py shows why).
* A held plan (glw_plan_hold, taken by glw_batch_create for every plan it
  captures) refuses what would move memory a captured graph replays into --
  autotuning, a larger split-row scratch -- and refuses destruction; after the
  batch is destroyed the plan works normally again.
"""

import ctypes

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _draw(n, order):
    from green_laplace_wave_reconstruction_amd import seastate as S

    return [d for d in S.batch_sea_states(256) if d.state.nx == n and d.order == order][0]


def test_fp32_gl_plan_is_refused(torch):
    from green_laplace_wave_reconstruction_amd import batch as B, plan as PL
    from green_laplace_wave_reconstruction_amd._native import GlwError, describe, load_library

    host, _ = B.state_plan(_draw(64, 3), PL.PREC_FP64)
    host.precision = PL.PREC_FP32
    lib = load_library()
    desc, keep = describe(host, 1)
    handle = ctypes.c_void_p()
    rc = lib.glw_plan_create(ctypes.byref(desc), ctypes.byref(handle))
    assert rc == 3 and not handle.value  # GLW_EUNSUPPORTED, no plan
    assert "refused" in lib.glw_last_error().decode()
    del keep
    import numpy as np

    from green_laplace_wave_reconstruction_amd._native import DevicePlan

    with pytest.raises(GlwError, match="refused"):
        DevicePlan(host, np.zeros((host.box1.wy, host.box1.wx), complex), max_instants=1)


def test_held_plan_refuses_moves_and_destruction(torch):
    from green_laplace_wave_reconstruction_amd import batch as B, plan as PL
    from green_laplace_wave_reconstruction_amd._native import DevicePlan, GlwError, GraphBatch

    host, s0 = B.state_plan(_draw(128, 3), PL.PREC_FP64)
    dp = DevicePlan(host, s0, max_instants=1)
    t = dp.times_tensor([0.0])
    a, b = dp.run("order3", t)
    ref = (a.clone(), b.clone())
    g = GraphBatch([(dp, "order3", t, a, b)], branches=1)
    with pytest.raises(GlwError, match="captured graph"):
        dp.autotune("order3", t, out=(a, b), reps=1)
    with pytest.raises(GlwError, match="captured graph"):
        dp.close()
    with pytest.raises(GlwError, match="captured graph"):  # the split path's scratch would have to grow
        dp.schedule = "o3_inner_row=0/2;o3_outer_row=0/2"
        dp.run("order3", t, out=(a, b))
    dp.schedule = ""
    a.zero_()
    g.launch()
    torch.cuda.synchronize()
    assert torch.equal(a, ref[0]) and torch.equal(b, ref[1])
    g.close()
    dp.hold(True)
    dp.hold(False)
    with pytest.raises(GlwError, match="already zero"):
        dp.hold(False)
    dp.autotune("order3", t, out=(a, b), reps=1)
    dp.close()
