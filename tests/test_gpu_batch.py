"""FFT schedules, the autotuner and the uneven batch graph on the GPU.

* Every schedule computes the same transforms: a plan run with forced
  schedules (greedy radices, one-row or half tiles, split row path) agrees
  with the default schedule to the order-3 parity bound 1e-7.  Schedules
  round differently (radix order; the compiler contracts multiply-adds per
  kernel instance), and order 3 amplifies FFT roundoff by up to e^49 before
  the roundoff floor (a 1e-16 input perturbation moves the 512^2 oracle
  output by 5e-10, an exact change of FFT size by 2.6e-8).  Where the oracle
  runs (N <= 256), every schedule is held to 1e-7 against it: the random
  batch states (kp h up to 5, steep, broad) are worse conditioned than the
  headline; the CPU emulation of the same reduced-grid decomposition is
  already 1e-9 from the oracle on the N=192 order-3 state.
* The graph launch of a batch writes exactly what the eager execute calls
  write (same kernels, same schedule, deterministic accumulation): bitwise,
  on every replay, with eager work between replays.
* Batch states against the oracle (small grids) at 1e-7; a 1024^2 order-3
  state (whose outer row tile exceeds LDS: split path) through the cubic
  homogeneity eta33(2 a) = 8 eta33(a).
"""

import math

import numpy as np
import pytest

from cases import rel
from oracle import gl_oracle as O

pytestmark = pytest.mark.gpu

O3_STAGES = ("o3_inner_colinv", "o3_inner_row", "o3_inner_colfwd", "o3_outer_colinv", "o3_outer_row",
             "o3_outer_colfwd", "synth_colinv")
O2_STAGES = ("o2_eta_colinv", "o2_eta_row", "o2_eta_colfwd", "o2_psi_colinv", "o2_psi_row", "o2_psi_colfwd",
             "o2_contact_colinv", "o2_contact_row", "o2_contact_colfwd", "synth_colinv")


@pytest.fixture(scope="module")
def torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _draws():
    from green_laplace_wave_reconstruction_amd import seastate as S

    return S.batch_sea_states(256)


def _pick(draws, n, order, k=0):
    return [d for d in draws if d.state.nx == n and d.order == order][k]


def _device(draw, nt=1):
    from green_laplace_wave_reconstruction_amd import batch as B, plan as PL
    from green_laplace_wave_reconstruction_amd._native import DevicePlan

    host, s0 = B.state_plan(draw, PL.PREC_FP64)
    return DevicePlan(host, s0, max_instants=nt)


def _oracle(draw):
    s = draw.state
    r = O.gl_spectral_surface(draw.order, s.gravity, s.depth, s.a, s.b, s.kx, s.ky, s.lx, s.ly, s.nx, s.ny, 0.0)
    c = r["components"]
    n = draw.order
    return c[f"eta{n}{n}_plus"], c[f"psi{n}{n}_plus"]


def _forced(stages, plan, tile):
    return ";".join(f"{s}={plan}/{tile}" for s in stages)


@pytest.mark.parametrize("n,order", [(192, 3), (256, 2), (512, 3)])
def test_forced_schedules_agree(torch, n, order):
    draw = _pick(_draws(), n, order)
    print(f"schedules N={n} order={order}", flush=True)
    dp = _device(draw)
    kind = f"order{order}"
    stages = O3_STAGES if order == 3 else O2_STAGES
    a0, b0 = [x.cpu().numpy() for x in dp.run(kind, [0.0])]
    ref = _oracle(draw) if n <= 256 else None
    for plan in (0, 1):
        for tile in (0, 1, 2):
            dp.schedule = _forced(stages, plan, tile)
            a, b = [x.cpu().numpy() for x in dp.run(kind, [0.0])]
            err = max(rel(a, a0), rel(b, b0))
            eo = max(rel(a[0], ref[0]), rel(b[0], ref[1])) if ref is not None else float("nan")
            print(f"plan {plan} tile {tile}: vs default {err:.2e}, vs oracle {eo:.2e}", flush=True)
            assert err < 1e-7, (plan, tile)
            if ref is not None:
                assert eo < 1e-7, (plan, tile)


def test_autotune_schedule_round_trip(torch):
    draw = _pick(_draws(), 384, 3)
    dp = _device(draw)
    a0, b0 = [x.cpu().numpy() for x in dp.run("order3", [0.0])]
    text = dp.autotune("order3", [0.0], reps=2)
    entries = dict(e.split("=") for e in text.split(";"))
    assert set(O3_STAGES) <= set(entries)
    for s, v in entries.items():
        p, t = map(int, v.split("/"))
        assert p in (0, 1) and t in (0, 1, 2)
    a, b = [x.cpu().numpy() for x in dp.run("order3", [0.0])]
    assert rel(a, a0) < 1e-7 and rel(b, b0) < 1e-7  # a greedy-radix stage rounds differently
    other = _device(draw)
    other.schedule = text
    assert other.schedule == text
    with pytest.raises(Exception):
        other.schedule = "o3_outer_row=2/0"


def test_batch_graph_matches_eager_and_oracle(torch):
    from green_laplace_wave_reconstruction_amd.batch import SeaStateBatch

    draws = [d for d in _draws() if d.state.nx <= 192][:12]
    assert {d.order for d in draws} == {2, 3}
    print(f"batch of {len(draws)} states", flush=True)
    batch = SeaStateBatch(draws, precision="fp64", branches=4, tune=True, reps=1)
    eager = [tuple(x.clone() for x in batch.fields(i)) for i in range(len(draws))]
    for e in batch.entries:
        e.a.zero_()
        e.b.zero_()
    for rep in range(3):
        batch.launch()
        torch.cuda.synchronize()
        for i, (ea, eb) in enumerate(eager):
            a, b = batch.fields(i)
            assert torch.equal(a, ea) and torch.equal(b, eb), (rep, i)
            a.zero_()
            b.zero_()
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        del junk
    batch.launch()
    torch.cuda.synchronize()
    for i in (0, 1, 2, 3):
        ra, rb = _oracle(draws[i])
        a, b = [x[0].cpu().numpy() for x in batch.fields(i)]
        assert rel(a, ra) < 1e-7 and rel(b, rb) < 1e-7, i
    assert batch.alg_bytes() > 0
    batch.close()


def test_order3_1024_split_path_homogeneity(torch):
    import dataclasses

    draw = _pick(_draws(), 1024, 3)
    print(f"1024^2 order 3: {draw.state.count} parents", flush=True)
    s2 = dataclasses.replace(draw.state, a=2 * draw.state.a, b=2 * draw.state.b)
    d2 = dataclasses.replace(draw, state=s2)
    r = []
    for d in (draw, d2):
        dp = _device(d)
        assert dp.host.g3[0] >= 512  # the fused outer row tile does not fit LDS there
        r.append([x.cpu().numpy() for x in dp.run("order3", [0.0])])
        dp.close()
    assert rel(r[1][0], 8 * r[0][0]) < 1e-10 and rel(r[1][1], 8 * r[0][1]) < 1e-10
    assert np.all(np.isfinite(r[0][0])) and np.linalg.norm(r[0][0]) > 0


def test_full_256_state_graph(torch):
    """The whole uneven batch (256 states, N = 64..1024, orders 2 and 3) in one graph launch: every
    state bitwise equal to its eager execute call; the N <= 128 states and every 8th N in {192, 256}
    state against the oracle at 3e-7 for order 3 (rounding at the 1e-16 level flips floor-threshold bins
    and moves these states by up to ~1e-7, tests/test_conditioning.py; measured worst 1.1e-7) and 1e-7
    for order 2."""
    from green_laplace_wave_reconstruction_amd.batch import SeaStateBatch

    draws = _draws()
    assert len(draws) == 256 and {d.state.nx for d in draws} == {64, 128, 192, 256, 384, 512, 1024}
    print("256-state batch: plans, tuning, eager runs, capture", flush=True)
    batch = SeaStateBatch(draws, precision="fp64", branches=4, tune=True, reps=1)
    eager = [tuple(x.clone() for x in batch.fields(i)) for i in range(len(draws))]
    for e in batch.entries:
        e.a.zero_()
        e.b.zero_()
    batch.launch()
    torch.cuda.synchronize()
    for i, (ea, eb) in enumerate(eager):
        a, b = batch.fields(i)
        assert torch.equal(a, ea) and torch.equal(b, eb), i
    del eager
    print("256-state batch: graph == eager for every state", flush=True)
    mid = [i for i, d in enumerate(draws) if d.state.nx in (192, 256)][::8]
    small = [i for i, d in enumerate(draws) if d.state.nx <= 128]
    worst = 0.0
    for i in small + mid:
        ra, rb = _oracle(draws[i])
        a, b = [x[0].cpu().numpy() for x in batch.fields(i)]
        err = max(rel(a, ra), rel(b, rb))
        worst = max(worst, err)
        assert err < (3e-7 if draws[i].order == 3 else 1e-7), (i, draws[i].state.nx, draws[i].order, err)
    print(f"256-state batch: {len(small) + len(mid)} states vs the oracle, worst rel-L2 {worst:.2e}", flush=True)
    batch.close()
