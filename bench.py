"""Headline benchmark: eta33+psi33 fields/s at 512^2 on one MI355X.

Workload (one "step"): the GL order-3 graph (inner GL4 x outer GL4) of one
seeded JONSWAP sea state -- 512^2, L=2000 m, h=30 m, Hs=6 m, Tp=12 s,
gamma=3.3, s=15, 3120 strict-forward parents with kh > 0.5 -- at 64 instants
uniform over [0, 10 Tp], from the t=0 analytic spectrum resident in HBM to
the eta33+ and psi33+ physical fields in HBM.  One unit = one (eta33, psi33)
pair at one instant.  FP64 throughout: FP32 transforms are refused (they miss
relative L2 1e-5 against FP64), see ``fp32_arm`` in the output.

Roofline: the library reports each stage's compulsory HBM bytes on the
reduced (alias-free) working grids -- every input element a launch consumes
and every output element it produces, once; tables shared by all tasks
excluded (glw_plan_stage_times).  ``roofline`` is the dominant (longest)
stage: its bytes per launch over its average launch time (HIP events on the
execute stream), against the 8 TB/s HBM peak.  ``pipeline`` is the same for
the whole step.  ``equiv_fullgrid_GBps`` is the reference-graph figure (358
full-grid transforms per field pair, 2 sweeps x (read + write) of N^2 complex
values each): an equivalent rate, not a roofline fraction.  The measured HBM
traffic of the dominant stage (rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/) is
reported beside it when the stored schedule matches.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload order3|order2|api|eta44|wit|batch|inverse|modal]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
INSTANTS = 64

FP32_ARM = {
    "status": "refused",
    "reason": "FP32 transforms with FP64 accumulation miss relative L2 1e-5 against FP64 on this sea state: the GL "
              "node factors amplify spectrum bins ~1e-7 below each spectrum's maximum, which FP32 cannot resolve",
    "emulated_rel_l2_vs_fp64": {
        "all stages FP32, reference FP64 floor": {"eta33": 3.0e15, "psi33": 4.2e14},
        "all stages FP32, floor 16 eps32 log2 N^2": {"eta33": 9.9e-3, "psi33": 3.3e-3},
        "outer stages FP32 (inner FP64), best floor 1e-7": {"eta33": 8.1e-6, "psi33": 1.02e-5},
        "FP64 arithmetic, floors raised to 1e-7": {"eta33": 1.4e-3, "psi33": 4.7e-4},
    },
    "source": "CPU emulation of the kernel sequence (tests/fp32_emulation.py, tests/test_precision.py), instant t=0; "
              "worst over 8 instants for the outer-only row",
}


def host_info() -> dict:
    """Host facts recorded with every CPU baseline."""
    import torch

    model = platform.processor() or ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"os_cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "cpu_model": model,
            "numpy": np.__version__, "torch": torch.__version__}


def stage_roofline(stages: dict, steps: int, ms_per_step: float, skip=()) -> tuple[dict, dict]:
    """(roofline of the dominant stage, whole-step pipeline figure) from stage_times records
    {name: (total ms, launches, compulsory bytes)} accumulated over `steps` steps."""
    timed = {k: v for k, v in stages.items() if v[2] > 0 and k not in skip}
    dom = max(timed, key=lambda k: timed[k][0])
    ms, launches, moved = timed[dom]
    avg_ms = ms / launches
    per_launch = moved / launches
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "alg_bytes_per_launch": per_launch,
            "avg_launch_ms": avg_ms}
    total = sum(v[2] for v in stages.values()) / steps
    pipe = {"compulsory_bytes_per_step": total, "GBps": total / (ms_per_step * 1e-3) / 1e9,
            "frac": total / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "stages": {k: {"ms_per_step": v[0] / steps, "GB_per_step": v[2] / steps / 1e9,
                           "GBps": (v[2] / (v[0] * 1e-3) / 1e9) if v[0] > 0 else None} for k, v in stages.items()}}
    return roof, pipe


def workload():
    from green_laplace_wave_reconstruction_amd import api, plan as PL, seastate

    s = seastate.order3_headline_512()
    import green_laplace_wave_reconstruction_amd as G

    c = G.gl_spectral_coefficients(3, s.gravity, s.depth, s.a, s.b, s.kx, s.ky, 0, 0)
    kxg, kyg = api._wavenumber_grid(s.lx, s.ly, s.nx, s.ny)
    spec, omega = api.deposit_analytic_input(c, s.lx, s.ly, s.nx, s.ny)
    return s, c, kxg, kyg, spec, omega


def stored_schedule(summary: str = "pmc_traffic.json"):
    """The FFT schedule a committed PMC summary was taken with, or None."""
    p = ROOT / "profiles" / summary
    try:
        return json.loads(p.read_text()).get("_schedule") or None
    except Exception:
        return None


def traffic_from_profiles(kernel: str, summary: str = "pmc_traffic.json", schedule: str | None = None):
    """Measured HBM bytes per dispatch of `kernel` from a committed PMC summary, or None.

    With `schedule` (stage=plan/tile;...), None unless the summary was taken with the same
    plan/tile for that stage (another tile moves other bytes)."""
    p = ROOT / "profiles" / summary
    if not p.exists():
        return None
    try:
        data = json.loads(p.read_text())
        if schedule is not None:
            def entry(text):
                return dict(e.split("=") for e in text.split(";") if "=" in e).get(kernel, "0/0")
            if entry(data.get("_schedule", "")) != entry(schedule):
                return None
        return data.get(kernel, {}).get("hbm_bytes_per_dispatch")
    except Exception:
        return None


def sq_from_profiles(kernel: str, schedule: str | None = None, summary: str = "sq_counters.json"):
    """SQ-counter figures of `kernel` (VALU busy, waves per SIMD, ...) from a committed summary, or None
    (None also when the summary was taken with another plan/tile for that stage)."""
    p = ROOT / "profiles" / summary
    if not p.exists():
        return None
    try:
        data = json.loads(p.read_text())
        if schedule is not None:
            def entry(text):
                return dict(e.split("=") for e in text.split(";") if "=" in e).get(kernel, "0/0")
            if entry(data.get("_schedule", "")) != entry(schedule):
                return None
        e = data.get(kernel)
        return {k: v for k, v in e.items() if k != "kernel"} if e else None
    except Exception:
        return None


def measured_hbm(traffic, avg_launch_ms):
    """Measured HBM rate of the dominant kernel (PMC bytes per dispatch / launch time) and its
    fraction of the 8 TB/s peak, next to the algorithmic figure; Nones without a PMC summary."""
    if not traffic or not avg_launch_ms:
        return {"traffic_GBps": None, "traffic_frac": None}
    gbps = traffic / (avg_launch_ms * 1e-3) / 1e9
    return {"traffic_GBps": gbps, "traffic_frac": gbps / HBM_PEAK_GBS}


def cpu_baseline(s, c, kxg, kyg, times_sample, gpu_fields=None):
    """FP64 NumPy oracle (1 core: numpy.fft is single-threaded) on a bounded sample of instants."""
    from oracle import gl_oracle as O

    k = np.hypot(s.kx, s.ky)
    om = np.sqrt(s.gravity * k * np.tanh(s.depth * k))
    qx, qy = s.depth * kxg, s.depth * kyg
    pot = s.depth * math.sqrt(s.gravity * s.depth)
    elapsed = 0.0
    errs = {"eta33": 0.0, "psi33": 0.0}
    for i, t in enumerate(times_sample):
        spec = O.deposit(s.a, s.b, s.kx, s.ky, om, s.depth, s.lx, s.ly, s.nx, s.ny, float(t))
        t0 = time.perf_counter()
        e, p, _ = O.order3(spec, qx, qy, peak_q=c["peak_depth"])
        elapsed += time.perf_counter() - t0
        if gpu_fields is not None:
            ge, gp = gpu_fields[i]
            errs["eta33"] = max(errs["eta33"], float(np.linalg.norm(ge - s.depth * e) / np.linalg.norm(s.depth * e)))
            errs["psi33"] = max(errs["psi33"], float(np.linalg.norm(gp - pot * p) / np.linalg.norm(pot * p)))
    return len(times_sample) / elapsed, errs


def bench_order2(args):
    """Plain order 2 at 128^2 (eta22 GL rank 6 + dual-branch psi22), FP64: (eta22, psi22) field pairs per
    second over 64 instants per step, the CPU restatement timed beside it on a bounded sample."""
    import torch

    from green_laplace_wave_reconstruction_amd import api, plan as PL, seastate
    from green_laplace_wave_reconstruction_amd._native import DevicePlan

    s = seastate.order2_plain_128()
    c = api.gl_spectral_coefficients(2, s.gravity, s.depth, s.a, s.b, s.kx, s.ky, 0, 0)
    kxg, kyg = api._wavenumber_grid(s.lx, s.ly, s.nx, s.ny)
    spec, omega = api.deposit_analytic_input(c, s.lx, s.ly, s.nx, s.ny)
    host = PL.build_plan(spec, s.depth * kxg, s.depth * kyg, depth=s.depth, gravity=s.gravity, order=2,
                         peak_q=c["peak_depth"])
    host.omega = host.box1.take(omega)
    nt = INSTANTS
    times = np.linspace(0.0, 10.0 * 10.0, nt)  # 64 instants over ten peak periods (Tp = 10 s)
    dp = DevicePlan(host, host.box1.take(spec), max_instants=nt)
    schedule = dp.autotune("order2", times, reps=3) if not args.no_tune else "default (plan 0, tile 0)"
    out = None
    for _ in range(args.warmup):
        out = dp.run("order2", times, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = dp.run("order2", times, out)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dp.profile(True)
    dp.stage_times()
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for _ in range(args.steps):
        dp.run("order2", times, out)
    torch.cuda.synchronize()
    wall_profiled = time.perf_counter() - tp
    stages = dp.stage_times()
    dp.profile(False)
    roof, pipe = stage_roofline(stages, args.steps, wall_profiled / args.steps * 1e3)
    cpu = None
    if not args.no_cpu:
        from oracle import gl_oracle as O

        sample = list(range(nt))  # every instant: about 5 s of CPU work
        k = np.hypot(s.kx, s.ky)
        om = np.sqrt(s.gravity * k * np.tanh(s.depth * k))
        qx, qy = s.depth * kxg, s.depth * kyg
        pot = s.depth * math.sqrt(s.gravity * s.depth)
        eta, psi = out[0].cpu().numpy(), out[1].cpu().numpy()
        elapsed, errs = 0.0, {"eta22": 0.0, "psi22": 0.0}
        for i in sample:
            sp = O.deposit(s.a, s.b, s.kx, s.ky, om, s.depth, s.lx, s.ly, s.nx, s.ny, float(times[i]))
            t1 = time.perf_counter()
            e22, _ = O.eta22(sp, qx, qy, 6, c["peak_depth"])
            p22, _ = O.psi22(sp, qx, qy, False, c["peak_depth"])
            elapsed += time.perf_counter() - t1
            errs["eta22"] = max(errs["eta22"], rel_l2(eta[i], s.depth * e22))
            errs["psi22"] = max(errs["psi22"], rel_l2(psi[i], pot * p22))
        cpu = {"value": len(sample) / elapsed, "unit": "field pairs/s", "cores": 1, "kind": "port",
               "sample": f"all {nt} instants of the step: FP64 NumPy oracle restatement of _green_laplace_eta22 "
                         "(rank 6) and _green_laplace_psi22, numpy.fft single-threaded",
               "gpu_rel_l2_vs_oracle": errs, **host_info()}
    return {
        "metric": "eta22+psi22 field pairs/s at 128^2 (plain order 2, FP64)", "value": nt * args.steps / wall,
        "unit": "field pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic seeded JONSWAP sea state (numpy default_rng seed 1)",
        "config": {"workload": f"GL order 2 (eta22 GL6 + psi22), 128x128, L=1000 m, h=50 m, Hs=4 m, Tp=10 s, "
                               f"gamma=3.3, s=10, {s.count} parents, {nt} instants over [0, 100 s]",
                   "instants_per_step": nt, "fft_schedule": schedule,
                   "working_grids": {"quadratic": list(host.g2[::-1]) if host.g2 else None}},
        "roofline": roof,
        "pipeline": pipe,
        "cpu_baseline": cpu,
    }


def bench_eta44(args):
    """eta44 GL rank J on the 256^2 dimensionless order-4 sea state (one field per step)."""
    import torch

    from green_laplace_wave_reconstruction_amd import api, plan as PL, seastate
    from green_laplace_wave_reconstruction_amd._native import DevicePlan

    s = seastate.order4_dimensionless_256()
    c = api.gl_spectral_coefficients(1, s.gravity, s.depth, s.a, s.b, s.kx, s.ky, 0, 0)
    kxg, kyg = api._wavenumber_grid(s.lx, s.ly, s.nx, s.ny)
    spec, _ = api.deposit_analytic_input(c, s.lx, s.ly, s.nx, s.ny)
    host = PL.build_order4_plan(spec, s.depth * kxg, s.depth * kyg, peak_q=c["peak_depth"], rank=args.rank)
    dp = DevicePlan(host, host.box1.take(spec), max_instants=1)
    t = np.array([0.0])
    for _ in range(args.warmup):
        dp.run("order4", t)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dp.run("order4", t)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dp.profile(True)
    dp.stage_times()
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for _ in range(args.steps):
        dp.run("order4", t)
    torch.cuda.synchronize()
    wall_profiled = time.perf_counter() - tp
    stages = dp.stage_times()
    roof, pipe = stage_roofline(stages, args.steps, wall_profiled / args.steps * 1e3)
    cpu = None
    if not args.no_cpu:
        from oracle import gl_oracle as O

        t1 = time.perf_counter()
        O.eta44(spec, s.depth * kxg, s.depth * kyg, c["peak_depth"], args.rank)
        cpu = {"value": 1.0 / (time.perf_counter() - t1), "unit": "fields/s", "cores": 1, "kind": "port",
               "sample": f"one GL{args.rank} eta44 field, FP64 NumPy oracle restatement, numpy.fft single-threaded",
               **host_info()}
    return {
        "metric": f"eta44 GL{args.rank} fields/s at 256^2", "value": args.steps / wall, "unit": "fields/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic seeded JONSWAP sea state (numpy default_rng seed 5)",
        "config": {"workload": f"eta44 GL{args.rank}, 256x256, g=h=1, kp h=1.5, kp Hs/2=0.12, L=5 peak wavelengths, "
                               f"{s.count} parents", "working_grids": host.audits["eta44"]["working_grids"],
                   "reference_fft_ifft_count": host.audits["eta44"]["reference_fft_ifft_count"]},
        "roofline": roof,
        "pipeline": pipe,
        "cpu_baseline": cpu,
    }


def bench_batch(args):
    """256 seeded uneven sea states (GL orders 2-3, N in 64..1024) per step, one graph launch per step."""
    import torch

    from green_laplace_wave_reconstruction_amd import batch as B, plan as PL, seastate

    draws = seastate.batch_sea_states(args.states)
    t0 = time.perf_counter()
    hosts = [B.state_plan(d, PL.PREC_FP64) for d in draws]
    host_s = time.perf_counter() - t0
    print(f"batch: {len(draws)} host plans built in {host_s:.1f} s", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    bt = B.SeaStateBatch(draws, precision="fp64", branches=args.branches, tune=True, reps=3, hosts=hosts)
    setup_s = time.perf_counter() - t1
    print(f"batch: plans, tuning and graph capture in {setup_s:.1f} s", file=sys.stderr, flush=True)
    # eager baseline: the same plans and schedules, one execute call after another on one stream
    for e in bt.entries:
        e.plan.run(e.kind, e.times, out=(e.a, e.b))
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for e in bt.entries:
        e.plan.run(e.kind, e.times, out=(e.a, e.b))
    torch.cuda.synchronize()
    eager_s = time.perf_counter() - t2
    for _ in range(args.warmup):
        bt.launch()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    for _ in range(args.steps):
        bt.launch()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t3
    ref_fields = [(e.a[0].cpu().numpy(), e.b[0].cpu().numpy()) for e in bt.entries]
    # per-stage compulsory bytes and times: one eager pass with event brackets (after capture)
    agg: dict = {}
    for e in bt.entries:
        e.plan.profile(True)
        e.plan.stage_times()
        e.plan.run(e.kind, e.times, out=(e.a, e.b))
        for k, (ms, n, moved) in e.plan.stage_times().items():
            prev = agg.get(k, (0.0, 0, 0.0))
            agg[k] = (prev[0] + ms, prev[1] + n, prev[2] + moved)
        e.plan.profile(False)
    torch.cuda.synchronize()
    ms_per_step = wall / args.steps * 1e3
    roof, pipe = stage_roofline(agg, 1, ms_per_step)
    roof["kernel"] += " (summed over the batch's plans, eager launches)"
    pipe["note"] = "compulsory bytes of every stage of every plan / graph-launch time per batch"
    sched = sorted({v for v in bt.schedules.values()})
    r64 = {"states_per_s": len(draws) * args.steps / wall, "ms_per_step": ms_per_step,
           "equiv_fullgrid_GBps": bt.alg_bytes() * args.steps / wall / 1e9,
           "eager_ms_per_batch": eager_s * 1e3, "setup_s": setup_s, "distinct_schedules": len(sched),
           "schedules": {f"N{k[2]}_order{k[0]}": v for k, v in bt.schedules.items()}}
    bt.close()
    del bt
    torch.cuda.empty_cache()
    cpu = None
    if not args.no_cpu:
        from oracle import gl_oracle as O

        # bounded sample: the first states of N <= 256 up to ~15 s of oracle time
        sample, el, errs = [], 0.0, []
        for i, d in enumerate(draws):
            if d.state.nx > 256:
                continue
            s = d.state
            t4 = time.perf_counter()
            r = O.gl_spectral_surface(d.order, s.gravity, s.depth, s.a, s.b, s.kx, s.ky, s.lx, s.ly, s.nx, s.ny, 0.0)
            el += time.perf_counter() - t4
            c = r["components"]
            ra, rb = c[f"eta{d.order}{d.order}_plus"], c[f"psi{d.order}{d.order}_plus"]
            errs.append(max(rel_l2(ref_fields[i][0], ra), rel_l2(ref_fields[i][1], rb)))
            sample.append(i)
            if el > 15.0:
                break
        cpu = {"value": len(sample) / el, "unit": "states/s", "cores": 1, "kind": "port",
               "sample": f"{len(sample)} batch states with N<=256 (indices {sample}), FP64 NumPy oracle "
                         "gl_spectral_surface (includes the lower orders it needs), numpy.fft single-threaded",
               "gpu_fp64_rel_l2_max": float(max(errs)) if errs else None, **host_info()}
    grids = sorted({(d.state.nx, d.order) for d in draws})
    return {
        "metric": f"batched uneven GL orders 2-3: sea states/s ({len(draws)} states, one graph launch)",
        "value": r64["states_per_s"], "unit": "states/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic seeded uneven JONSWAP batch (numpy default_rng seed 3)",
        "config": {"workload": f"{len(draws)} sea states, N in 64..1024, orders 2-3, t=0, graph with "
                               f"{args.branches} branches, autotuned FFT schedules",
                   "grid_orders": [list(g) for g in grids], "host_plan_build_s": host_s},
        "roofline": roof, "pipeline": pipe,
        "fp64": r64, "fp32": {"status": "refused", "reason": FP32_ARM["reason"]},
        "cpu_baseline": cpu,
    }


def bench_inverse(args):
    """Inverse reconstruction at 256^2: one step = misfit gradient (forward eta11+eta22+eta33 + adjoint)."""
    import torch

    from green_laplace_wave_reconstruction_amd import seastate
    from green_laplace_wave_reconstruction_amd.inverse import GLElevationModel, gl_elevation

    s = seastate.inverse_256()
    m = GLElevationModel(s.kx, s.ky, s.a, s.b, lx=s.lx, ly=s.ly, nx=s.nx, ny=s.ny, depth=s.depth,
                         gravity=s.gravity, order=3)
    m_graph = GLElevationModel(s.kx, s.ky, s.a, s.b, lx=s.lx, ly=s.ly, nx=s.nx, ny=s.ny, depth=s.depth,
                               gravity=s.gravity, order=3, graphs=True)
    at = torch.tensor(s.a, device=m.device)
    bt = torch.tensor(s.b, device=m.device)
    target = m.elevation(at, bt)
    rng = np.random.default_rng(14)
    a = (at * (1 + 0.2 * torch.tensor(rng.standard_normal(s.count), device=m.device))).float().requires_grad_()
    b = (bt * (1 + 0.2 * torch.tensor(rng.standard_normal(s.count), device=m.device))).float().requires_grad_()

    def step():
        if a.grad is not None:
            a.grad = None
            b.grad = None
        loss = 0.5 * ((gl_elevation(a, b, m) - target) ** 2).sum()
        loss.backward()
        return loss

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the same step with the forward and adjoint replayed as HIP graphs, for comparison
    model_eager = m
    m = m_graph
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0g = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    graph_rate = args.steps / (time.perf_counter() - t0g)
    m = model_eager
    # per-stage compulsory bytes and times of the plan's launches (forward and adjoint), further eager steps
    m.plan.profile(True)
    m.plan.stage_times()
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall_profiled = time.perf_counter() - tp
    roof, pipe = stage_roofline(m.plan.stage_times(), args.steps, wall_profiled / args.steps * 1e3)
    m.plan.profile(False)
    pipe["note"] = "plan launches only (forward eta11+eta22+eta33 and adjoint); the step's torch work is not counted"
    # the 200-step Adam fit (timed as a whole)
    a.data.copy_((at * (1 + 0.2 * torch.tensor(np.random.default_rng(14).standard_normal(s.count),
                                               device=m.device))).float())
    b.data.copy_((bt * (1 + 0.2 * torch.tensor(np.random.default_rng(15).standard_normal(s.count),
                                               device=m.device))).float())
    amp = float(torch.sqrt((at**2 + bt**2).max()))
    opt = torch.optim.Adam([a, b], lr=3e-2 * amp)
    sched = torch.optim.lr_scheduler.ExponentialLR(opt, 0.97)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    first = None
    for _ in range(200):
        opt.zero_grad()
        loss = step()
        opt.step()
        sched.step()
        first = float(loss.detach()) if first is None else first
    torch.cuda.synchronize()
    adam_s = time.perf_counter() - t1
    final = float(0.5 * ((m.elevation(a.detach().double(), b.detach().double()) - target) ** 2).sum())
    err = float(torch.sqrt(((a.detach().double() - at) ** 2 + (b.detach().double() - bt) ** 2).sum()
                           / (at**2 + bt**2).sum()))
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        import math as _m

        from oracle import gl_oracle as O
        from torch_gl import TorchGL, deposit

        k = np.hypot(s.kx, s.ky)
        om = np.sqrt(s.gravity * k * np.tanh(s.depth * k))
        spec = O.deposit(s.a, s.b, s.kx, s.ky, om, s.depth, s.lx, s.ly, s.nx, s.ny, 0.0)
        kxg, kyg = O.wavenumber_grid(s.lx, s.ly, s.nx, s.ny)
        tg = TorchGL(spec, s.depth * kxg, s.depth * kyg, depth=s.depth, order=3, peak_q=m.host.peak_q)
        mx = np.rint(s.kx * s.lx / (2 * _m.pi)).astype(int)
        my = np.rint(s.ky * s.ly / (2 * _m.pi)).astype(int)
        tgt = target.cpu()
        ac = torch.tensor(s.a, requires_grad=True)
        bc = torch.tensor(s.b, requires_grad=True)
        t2 = time.perf_counter()
        lc = 0.5 * ((tg.elevation(deposit(ac, bc, mx, my, s.nx, s.ny, s.depth)) - tgt) ** 2).sum()
        lc.backward()
        el = time.perf_counter() - t2
        cpu = {"value": 1.0 / el, "unit": "gradients/s", "cores": torch.get_num_threads(), "kind": "port",
               "sample": "one misfit gradient by torch-CPU autograd of the full-grid FP64 restatement "
                         "(tests/torch_gl.py), torch.fft on the host threads", **host_info()}
    return {
        "metric": "inverse reconstruction: misfit gradients/s at 256^2 (eta11+eta22+eta33, hand-written adjoint)",
        "value": args.steps / wall, "unit": "gradients/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic seeded JONSWAP sea state (numpy default_rng seed 4), hidden amplitudes",
        "config": {"workload": f"256x256, h=40 m, Hs=5 m, Tp=11 s, L=1700 m, {s.count} parents; amplitudes float32, "
                               "forward and adjoint FP64"},
        "roofline": roof, "pipeline": pipe,
        "graph_gradients_per_s": graph_rate,
        "adam_200_steps": {"seconds": adam_s, "misfit_first": first, "misfit_final": final,
                           "amplitude_rel_error_final": err, "amplitude_rel_error_initial": 0.2},
        "cpu_baseline": cpu,
    }


def bench_api(args):
    """The whole public call at the headline sea state: gl_spectral_surface(coeffs, 2000, 2000, 512, 512, t)
    for the 64 instants -- host validation, deposit and FP64 plan (NumPy), device plan, eta11/psi11,
    eta22/psi22 and eta33/psi33 on the GPU, every component copied to host arrays."""
    import torch

    import green_laplace_wave_reconstruction_amd as G
    from green_laplace_wave_reconstruction_amd import seastate

    s = seastate.order3_headline_512()
    times = seastate.headline_instants()
    c = G.gl_spectral_coefficients(3, s.gravity, s.depth, s.a, s.b, s.kx, s.ky, 0, 0)
    for _ in range(args.warmup):
        G.gl_spectral_surface(c, s.lx, s.ly, s.nx, s.ny, times)
    torch.cuda.synchronize()
    walls = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        eta, psi, X, Y, comps, audit = G.gl_spectral_surface(c, s.lx, s.ly, s.nx, s.ny, times)
        walls.append(time.perf_counter() - t0)
    wall = sum(walls)
    return {
        "metric": "gl_spectral_surface calls at 512^2, order 3, 64 instants: instants/s end to end "
                  "(host plan + GPU + copies of every component to host)",
        "value": INSTANTS * args.steps / wall, "unit": "instants/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic seeded JONSWAP sea state (numpy default_rng seed 2)",
        "config": {"workload": "gl_spectral_surface(coeffs, 2000, 2000, 512, 512, t) with t = the 64 headline "
                               "instants; returns 14 host arrays [64][512][512] (eta, psi and the *_plus / real "
                               "components of orders 1-3)",
                   "min_ms": min(walls) * 1e3, "max_ms": max(walls) * 1e3,
                   "host_bytes_returned": int(sum(v.nbytes for v in comps.values()))},
        "roofline": None,
        "cpu_baseline": None,
    }


MODAL_POINTS = 8192   # default modal grid of the bench (--modal-points)


def _modal_alg_bytes(d) -> dict:
    """Algorithmic HBM bytes per launch of each modal stage: every x-space column written once and read
    once (16 B per complex value): sources 2 x W2 x Q, pair fields 5 x W2 x Q, forcing 2 x W3 x Q."""
    q, w2, w3 = d.q_len, 2 * d.span + 1, 3 * d.span + 1
    return {"time_pass_o2": 16 * 2 * w2 * q, "modal_pass_eta22": 16 * (2 + 5) * w2 * q,
            "time_pass_o3": 16 * (5 * w2 + 2 * w3) * q, "modal_pass_eta33": 16 * 2 * w3 * q}


def bench_modal(args):
    """Single-direction modal time path: one step = one modal level (8192 points by default) (eta22 + eta33, GL8) and
    eta20 (GL16, 128 x 16) of a 2048-sample k_p h = 1 JONSWAP record over 50 T_p, deposits resident."""
    import torch

    from green_laplace_wave_reconstruction_amd import modal as MD
    from green_laplace_wave_reconstruction_amd import seastate

    points = int(args.modal_points)
    eta1, t, h = seastate.unidirectional_record()
    g = 9.81
    band = MD._select_band(eta1, t, depth=h, gravity=g, order=3, omega_max=None, energy_fraction=0.999)
    dev = MD.ModalDevice(band.count)
    lp = MD.modal_level_plan(band, depth=h, gravity=g, rank=8, order=3, points=points, modal_limit=None)
    dep = dev.upload(lp.deposit)
    d20, p20, _ = MD.eta20_plan(band.amplitudes, band.omega, band.q, band.count, band.dt, depth=h, gravity=g,
                                rank=16, nx=128, ny=16)
    dep20 = dev.upload(p20)

    def step():
        out = dev.modal(lp.desc, dep, eta11=True)
        return out, dev.eta20(d20, dep20)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # per-stage durations (HIP events on the launch stream)
    stages = []
    for _ in range(5):
        dev.modal(lp.desc, dep, eta11=True, stage_ms=stages)
    names = ("twiddles", "time_pass_o2", "modal_pass_eta22", "time_pass_o3", "modal_pass_eta33", "synthesis")
    mean_ms = {n: float(np.mean([st[i] for st in stages])) for i, n in enumerate(names)}
    alg = _modal_alg_bytes(lp.desc)
    dominant = max(alg, key=lambda n: mean_ms[n])
    achieved = alg[dominant] / (mean_ms[dominant] * 1e-3) / 1e9
    # counter-measured bytes exist for the default grid only
    traffic = traffic_from_profiles(dominant, "pmc_traffic_modal.json") if points == MODAL_POINTS else None
    # end-to-end public calls (host analysis, planning, device, copies back), after one untimed call
    api = None
    if not args.no_api:
        MD.reconstruct_unidirectional_modal_timeseries(eta1, t, depth=h, modal_points=points)
        t1 = time.perf_counter()
        MD.reconstruct_unidirectional_modal_timeseries(eta1, t, depth=h, modal_points=points)
        api_s = time.perf_counter() - t1
        t2 = time.perf_counter()
        lad = MD.reconstruct_unidirectional_modal_timeseries(eta1, t, depth=h, modal_points=(2048, 4096, 8192))
        ladder_s = time.perf_counter() - t2
        api = {"one_level": api_s, "ladder_2048_4096_8192": ladder_s, "ladder_converged_at": lad.audit["modal_points"]}
    cpu = None
    if not args.no_cpu:
        from oracle import modal_oracle as M

        ob = M.select_band(eta1, t, depth=h, gravity=g, order=3, omega_max=None, energy_fraction=0.999)
        t3 = time.perf_counter()
        M.modal_level(ob, depth=h, gravity=g, rank=8, order=3, points=points, limit=None)
        M.eta20(ob.amplitudes, ob.omega, ob.q, t, depth=h, gravity=g, rank=16, nx=128, ny=16)
        el = time.perf_counter() - t3
        cpu = {"value": 1.0 / el, "unit": "records/s", "cores": 1, "kind": "port",
               "sample": f"one {points}-point level + eta20 of the same record by the NumPy oracle "
                         "(oracle/modal_oracle.py, pocketfft, one thread)", **host_info()}
    return {
        "metric": f"single-direction modal time path: records/s at {points} modal points "
                  "(eta22 + eta33 GL8, eta20 GL16)",
        "value": args.steps / wall, "unit": "records/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic seeded unidirectional JONSWAP record (numpy default_rng seed 11)",
        "config": {"workload": f"{band.count} samples over 50 T_p, k_p h = 1, h = {h} m, Hs = 1.5 m, gamma = 3.3, "
                               f"{band.bins.size} bins (99.9% energy), order 3, {points} modal points",
                   "reduced_grid": {"q_len": int(lp.desc.q_len), "t_len": int(lp.desc.t_len),
                                    "reference_grid": [points, lp.work_count]},
                   "eta20_grid": {"q_len": int(d20.q_len), "t_len": int(d20.t_len)}},
        "stage_ms": mean_ms,
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": traffic,
                     "alg_bytes_per_launch": alg[dominant], "avg_launch_ms": mean_ms[dominant],
                     **measured_hbm(traffic, mean_ms[dominant])},
        "api_seconds": api,
        "reference_published": {"seconds_8192_points": [6.78, 6.91], "ladder_median_seconds": 12.07,
                                "source": "python/docs/random_hos_modal_comparison.md (HOS records, other machine)"}
        if points == MODAL_POINTS else None,
        "cpu_baseline": cpu,
    }


def rel_l2(a, b) -> float:
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def bench_wit(args):
    """WIT sector-44 quartet sum on the reference's frozen DIR4IN01 inputs (512^2)."""
    import torch

    from green_laplace_wave_reconstruction_amd import formats, wit

    name = f"n{args.modes:04d}"
    inp = formats.read_dir4in(ROOT / "tests" / "golden" / f"wit_{name}_input.bin")
    plan = wit.WitPlan(inp["modes"], nx=inp["nx"], ny=inp["ny"], gravity=inp["gravity"], depth=inp["depth"],
                       dkx=inp["dkx"], dky=inp["dky"], threshold=inp["threshold"], precision=args.precision)
    out = None
    for _ in range(args.warmup):
        out, cnt = plan.sector44_device(inp["amplitudes"], out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, cnt = plan.sector44_device(inp["amplitudes"], out)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    c = cnt.cpu().numpy()
    vs64 = None
    if args.precision == "fp32":  # the FP64 sum of the same quartets, untimed
        p64 = wit.WitPlan(inp["modes"], nx=inp["nx"], ny=inp["ny"], gravity=inp["gravity"], depth=inp["depth"],
                          dkx=inp["dkx"], dky=inp["dky"], threshold=inp["threshold"])
        o64, _ = p64.sector44_device(inp["amplitudes"])
        torch.cuda.synchronize()
        vs64 = {"eta": rel_l2(out[0].cpu().numpy(), o64[0].cpu().numpy()),
                "psi": rel_l2(out[1].cpu().numpy(), o64[1].cpu().numpy())}
        p64.close()
    cpu = None
    if not args.no_cpu:
        from oracle import wit as ow

        m = inp["modes"].shape[0]
        cores = len(os.sched_getaffinity(0))
        # bounded sample: the last n first indices hold C(n+3, 4) quartets; aim at ~10 s of CPU work
        target = 10.0 * cores / 40e-6
        n = 1
        while n < m and (n + 1) * (n + 2) * (n + 3) * (n + 4) / 24 <= target:
            n += 1
        ow.sector44(dict(inp, modes=inp["modes"][:4], amplitudes=inp["amplitudes"][:, :4]))  # OpenMP warm-up
        t1 = time.perf_counter()
        r = ow.sector44(inp, (m - n, m))
        el = time.perf_counter() - t1
        cpu = {"value": r["kept"] / el, "unit": "quartets/s", "cores": cores, "kind": "port",
               "sample": f"quartets with first index in [{m - n}, {m}) ({int(r['kept'])} of {int(c[0])} quartets), "
                         "C++/OpenMP restatement of DirectionalKernel::evaluate", **host_info()}
    return {
        "metric": f"WIT sector-44 quartets/s, M={inp['modes'].shape[0]}", "value": c[1] * args.steps / wall,
        "unit": "quartets/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 evaluation, f64 accumulation",
        "rel_l2_vs_fp64": vs64,
        "data": f"reference frozen input paper/order4/inputs/runtime/{name} (copied as a fixture)",
        "config": {"workload": f"WIT sector 44, M={inp['modes'].shape[0]}, {inp['nx']}x{inp['ny']}",
                   "quartets": int(c[0])},
        "roofline": _wit_roofline(args.precision, c[1] * args.steps / wall),
        "cpu_baseline": cpu,
    }


def bench_wit_signed(args):
    """WIT sector 40 (signs + + - -) or 42 (+ + + -) on the reference's frozen DIR4IN01 inputs
    (512^2): every candidate of the sector's index space (P^2 pairs of pairs or T x M triples and
    modes), FP64. The CPU leg runs the C++/OpenMP restatement (oracle/wit_oracle.cpp) on a bounded
    slice of the same input's outer candidate loop."""
    import torch

    from green_laplace_wave_reconstruction_amd import formats, wit

    if args.precision != "fp64":
        raise SystemExit("sectors 40 and 42 run in FP64 only")
    name = f"n{args.modes:04d}"
    inp = formats.read_dir4in(ROOT / "tests" / "golden" / f"wit_{name}_input.bin")
    plan = wit.WitPlan(inp["modes"], nx=inp["nx"], ny=inp["ny"], gravity=inp["gravity"], depth=inp["depth"],
                       dkx=inp["dkx"], dky=inp["dky"], threshold=inp["threshold"])
    out = cnt = None
    for _ in range(args.warmup):
        out, cnt = plan.sector_device(args.sector, inp["amplitudes"], out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, cnt = plan.sector_device(args.sector, inp["amplitudes"], out)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    c = cnt.cpu().numpy()
    rate = c[1] * args.steps / wall
    cpu = None
    if not args.no_cpu:
        from oracle import wit as ow

        m = inp["modes"].shape[0]
        cores = len(os.sched_getaffinity(0))
        npair, ntrip = m * (m + 1) // 2, m * (m + 1) * (m + 2) // 6
        inner = npair if args.sector == 40 else ntrip  # candidates per outer loop index
        # bounded sample: the first n outer indices, ~10 s of CPU work at ~18 core-us per candidate
        n = max(1, min(npair if args.sector == 40 else m, int(10.0 * cores / 18e-6 / inner)))
        small = dict(inp, modes=inp["modes"][:8], amplitudes=inp["amplitudes"][:, :8])
        ow.signed(small, args.sector)  # OpenMP warm-up
        t1 = time.perf_counter()
        r = ow.signed(inp, args.sector, (0, n))
        el = time.perf_counter() - t1
        loop = "minus pairs" if args.sector == 40 else "minus modes"
        # kept candidates per second, the unit of the GPU value (the zero-output, half-plane and alias
        # filters drop the rest before any evaluation)
        cpu = {"value": r["kept"] / el, "unit": "candidates/s", "cores": cores, "kind": "port",
               "sample": f"the first {n} {loop} of the M={m} sector-{args.sector} loop ({n * inner} candidates, "
                         f"{int(r['kept'])} kept; kept per second), C++/OpenMP restatement oracle/wit_oracle.cpp",
               **host_info()}
    return {
        "metric": f"WIT sector-{args.sector} candidates/s, M={inp['modes'].shape[0]}", "value": rate,
        "unit": "candidates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": f"reference frozen input paper/order4/inputs/runtime/{name} (copied as a fixture)",
        "config": {"workload": f"WIT sector {args.sector}, M={inp['modes'].shape[0]}, {inp['nx']}x{inp['ny']}",
                   "candidates": int(c[0]), "kept": int(c[1]), "near_resonant": int(c[2])},
        "roofline": _wit_roofline("fp64", rate, key=f"sector{args.sector} fp64", kernel=f"k_wit_signed_tab<{args.sector}>"),
        "cpu_baseline": cpu,
    }


def _wit_roofline(precision: str, quartets_per_s: float, key=None, kernel="k_wit_quartets"):
    """Vector-ALU roofline of the quartet kernel: FLOP per quartet measured with rocprofv3 (F64/F32
    add, mul and FMA wave instructions x 64 lanes, profiles/wit_m400_valu_counters.md) times the live
    quartet rate, against the 78.6 TFLOP/s FP64 (= non-packed FP32) vector peak."""
    p = ROOT / "profiles" / "wit_flops.json"
    if not p.exists():
        return None
    rec = json.loads(p.read_text()).get(key or ("fp64" if precision == "fp64" else "fp32 eval"))
    if not rec:
        return None
    achieved = rec["flop_per_quartet"] * quartets_per_s / 1e12
    return {"bound": "fp64 vector ALU", "kernel": kernel, "achieved": achieved, "peak": 78.6,
            "unit": "TFLOP/s", "frac": achieved / 78.6, "traffic": None,
            "flop_per_quartet": rec["flop_per_quartet"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="order3",
                    choices=("order3", "order2", "api", "eta44", "wit", "batch", "inverse", "modal"))
    ap.add_argument("--states", type=int, default=256, help="batch workload: sea states per step")
    ap.add_argument("--branches", type=int, default=4, help="batch workload: parallel graph branches")
    ap.add_argument("--rank", type=int, default=6, help="eta44 GL rank")
    ap.add_argument("--modes", type=int, default=200, help="WIT mode count (200, 400, 800)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 20 for order3, whose step is ~4 ms, 10 for the other workloads)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps first (default: 5 for order3, else 2)")
    ap.add_argument("--precision", default="fp64", choices=("fp64", "fp32"))
    ap.add_argument("--sector", type=int, default=44, choices=(40, 42, 44), help="wit: WIT sector")
    ap.add_argument("--cpu-sample", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="order3: skip the FFT schedule autotuner")
    ap.add_argument("--tune", action="store_true",
                    help="order3: autotune the FFT schedule now instead of using the stored (profiled) one")
    ap.add_argument("--modal-points", type=int, default=MODAL_POINTS,
                    help="modal: modal grid points of the level (8192; up to 65536, split into sub-transforms)")
    ap.add_argument("--no-api", action="store_true", help="modal: skip the end-to-end API timings (profiling runs)")
    ap.add_argument("--chunk", type=int, default=INSTANTS,
                    help="order3: instants per plan evaluation (the 64 instants run in chunks of this size)")
    ap.add_argument("--schedule", default="", help="order3: use this FFT schedule (stage=plan/tile;...) instead of tuning")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (HBM counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--calibrate-hbm", action="store_true",
                    help="time plain 4 GiB copy (read + write), fill (write) and sum (read) kernels: the HBM rates "
                         "a streaming kernel reaches on this device")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 20 if args.workload == "order3" else 10
    if args.warmup is None:
        args.warmup = 5 if args.workload == "order3" else 2
    if args.calibrate_hbm:
        import torch

        x = torch.ones(1 << 29, dtype=torch.float64, device="cuda")  # 4 GiB
        y = torch.empty_like(x)
        n = x.numel() * 8
        rates = {}
        for name, fn, nbytes in (("copy (read + write)", lambda: y.copy_(x), 2 * n), ("fill (write)", lambda: y.fill_(2.0), n),
                                 ("sum (read)", lambda: x.sum(), n)):
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(10):
                fn()
            e1.record()
            torch.cuda.synchronize()
            rates[name] = nbytes / (e0.elapsed_time(e1) / 10 * 1e-3) / 1e9
        # the library's own streaming kernels (16 bytes per thread, many short workgroups)
        import ctypes

        from green_laplace_wave_reconstruction_amd import _native

        lib = _native.load_library()
        probe = {}
        for name, kind, nbytes in (("copy (read + write)", 0, 2 * n), ("fill (write)", 1, n), ("read", 2, n)):
            ms = ctypes.c_double()
            _native._check(lib.glw_hbm_probe(kind, x.data_ptr(), y.data_ptr(), n, 10, ctypes.byref(ms)))
            probe[name] = nbytes / (ms.value * 1e-3) / 1e9
        print(json.dumps({"calibration": "hbm", "GBps_torch": rates, "GBps_glw_probe": probe,
                          "frac_of_8TBps": {"torch": {k: v / HBM_PEAK_GBS for k, v in rates.items()},
                                            "glw_probe": {k: v / HBM_PEAK_GBS for k, v in probe.items()}}}))
        return
    if args.calibrate_copy:
        import torch

        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read_per_dispatch": src.numel() * 8,
                          "bytes_written_per_dispatch": src.numel() * 8, "dispatches": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU only: the GL graph of one sea state is timed on one "
                                   "MI355X; --gpus must be 1"}))
        sys.exit(2)

    import torch

    from green_laplace_wave_reconstruction_amd import plan as PL
    from green_laplace_wave_reconstruction_amd._native import DevicePlan

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    if args.workload == "eta44":
        print(json.dumps(bench_eta44(args)))
        return
    if args.workload == "order2":
        print(json.dumps(bench_order2(args)))
        return
    if args.workload == "wit":
        print(json.dumps(bench_wit(args) if args.sector == 44 else bench_wit_signed(args)))
        return
    if args.workload == "batch":
        print(json.dumps(bench_batch(args)))
        return
    if args.workload == "modal":
        print(json.dumps(bench_modal(args)), flush=True)
        return
    if args.workload == "inverse":
        print(json.dumps(bench_inverse(args)))
        return
    if args.workload == "api":
        print(json.dumps(bench_api(args)))
        return
    s, c, kxg, kyg, spec, omega = workload()
    if args.precision != "fp64":
        print(json.dumps({"error": "order3: " + FP32_ARM["reason"], "fp32_arm": FP32_ARM}))
        sys.exit(2)
    host = PL.build_plan(spec, s.depth * kxg, s.depth * kyg, depth=s.depth, gravity=s.gravity, order=3,
                         peak_q=c["peak_depth"])
    host.omega = host.box1.take(omega)
    dp = DevicePlan(host, host.box1.take(spec), max_instants=args.chunk)
    times = torch.linspace(0.0, 10.0 * 12.0, INSTANTS, dtype=torch.float64, device=dp.device)
    shape = (INSTANTS, s.ny, s.nx)
    eta = torch.empty(shape, dtype=torch.complex128, device=dp.device)
    psi = torch.empty(shape, dtype=torch.complex128, device=dp.device)
    schedule = "default (plan 0, tile 0)"
    stored = stored_schedule()
    if args.schedule:  # a schedule from an earlier autotuned run (profiling runs replay it)
        dp.schedule = args.schedule
        schedule = dp.schedule
        source = "given"
    elif not args.no_tune and stored and not args.tune:
        # the autotuned schedule the committed HBM-traffic profile was taken with (FFT "wisdom"):
        # the measured traffic then belongs to the kernels timed here
        dp.schedule = stored
        schedule = dp.schedule
        source = "stored (profiles/pmc_traffic.json, autotuned on MI355X)"
    elif not args.no_tune:  # setup, outside the timed region
        schedule = dp.autotune("order3", times, out=(eta, psi), reps=3)
        source = "autotuned"
    else:
        source = "default"
    for _ in range(args.warmup):
        dp.run("order3", times, out=(eta, psi))
    torch.cuda.synchronize()
    # timed region: plain execute calls, no event brackets
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dp.run("order3", times, out=(eta, psi))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # per-stage times (HIP events around every launch, on the execute stream) and compulsory bytes,
    # from the same number of further steps
    # (one step holds 64 / chunk x ~16 launches; the event pool times up to 4096 launches per read)
    prof_steps = args.steps if args.chunk >= INSTANTS else max(1, args.steps * args.chunk // INSTANTS)
    dp.profile(True)
    dp.stage_times()
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for _ in range(prof_steps):
        dp.run("order3", times, out=(eta, psi))
    torch.cuda.synchronize()
    wall_profiled = time.perf_counter() - tp
    stages = dp.stage_times()
    dp.profile(False)
    # PCIe-inclusive rate (outside `value`): the step's eta and psi fields copied to pinned host memory
    pinned = [torch.empty(eta.shape, dtype=eta.dtype, pin_memory=True) for _ in range(2)]
    for _ in range(2):
        pinned[0].copy_(eta, non_blocking=True)
        pinned[1].copy_(psi, non_blocking=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        pinned[0].copy_(eta, non_blocking=True)
        pinned[1].copy_(psi, non_blocking=True)
    torch.cuda.synchronize()
    d2h_s = (time.perf_counter() - t1) / reps
    pcie = {"d2h_ms_per_step": d2h_s * 1e3, "bytes_per_step": 2 * eta.numel() * eta.element_size(),
            "fields_per_s_with_d2h": INSTANTS / (wall / args.steps + d2h_s),
            "note": "eta+psi of the 64 instants copied to pinned host memory after each step (serial)"}
    ms_per_step = wall / args.steps * 1e3
    value = INSTANTS * args.steps / wall
    roof, pipe = stage_roofline(stages, prof_steps, wall_profiled / prof_steps * 1e3)
    traffic = traffic_from_profiles(roof["kernel"], schedule=schedule)
    roof["traffic"] = traffic
    roof.update(measured_hbm(traffic, roof["avg_launch_ms"]))
    # the dominant stage's SQ counters (profiles/sq_counters.json, same schedule): a row kernel is
    # compute/latency-bound (VALU half busy at 2-4 waves per SIMD), not HBM-bound
    roof["sq_counters"] = sq_from_profiles(roof["kernel"], schedule=schedule)
    n2 = s.nx * s.ny

    sample_idx = np.linspace(0, INSTANTS - 1, max(1, args.cpu_sample)).round().astype(int)
    cpu = None
    errs = None
    if not args.no_cpu:
        tt = times.cpu().numpy()
        gpu_fields = [(eta[i].cpu().numpy(), psi[i].cpu().numpy()) for i in sample_idx]
        cpu_rate, errs = cpu_baseline(s, c, kxg, kyg, tt[sample_idx], gpu_fields)
        cpu = {"value": cpu_rate, "unit": "fields/s", "cores": 1, "kind": "port",
               "sample": f"{len(sample_idx)} of the {INSTANTS} instants (indices {sample_idx.tolist()}), "
                         "FP64 NumPy oracle restatement of green_laplace_order3, numpy.fft single-threaded",
               **host_info()}
    result = {
        "metric": "eta33+psi33 fields/s at 512², FP32-acc vs FP64 at rel-L2≤1e-5; %HBM roofline",
        "value": value,
        "unit": "fields/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic seeded JONSWAP sea state (numpy default_rng seed 2)",
        "config": {
            "workload": "GL order 3 (inner GL4 x outer GL4) eta33+psi33, 512x512, L=2000 m, h=30 m, Hs=6 m, "
                        "Tp=12 s, gamma=3.3, s=15, 3120 parents kh>0.5, 64 instants over [0, 120 s]",
            "instants_per_step": INSTANTS, "instants_per_chunk": args.chunk, "grid": [s.ny, s.nx],
            "precision": "fp64 (the FP32-transform arm is refused: fp32_arm)",
            "working_grids": {"quadratic": list(host.g2[::-1]), "cubic": list(host.g3[::-1])},
            "fft_schedule": schedule, "fft_schedule_source": source,
        },
        "roofline": roof,
        "pipeline": pipe,
        "equiv_fullgrid_GBps": 716 * 2 * n2 * 16 * value / 1e9,
        "rel_l2_vs_oracle": errs,
        "fp32_arm": FP32_ARM,
        "cpu_baseline": cpu,
        "pcie_inclusive": pcie,
    }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
