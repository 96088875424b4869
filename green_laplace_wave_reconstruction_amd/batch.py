"""Uneven batches of sea states evaluated by one HIP graph launch.

Each state of the batch has its own grid (any supported N), GL order (2: eta22
and psi22; 3: eta33 and psi33) and plan.  The FFT schedule of every plan comes
from the autotuner (``glw_plan_autotune``): one representative plan per
distinct shape (order, precision, output grid, working grids, GL ranks) is
tuned on the device and its per-stage schedule is applied to the others.
All execute calls are then captured into one graph (``glw_batch_create``),
which ``launch`` replays with a single ``hipGraphLaunch``.

The per-state physics is exactly that of ``gl_spectral_surface``
(src/gl_spectral_surface.m:1; python/src/green_laplace/spatial.py:90-131):
the batch only changes how the launches are issued.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import api
from . import plan as PL
from ._native import DevicePlan, GraphBatch
from .seastate import BatchState

#: full-grid transforms per field pair of the reference graphs, for the
#: algorithmic-byte count B_alg = T x 4 N^2 x (bytes per complex value):
#: order 2 = eta22 at GL rank 6 (61) + psi22 (100); order 3 = 358
#: (audit core_fft_ifft_count, src/internal/gl_no_stokes_eta33.m:251)
REFERENCE_TRANSFORMS = {2: 161, 3: 358}


@dataclass
class BatchEntry:
    draw: BatchState
    host: PL.HostPlan
    plan: DevicePlan
    times: object
    a: object
    b: object
    signature: tuple = field(default=())

    @property
    def kind(self) -> str:
        return f"order{self.draw.order}"


def state_plan(draw: BatchState, precision: int) -> tuple[PL.HostPlan, np.ndarray]:
    """Host plan and t=0 parent-box spectrum of one batch state."""
    s = draw.state
    c = api.gl_spectral_coefficients(draw.order, s.gravity, s.depth, s.a, s.b, s.kx, s.ky, 0, 0)
    kxg, kyg = api._wavenumber_grid(s.lx, s.ly, s.nx, s.ny)
    spec, omega = api.deposit_analytic_input(c, s.lx, s.ly, s.nx, s.ny)
    host = PL.build_plan(spec, s.depth * kxg, s.depth * kyg, depth=s.depth, gravity=s.gravity, order=draw.order,
                         peak_q=c["peak_depth"], precision=precision)
    host.omega = host.box1.take(omega)
    return host, host.box1.take(spec)


def _signature(draw: BatchState, host: PL.HostPlan) -> tuple:
    return (draw.order, host.precision, host.nx, host.ny, tuple(host.g2 or ()), tuple(host.g3 or ()))


class SeaStateBatch:
    """Plans, tuned schedules and the captured graph of an uneven batch.

    ``precision``: "fp64" (FP32 transforms are refused, as by
    ``gl_spectral_surface``).  ``times``: the instants evaluated for every
    state.  ``branches``: parallel chains in the graph.
    """

    def __init__(self, draws: list[BatchState], *, precision: str = "fp64", times=(0.0,), branches: int = 4,
                 tune: bool = True, reps: int = 3, hosts=None):
        import dataclasses

        import torch

        self.precision = precision
        prec = api._precision(precision)
        nt = len(times)
        self.entries: list[BatchEntry] = []
        for i, d in enumerate(draws):
            if hosts is not None:  # host plans built once serve both precisions
                host, s0 = hosts[i]
                host = dataclasses.replace(host, precision=prec)
            else:
                host, s0 = state_plan(d, prec)
            dp = DevicePlan(host, s0, max_instants=nt)
            t = dp.times_tensor(np.asarray(times, dtype=float))
            shape = (nt, host.ny, host.nx)
            a = torch.empty(shape, dtype=torch.complex128, device=dp.device)
            b = torch.empty(shape, dtype=torch.complex128, device=dp.device)
            self.entries.append(BatchEntry(d, host, dp, t, a, b, _signature(d, host)))
        self.schedules: dict[tuple, str] = {}
        if tune:
            for e in self.entries:
                if e.signature not in self.schedules:
                    self.schedules[e.signature] = e.plan.autotune(e.kind, e.times, out=(e.a, e.b), reps=reps)
                e.plan.schedule = self.schedules[e.signature]
        for e in self.entries:  # one eager run per plan sizes its scratch before capture
            e.plan.run(e.kind, e.times, out=(e.a, e.b))
        torch.cuda.synchronize()
        self.graph = GraphBatch([(e.plan, e.kind, e.times, e.a, e.b) for e in self.entries], branches=branches)

    def launch(self) -> None:
        """Evaluate every state (one graph launch on the current stream)."""
        self.graph.launch()

    def fields(self, i: int):
        """(a, b) device tensors [nt, N, N] of state i: (eta22, psi22) or (eta33, psi33)."""
        e = self.entries[i]
        return e.a, e.b

    def alg_bytes(self) -> int:
        """Sum over states of B_alg = T x 4 N^2 x 16 B per instant (the reference's full-grid FP64 transforms)."""
        b = 16
        return int(sum(REFERENCE_TRANSFORMS[e.draw.order] * 4 * e.host.nx * e.host.ny * b * int(e.times.numel())
                       for e in self.entries))

    def close(self) -> None:
        self.graph.close()
        for e in self.entries:
            e.plan.close()
