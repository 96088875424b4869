"""Differentiable GL elevation for inverse reconstruction.

``GLElevationModel`` evaluates the physical elevation
``eta = Re(eta11+ + eta22+ + eta33+)`` of MF12-style amplitudes ``(a, b)`` on
fixed modes (src/gl_spectral_surface.m:1-102 at one instant) and its exact
gradient through the hand-written adjoint (``glw_adjoint_execute``), and
``gl_elevation`` wraps both in a ``torch.autograd.Function``.  With ``graphs=True``
the forward (phase, eta11, eta22, eta33) and the adjoint are captured once
into two HIP graphs (``torch.cuda.CUDAGraph`` on the capture stream, which the
library launches on) and replayed: the same kernels, one launch each.

The plan (parent support, pair/triple supports, GL nodes, balance shifts,
peak wavenumber) is built once from the initial amplitudes and held fixed:
those are piecewise-constant functions of the amplitudes, and the roundoff
floors and masks of each forward evaluation are held fixed in its adjoint.
The reference has no gradient; the adjoint is pinned by central finite
differences, by a plain-torch autograd restatement of the same graph and by
Euler homogeneity (degree n per order n).
"""

from __future__ import annotations

import math

import numpy as np

from . import api
from . import plan as PL
from ._native import DevicePlan

ORDER_MASK = {1: 1, 2: 3, 3: 7}  # components eta11 (1), eta22 (2), eta33 (4) up to an order


class GLElevationModel:
    """Fixed-plan map (a, b) -> eta on an ``[ny, nx]`` grid, with its adjoint on the GPU."""

    def __init__(self, kx, ky, a0, b0, *, lx: float, ly: float, nx: int, ny: int, depth: float,
                 gravity: float = 9.81, order: int = 3, time: float = 0.0, peak_q: float | None = None,
                 graphs: bool = False):
        import torch

        if order not in (1, 2, 3):
            raise ValueError("order must be 1, 2 or 3")
        self.kx = np.asarray(kx, dtype=float).ravel()
        self.ky = np.asarray(ky, dtype=float).ravel()
        self.order, self.depth, self.gravity = int(order), float(depth), float(gravity)
        self.nx, self.ny, self.time = int(nx), int(ny), float(time)
        coeffs = api.gl_spectral_coefficients(order, gravity, depth, np.asarray(a0, dtype=float).ravel(),
                                              np.asarray(b0, dtype=float).ravel(), self.kx, self.ky, 0, 0)
        kxg, kyg = api._wavenumber_grid(lx, ly, nx, ny)
        spec, omega = api.deposit_analytic_input(coeffs, lx, ly, nx, ny)
        host = PL.build_plan(spec, depth * kxg, depth * kyg, depth=depth, gravity=gravity, order=order,
                             peak_q=coeffs["peak_depth"] if peak_q is None else peak_q)
        host.omega = host.box1.take(omega)
        self.host = host
        self.plan = DevicePlan(host, host.box1.take(spec), max_instants=1)
        self.device = self.plan.device
        box = host.box1
        mx = np.rint(self.kx * lx / (2.0 * math.pi)).astype(np.int64)
        my = np.rint(self.ky * ly / (2.0 * math.pi)).astype(np.int64)
        if np.any(mx < box.mx0) or np.any(mx >= box.mx0 + box.wx) or np.any(my < box.my0) or \
                np.any(my >= box.my0 + box.wy):
            raise ValueError("every mode must lie in the plan's parent box (nonzero initial amplitudes)")
        self.index = torch.as_tensor((my - box.my0) * box.wx + (mx - box.mx0), device=self.device)
        self.scale = nx * ny / depth  # S0 = N^2 (a + i b) / h  (src/gl_spectral_surface.m:145-147)
        self.times = torch.tensor([self.time], dtype=torch.float64, device=self.device)
        self.orders = ORDER_MASK[self.order]
        self.generation = 0
        shape = (1, ny, nx)
        self._out = [torch.empty(shape, dtype=torch.complex128, device=self.device) for _ in range(2)]
        self.graphs = bool(graphs)
        self._cap = None  # captured forward / adjoint graphs and their static buffers

    # ---- HIP-graph mode: the forward (phase, eta11, eta22, eta33) and the adjoint as two graph launches
    def _capture(self) -> None:
        """Capture the forward and the adjoint (recompute=False) once, after one eager run of each."""
        import torch

        n = self.kx.size
        a_in = torch.zeros(n, dtype=torch.float64, device=self.device)
        b_in = torch.zeros(n, dtype=torch.float64, device=self.device)
        g_in = torch.zeros((self.ny, self.nx), dtype=torch.float64, device=self.device)

        def forward():
            self._load(a_in, b_in)
            return self._evaluate()

        def backward():
            return self.gradient(g_in, recompute=False)

        forward()  # sizes the plan scratch eagerly (nothing may grow under capture)
        backward()
        torch.cuda.synchronize()
        fwd_graph, bwd_graph = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
        with torch.cuda.graph(fwd_graph):
            eta = forward()
        with torch.cuda.graph(bwd_graph):
            ga, gb = backward()
        self._cap = {"a": a_in, "b": b_in, "g": g_in, "fwd": fwd_graph, "bwd": bwd_graph, "eta": eta, "ga": ga,
                     "gb": gb}
        self.plan.hold(True)  # the graphs replay into the plan's workspace: nothing of it may move

    def close(self) -> None:
        """Drop the captured graphs (releasing the plan's hold) and destroy the plan."""
        if self._cap is not None:
            self._cap = None
            self.plan.hold(False)
        self.plan.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _elevation_graph(self, a, b):
        cap = self._cap
        cap["a"].copy_(a.detach().to(self.device, torch_float64()))
        cap["b"].copy_(b.detach().to(self.device, torch_float64()))
        cap["fwd"].replay()
        self.generation += 1
        return cap["eta"].clone()

    def _gradient_graph(self, grad_eta):
        cap = self._cap
        cap["g"].copy_(grad_eta.to(torch_float64()))
        cap["bwd"].replay()
        return cap["ga"].clone(), cap["gb"].clone()

    def _evaluate(self):
        eta = None
        for kind, n in (("linear", 1), ("order2", 2), ("order3", 3)):
            if n > self.order:
                break
            e, _ = self.plan.run(kind, self.times, out=self._out)
            eta = e[0].real.clone() if eta is None else eta + e[0].real
        return eta

    def _load(self, a, b) -> None:
        import torch

        s = torch.zeros(self.host.box1.wy * self.host.box1.wx, dtype=torch.complex128, device=self.device)
        amp = torch.complex(a.detach().to(self.device, torch.float64), b.detach().to(self.device, torch.float64))
        s.index_add_(0, self.index, amp * self.scale)
        self.plan.s0.copy_(s.view_as(self.plan.s0))
        self.generation += 1

    def elevation(self, a, b):
        """Re(eta11 + eta22 + eta33) (up to the model order), float64 [ny, nx] device tensor."""
        if self.graphs:
            if self._cap is None:
                self._capture()
            return self._elevation_graph(a, b)
        self._load(a, b)
        return self._evaluate()

    def component(self, a, b, n: int):
        """Re(eta_nn+) alone (n = 1, 2, 3)."""
        self._load(a, b)
        kind = {1: "linear", 2: "order2", 3: "order3"}[n]
        e, _ = self.plan.run(kind, self.times, out=self._out)
        return e[0].real.clone()

    def gradient(self, grad_eta, *, orders: int | None = None, recompute: bool = True):
        """(dL/da, dL/db) float64 device tensors for dL/deta = ``grad_eta`` (real [ny, nx])."""
        g0 = self.plan.adjoint(grad_eta, self.times, orders=self.orders if orders is None else orders,
                               recompute=recompute).reshape(-1)
        gm = g0[self.index] * self.scale
        return gm.real.contiguous(), gm.imag.contiguous()


def torch_float64():
    import torch

    return torch.float64


def _function():
    import torch

    class GLElevation(torch.autograd.Function):
        @staticmethod
        def forward(ctx, a, b, model: GLElevationModel):
            eta = model.elevation(a, b)
            ctx.model = model
            ctx.generation = model.generation
            ctx.save_for_backward(a.detach().clone(), b.detach().clone())
            ctx.dtypes = (a.dtype, b.dtype)
            return eta

        @staticmethod
        def backward(ctx, grad_eta):
            model = ctx.model
            a, b = ctx.saved_tensors
            fresh = model.generation == ctx.generation
            if not fresh:  # another evaluation ran since: restore this one's spectrum
                model._load(a, b)
            if model.graphs and fresh:
                ga, gb = model._gradient_graph(grad_eta)
            else:
                ga, gb = model.gradient(grad_eta.to(torch.float64), recompute=not fresh)
            return ga.to(ctx.dtypes[0]), gb.to(ctx.dtypes[1]), None

    return GLElevation


_GLElevation = None


def gl_elevation(a, b, model: GLElevationModel):
    """Differentiable ``Re(eta11 + eta22 + eta33)`` of amplitudes ``a, b`` (torch tensors)."""
    global _GLElevation
    if _GLElevation is None:
        _GLElevation = _function()
    return _GLElevation.apply(a, b, model)
