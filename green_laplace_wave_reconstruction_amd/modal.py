"""Single-direction time records on the GPU: the modal Green--Laplace resolvent.

``reconstruct_unidirectional_modal_timeseries`` mirrors
python/src/green_laplace/modal_time.py:26-184 (same options, validation
messages, ladder logic, components and audit keys) and
``reconstruct_unidirectional_timeseries`` mirrors
python/src/green_laplace/temporal.py:18-60.  The record analysis (one FFT of
the input), band selection, dispersion solve and the bracketing of each
wavenumber on the reference's modal axis run on the host; every level of the
modal-points ladder (eta22, eta33) is one ``glw_modal_execute`` call and
eta20 one ``glw_eta20_execute`` call on the device.

The device path evaluates the reference's [P, W] modal arrays on band-shifted
grids: q_len >= order * kspan + 1 modal points (the deposit occupies
kspan + 1 nodes) and t_len >= order * span + 1 frequencies, the smallest
supported FFT lengths.  No product wraps on those grids, so the spectra the
reference reads are reproduced up to rounding.
"""

from __future__ import annotations

import ctypes
import math
from collections.abc import Sequence
from ctypes import POINTER, c_double, c_int, c_int32, c_size_t, c_void_p
from dataclasses import dataclass
from time import perf_counter
from typing import Any

import numpy as np

from ._native import GlwError, _check, load_library
from .plan import gauss_laguerre_rule

TINY = np.finfo(float).tiny
EPS = np.finfo(float).eps
MAX_RANK = 16


class CModalDesc(ctypes.Structure):
    _fields_ = [("order", c_int32), ("count", c_int32), ("base_bin", c_int32), ("span", c_int32),
                ("q_len", c_int32), ("t_len", c_int32), ("k0", c_int32), ("kspan", c_int32), ("rank", c_int32),
                ("dq", c_double), ("fstep", c_double), ("depth", c_double), ("lambda2", c_double),
                ("lambda_inner", c_double), ("lambda3", c_double), ("nodes", c_double * MAX_RANK),
                ("weights", c_double * MAX_RANK)]


class CEta20Desc(ctypes.Structure):
    _fields_ = [("count", c_int32), ("base_bin", c_int32), ("span", c_int32), ("q_len", c_int32),
                ("t_len", c_int32), ("k0", c_int32), ("kspan", c_int32), ("rank", c_int32), ("dq", c_double),
                ("fstep", c_double), ("depth", c_double), ("delta_q", c_double),
                ("nodes", c_double * MAX_RANK), ("weights", c_double * MAX_RANK)]


def _bind(lib):
    lib.glw_modal_length_supported.argtypes = [c_int32]
    lib.glw_modal_length_supported.restype = c_int
    lib.glw_modal_workspace_size.argtypes = [POINTER(CModalDesc), POINTER(c_size_t)]
    lib.glw_modal_workspace_size.restype = c_int
    lib.glw_modal_execute.argtypes = [POINTER(CModalDesc), c_void_p, c_void_p, c_size_t, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_void_p]
    lib.glw_modal_execute.restype = c_int
    lib.glw_modal_stage_times.argtypes = [POINTER(CModalDesc), c_void_p, c_void_p, c_size_t, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p, POINTER(ctypes.c_float)]
    lib.glw_modal_stage_times.restype = c_int
    lib.glw_modal_execute_noalloc.argtypes = [POINTER(CModalDesc), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                              c_void_p]
    lib.glw_modal_execute_noalloc.restype = c_int
    lib.glw_eta20_workspace_size.argtypes = [POINTER(CEta20Desc), POINTER(c_size_t)]
    lib.glw_eta20_workspace_size.restype = c_int
    lib.glw_eta20_execute.argtypes = [POINTER(CEta20Desc), c_void_p, c_void_p, c_size_t, c_void_p, c_void_p]
    lib.glw_eta20_execute.restype = c_int
    return lib


_LENGTHS: dict[int, list[int]] = {}


def grid_length(n: int, which: str) -> int:
    """Smallest supported t (``which='t'``), modal (``'q'``) or eta20 modal (``'q20'``: one
    workgroup's LDS transform) grid length >= n."""
    lib = _bind(load_library())
    bit = {"t": 1, "q": 2, "q20": 4}[which]
    if bit not in _LENGTHS:
        _LENGTHS[bit] = [m for m in range(16, 65537, 16) if lib.glw_modal_length_supported(m) & bit]
    for m in _LENGTHS[bit]:
        if m >= n:
            return m
    raise GlwError(f"no supported modal {which} grid length >= {n} (largest {_LENGTHS[bit][-1]}); "
                   "lower modal_points or the frequency band")


@dataclass(frozen=True)
class ModalTimeSeriesResult:
    time: np.ndarray
    eta: np.ndarray
    components: dict
    audit: dict


TimeSeriesResult = ModalTimeSeriesResult


def dimensionless_wavenumber_from_frequency(omega, *, depth: float, gravity: float = 9.81) -> np.ndarray:
    """q >= 0 with q tanh q = omega^2 h / g (python/src/green_laplace/dispersion.py:19-45).

    Newton iteration from sqrt(target) (shallow) or target (deep) with the
    step limited to halve q at most; accepted at relative residual 2e-13.
    """
    w = np.asarray(omega, dtype=float)
    if depth <= 0 or gravity <= 0 or np.any(w < 0):
        raise ValueError("depth and gravity must be positive and omega nonnegative")
    target = w * w * depth / gravity
    q = np.maximum(np.where(target < 1.0, np.sqrt(target), target), TINY)
    for _ in range(60):
        t = np.tanh(q)
        dq = (q * t - target) / np.maximum(t + q * (1.0 - t * t), TINY)
        nxt = np.maximum(q - dq, 0.5 * q)
        moved = float(np.max(np.abs(nxt - q) / np.maximum(1.0, q))) if q.size else 0.0
        q = nxt
        if moved < 5.0e-15:
            break
    if not np.allclose(q * np.tanh(q), target, rtol=2.0e-13, atol=2.0e-15):
        raise ArithmeticError("finite-depth dispersion solve did not converge")
    return q


def _modal_axis(points: int, qmax: float, order: float, limit: float | None):
    """Axis (-P/2 .. P/2-1) dq reaching (order + 0.05) qmax (modal_time.py:344-352)."""
    if limit is None:
        lim = (order + 0.05) * qmax / (1 - 2 / points)
    else:
        lim = float(limit)
        if lim <= order * qmax:
            raise ValueError("modal_limit must exceed the requested nonlinear output support")
    dq = 2 * lim / points
    return np.arange(-points / 2, points / 2) * dq, dq


def _bracket(value, axis, dq):
    pos = (value - axis[0]) / dq
    lo = np.clip(np.floor(pos).astype(int), 0, axis.size - 2)
    wu = np.clip(pos - lo, 0.0, 1.0)
    return lo, lo + 1, 1.0 - wu, wu


@dataclass
class _Band:
    count: int
    dt: float
    bins: np.ndarray
    omega: np.ndarray
    amplitudes: np.ndarray
    q: np.ndarray
    peak_q: float
    energy: np.ndarray
    total_energy: float


def _select_band(values, times, *, depth, gravity, order, omega_max, energy_fraction) -> _Band:
    """Positive-frequency band (modal_time.py:66-96)."""
    count = values.size
    dt = float(np.mean(np.diff(times)))
    expected = times[0] + np.arange(count) * dt
    if dt <= 0 or float(np.max(np.abs(times - expected))) > max(1e-10, 1024 * EPS * max(1.0, times[-1] - times[0])):
        raise ValueError("a uniformly increasing time record is required")
    positive = (count - 1) // 2
    coeff = np.fft.fft(values) / count
    energy = np.abs(coeff[1:positive + 1]) ** 2
    total = float(np.sum(energy))
    if total == 0:
        raise ValueError("eta1 has no positive-frequency content")
    idx = min(int(np.searchsorted(np.cumsum(energy) / total, energy_fraction)), positive - 1)
    top = min(idx + 1, positive // order)
    if omega_max is not None:
        top = min(top, int(np.floor(omega_max * count * dt / (2 * np.pi))))
    if top < 1:
        raise ValueError("the selected eta1 frequency band is empty")
    bins = np.arange(1, top + 1)
    omega = 2 * np.pi * bins / (count * dt)
    amps = 2 * np.conj(coeff[bins])
    q = dimensionless_wavenumber_from_frequency(omega, depth=depth, gravity=gravity)
    return _Band(count, dt, bins, omega, amps, q, float(q[int(np.argmax(np.abs(amps)))]), energy, total)


class ModalDevice:
    """Device state of one record: workspace, outputs and the two C entries."""

    def __init__(self, count: int, device=None):
        import torch

        if not torch.cuda.is_available():
            raise GlwError("a HIP device is required")
        self.lib = _bind(load_library())
        self.device = torch.device(device or "cuda")
        self.count = int(count)
        self.ws = torch.empty(0, dtype=torch.uint8, device=self.device)

    def workspace(self, nbytes: int):
        import torch

        if self.ws.numel() < nbytes:
            self.ws = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        return self.ws

    def stream(self):
        import torch

        return c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def upload(self, deposit: np.ndarray):
        """Device copy of a [bins][5] deposit table (Re A, Im A, w_lo, w_hi, k_lo)."""
        import torch

        return torch.as_tensor(np.ascontiguousarray(deposit, dtype=np.float64), device=self.device)

    def modal(self, desc: CModalDesc, dep, *, eta11: bool, stage_ms: list | None = None):
        """One level: (eta11, eta22, eta33, audit) device tensors.  With ``stage_ms`` (a list), the
        call waits and appends the six stage durations in ms (glw_modal_stage_times)."""
        import torch

        size = c_size_t()
        _check(self.lib.glw_modal_workspace_size(ctypes.byref(desc), ctypes.byref(size)))
        ws = self.workspace(size.value)
        n = self.count
        e11 = torch.empty(n, dtype=torch.float64, device=self.device) if eta11 else None
        e22 = torch.empty(n, dtype=torch.complex128, device=self.device)
        e33 = torch.empty(n, dtype=torch.complex128, device=self.device) if desc.order == 3 else None
        audit = torch.empty(4, dtype=torch.float64, device=self.device)
        args = (ctypes.byref(desc), c_void_p(dep.data_ptr()), c_void_p(ws.data_ptr()), ws.numel(),
                None if e11 is None else c_void_p(e11.data_ptr()), c_void_p(e22.data_ptr()),
                None if e33 is None else c_void_p(e33.data_ptr()), c_void_p(audit.data_ptr()), self.stream())
        if stage_ms is None:
            _check(self.lib.glw_modal_execute(*args))
        else:
            ms = (ctypes.c_float * 6)()
            _check(self.lib.glw_modal_stage_times(*args, ms))
            stage_ms.append(list(ms))
        return e11, e22, e33, audit

    def eta20(self, desc: CEta20Desc, dep):
        import torch

        size = c_size_t()
        _check(self.lib.glw_eta20_workspace_size(ctypes.byref(desc), ctypes.byref(size)))
        ws = self.workspace(size.value)
        out = torch.empty(self.count, dtype=torch.float64, device=self.device)
        _check(self.lib.glw_eta20_execute(ctypes.byref(desc), c_void_p(dep.data_ptr()), c_void_p(ws.data_ptr()),
                                          ws.numel(), c_void_p(out.data_ptr()), self.stream()))
        return out


def _rule(rank: int):
    nodes, weights = gauss_laguerre_rule(rank)
    if rank > MAX_RANK:
        raise ValueError(f"GL rank above {MAX_RANK}")
    pad = np.zeros(MAX_RANK)
    n = pad.copy()
    w = pad.copy()
    n[:rank] = nodes
    w[:rank] = weights
    return (c_double * MAX_RANK)(*n), (c_double * MAX_RANK)(*w)


@dataclass
class ModalLevelPlan:
    """Host description of one ladder level (what glw_modal_execute receives)."""

    desc: CModalDesc
    deposit: np.ndarray  # [span + 1][5]
    projection: float
    modal_limit: float
    work_count: int


def modal_level_plan(band: _Band, *, depth, gravity, rank, order, points, modal_limit) -> ModalLevelPlan:
    """Deposit and reduced grids of one level (modal_time.py:187-219, 344-368)."""
    base = int(band.bins.min())
    work = band.bins - base
    span = int(work.max())
    width = order * span + 1
    if width > band.count:
        raise ValueError("the nonlinear output band does not fit in the time record")
    axis, dq = _modal_axis(points, float(np.max(np.abs(band.q))), order, modal_limit)
    lo, hi, wl, wu = _bracket(band.q, axis, dq)
    mass = np.abs(band.amplitudes) ** 2
    errors = np.concatenate((np.abs(axis[lo] - band.q) / band.q, np.abs(axis[hi] - band.q) / band.q))
    weights = np.concatenate((mass * wl, mass * wu))
    projection = float(np.sqrt(np.sum(weights * errors**2) / max(np.sum(weights), TINY)))
    klo = lo - points // 2  # signed node index: q = k dq
    if np.any(work != np.arange(span + 1)):
        raise ValueError("the input band must be contiguous")
    k0 = int(klo.min())
    kspan = int(klo.max()) + 1 - k0
    pq = band.peak_q
    nu_p = math.sqrt(pq * math.tanh(pq))
    nodes, wts = _rule(rank)
    desc = CModalDesc(order=order, count=band.count, base_bin=base, span=span,
                      q_len=grid_length(order * kspan + 1, "q"), t_len=grid_length(order * span + 1, "t"),
                      k0=k0, kspan=kspan, rank=rank, dq=dq,
                      fstep=2 * np.pi / (band.count * band.dt) * np.sqrt(depth / gravity), depth=depth,
                      lambda2=math.sqrt(4 * pq * math.tanh(pq) - 2 * pq * math.tanh(2 * pq)),
                      lambda_inner=2 * nu_p - math.sqrt(2 * pq * math.tanh(2 * pq)),
                      lambda3=3 * nu_p - math.sqrt(3 * pq * math.tanh(3 * pq)), nodes=nodes, weights=wts)
    dep = np.stack([band.amplitudes.real, band.amplitudes.imag, wl, wu, klo.astype(float)], axis=1)
    return ModalLevelPlan(desc, dep, projection, float(abs(axis[0])), width)


def eta20_plan(amplitudes, omega, q, count: int, dt: float, *, depth, gravity, rank, nx, ny):
    """Host description of eta20 (_eta20.py:13-101), or (None, audit) when degenerate."""
    if rank not in (6, 12, 16):
        raise ValueError("eta20 rank must be 6, 12, or 16")
    if nx < 16 or ny < 16 or nx % 2 or ny % 2:
        raise ValueError("eta20 modal-grid sizes must be even and at least 16")
    bins = np.rint(omega * count * dt / (2.0 * np.pi)).astype(int)
    if np.any(bins < 1) or not np.allclose(omega, 2.0 * np.pi * bins / (count * dt), rtol=1e-10, atol=1e-12):
        raise ValueError("eta20 input frequencies must lie on temporal DFT bins")
    base = int(bins.min())
    work = bins - base
    span = int(work.max())
    if np.any(work != np.arange(span + 1)):
        raise ValueError("the input band must be contiguous")
    qmax = float(np.max(q))
    if qmax == 0.0:
        raise ValueError("a nonzero modal-axis limit is required")
    xaxis, dqx = _modal_axis(nx, qmax, 2, None)
    yaxis, dqy = _modal_axis(ny, 0.0, 2, (2.05 * qmax) / (1.0 - 2.0 / ny))
    lo, _, wl, wu = _bracket(q, xaxis, dqx)
    ylo, yhi, yl, yu = _bracket(np.zeros(1), yaxis, dqy)
    # the qy = 0 plane carries the whole deposit: weight on node ny/2, nothing elsewhere
    on = np.where(ylo == ny // 2, yl, 0.0) + np.where(yhi == ny // 2, yu, 0.0)
    off = (yl + yu) - on
    if float(off[0]) > 1e-12:
        raise ValueError("eta20: the qy bracket leaves the zero plane")
    energy = np.abs(amplitudes) ** 2
    esum = float(np.sum(energy))
    mean_q = float(np.sum(q * energy) / esum)
    delta_q = float(np.sqrt(2.0 * np.sum((q - mean_q) ** 2 * energy) / esum))
    audit = {"implementation": "modal-grid-green-laplace-eta20", "quadrature_rank": rank}
    if delta_q <= 64.0 * EPS * max(1.0, qmax):
        audit.update(status="degenerate-zero-difference-output", interaction_enumeration=False)
        return None, None, audit
    klo = lo - nx // 2
    k0 = int(klo.min())
    kspan = int(klo.max()) + 1 - k0
    nodes, wts = _rule(rank)
    desc = CEta20Desc(count=count, base_bin=base, span=span, q_len=grid_length(2 * kspan + 1, "q20"),
                      t_len=grid_length(2 * span + 1, "t"), k0=k0, kspan=kspan, rank=rank, dq=dqx,
                      fstep=2.0 * np.pi / (count * dt) * np.sqrt(depth / gravity), depth=depth, delta_q=delta_q,
                      nodes=nodes, weights=wts)
    amp = amplitudes * float(on[0])
    dep = np.stack([amp.real, amp.imag, wl, wu, klo.astype(float)], axis=1)
    audit.update({"modal_grid": [ny, nx], "delta_q_rms_pair": delta_q, "strict_zero_spatial_mode": "excluded",
                  "interaction_enumeration": False})
    return desc, dep, audit


def reconstruct_unidirectional_modal_timeseries(
    eta1,
    time,
    *,
    depth: float,
    gravity: float = 9.81,
    omega_max: float | None = None,
    energy_fraction: float = 0.999,
    quadrature_rank: int = 8,
    order: int = 3,
    modal_points: int | Sequence[int] = (2048, 4096, 8192, 16384, 32768, 65536),
    modal_limit: float | None = None,
    relative_tolerance: float = 1.0e-3,
    include_eta20: bool = True,
    eta20_rank: int = 16,
    eta20_grid: tuple[int, int] = (128, 16),
    device=None,
    return_device: bool = False,
) -> ModalTimeSeriesResult:
    """Reconstruct a single-direction record with the modal GL resolvent on the GPU.

    Same options and results as the reference (modal_time.py:26-184).
    ``return_device=True`` leaves the component arrays as device tensors.
    """
    import torch

    started = perf_counter()
    values = np.asarray(eta1, dtype=float).reshape(-1)
    times = np.asarray(time, dtype=float).reshape(-1)
    if values.size != times.size or values.size < 9:
        raise ValueError("eta1 and time must have the same length of at least nine")
    if depth <= 0 or gravity <= 0 or order not in (2, 3):
        raise ValueError("depth and gravity must be positive and order must be 2 or 3")
    ladder = (modal_points,) if isinstance(modal_points, int) else tuple(int(v) for v in modal_points)
    if not ladder or any(v < 16 or v % 2 for v in ladder) or any(b <= a for a, b in zip(ladder, ladder[1:])):
        raise ValueError("modal_points must be an increasing sequence of even integers")
    if relative_tolerance <= 0:
        raise ValueError("relative_tolerance must be positive")
    if order == 3 and quadrature_rank not in (4, 6, 8, 10):
        raise ValueError("third-order modal reconstruction supports rank 4, 6, 8, or 10")
    band = _select_band(values, times, depth=depth, gravity=gravity, order=order, omega_max=omega_max,
                        energy_fraction=energy_fraction)
    dev = ModalDevice(band.count, device)

    previous = None
    levels: list[dict[str, Any]] = []
    converged = False
    eta11 = eta22 = eta33 = None
    modal_audit: dict[str, Any] = {}
    for points in ladder:
        t0 = perf_counter()
        lp = modal_level_plan(band, depth=depth, gravity=gravity, rank=quadrature_rank, order=order,
                              points=points, modal_limit=modal_limit)
        dep = dev.upload(lp.deposit)
        e11, eta22, eta33, audit = dev.modal(lp.desc, dep, eta11=eta11 is None)
        if e11 is not None:
            eta11 = e11
        current = (eta22,) if eta33 is None else (eta22, eta33)
        changes = None
        if previous is not None:
            changes = tuple(float(torch.linalg.vector_norm(c - p) / torch.linalg.vector_norm(c))
                            for c, p in zip(current, previous))
            converged = max(changes) <= relative_tolerance
        au = audit.cpu().numpy()
        modal_audit = {
            "modal_points": points,
            "modal_limit": lp.modal_limit,
            "quadrature_rank": quadrature_rank,
            "base_bin": int(lp.desc.base_bin),
            "work_frequency_count": lp.work_count,
            "wavevector_projection_rms_relative": lp.projection,
            "eta22_invalid_source_energy_fraction": float(au[0] / max(au[1], TINY)),
            "eta33_invalid_source_energy_fraction": float(au[2] / max(au[3], TINY)) if order == 3 else None,
            "reduced_grid": {"q_len": int(lp.desc.q_len), "t_len": int(lp.desc.t_len)},
        }
        levels.append({
            "modal_points": points,
            "seconds": perf_counter() - t0,
            "eta22_change": None if changes is None else changes[0],
            "eta33_change": None if changes is None or eta33 is None else changes[1],
            "wavevector_projection_rms_relative": lp.projection,
        })
        previous = current
        if converged:
            break
    comps: dict[str, Any] = {"eta11": eta11, "eta22": eta22.real, "eta22_analytic": eta22}
    total = eta11 + eta22.real
    if eta33 is not None:
        comps["eta33"] = eta33.real
        comps["eta33_analytic"] = eta33
        total = total + eta33.real
    if include_eta20:
        desc20, dep20, audit20 = eta20_plan(band.amplitudes, band.omega, band.q, band.count, band.dt, depth=depth,
                                            gravity=gravity, rank=eta20_rank, nx=eta20_grid[0], ny=eta20_grid[1])
        if desc20 is None:
            e20 = torch.zeros(band.count, dtype=torch.float64, device=dev.device)
        else:
            e20 = dev.eta20(desc20, dev.upload(dep20))
        comps["eta20"] = e20
        total = total + e20
        modal_audit["eta20"] = audit20
    modal_audit.update({
        "implementation": "one-dimensional-modal-green-laplace-time",
        "retained_positive_frequency_energy": float(np.sum(band.energy[:band.bins.size]) / band.total_energy),
        "input_projection_relative": float(
            torch.linalg.vector_norm(eta11 - torch.as_tensor(values, device=dev.device))
            / max(float(np.linalg.norm(values)), TINY)),
        "solve_seconds": perf_counter() - started,
        "levels": levels,
        "converged": converged,
        "relative_tolerance": relative_tolerance,
        "auxiliary_space_domain": False,
        "interaction_enumeration": False,
        "stokes_correction": False,
        "device": str(dev.device),
    })
    if not return_device:
        comps = {k: v.cpu().numpy() for k, v in comps.items()}
        total = total.cpu().numpy()
    return ModalTimeSeriesResult(times.copy(), total, comps, modal_audit)


def reconstruct_unidirectional_timeseries(
    eta1,
    time,
    *,
    depth: float,
    gravity: float = 9.81,
    omega_max: float | None = None,
    energy_fraction: float = 0.999,
    quadrature_rank: int = 8,
    order: int = 3,
    modal_points: int | Sequence[int] = (2048, 4096, 8192, 16384, 32768, 65536),
    modal_limit: float | None = None,
    relative_tolerance: float = 1.0e-3,
    include_eta20: bool = True,
    eta20_rank: int = 16,
    eta20_grid: tuple[int, int] = (128, 16),
    **unknown: Any,
) -> TimeSeriesResult:
    """Public single-direction interface (temporal.py:18-60): the modal resolvent."""
    if unknown:
        raise TypeError(f"unknown time-reconstruction options: {', '.join(sorted(unknown))}")
    return reconstruct_unidirectional_modal_timeseries(
        eta1, time, depth=depth, gravity=gravity, omega_max=omega_max, energy_fraction=energy_fraction,
        quadrature_rank=quadrature_rank, order=order, modal_points=modal_points, modal_limit=modal_limit,
        relative_tolerance=relative_tolerance, include_eta20=include_eta20, eta20_rank=eta20_rank,
        eta20_grid=eta20_grid)
