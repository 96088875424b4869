"""Drop-in public API of the Green--Laplace reconstruction on MI355X.

Mirrors the reference's two entry layers:

* ``gl_spectral_coefficients`` / ``gl_spectral_surface`` -- the MATLAB
  two-stage API (src/gl_spectral_coefficients.m:1-112,
  src/gl_spectral_surface.m:1-149), same inputs, outputs and error
  identifiers (``GreenLaplaceError.identifier``).
* ``reconstruct_directional_field`` -- the Python package entry point
  (python/src/green_laplace/spatial.py:21-153), same arguments, dataclass and
  ``ValueError`` messages.

Host work is input validation, the analytic-spectrum deposit and the FP64
plan (``plan.build_plan``); every field is computed by the glw HIP kernels.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np

from . import plan as _plan
from ._native import DevicePlan, to_host

API_ID = "green-laplace-spectral-v1"


class GreenLaplaceError(ValueError):
    """Reference error with its MATLAB identifier (``green_laplace:...``)."""

    def __init__(self, identifier: str, message: str):
        super().__init__(message)
        self.identifier = identifier


@dataclass(frozen=True)
class DirectionalFieldResult:
    """Fields and execution information returned by a spatial reconstruction."""

    x: np.ndarray
    y: np.ndarray
    eta: np.ndarray
    psi: np.ndarray
    components: dict
    audit: dict


_DEFAULT_OPTIONS = {
    "sector": "pure-sum",
    "eta22_rank": 6,
    "peak_wavenumber": None,
    "allow_below_model_parent_domain": False,
    "grid_tolerance": 1e-10,
}


def _is_int(x) -> bool:
    return isinstance(x, (int, np.integer)) or (isinstance(x, float) and float(x).is_integer())


def gl_spectral_coefficients(order, g, h, a, b, kx, ky, Ux, Uy, options: dict | None = None) -> dict:
    """Validate and pack an MF12-style component set (src/gl_spectral_coefficients.m:1-88)."""
    opts = dict(_DEFAULT_OPTIONS)
    if options is not None:
        if not isinstance(options, dict):
            raise GreenLaplaceError("green_laplace:Options", "The optional input must be one options struct.")
        for key, value in options.items():
            if key not in opts:
                raise GreenLaplaceError("green_laplace:UnknownOption", f"Unknown Green--Laplace option: {key}")
            opts[key] = value
    opts["sector"] = str(opts["sector"])
    opts["allow_below_model_parent_domain"] = bool(opts["allow_below_model_parent_domain"])
    if not _is_int(order) or not 1 <= int(order) <= 3:
        raise ValueError("order must be an integer from 1 to 3")
    for name, val in (("g", g), ("h", h)):
        if not np.isscalar(val) or not np.isfinite(val) or val <= 0:
            raise ValueError(f"{name} must be a finite positive scalar")
    for name, val in (("Ux", Ux), ("Uy", Uy)):
        if not np.isscalar(val) or not np.isfinite(val):
            raise ValueError(f"{name} must be a finite scalar")
    if Ux != 0 or Uy != 0:
        raise GreenLaplaceError("green_laplace:UniformCurrent",
                                "The released API preserves the MF12-compatible Ux, Uy input slots but has only "
                                "been certified for Ux=Uy=0.")
    a, b, kx, ky = (np.asarray(v, dtype=float).ravel() for v in (a, b, kx, ky))
    if a.size == 0 or not (a.size == b.size == kx.size == ky.size):
        raise GreenLaplaceError("green_laplace:ComponentSize", "a, b, kx, and ky must be nonempty vectors of equal length.")
    if not np.all(np.isfinite(np.concatenate([a, b, kx, ky]))) or np.any(np.hypot(kx, ky) <= 0):
        raise GreenLaplaceError("green_laplace:Components", "Component amplitudes and nonzero wavenumbers must be finite.")
    if np.any(kx <= 0):
        raise GreenLaplaceError("green_laplace:Support", "The released GL graphs require strict-forward support kx>0.")
    if opts["sector"] != "pure-sum":
        raise GreenLaplaceError("green_laplace:Sector",
                                "The released total-field API supports the positive pure-sum sector only.")
    r = opts["eta22_rank"]
    if not _is_int(r) or not 1 <= int(r) <= 12:
        raise ValueError("eta22_rank must be a positive integer no larger than 12")
    amp = np.hypot(a, b)
    if not np.any(amp > 0):
        raise GreenLaplaceError("green_laplace:ZeroInput", "At least one input amplitude is required.")
    if opts["peak_wavenumber"] is None:
        peak_wavenumber = float(np.hypot(kx, ky)[int(np.argmax(amp**2))])
    else:
        pw = opts["peak_wavenumber"]
        if not np.isscalar(pw) or not np.isfinite(pw) or pw <= 0:
            raise ValueError("peak_wavenumber must be a finite positive scalar")
        peak_wavenumber = float(pw)
    parent_q = h * np.hypot(kx, ky)
    if order >= 2 and np.any(parent_q < 0.3) and not opts["allow_below_model_parent_domain"]:
        raise GreenLaplaceError("green_laplace:Order2Domain",
                                "Order-two surface-potential reconstruction requires kh>=0.3. A labelled tail "
                                "extrapolation can be enabled in options.")
    if order >= 3 and np.any(parent_q <= 0.5):
        raise GreenLaplaceError("green_laplace:Order3Domain", "Order-three reconstruction requires every parent kh>0.5.")
    kabs = np.hypot(kx, ky)
    intrinsic = np.sqrt(g * kabs * np.tanh(h * kabs))
    return {
        "api": API_ID, "order": int(order), "g": float(g), "h": float(h), "a": a, "b": b, "kx": kx, "ky": ky,
        "Ux": float(Ux), "Uy": float(Uy), "intrinsic_omega": intrinsic, "omega": kx * Ux + ky * Uy + intrinsic,
        "peak_wavenumber": peak_wavenumber, "peak_depth": h * peak_wavenumber, "options": opts,
        "information_boundary": "first-order surface elevation only",
        "surface_potential_definition": "psi=phi(x,z=eta,t)", "flat_potential_is_public_output": False,
    }


def _wavenumber_grid(lx, ly, nx, ny):
    kx1 = 2.0 * np.pi * np.fft.fftfreq(nx, d=lx / nx)
    ky1 = 2.0 * np.pi * np.fft.fftfreq(ny, d=ly / ny)
    return np.meshgrid(kx1, ky1)


def deposit_analytic_input(coeffs: dict, lx: float, ly: float, nx: int, ny: int):
    """t=0 analytic spectrum and per-bin angular frequency (src/gl_spectral_surface.m:123-149).

    Vectorised over the components with the reference loop's results: the
    first offending component raises (off-grid before Nyquist), components on
    one bin add in input order, and the last one sets the bin's frequency.
    """
    dkx, dky = 2.0 * np.pi / lx, 2.0 * np.pi / ly
    kx = np.asarray(coeffs["kx"], dtype=float).ravel()
    ky = np.asarray(coeffs["ky"], dtype=float).ravel()
    mx = np.round(kx / dkx).astype(np.int64)
    my = np.round(ky / dky).astype(np.int64)
    tol = coeffs["options"]["grid_tolerance"] * np.maximum(1.0, np.hypot(kx, ky))
    off = (np.abs(kx - mx * dkx) > tol) | (np.abs(ky - my * dky) > tol)
    nyq = (np.abs(mx) >= nx / 2) | (np.abs(my) >= ny / 2)
    bad = off | nyq
    if bad.any():
        if off[int(np.argmax(bad))]:
            raise GreenLaplaceError("green_laplace:OffGridComponent", "Every component must lie on the requested FFT grid.")
        raise GreenLaplaceError("green_laplace:Nyquist", "An input component reaches or exceeds a Nyquist boundary.")
    flat = (my % ny) * nx + (mx % nx)
    # Real and imaginary parts divided apart: numpy's complex-by-real division
    # multiplies by a reciprocal, one ulp away from the per-component quotient.
    amp = np.empty(flat.size, dtype=np.complex128)
    amp.real = nx * ny * np.asarray(coeffs["a"], dtype=float).ravel() / coeffs["h"]
    amp.imag = nx * ny * np.asarray(coeffs["b"], dtype=float).ravel() / coeffs["h"]
    spec = np.zeros(ny * nx, dtype=np.complex128)
    np.add.at(spec, flat, amp)
    omega = np.zeros(ny * nx)
    last = flat.size - 1 - np.unique(flat[::-1], return_index=True)[1]  # last component on each bin
    omega[flat[last]] = np.asarray(coeffs["omega"], dtype=float).ravel()[last]
    return spec.reshape(ny, nx), omega.reshape(ny, nx)


def _validate_coefficients(coeffs) -> None:
    required = ("api", "order", "g", "h", "a", "b", "kx", "ky", "omega", "peak_depth", "options")
    if not isinstance(coeffs, dict) or any(k not in coeffs for k in required) or coeffs["api"] != API_ID:
        raise GreenLaplaceError("green_laplace:Coefficients", "Input must come from gl_spectral_coefficients.")


def _evaluate(host: _plan.HostPlan, s0_box: np.ndarray, times: np.ndarray, order: int, *, max_instants: int = 16):
    """Device tensors of every *_plus component of orders 1..order at the given instants."""
    dp = DevicePlan(host, s0_box, max_instants=max(1, min(max_instants, len(times))))
    comps = {}
    try:
        names = [("linear", "eta11_plus", "psi11_plus")]
        if order >= 2:
            names.append(("order2", "eta22_plus", "psi22_plus"))
        if order >= 3:
            names.append(("order3", "eta33_plus", "psi33_plus"))
        for kind, ke, kp in names:
            e, p = dp.run(kind, times)
            comps[ke], comps[kp] = e, p
    finally:
        import torch

        torch.cuda.synchronize()
        dp.close()
    return comps


def _all_finite(x) -> bool:
    if isinstance(x, np.ndarray):
        return bool(np.all(np.isfinite(x)))
    import torch

    return bool(torch.isfinite(x).all())


def _check_order3_finite(comps: dict) -> None:
    """Non-finite order-3 outputs are an error (src/internal/gl_no_stokes_eta33.m:223-228).

    The reference raises these without an identifier, so ``identifier`` is empty.
    """
    if "eta33_plus" in comps and not _all_finite(comps["eta33_plus"]):
        raise GreenLaplaceError("", "Nested Green--Laplace eta33 output is not finite.")
    if "psi33_plus" in comps and not _all_finite(comps["psi33_plus"]):
        raise GreenLaplaceError("", "Direct Green--Laplace Psi33 output is not finite.")


def _finish(comps: dict, order: int, single: bool, to_numpy: bool) -> dict:
    """Output dictionary: every *_plus component, its real part, and the summed eta/psi.

    The sums run on the device in the reference's order (orders 1, 2, 3 added
    left to right), then every field is copied to host once when ``to_numpy``.
    """
    keys = ["eta11", "psi11"] + (["eta22", "psi22"] if order >= 2 else []) + (["eta33", "psi33"] if order >= 3 else [])
    plus: dict = {}
    eta_plus = None
    psi_plus = None
    for k in keys:
        v = comps[k + "_plus"]
        plus[k + "_plus"] = v
        if k.startswith("eta"):
            eta_plus = v if eta_plus is None else eta_plus + v
        else:
            psi_plus = v if psi_plus is None else psi_plus + v
    plus["eta_plus"], plus["psi_plus"] = eta_plus, psi_plus
    out: dict = {}
    for k, v in plus.items():
        if single:
            v = v[0]
        out[k] = to_host(v) if to_numpy else v
    for k in keys:
        out[k] = out[k + "_plus"].real
    out["eta"], out["psi"] = out["eta_plus"].real, out["psi_plus"].real
    return out


def gl_spectral_surface(coeffs: dict, Lx: float, Ly: float, Nx: int, Ny: int, t, *, precision: str = "fp64",
                        return_device: bool = False):
    """Reconstruct eta and psi on an Nx x Ny grid (src/gl_spectral_surface.m:1-102).

    Returns ``(eta, psi, X, Y, components, audit)`` like the reference.  ``t``
    may also be a 1-D array of instants; every field then gains a leading time
    axis (one plan, one batched evaluation).  ``precision`` must be
    ``"fp64"``: ``"fp32"`` transforms are refused (``_precision``).
    """
    _validate_coefficients(coeffs)
    for name, val in (("Lx", Lx), ("Ly", Ly)):
        if not np.isscalar(val) or not np.isfinite(val) or val <= 0:
            raise ValueError(f"{name} must be a finite positive scalar")
    for name, val in (("Nx", Nx), ("Ny", Ny)):
        if not _is_int(val) or int(val) <= 0 or int(val) % 2:
            raise ValueError(f"{name} must be a positive even integer")
    Nx, Ny = int(Nx), int(Ny)
    times = np.atleast_1d(np.asarray(t, dtype=float))
    single = np.ndim(t) == 0
    if times.ndim != 1 or not np.all(np.isfinite(times)):
        raise ValueError("t must be a finite real scalar or 1-D array")
    h, g = coeffs["h"], coeffs["g"]
    kxg, kyg = _wavenumber_grid(Lx, Ly, Nx, Ny)
    spec, omega = deposit_analytic_input(coeffs, Lx, Ly, Nx, Ny)
    order = coeffs["order"]
    host = _plan.build_plan(spec, h * kxg, h * kyg, depth=h, gravity=g, order=order,
                            eta22_rank=int(coeffs["options"]["eta22_rank"]), peak_q=coeffs["peak_depth"],
                            allow_below_model_parent_domain=coeffs["options"]["allow_below_model_parent_domain"],
                            precision=_precision(precision))
    host.omega = host.box1.take(omega)
    comps = _evaluate(host, host.box1.take(spec), times, order)
    _check_order3_finite(comps)
    out = _finish(comps, order, single, to_numpy=not return_device)
    X, Y = np.meshgrid(np.arange(Nx) * Lx / Nx, np.arange(Ny) * Ly / Ny)
    included = ["eta11", "psi11"] + (["eta22", "psi22"] if order >= 2 else []) + (["eta33", "psi33"] if order >= 3 else [])
    audit = {
        "api": API_ID, "order": order, "sector": coeffs["options"]["sector"], "included_components": included,
        "surface_potential_definition": "psi=phi(x,z=eta,t)", "flat_potential_exposed": False,
        "rescaling_alignment_or_gain": False, "grid": [Ny, Nx], "domain": [Ly, Lx],
        "time": float(t) if single else times.tolist(), "component_audits": host.audits,
        "precision": precision, "device": "gfx950",
    }
    return out["eta"], out["psi"], X, Y, out, audit


#: Why GL reconstructions refuse FP32 transforms (also the C ABI's refusal, glw_plan_create).
FP32_REFUSAL = ("precision='fp32' is refused for GL reconstructions: FP32 rounding of the intermediate spectra, "
                "amplified by the GL node factors, exceeds relative L2 1e-5 against FP64 for eta22, eta33, psi33 "
                "and eta44; use precision='fp64'")


def _precision(p: str) -> int:
    """Arithmetic of the GL transforms and products: FP64 only.

    FP32 transforms lose the spectrum bins ~1e-7 below each spectrum's
    maximum; the reference keeps them (its floor is 16 eps64 log2 N^2,
    python/src/green_laplace/_spectral.py:52-64) and the GL node factors
    amplify them into the result: at the 512^2 order-3 sea state, FP32 with
    the FP64 floor gives relative L2 ~1e15, with an FP32 floor 1e-2, and even
    FP32 outer stages alone (inner pair stage in FP64) 1e-5..3e-3 depending
    on the floor level.
    """
    if p == "fp64":
        return _plan.PREC_FP64
    if p == "fp32":
        raise ValueError(FP32_REFUSAL)
    raise ValueError("precision must be 'fp64'")


def analytic_spectrum_from_real_field(eta1: np.ndarray, *, depth: float, kx: np.ndarray, tolerance: float = 1.0e-11):
    """Strict-forward analytic spectrum of a real snapshot (python/src/green_laplace/_spectral.py:19-49)."""
    if eta1.ndim != 2 or eta1.shape != kx.shape or not np.all(np.isfinite(eta1)):
        raise ValueError("eta1 must be a finite two-dimensional grid")
    if depth <= 0:
        raise ValueError("depth must be positive")
    raw = np.fft.fft2(eta1 / depth)
    scale = max(float(np.max(np.abs(raw))), np.finfo(float).tiny)
    zero_line = np.abs(kx) <= 32.0 * np.finfo(float).eps * max(1.0, float(np.max(np.abs(kx))))
    if np.max(np.abs(raw[zero_line]), initial=0.0) > tolerance * scale:
        raise ValueError("eta1 contains a non-negligible kx=0 component")
    spectrum = np.zeros_like(raw, dtype=np.complex128)
    spectrum[kx > 0] = 2.0 * raw[kx > 0]
    back = np.real(depth * np.fft.ifft2(spectrum))
    err = float(np.linalg.norm(back - eta1) / max(np.linalg.norm(eta1), np.finfo(float).tiny))
    if err > tolerance:
        raise ValueError("eta1 is not consistent with the declared strict-forward x propagation")
    return spectrum, err


def reconstruct_directional_field(eta1, *, lx: float, ly: float, depth: float, gravity: float = 9.81,
                                  time: float = 0.0, order: int = 2, eta22_rank: int = 6,
                                  allow_below_model_parent_domain: bool = False,
                                  precision: str = "fp64") -> DirectionalFieldResult:
    """Reconstruct a strict-forward directional field from a real ``eta1`` grid
    (python/src/green_laplace/spatial.py:33-153)."""
    field = np.asarray(eta1, dtype=float)
    if field.ndim != 2:
        raise ValueError("eta1 must be a two-dimensional array")
    if not np.isfinite(time) or depth <= 0 or gravity <= 0:
        raise ValueError("time must be finite; depth and gravity must be positive")
    if order not in (1, 2, 3):
        raise ValueError("order must be 1, 2, or 3")
    if not isinstance(eta22_rank, int) or eta22_rank < 1 or eta22_rank > 16:
        raise ValueError("eta22_rank must be an integer from 1 to 16")
    ny, nx = field.shape
    if lx <= 0 or ly <= 0 or nx < 2 or ny < 2 or nx % 2 or ny % 2:
        raise ValueError("positive domain lengths and positive even grid sizes are required")
    kx, ky = _wavenumber_grid(lx, ly, nx, ny)
    spec, in_err = analytic_spectrum_from_real_field(field, depth=depth, kx=kx)
    host = _plan.build_plan(spec, depth * kx, depth * ky, depth=depth, gravity=gravity, order=order,
                            eta22_rank=eta22_rank, allow_below_model_parent_domain=allow_below_model_parent_domain,
                            precision=_precision(precision))
    comps = _evaluate(host, host.box1.take(spec), np.array([float(time)]), order)
    out = _finish(comps, order, True, to_numpy=True)
    x, y = np.meshgrid(np.arange(nx, dtype=float) * lx / nx, np.arange(ny, dtype=float) * ly / ny)
    audit = {
        "implementation": "green-laplace-mi355x-directional-spatial", "order": order, "grid": [ny, nx],
        "domain": [ly, lx], "time": float(time), "input": "real eta1(x,y) strict-forward in x",
        "input_reconstruction_relative": in_err, "phase_convention": "exp(i*(kx*x+ky*y)-i*omega*t)",
        "stokes_correction": False, "component_audits": host.audits, "precision": precision,
    }
    return DirectionalFieldResult(x=x, y=y, eta=out["eta"], psi=out["psi"], components=out, audit=audit)


def gl_pure_sum_order4(s, qx, qy, qp: float, rank: int = 4, *, precision: str = "fp64"):
    """Experimental dimensionless positive pure-sum order-four field (src/gl_pure_sum_order4.m:1-15).

    ``s`` is ``fft2(eta11_plus/h)`` on the full grid, ``qx = kx h``, ``qy = ky h``,
    ``qp = kp h``.  Returns ``(eta44, psi44, audit)`` with ``eta44`` the analytic
    eta44/h (take the real part for the physical field).  The reference's
    C++ runtime (paper/order4/cpp/gl_eta44_fftw.cpp) computes eta44 only; so
    does this build: ``psi44`` is ``None``.
    """
    eta, audit = _order4_field(s, qx, qy, qp, int(rank), precision)
    audit["release_status"] = "experimental no-Stokes; eta44 only"
    return eta, None, audit


def _order4_field(s, qx, qy, qp, rank: int, precision: str, outer: int = -1, part: int = -2):
    """eta44/h on the GPU; ``outer >= 0`` keeps one outer (quartic) GL node and ``part`` selects,
    within it, the 2+2/2+1+1/1+1+1+1 partitions (-1) or the 3+1 partition of one third-order
    node (>= 0), as the reference's split runs do (paper/order4/cpp/gl_eta44_fftw.cpp:694-716)."""
    import torch

    host = _plan.build_order4_plan(np.asarray(s), np.asarray(qx, float), np.asarray(qy, float), peak_q=float(qp),
                                   rank=rank, precision=_precision(precision))
    o4 = host.o4
    if outer >= 0:  # the other outer nodes weigh zero in the quartic sum
        keep = np.zeros(rank)
        keep[outer] = 1.0
        o4["quartic_coef"] = o4["quartic_coef"] * keep
    if part == -1:  # no 3+1 partition: zero third-order state
        o4["third_coef"] = np.zeros_like(o4["third_coef"])
    elif part >= 0:  # the 3+1 partition of third-order node `part` only
        keep = np.zeros(rank)
        keep[part] = 1.0
        o4["third_coef"] = o4["third_coef"] * keep
        o4["forcing_terms"] = 1
    dp = DevicePlan(host, host.box1.take(np.asarray(s, dtype=np.complex128)), max_instants=1)
    try:
        e, _ = dp.run("order4", np.array([0.0]))
        torch.cuda.synchronize()
        eta = e[0].cpu().numpy()
    finally:
        dp.close()
    audit = dict(host.audits["eta44"])
    if outer >= 0:
        audit["outer_index"], audit["part_index"] = outer, part
    return eta, audit


def run_glrti(input_path, output_path=None, rank: int = 6, *, outer: int | None = None, part: int | None = None,
              precision: str = "fp64") -> dict:
    """Drop-in for ``gl_eta44 INPUT OUTPUT RANK THREADS [OUTER4_INDEX [PART_INDEX]]``
    (paper/order4/cpp/gl_eta44_fftw.cpp:639-764).

    Reads a GLRTI001 file (spectrum, dk, depth, peak kh, GL rules), evaluates
    the pure-sum GL eta44 on the GPU and optionally writes GLE44O01 (the
    dimensionless analytic eta44 on the n x n grid).  The GL rule of the
    requested rank must be present in the file, as the reference requires;
    this build regenerates it (Golub--Welsch) and checks it matches.
    ``outer`` (one outer GL node; negative: all) and ``part`` (-1: the
    partitions without the third-order state, p >= 0: the 3+1 partition of
    third-order node p; needs one outer node) give the reference's split-run
    partial fields, which ``formats.merge_gle44`` sums.
    """
    import time

    from . import formats

    o = -1 if outer is None else int(outer)
    pt = -2 if part is None else int(part)
    if o >= int(rank):
        raise ValueError("outer index")  # reference messages (gl_eta44_fftw.cpp:643-646)
    if pt >= int(rank) or pt < -2:
        raise ValueError("part index")
    if part is not None and o < 0:
        raise ValueError("split mode requires one outer4 index")
    inp = formats.read_glrti(input_path)
    if int(rank) not in inp["rules"]:
        raise ValueError("rank")  # reference message (gl_eta44_fftw.cpp:690)
    x, w = _plan.gauss_laguerre_rule(int(rank))
    fx, fw = inp["rules"][int(rank)]
    if not (np.allclose(x, fx, rtol=1e-12, atol=0) and np.allclose(w, fw, rtol=1e-9, atol=0)):
        raise ValueError("GL rule in the input differs from the Golub--Welsch rule of that rank")
    n = int(inp["n"])
    mode = np.rint(np.fft.fftfreq(n) * n)
    mx, my = np.meshgrid(mode, mode)
    h, dk = inp["depth"], inp["dk"]
    t0 = time.perf_counter()
    eta, audit = _order4_field(inp["spectrum"], h * dk * mx, h * dk * my, inp["peak_kh"], int(rank), precision,
                               outer=o, part=pt if o >= 0 else -2)
    elapsed = time.perf_counter() - t0
    if output_path is not None:
        formats.write_gle44(output_path, eta, rank=int(rank), threads=1, elapsed=elapsed)
    return {"field": eta, "elapsed": elapsed, "audit": audit}
