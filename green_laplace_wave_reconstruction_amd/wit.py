"""WIT quartet sums on the GPU (reference paper/order4/cpp/directional_eta4_openmp.cpp).

``WitPlan(modes, ...)`` tabulates the single/pair/triple subset states of a
mode set once; ``sector44(amplitudes)`` sums every unordered quartet
i<=j<=k<=l into the five DIR44O01 output arrays and the reference's counters.
``sector(40 | 42, amplitudes)`` runs the signed-parent sectors (plus pair x
minus pair; plus triple x minus mode), evaluated per candidate in FP64.
``run_dir4in(input, output, sector=...)`` is the drop-in for
``wit_eta4 INPUT OUTPUT --sector 40|42|44``.
"""

from __future__ import annotations

import ctypes
import time
from ctypes import POINTER, c_double, c_int, c_int32, c_void_p

import numpy as np

from . import formats
from ._native import GlwError, _check, load_library

FIELDS = ("eta", "psi", "eta_coordinate", "psi_coordinate", "retained")
COUNTERS = ("total", "kept", "near", "strict_zero_output", "strict_zero_intermediate", "max_residual",
            "min_detuning", "max_strict_zero_source")


class CWitDesc(ctypes.Structure):
    _fields_ = [("modes_count", c_int32), ("modes", POINTER(c_int32)), ("nx", c_int32), ("ny", c_int32),
                ("gravity", c_double), ("depth", c_double), ("dkx", c_double), ("dky", c_double),
                ("threshold", c_double), ("precision", c_int32)]


def _bind(lib):
    lib.glw_wit_create.argtypes = [POINTER(CWitDesc), POINTER(c_void_p)]
    lib.glw_wit_create.restype = c_int
    lib.glw_wit_destroy.argtypes = [c_void_p]
    lib.glw_wit_destroy.restype = c_int
    lib.glw_wit_sector44.argtypes = [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p]
    lib.glw_wit_sector44.restype = c_int
    lib.glw_wit_sector44_range.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_int32, c_void_p, c_void_p,
                                           c_void_p]
    lib.glw_wit_sector44_range.restype = c_int
    lib.glw_wit_sector.argtypes = [c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_void_p]
    lib.glw_wit_sector.restype = c_int
    return lib


class WitPlan:
    def __init__(self, modes, *, nx: int, ny: int, gravity: float, depth: float, dkx: float, dky: float,
                 threshold: float, precision: str = "fp64"):
        import torch

        if not torch.cuda.is_available():
            raise GlwError("a HIP device is required")
        self.lib = _bind(load_library())
        m = np.ascontiguousarray(modes, dtype=np.int32).reshape(-1, 2)
        if m.shape[0] > 4096:
            raise ValueError("at most 4096 modes (triple table size)")
        self.modes = m
        self.nx, self.ny = int(nx), int(ny)
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.precision = precision
        d = CWitDesc(m.shape[0], m.ctypes.data_as(POINTER(c_int32)), self.nx, self.ny, gravity, depth, dkx, dky,
                     threshold, 0 if precision == "fp64" else 1)
        h = c_void_p()
        _check(self.lib.glw_wit_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.device = torch.device("cuda", torch.cuda.current_device())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.glw_wit_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sector44_device(self, amplitudes, out=None, counters=None):
        """Device run: amplitudes complex128 [cases][M] (numpy or torch); returns (out, counters) tensors."""
        return self.sector_device(44, amplitudes, out, counters)

    def sector_device(self, sector: int, amplitudes, out=None, counters=None, *, first_range=None,
                      global_atomics: bool = False):
        """Device run of sector 40, 42 or 44 (see sector44_device).  Sector 44 only: ``first_range``
        (begin, end) restricts the sum to quartets whose first index lies in [begin, end) (the outermost
        loop of the reference's sector-44 sum); ``global_atomics`` adds every quartet to the output with
        global atomics instead of the workgroup's LDS window."""
        import torch

        if sector not in (40, 42, 44):
            raise ValueError("sector must be 40, 42, or 44")  # reference message (:511-513)

        a = torch.as_tensor(np.atleast_2d(amplitudes) if isinstance(amplitudes, np.ndarray) else amplitudes)
        a = a.to(device=self.device, dtype=torch.complex128).contiguous()
        if a.dim() != 2 or a.shape[1] != self.modes.shape[0]:
            raise ValueError("amplitudes must be [cases][M]")
        cases = a.shape[0]
        if out is None:
            out = torch.empty((5, cases, self.ny, self.nx), dtype=torch.complex128, device=self.device)
        if counters is None:
            counters = torch.empty(8, dtype=torch.float64, device=self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        if first_range is not None or global_atomics:
            if sector != 44:
                raise ValueError("first-index ranges and global atomics apply to sector 44")
            lo, hi = first_range if first_range is not None else (0, self.modes.shape[0])
            _check(self.lib.glw_wit_sector44_range(self._h, int(lo), int(hi), 1 if global_atomics else 0,
                                                   c_void_p(a.data_ptr()), cases, c_void_p(out.data_ptr()),
                                                   c_void_p(counters.data_ptr()), c_void_p(stream)))
            return out, counters
        _check(self.lib.glw_wit_sector(self._h, int(sector), c_void_p(a.data_ptr()), cases, c_void_p(out.data_ptr()),
                                       c_void_p(counters.data_ptr()), c_void_p(stream)))
        return out, counters

    def sector44(self, amplitudes, *, first_range=None, global_atomics: bool = False) -> dict:
        return self.sector(44, amplitudes, first_range=first_range, global_atomics=global_atomics)

    def sector(self, sector: int, amplitudes, **kw) -> dict:
        """Host run of one sector: the five output arrays and the counters, plus the wall time."""
        import torch

        t0 = time.perf_counter()
        out, cnt = self.sector_device(sector, amplitudes, **kw)
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        o = out.cpu().numpy()
        c = cnt.cpu().numpy()
        res = {k: o[i] for i, k in enumerate(FIELDS)}
        res.update({k: float(c[i]) for i, k in enumerate(COUNTERS)})
        for k in ("total", "kept", "near", "strict_zero_output", "strict_zero_intermediate"):
            res[k] = int(round(res[k]))
        res["seconds"] = secs
        return res


def run_dir4in(input_path, output_path=None, *, sector: int = 44) -> dict:
    """Drop-in for ``wit_eta4 INPUT OUTPUT --sector 40|42|44``: read DIR4IN01, sum, optionally write DIR4xO01."""
    if sector not in (40, 42, 44):
        raise ValueError("sector must be 40, 42, or 44")  # reference message (:511-513)
    inp = formats.read_dir4in(input_path)
    plan = WitPlan(inp["modes"], nx=inp["nx"], ny=inp["ny"], gravity=inp["gravity"], depth=inp["depth"],
                   dkx=inp["dkx"], dky=inp["dky"], threshold=inp["threshold"])
    try:
        res = plan.sector(sector, inp["amplitudes"])
    finally:
        plan.close()
    if output_path is not None:
        formats.write_dir4out(output_path, sector, res, res)
    return res
