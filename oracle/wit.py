"""ctypes wrapper of the oracle's WIT restatement (oracle/wit_oracle.cpp).

TEST INFRASTRUCTURE ONLY.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "_ref" / "libwit_oracle.so"
REF_EXE = Path(__file__).resolve().parent / "_ref" / "wit_eta4"


def _lib():
    if not LIB.exists():
        from . import build_oracle

        build_oracle.build_restatement()
    lib = ctypes.CDLL(str(LIB))
    f = lib.wit_oracle_sector44
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                  ctypes.c_double, ctypes.c_double, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_int64,
                  ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    g = lib.wit_oracle_signed
    g.restype = ctypes.c_int
    g.argtypes = [ctypes.c_int32] + f.argtypes
    return lib


def sector44(inp: dict, first_range=None) -> dict:
    """Sector-44 WIT sum of a DIR4IN01 dict; optionally only first indices in [a, b)."""
    modes = np.ascontiguousarray(inp["modes"], dtype=np.int32)
    amps = np.ascontiguousarray(inp["amplitudes"], dtype=np.complex128)
    m = modes.shape[0]
    cases = amps.shape[0]
    nx, ny = int(inp["nx"]), int(inp["ny"])
    out = np.zeros((5, cases, ny, nx), dtype=np.complex128)
    cnt = np.zeros(8)
    a, b = first_range if first_range else (0, m)
    _lib().wit_oracle_sector44(m, modes.ctypes.data, amps.ctypes.data, cases, inp["gravity"], inp["depth"],
                               inp["dkx"], inp["dky"], nx, ny, inp["threshold"], a, b, out.ctypes.data,
                               cnt.ctypes.data)
    names = ("eta", "psi", "eta_coordinate", "psi_coordinate", "retained")
    res = {k: out[i] for i, k in enumerate(names)}
    res.update(total=cnt[0], kept=cnt[1], near=cnt[2], strict_zero_output=cnt[3], strict_zero_intermediate=cnt[4],
               max_residual=cnt[5], min_detuning=cnt[6], max_strict_zero_source=cnt[7])
    return res


def signed(inp: dict, sector: int, outer_range=None) -> dict:
    """Signed sector 40 or 42 of a DIR4IN01 dict; optionally only the outer candidate loop indices
    [a, b) (sector 40: minus pairs i <= j in row-major order; sector 42: minus modes)."""
    modes = np.ascontiguousarray(inp["modes"], dtype=np.int32)
    amps = np.ascontiguousarray(inp["amplitudes"], dtype=np.complex128)
    m = modes.shape[0]
    cases = amps.shape[0]
    nx, ny = int(inp["nx"]), int(inp["ny"])
    out = np.zeros((5, cases, ny, nx), dtype=np.complex128)
    cnt = np.zeros(8)
    a, b = outer_range if outer_range else (0, m * (m + 1) // 2 if sector == 40 else m)
    rc = _lib().wit_oracle_signed(int(sector), m, modes.ctypes.data, amps.ctypes.data, cases, inp["gravity"],
                                  inp["depth"], inp["dkx"], inp["dky"], nx, ny, inp["threshold"], a, b,
                                  out.ctypes.data, cnt.ctypes.data)
    if rc:
        raise ValueError("sector must be 40 or 42")
    names = ("eta", "psi", "eta_coordinate", "psi_coordinate", "retained")
    res = {k: out[i] for i, k in enumerate(names)}
    res.update(total=cnt[0], kept=cnt[1], near=cnt[2], strict_zero_output=cnt[3], strict_zero_intermediate=cnt[4],
               max_residual=cnt[5], min_detuning=cnt[6], max_strict_zero_source=cnt[7])
    return res
