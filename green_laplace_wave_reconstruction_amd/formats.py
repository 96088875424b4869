"""Readers and writers of the reference's order-4 binary formats.

* DIR4IN01 -- WIT input (paper/order4/cpp/directional_eta4_openmp.cpp:518-557):
  magic8, u64 M, nx, ny, cases, f64 g, h, dkx, dky, threshold,
  i32 modes[M][2], c128 amplitudes[cases][M].
* DIR40O01 / DIR42O01 / DIR44O01 -- WIT output (:747-773): magic8, u64 nx, ny,
  cases, total, kept, near, strictZeroOutput, strictZeroIntermediate,
  f64 seconds, maxResidual, minDetuning, maxStrictZeroSource, then five
  c128[cases][ny][nx] arrays (x fastest): eta, psi, eta coordinate,
  psi coordinate, retained.
* GLRTI001 -- GL eta44 input (paper/order4/cpp/gl_eta44_fftw.cpp:648-691):
  magic8, u64 n, f64 dk, h, peak_kh, u64 nrules, {u64 J, f64 x[J], w[J]}...,
  c128 S[n*n] (linear index iy + n*ix).
* GLE44O01 -- GL eta44 output (:753-763): magic8, u64 n, rank, threads,
  f64 elapsed, c128 field[n*n].

All little-endian, as the reference writes them.
"""

from __future__ import annotations

import struct
from pathlib import Path

import numpy as np


def read_dir4in(path) -> dict:
    b = Path(path).read_bytes()
    if b[:8] != b"DIR4IN01":
        raise ValueError("directional input magic mismatch")
    m, nx, ny, cases = struct.unpack_from("<4Q", b, 8)
    g, h, dkx, dky, thr = struct.unpack_from("<5d", b, 40)
    off = 80
    modes = np.frombuffer(b, dtype="<i4", count=2 * m, offset=off).reshape(m, 2).copy()
    off += 8 * m
    amps = np.frombuffer(b, dtype="<c16", count=cases * m, offset=off).reshape(cases, m).copy()
    return {"modes": modes, "amplitudes": amps, "nx": nx, "ny": ny, "cases": cases, "gravity": g, "depth": h,
            "dkx": dkx, "dky": dky, "threshold": thr}


def write_dir4in(path, modes, amplitudes, *, nx, ny, gravity, depth, dkx, dky, threshold) -> None:
    modes = np.ascontiguousarray(modes, dtype="<i4").reshape(-1, 2)
    amps = np.ascontiguousarray(np.atleast_2d(amplitudes), dtype="<c16")
    if amps.shape[1] != modes.shape[0]:
        raise ValueError("amplitudes must be [cases][M]")
    head = b"DIR4IN01" + struct.pack("<4Q5d", modes.shape[0], nx, ny, amps.shape[0], gravity, depth, dkx, dky,
                                     threshold)
    Path(path).write_bytes(head + modes.tobytes() + amps.tobytes())


_OUT_FIELDS = ("eta", "psi", "eta_coordinate", "psi_coordinate", "retained")


def read_dir4out(path) -> dict:
    b = Path(path).read_bytes()
    magic = b[:8].decode()
    if magic not in ("DIR40O01", "DIR42O01", "DIR44O01"):
        raise ValueError("directional output magic mismatch")
    nx, ny, cases, total, kept, near, zo, zi = struct.unpack_from("<8Q", b, 8)
    secs, maxres, mindet, maxz = struct.unpack_from("<4d", b, 72)
    off = 104
    n = nx * ny * cases
    out = {"sector": int(magic[3:5]), "nx": nx, "ny": ny, "cases": cases, "total": total, "kept": kept,
           "near": near, "strict_zero_output": zo, "strict_zero_intermediate": zi, "seconds": secs,
           "max_residual": maxres, "min_detuning": mindet, "max_strict_zero_source": maxz}
    for name in _OUT_FIELDS:
        out[name] = np.frombuffer(b, dtype="<c16", count=n, offset=off).reshape(cases, ny, nx).copy()
        off += 16 * n
    return out


def write_dir4out(path, sector: int, arrays: dict, counters: dict) -> None:
    eta = np.asarray(arrays["eta"])
    cases, ny, nx = eta.shape
    head = (f"DIR{sector}O01".encode() +
            struct.pack("<8Q4d", nx, ny, cases, int(counters["total"]), int(counters["kept"]), int(counters["near"]),
                        int(counters["strict_zero_output"]), int(counters["strict_zero_intermediate"]),
                        float(counters.get("seconds", 0.0)), float(counters["max_residual"]),
                        float(counters["min_detuning"]), float(counters["max_strict_zero_source"])))
    body = b"".join(np.ascontiguousarray(arrays[k], dtype="<c16").tobytes() for k in _OUT_FIELDS)
    Path(path).write_bytes(head + body)


def read_glrti(path) -> dict:
    b = Path(path).read_bytes()
    if b[:8] != b"GLRTI001":
        raise ValueError("GL runtime input magic mismatch")
    (n,) = struct.unpack_from("<Q", b, 8)
    dk, h, pkh = struct.unpack_from("<3d", b, 16)
    (nr,) = struct.unpack_from("<Q", b, 40)
    off = 48
    rules = {}
    for _ in range(nr):
        (j,) = struct.unpack_from("<Q", b, off)
        off += 8
        x = np.frombuffer(b, "<f8", j, off).copy()
        off += 8 * j
        w = np.frombuffer(b, "<f8", j, off).copy()
        off += 8 * j
        rules[int(j)] = (x, w)
    s = np.frombuffer(b, "<c16", n * n, off).reshape(n, n).T.copy()  # stored iy + n*ix -> [iy][ix]
    return {"n": n, "dk": dk, "depth": h, "peak_kh": pkh, "rules": rules, "spectrum": s}


def write_glrti(path, spectrum, *, dk, depth, peak_kh, rules: dict) -> None:
    s = np.asarray(spectrum, dtype=np.complex128)
    n = s.shape[0]
    parts = [b"GLRTI001", struct.pack("<Q3dQ", n, dk, depth, peak_kh, len(rules))]
    for j, (x, w) in rules.items():
        parts += [struct.pack("<Q", j), np.asarray(x, "<f8").tobytes(), np.asarray(w, "<f8").tobytes()]
    parts.append(np.ascontiguousarray(s.T, dtype="<c16").tobytes())
    Path(path).write_bytes(b"".join(parts))


def read_gle44(path) -> dict:
    b = Path(path).read_bytes()
    if b[:8] != b"GLE44O01":
        raise ValueError("GL eta44 output magic mismatch")
    n, rank, threads = struct.unpack_from("<3Q", b, 8)
    (elapsed,) = struct.unpack_from("<d", b, 32)
    f = np.frombuffer(b, "<c16", n * n, 40).reshape(n, n).T.copy()
    return {"n": n, "rank": rank, "threads": threads, "elapsed": elapsed, "field": f}


def merge_gle44(output_path, input_paths) -> dict:
    """Sum partial GLE44O01 fields (paper/order4/cpp/merge_eta44_outer.cpp): same n and rank
    required; the header records the partial count as threads and the largest elapsed time."""
    paths = list(input_paths)
    if len(paths) < 2:
        raise ValueError("usage OUTPUT INPUT1 INPUT2 ...")
    total = None
    for p in paths:
        r = read_gle44(p)
        if total is None:
            n, rank, wall, total = r["n"], r["rank"], r["elapsed"], r["field"].copy()
        else:
            if r["n"] != n or r["rank"] != rank:
                raise ValueError("incompatible partial outputs")
            wall = max(wall, r["elapsed"])
            total += r["field"]
    write_gle44(output_path, total, rank=rank, threads=len(paths), elapsed=wall)
    return {"n": n, "rank": rank, "partials": len(paths), "elapsed": wall, "field": total}


def write_gle44(path, field, *, rank: int, threads: int = 1, elapsed: float = 0.0) -> None:
    f = np.asarray(field, dtype=np.complex128)
    n = f.shape[0]
    Path(path).write_bytes(b"GLE44O01" + struct.pack("<3Qd", n, rank, threads, elapsed) +
                           np.ascontiguousarray(f.T, dtype="<c16").tobytes())
