"""Shared seeded inputs for the parity tests (CPU-side construction only)."""

from __future__ import annotations

from pathlib import Path

import numpy as np

from green_laplace_wave_reconstruction_amd import seastate

GOLDEN = Path(__file__).resolve().parent / "golden"


def fixture(name: str) -> dict:
    d = np.load(GOLDEN / f"reference_spatial_{name}.npz")
    return {k: d[k] for k in d.files}


def rel(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def components(state: seastate.SeaState):
    return state.a, state.b, state.kx, state.ky


def headline():
    return seastate.order3_headline_512()


def order2_plain():
    return seastate.order2_plain_128()
