"""Build the in-tree HIP library ``_lib/libglw.so`` for gfx950 with hipcc."""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
ROW_SETS = 8  # operator sets of glw_launch_rows.hip (see GLW_ROW_SET there)
# (source, extra defines, object name): one hipcc process per unit, run in parallel
UNITS = [
    (CSRC / "glw_plan.hip", [], "glw_plan.o"),
    (CSRC / "glw_wit.hip", [], "glw_wit.o"),
    (CSRC / "glw_batch.hip", [], "glw_batch.o"),
    (CSRC / "glw_modal.hip", [], "glw_modal.o"),
    *[(CSRC / "glw_modal_launch.hip", [f"-DGLW_MODAL_SET={i}"], f"glw_modal_launch{i}.o") for i in range(4)],
    *[(CSRC / "glw_launch_cols.hip", [f"-DGLW_COL_SET={i}"], f"glw_launch_cols{i}.o") for i in range(2)],
    *[(CSRC / "glw_launch_rows.hip", [f"-DGLW_ROW_SET={i}"], f"glw_launch_rows{i}.o") for i in range(ROW_SETS)],
]
SOURCES = sorted({u[0] for u in UNITS})
OUT = PKG / "_lib" / "libglw.so"
OBJ = PKG / "_lib" / "obj"

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _compile(unit, verbose: bool) -> Path:
    src, defines, name = unit
    obj = OBJ / name
    cmd = [hipcc(), *HIPCC_FLAGS, *defines, "-I", str(ROOT / "include"), "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every translation unit (in parallel) and link ``libglw.so``."""
    from concurrent.futures import ThreadPoolExecutor

    OBJ.mkdir(parents=True, exist_ok=True)
    jobs = jobs or min(len(UNITS), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(lambda u: _compile(u, verbose), UNITS))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *[str(o) for o in objs], "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
