"""Build the oracle's native pieces (test infrastructure only).

* oracle/_ref/wit_eta4 -- the reference WIT executable, compiled from its own
  source where it lies (paper/order4/cpp/directional_eta4_openmp.cpp under the
  reference tree) with g++ -O3 -std=c++17 -fopenmp, only when that tree is
  present.  Never copied into this repository.
* oracle/_ref/libwit_oracle.so -- this project's C++ restatement of the WIT
  quartet kernel (oracle/wit_oracle.cpp), built with g++.
"""

from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
REF = Path(os.environ.get("GLW_REFERENCE_ROOT", "/root/reference"))
OUT = HERE / "_ref"


def build_reference() -> Path | None:
    src = REF / "paper" / "order4" / "cpp" / "directional_eta4_openmp.cpp"
    if not src.exists():
        return None
    OUT.mkdir(exist_ok=True)
    exe = OUT / "wit_eta4"
    subprocess.run([shutil.which("g++") or "g++", "-O3", "-std=c++17", "-fopenmp", str(src), "-o", str(exe)],
                   check=True)
    return exe


def build_restatement() -> Path | None:
    src = HERE / "wit_oracle.cpp"
    if not src.exists():
        return None
    OUT.mkdir(exist_ok=True)
    lib = OUT / "libwit_oracle.so"
    subprocess.run([shutil.which("g++") or "g++", "-O3", "-std=c++17", "-fopenmp", "-shared", "-fPIC", str(src),
                    "-o", str(lib)], check=True)
    return lib


if __name__ == "__main__":
    print(build_reference(), build_restatement())
