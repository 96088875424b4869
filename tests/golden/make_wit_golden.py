"""WIT golden vectors from the reference (run in the build container only).

* copies the reference's DIR4IN01 inputs paper/order4/inputs/runtime/n0050,
  n0100, n0200, n0400, n0800 and paper/order4/inputs/field_n0382 (data files)
  to tests/golden/ (n0200..n0800 are timing inputs: no golden output);
* runs the reference executable (oracle/_ref/wit_eta4, built from the
  reference source by oracle/build_oracle.py) on n0050 and n0100, sector 44;
* converts those outputs and the reference's shipped n0382 output
  (paper/order4/data/wit_field_n0382.bin) to npz: eta and psi on their
  nonzero bins plus the header counters (the other three arrays are zero in
  these runs: near = 0);
* runs it on n0050 in sectors 40 and 42 (signed parents) and keeps all five
  arrays on their nonzero bins (eta and psi coordinates, and in sector 42
  two near-resonant retained bins);
* with ``--m200 [DIR] [--sectors=44,40,42]`` only: runs it on n0200 (the 512^2, 200-mode runtime
  input) in sectors 44, 40 and 42 -- about 3, 8 and 11 minutes on 8 cores --
  or reuses DIR/out_s{44,40,42}.bin from an earlier run of
  ``wit_eta4 tests/golden/wit_n0200_input.bin DIR/out_sS.bin --sector S``,
  and keeps all five arrays on their nonzero bins;
* with ``--m400 [DIR]`` only: the same for n0400 (the 400-mode runtime input) in
  sector 44 -- 1.08e9 quartets, about 71 minutes on 6 cores -- or reuses
  DIR/out_s44.bin from an earlier ``wit_eta4 tests/golden/wit_n0400_input.bin
  DIR/out_s44.bin --sector 44``.
"""

import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
from green_laplace_wave_reconstruction_amd import formats  # noqa: E402

ARGS = [x for x in sys.argv[1:] if not x.startswith("--")]
M200 = "--m200" in sys.argv
M400 = "--m400" in sys.argv
SECTORS = [int(x) for a in sys.argv if a.startswith("--sectors=") for x in a.split("=", 1)[1].split(",")] or [44, 40, 42]
REF = Path(ARGS[0] if ARGS and not (M200 or M400) else "/root/reference")
EXE = ROOT / "oracle" / "_ref" / "wit_eta4"


def save(name, out):
    for k in ("retained", "eta_coordinate", "psi_coordinate"):
        assert not np.any(out[k]), k
    idx = np.flatnonzero((out["eta"] != 0) | (out["psi"] != 0))
    np.savez_compressed(HERE / f"wit_{name}_sector44.npz", index=idx, eta=out["eta"].ravel()[idx],
                        psi=out["psi"].ravel()[idx], shape=np.array(out["eta"].shape),
                        counters=np.array([out[k] for k in ("total", "kept", "near", "strict_zero_output",
                                                            "strict_zero_intermediate", "max_residual",
                                                            "min_detuning", "max_strict_zero_source")], float))
    print(name, "nonzero bins", idx.size, "kept", out["kept"])


ARRAYS = ("eta", "psi", "eta_coordinate", "psi_coordinate", "retained")
COUNTERS = ("total", "kept", "near", "strict_zero_output", "strict_zero_intermediate", "max_residual",
            "min_detuning", "max_strict_zero_source")


def save_all(path, out):
    idx = np.flatnonzero(np.any([out[k] != 0 for k in ARRAYS], axis=0))
    np.savez_compressed(path, index=idx, shape=np.array(out["eta"].shape),
                        counters=np.array([out[k] for k in COUNTERS], float),
                        **{k: out[k].ravel()[idx] for k in ARRAYS})
    return idx.size


if M400:
    given = Path(ARGS[0]) if ARGS else None
    reuse = given / "out_s44.bin" if given else None
    with tempfile.TemporaryDirectory() as td:
        out_path = reuse if reuse and reuse.exists() else Path(td) / "out.bin"
        if out_path != reuse:
            subprocess.run([str(EXE), str(HERE / "wit_n0400_input.bin"), str(out_path), "--sector", "44"], check=True)
        out = formats.read_dir4out(out_path)
    n = save_all(HERE / "wit_n0400_sector44.npz", out)
    print("n0400 sector 44 nonzero bins", n, "kept", out["kept"], "near", out["near"])
    sys.exit(0)

if M200:
    given = Path(ARGS[0]) if ARGS else None
    for sector in SECTORS:
        reuse = given / f"out_s{sector}.bin" if given else None
        with tempfile.TemporaryDirectory() as td:
            out_path = reuse if reuse and reuse.exists() else Path(td) / "out.bin"
            if out_path != reuse:
                subprocess.run([str(EXE), str(HERE / "wit_n0200_input.bin"), str(out_path), "--sector", str(sector)],
                               check=True)
            out = formats.read_dir4out(out_path)
        n = save_all(HERE / f"wit_n0200_sector{sector}.npz", out)
        print("n0200 sector", sector, "nonzero bins", n, "kept", out["kept"], "near", out["near"])
    sys.exit(0)

for name, rel in (("n0050", "inputs/runtime/n0050"), ("n0100", "inputs/runtime/n0100"),
                  ("n0200", "inputs/runtime/n0200"), ("n0400", "inputs/runtime/n0400"),
                  ("n0800", "inputs/runtime/n0800"), ("n0382", "inputs/field_n0382")):
    shutil.copy(REF / "paper" / "order4" / rel / "directional_eta44_input.bin", HERE / f"wit_{name}_input.bin")
for name in ("n0050", "n0100"):
    with tempfile.TemporaryDirectory() as td:
        out_path = Path(td) / "out.bin"
        subprocess.run([str(EXE), str(HERE / f"wit_{name}_input.bin"), str(out_path), "--sector", "44"], check=True)
        save(name, formats.read_dir4out(out_path))
save("n0382", formats.read_dir4out(REF / "paper" / "order4" / "data" / "wit_field_n0382.bin"))


for sector in (40, 42):
    with tempfile.TemporaryDirectory() as td:
        out_path = Path(td) / "out.bin"
        subprocess.run([str(EXE), str(HERE / "wit_n0050_input.bin"), str(out_path), "--sector", str(sector)],
                       check=True)
        out = formats.read_dir4out(out_path)
    n = save_all(HERE / f"wit_n0050_sector{sector}.npz", out)
    print("n0050 sector", sector, "nonzero bins", n, "kept", out["kept"], "near", out["near"])

