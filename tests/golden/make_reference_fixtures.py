"""Convert the reference's MATLAB-generated fixtures to compact npz.

Reads python/tests/fixtures/directional_spatial_{eta22,order3,nonzero_time}.json
and unidirectional_{modal_order3,time_eta20}.json from the reference tree
(generated by python/tools/generate_matlab_fixtures.m) and writes
tests/golden/reference_spatial_*.npz and reference_unidirectional_*.npz with the
same arrays (float64).  These are data fixtures (inputs and expected outputs),
not source.
"""
import json
import sys
from pathlib import Path

import numpy as np

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/python/tests/fixtures")
DST = Path(__file__).parent

NAMES = {
    "directional_spatial_eta22": "reference_spatial_eta22",
    "directional_spatial_order3": "reference_spatial_order3",
    "directional_spatial_nonzero_time": "reference_spatial_nonzero_time",
    "unidirectional_modal_order3": "reference_unidirectional_modal_order3",
    "unidirectional_time_eta20": "reference_unidirectional_time_eta20",
}

for name, target in NAMES.items():
    data = json.loads((SRC / f"{name}.json").read_text(encoding="utf-8"))
    out = {}
    for key, value in data.items():
        if isinstance(value, list):
            out[key] = np.asarray(value, dtype=float)
        elif isinstance(value, (int, float)):
            out[key] = np.asarray(value, dtype=float)
        else:
            out[key] = np.asarray(str(value))
    np.savez_compressed(DST / f"{target}.npz", **out)
    print(name, sorted(out))

# Frozen GL8 nodes and weights of the reference's symbolic interface
# (symbolic/generated/finite_depth_directional_order2_eta22_pure_gl8.json).

frozen = json.loads((Path("/root/reference") / "symbolic" / "generated" /
                     "finite_depth_directional_order2_eta22_pure_gl8.json").read_text(encoding="utf-8"))
(DST / "frozen_gl8.json").write_text(json.dumps({
    "source": "symbolic/generated/finite_depth_directional_order2_eta22_pure_gl8.json",
    "quadrature_rank": frozen["quadrature_rank"], "nodes": frozen["nodes"], "weights": frozen["weights"]},
    indent=1) + "\n")
print("frozen_gl8.json")
