"""GPU WIT sums (sector 44; signed sectors 40, 42) vs the reference executable's outputs and the oracle restatement.

Tolerance 1e-12 relative L2 (measured ~1e-16 for the restatement): the WIT sum
is a well-conditioned sum of positive-definite-denominator terms; the GPU adds
contributions to each bin with FP64 atomics in a schedule-dependent order.
"""

import numpy as np
import pytest

from cases import GOLDEN, rel
from green_laplace_wave_reconstruction_amd import formats
from test_wit_oracle import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def witmod():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from green_laplace_wave_reconstruction_amd import wit as W

    return W


@pytest.mark.parametrize("name", ["n0050", "n0100", "n0382"])
def test_wit_matches_reference_outputs(witmod, name):
    print(f"WIT {name}", flush=True)
    r = witmod.run_dir4in(GOLDEN / f"wit_{name}_input.bin")
    eta, psi, c = golden(name)
    print(f"WIT {name}: {r['kept']} quartets in {r['seconds']:.3f} s", flush=True)
    assert rel(r["eta"], eta) < 1e-12
    assert rel(r["psi"], psi) < 1e-12
    for k in ("total", "kept", "near", "strict_zero_output", "strict_zero_intermediate"):
        assert r[k] == c[k], k
    assert r["max_residual"] < 10 * c["max_residual"] + 1e-15
    assert r["min_detuning"] == pytest.approx(c["min_detuning"], rel=1e-12)
    assert not np.any(r["eta_coordinate"]) and not np.any(r["retained"])


SIGNED_ARRAYS = ("eta", "psi", "eta_coordinate", "psi_coordinate", "retained")


def _check_all_arrays(witmod, name, sector):
    g = np.load(GOLDEN / f"wit_{name}_sector{sector}.npz")
    c = dict(zip(("total", "kept", "near", "strict_zero_output", "strict_zero_intermediate", "max_residual",
                  "min_detuning", "max_strict_zero_source"), g["counters"]))
    print(f"WIT {name} sector {sector}", flush=True)
    r = witmod.run_dir4in(GOLDEN / f"wit_{name}_input.bin", sector=sector)
    print(f"WIT {name} sector {sector}: {r['kept']} candidates in {r['seconds']:.3f} s", flush=True)
    for k in ("total", "kept", "near", "strict_zero_output", "strict_zero_intermediate"):
        assert r[k] == int(c[k]), k
    # a term with normalized detuning rho is solved through a denominator ~rho: rounding moves it by
    # ~eps / rho (sector 42 reaches rho = 6.6e-7, sector 40 8e-5)
    tol = max(1e-12, 1e-15 / c["min_detuning"])
    for k in SIGNED_ARRAYS:
        ref = np.zeros(int(np.prod(g["shape"])), complex)
        ref[g["index"]] = g[k]
        if not np.any(ref):  # no near-resonant quartet: nothing retained
            assert not np.any(r[k]), k
            continue
        err = rel(r[k], ref.reshape(tuple(g["shape"])))
        print(f"  {k}: {err:.2e} (tolerance {tol:.1e})", flush=True)
        assert err < tol, k
    assert r["max_residual"] < 10 * c["max_residual"] + 1e-15
    assert r["min_detuning"] == pytest.approx(c["min_detuning"], rel=1e-12)
    assert r["max_strict_zero_source"] == pytest.approx(c["max_strict_zero_source"], rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("sector", [40, 42])
def test_wit_signed_sectors_match_reference_outputs(witmod, sector):
    """Sectors 40 and 42 (signed parents: retained-V3 push, y31 coordinate) on n0050 against the
    reference executable's own outputs (tests/golden/make_wit_golden.py): all five arrays, the
    counters; sector 42 has two near-resonant quartets (retained bins)."""
    _check_all_arrays(witmod, "n0050", sector)


@pytest.mark.parametrize("sector", [44, 40, 42])
def test_wit_m200_matches_reference_outputs(witmod, sector):
    """The 512^2, 200-mode runtime input (the reference's M = 200 timing case) in every sector against
    the reference executable's outputs: 6.9e7 quartets (44), 4.0e8 / 2.7e8 candidates (40 / 42) --
    the kernels and launch shapes of the timed configuration, including the four-wave signed kernel."""
    _check_all_arrays(witmod, "n0200", sector)


def test_wit_m400_matches_reference_outputs(witmod):
    """The 512^2, 400-mode runtime input (the reference's M = 400 timing case, 1.08e9 quartets, 71 min on
    6 cores for the reference executable in the build container) in sector 44 against its outputs: every
    array and counter, in FP64 and with FP32 evaluation (bound 1e-6, as for the smaller inputs)."""
    _check_all_arrays(witmod, "n0400", 44)
    g = np.load(GOLDEN / "wit_n0400_sector44.npz")
    inp = formats.read_dir4in(GOLDEN / "wit_n0400_input.bin")
    plan = witmod.WitPlan(inp["modes"], nx=inp["nx"], ny=inp["ny"], gravity=inp["gravity"], depth=inp["depth"],
                          dkx=inp["dkx"], dky=inp["dky"], threshold=inp["threshold"], precision="fp32")
    r = plan.sector44(inp["amplitudes"])
    plan.close()
    for k in ("eta", "psi"):
        ref = np.zeros(int(np.prod(g["shape"])), complex)
        ref[g["index"]] = g[k]
        err = rel(r[k], ref.reshape(tuple(g["shape"])))
        print(f"WIT n0400 FP32 {k}: {err:.2e}", flush=True)
        assert err < 1e-6, k
    assert r["kept"] == int(g["counters"][1])


@pytest.mark.parametrize("name", ["n0050", "n0100"])
def test_wit_global_atomic_path(witmod, name):
    """Every quartet added with global atomics (the path of LDS windows above 128 KB) against the
    reference outputs."""
    inp = formats.read_dir4in(GOLDEN / f"wit_{name}_input.bin")
    plan = witmod.WitPlan(inp["modes"], nx=inp["nx"], ny=inp["ny"], gravity=inp["gravity"], depth=inp["depth"],
                          dkx=inp["dkx"], dky=inp["dky"], threshold=inp["threshold"])
    r = plan.sector44(inp["amplitudes"], global_atomics=True)
    plan.close()
    eta, psi, c = golden(name)
    assert rel(r["eta"], eta) < 1e-12 and rel(r["psi"], psi) < 1e-12
    for k in ("total", "kept", "near", "strict_zero_output", "strict_zero_intermediate"):
        assert r[k] == c[k], k


def test_wit_m800_first_index_slice_vs_oracle(witmod):
    """M = 800 (the reference's 21,857 s case, beyond a full CPU oracle here): the quartets with first
    index in [703, 800) -- 3.9e6 of 1.7e10 -- against the oracle restatement over the same slice."""
    from oracle import wit as ow

    inp = formats.read_dir4in(GOLDEN / "wit_n0800_input.bin")
    rng = (703, 800)
    print("WIT n0800 slice: oracle", flush=True)
    ref = ow.sector44(inp, rng)
    print("WIT n0800 slice: GPU", flush=True)
    plan = witmod.WitPlan(inp["modes"], nx=inp["nx"], ny=inp["ny"], gravity=inp["gravity"], depth=inp["depth"],
                          dkx=inp["dkx"], dky=inp["dky"], threshold=inp["threshold"])
    r = plan.sector44(inp["amplitudes"], first_range=rng)
    plan.close()
    assert r["total"] == 3921225  # sum over i of C(800 - i + 2, 3) = C(100, 4)
    for k in ("kept", "near", "strict_zero_output", "strict_zero_intermediate"):
        assert r[k] == int(ref[k]), k
    for k in ("eta", "psi"):
        assert rel(r[k], ref[k]) < 1e-12, k
    assert r["max_residual"] < 10 * ref["max_residual"] + 1e-15


def test_wit_ranges_add_up_and_spread(witmod):
    """Linearity over the first index: two ranges sum to the whole sector-44 sum, and the run-to-run
    spread of the FP64 atomic accumulation (schedule-dependent summation order) is at rounding level."""
    inp = formats.read_dir4in(GOLDEN / "wit_n0100_input.bin")
    plan = witmod.WitPlan(inp["modes"], nx=inp["nx"], ny=inp["ny"], gravity=inp["gravity"], depth=inp["depth"],
                          dkx=inp["dkx"], dky=inp["dky"], threshold=inp["threshold"])
    full = plan.sector44(inp["amplitudes"])
    again = plan.sector44(inp["amplitudes"])
    lo = plan.sector44(inp["amplitudes"], first_range=(0, 37))
    hi = plan.sector44(inp["amplitudes"], first_range=(37, 100))
    plan.close()
    assert lo["total"] + hi["total"] == full["total"] and lo["kept"] + hi["kept"] == full["kept"]
    for k in ("eta", "psi"):
        assert rel(lo[k] + hi[k], full[k]) < 1e-14, k
    nz = full["eta"] != 0
    spread = float(np.max(np.abs(again["eta"][nz] - full["eta"][nz]) / np.abs(full["eta"][nz])))
    print(f"WIT n0100 run-to-run: max relative bin difference {spread:.2e}, rel-L2 "
          f"{rel(again['eta'], full['eta']):.2e}", flush=True)
    assert rel(again["eta"], full["eta"]) < 1e-15 and spread < 1e-12


def test_wit_near_resonant_branch_vs_restatement(witmod):
    """Force the retained-V3 branch with a large threshold; compare with the CPU restatement."""
    from oracle import wit as ow

    inp = formats.read_dir4in(GOLDEN / "wit_n0050_input.bin")
    inp["modes"] = inp["modes"][:24]
    inp["amplitudes"] = inp["amplitudes"][:, :24]
    inp["threshold"] = 0.6
    ref = ow.sector44(inp)
    assert ref["near"] > 0
    plan = witmod.WitPlan(inp["modes"], nx=inp["nx"], ny=inp["ny"], gravity=inp["gravity"], depth=inp["depth"],
                          dkx=inp["dkx"], dky=inp["dky"], threshold=inp["threshold"])
    r = plan.sector44(inp["amplitudes"])
    plan.close()
    assert r["near"] == ref["near"] and r["kept"] == ref["kept"]
    for k in ("eta", "psi", "retained"):
        assert rel(r[k], ref[k]) < 1e-12, k


def test_wit_homogeneity_degree_four(witmod):
    inp = formats.read_dir4in(GOLDEN / "wit_n0100_input.bin")
    plan = witmod.WitPlan(inp["modes"], nx=inp["nx"], ny=inp["ny"], gravity=inp["gravity"], depth=inp["depth"],
                          dkx=inp["dkx"], dky=inp["dky"], threshold=inp["threshold"])
    a = plan.sector44(inp["amplitudes"])
    b = plan.sector44(2.0 * inp["amplitudes"])
    plan.close()
    assert rel(b["eta"], 16.0 * a["eta"]) < 1e-14


@pytest.mark.parametrize("name", ["n0050", "n0100", "n0382"])
def test_wit_fp32_evaluation_vs_reference_outputs(witmod, name):
    """FP32 degree-4 evaluation (FP32 copies of the FP64 subset tables), FP64 weights and accumulation.

    Bound 1e-6 relative L2 (measured 2-3e-8): each quartet term carries FP32
    rounding (~1e-7 relative); the sum itself is exact to FP64 rounding.
    """
    inp = formats.read_dir4in(GOLDEN / f"wit_{name}_input.bin")
    plan = witmod.WitPlan(inp["modes"], nx=inp["nx"], ny=inp["ny"], gravity=inp["gravity"], depth=inp["depth"],
                          dkx=inp["dkx"], dky=inp["dky"], threshold=inp["threshold"], precision="fp32")
    r = plan.sector44(inp["amplitudes"])
    plan.close()
    eta, psi, c = golden(name)
    e1, e2 = rel(r["eta"], eta), rel(r["psi"], psi)
    print(f"WIT {name} FP32: eta {e1:.2e} psi {e2:.2e}", flush=True)
    assert e1 < 1e-6 and e2 < 1e-6
    for k in ("total", "kept", "near", "strict_zero_output"):
        assert r[k] == c[k], k
