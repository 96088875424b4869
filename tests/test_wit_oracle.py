"""WIT oracle (CPU): the C++ restatement vs the reference executable's outputs, and the formats."""

import numpy as np
import pytest

from cases import GOLDEN, rel
from green_laplace_wave_reconstruction_amd import formats
from oracle import wit

COUNTERS = ("total", "kept", "near", "strict_zero_output", "strict_zero_intermediate", "max_residual",
            "min_detuning", "max_strict_zero_source")


def golden(name):
    g = np.load(GOLDEN / f"wit_{name}_sector44.npz")
    shape = tuple(g["shape"])
    eta = np.zeros(int(np.prod(shape)), complex)
    psi = np.zeros_like(eta)
    eta[g["index"]] = g["eta"]
    psi[g["index"]] = g["psi"]
    return eta.reshape(shape), psi.reshape(shape), dict(zip(COUNTERS, g["counters"]))


def test_restatement_matches_reference_n0050():
    inp = formats.read_dir4in(GOLDEN / "wit_n0050_input.bin")
    r = wit.sector44(inp)
    eta, psi, c = golden("n0050")
    assert rel(r["eta"], eta) < 1e-13
    assert rel(r["psi"], psi) < 1e-13
    for k in ("total", "kept", "near", "strict_zero_output", "strict_zero_intermediate"):
        assert r[k] == c[k], k
    assert r["max_residual"] == pytest.approx(c["max_residual"], rel=1e-6)
    assert r["min_detuning"] == pytest.approx(c["min_detuning"], rel=1e-12)


def test_restatement_first_index_split_is_additive():
    inp = formats.read_dir4in(GOLDEN / "wit_n0050_input.bin")
    inp["modes"] = inp["modes"][:20]
    inp["amplitudes"] = inp["amplitudes"][:, :20]
    full = wit.sector44(inp)
    a = wit.sector44(inp, (0, 7))
    b = wit.sector44(inp, (7, 20))
    assert rel(a["eta"] + b["eta"], full["eta"]) < 1e-14
    assert a["kept"] + b["kept"] == full["kept"]


def test_dir4_formats_roundtrip(tmp_path):
    inp = formats.read_dir4in(GOLDEN / "wit_n0050_input.bin")
    p = tmp_path / "in.bin"
    formats.write_dir4in(p, inp["modes"], inp["amplitudes"], nx=inp["nx"], ny=inp["ny"], gravity=inp["gravity"],
                         depth=inp["depth"], dkx=inp["dkx"], dky=inp["dky"], threshold=inp["threshold"])
    assert p.read_bytes() == (GOLDEN / "wit_n0050_input.bin").read_bytes()
    eta, psi, c = golden("n0050")
    arrays = {"eta": eta, "psi": psi, "eta_coordinate": 0 * eta, "psi_coordinate": 0 * eta, "retained": 0 * eta}
    q = tmp_path / "out.bin"
    formats.write_dir4out(q, 44, arrays, dict(c, seconds=1.0))
    back = formats.read_dir4out(q)
    assert back["sector"] == 44 and back["kept"] == c["kept"] and np.array_equal(back["eta"], eta)


def test_glrti_gle44_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    s = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    rules = {6: (np.arange(6.0), np.ones(6)), 8: (np.arange(8.0), np.ones(8))}
    formats.write_glrti(tmp_path / "g.bin", s, dk=0.2, depth=1.0, peak_kh=1.5, rules=rules)
    r = formats.read_glrti(tmp_path / "g.bin")
    assert np.array_equal(r["spectrum"], s) and set(r["rules"]) == {6, 8} and r["peak_kh"] == 1.5
    formats.write_gle44(tmp_path / "o.bin", s, rank=8)
    o = formats.read_gle44(tmp_path / "o.bin")
    assert np.array_equal(o["field"], s) and o["rank"] == 8
