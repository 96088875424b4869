"""The C-ABI library loads and exports every symbol include/glw.h declares (CPU, no compute)."""

import re
from pathlib import Path

from green_laplace_wave_reconstruction_amd import _native

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "glw.h").read_text()
    return sorted(set(re.findall(r"^\S.*?\b(glw_\w+)\(", text, flags=re.M)))


def test_header_matches_binding_list():
    assert _declared() == sorted(_native.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not _native.LIB_PATH.exists():
        from green_laplace_wave_reconstruction_amd import build

        build.build()
    lib = _native.load_library(_native.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.glw_version().decode().startswith("glw ")
    assert lib.glw_fft_length_supported(512) == 1 and lib.glw_fft_length_supported(100) == 0
