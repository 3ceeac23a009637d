"""The C ABI library loads and exports every symbol include/permhip.h
declares (no device calls: CPU only)."""

import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "permhip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(phip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for required in (
        "phip_available",
        "phip_permanent_residues",
        "phip_permanent_modp",
        "phip_schur_permanent_modp",
        "phip_binary_permanent_modp",
        "phip_permanent_mod3_matrix",
        "phip_permanent_mod3",
        "phip_batch_create",
        "phip_batch_run",
    ):
        assert required in names


def test_library_exports_every_declared_symbol():
    from permhip import _native

    lib = ctypes.CDLL(str(_native.LIBRARY_PATH))
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert not missing, missing
    # The Python binding covers exactly the declared set.
    assert set(_native.exported_symbols()) == set(declared_functions())


def test_no_device_means_loud_failure(monkeypatch):
    from permhip import _native

    class NoDevice:
        def phip_available(self):
            return 0

    monkeypatch.setattr(_native, "load", lambda path=None: NoDevice())
    import pytest

    with pytest.raises(RuntimeError, match="no HIP device"):
        _native.permanent_residues([[1, 2], [3, 4]], [101])
