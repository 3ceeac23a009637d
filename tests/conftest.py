"""Shared test set-up: import paths, the ``gpu`` marker and fixtures."""

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PACKAGE_ROOT = ROOT / "rivin-perm_amd"
ORACLE_ROOT = ROOT / "oracle"
for path in (str(PACKAGE_ROOT), str(ORACLE_ROOT)):
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libpermhip.so")


@pytest.fixture(scope="session")
def native():
    """The loaded C ABI; GPU tests fail (not skip) when it cannot run."""
    from permhip import _native

    lib = _native.load()
    assert lib.phip_available() == 1, "no HIP device visible to libpermhip"
    return _native


@pytest.fixture(scope="session")
def oracle_c():
    import oracle

    return oracle.C
