"""Device calls in a process forked after the parent initialised HIP must
raise a clear error instead of entering the library (the reference's
matrix_permanent sends small requests through multiprocessing.Pool,
perm.py:1188-1197 and 871-878).  Runs on the CPU with a fake library."""

import os

import pytest


class FakeLib:
    def __init__(self):
        self.calls = 0

    def phip_available(self):
        self.calls += 1
        return 1


def test_forked_child_refuses_device_calls(monkeypatch):
    from permhip import _native

    lib = FakeLib()
    monkeypatch.setattr(_native, "load", lambda path=None: lib)
    monkeypatch.setattr(_native, "_device_pid", None)
    parent = os.getpid()
    monkeypatch.setattr(_native.os, "getpid", lambda: parent)
    assert _native.require_device() is lib
    assert _native.available()
    assert lib.calls == 2
    # The same module state seen from a forked child.
    monkeypatch.setattr(_native.os, "getpid", lambda: parent + 1)
    with pytest.raises(_native.ForkedProcessError, match="spawn"):
        _native.require_device()
    assert not _native.available()
    with pytest.raises(RuntimeError, match="parallel=False"):
        _native.permanent_residues([[1, 2], [3, 4]], [7])
    assert lib.calls == 2  # the library was never entered from the child


def test_real_fork_child_raises(monkeypatch):
    """An actual os.fork after a (fake) initialisation: the child reports
    the refusal through its exit status."""
    from permhip import _native

    lib = FakeLib()
    monkeypatch.setattr(_native, "load", lambda path=None: lib)
    monkeypatch.setattr(_native, "_device_pid", None)
    _native.require_device()
    pid = os.fork()
    if pid == 0:  # child
        code = 3
        try:
            _native.require_device()
        except _native.ForkedProcessError:
            code = 0
        os._exit(code)
    _, status = os.waitpid(pid, 0)
    assert os.WEXITSTATUS(status) == 0


def test_wrapper_modules_check_fork(monkeypatch):
    import perm_cuda
    from permhip import _native

    monkeypatch.setattr(_native, "_device_pid", os.getpid() + 1)
    with pytest.raises(_native.ForkedProcessError):
        perm_cuda._ModpLibrary(FakeLib()).schur_permanent_modp_cuda(5, 1, 11)
    with pytest.raises(_native.ForkedProcessError):
        perm_cuda._ModpLibrary(FakeLib()).cuda_runtime_available()
