"""ctypes binding of ``libpermhip.so`` (the C ABI in ``include/permhip.h``).

The library sits next to this module and carries the ``lib`` prefix so it
never shadows a Python module (the rule of perm_binary_cuda.py:42-60).  There
is no CPU fallback: every entry point raises ``RuntimeError`` when the
library is missing, when no HIP device is present, or when a call returns the
negative sentinel (perm_cuda.py:125-129).
"""

from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

import numpy as np

LIBRARY_PATH = Path(os.environ.get("PERMHIP_LIBRARY", Path(__file__).resolve().parent / "libpermhip.so"))

_lock = threading.Lock()
_lib = None

_I64P = ctypes.POINTER(ctypes.c_int64)
_U64P = ctypes.POINTER(ctypes.c_uint64)
_U8P = ctypes.POINTER(ctypes.c_uint8)
_INTP = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes)
_SIGNATURES = {
    "phip_available": (ctypes.c_int, []),
    "phip_last_error": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "phip_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "phip_permanent_residues": (
        ctypes.c_int,
        [_I64P, ctypes.c_int64, ctypes.c_int, _I64P, ctypes.c_int, _I64P],
    ),
    "phip_permanent_modp": (ctypes.c_int64, [_I64P, ctypes.c_int, ctypes.c_int64]),
    "phip_schur_permanent_modp": (ctypes.c_int64, [ctypes.c_int, ctypes.c_int64, ctypes.c_int64]),
    "phip_binary_permanent_residues": (
        ctypes.c_int,
        [_U8P, ctypes.c_int, _I64P, ctypes.c_int, _I64P],
    ),
    "phip_binary_permanent_modp": (ctypes.c_int64, [_U8P, ctypes.c_int, ctypes.c_int64]),
    "phip_permanent_mod3_matrix": (ctypes.c_int, [_INTP, ctypes.c_int]),
    "phip_permanent_mod3": (ctypes.c_int, [_U64P, _U64P, ctypes.c_int]),
}


def exported_symbols():
    """Names the header declares and the library must export."""
    return tuple(_SIGNATURES)


def load(path: Path | None = None):
    """Load the library once; raises when it is absent."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        target = Path(path) if path is not None else LIBRARY_PATH
        if not target.exists():
            raise RuntimeError(
                f"libpermhip.so not found at {target}; build it with "
                "`python -m permhip.build` (there is no CPU fallback)"
            )
        lib = ctypes.CDLL(str(target))
        for name, (restype, argtypes) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        if path is None:
            _lib = lib
        return lib


def last_error() -> str:
    lib = load()
    buf = ctypes.create_string_buffer(1024)
    lib.phip_last_error(buf, len(buf))
    return buf.value.decode("utf-8", "replace")


def available() -> bool:
    """True when the library loads and a HIP device is usable."""
    try:
        return bool(load().phip_available())
    except (OSError, RuntimeError):
        return False


def require_device():
    lib = load()
    if not lib.phip_available():
        raise RuntimeError("no HIP device is available to libpermhip (there is no CPU fallback)")
    return lib


def device_info() -> str:
    lib = require_device()
    buf = ctypes.create_string_buffer(256)
    if lib.phip_device_info(buf, len(buf)) != 0:
        raise RuntimeError(last_error())
    return buf.value.decode()


def _check(status: int, what: str) -> None:
    if status < 0:
        raise RuntimeError(f"{what} failed: {last_error()}")


def _i64ptr(a: np.ndarray):
    return a.ctypes.data_as(_I64P)


def permanent_residues(matrix: np.ndarray, moduli) -> list:
    """perm(matrix) mod each modulus, one device pass.

    ``matrix`` is either an (n, n) int64 array used for every modulus or a
    (k, n, n) int64 array holding one already-reduced copy per modulus.
    """
    lib = require_device()
    a = np.ascontiguousarray(matrix, dtype=np.int64)
    mods = np.ascontiguousarray(np.asarray(list(moduli), dtype=np.int64))
    k = int(mods.shape[0])
    if a.ndim == 2:
        n = int(a.shape[0])
        stride = 0
    elif a.ndim == 3:
        if a.shape[0] != k:
            raise ValueError("one matrix per modulus is required")
        n = int(a.shape[1])
        stride = n * n
    else:
        raise ValueError("matrix must be 2- or 3-dimensional")
    out = np.zeros(k, dtype=np.int64)
    status = lib.phip_permanent_residues(_i64ptr(a), stride, n, _i64ptr(mods), k, _i64ptr(out))
    _check(status, "phip_permanent_residues")
    return [int(v) for v in out]


def schur_permanent_modp(n: int, omega: int, p: int) -> int:
    lib = require_device()
    result = lib.phip_schur_permanent_modp(int(n), int(omega), int(p))
    _check(result, "phip_schur_permanent_modp")
    return int(result)


def binary_permanent_residues(matrix: np.ndarray, primes) -> list:
    lib = require_device()
    a = np.ascontiguousarray(matrix, dtype=np.uint8)
    n = int(a.shape[0]) if a.ndim == 2 else 0
    ps = np.ascontiguousarray(np.asarray(list(primes), dtype=np.int64))
    out = np.zeros(len(ps), dtype=np.int64)
    ptr = a.ctypes.data_as(_U8P) if n else ctypes.cast(None, _U8P)
    status = lib.phip_binary_permanent_residues(ptr, n, _i64ptr(ps), len(ps), _i64ptr(out))
    _check(status, "phip_binary_permanent_residues")
    return [int(v) for v in out]


def permanent_mod3(matrix: np.ndarray) -> int:
    """Takes an (n, n) matrix already reduced modulo 3 (values 0, 1, 2)."""
    lib = require_device()
    a = np.ascontiguousarray(matrix, dtype=np.int32)
    n = int(a.shape[0])
    result = lib.phip_permanent_mod3_matrix(a.ctypes.data_as(_INTP), n)
    _check(result, "phip_permanent_mod3_matrix")
    return int(result)
