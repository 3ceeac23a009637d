"""Device 0/1 permanents behind the interface of the reference's
``perm_binary_cuda`` (perm_binary_cuda.py), computed on the MI355X by
``libpermhip.so``.

Exact results reuse the host preparation of :mod:`perm_binary` (strict 0/1
normalisation, matching feasibility, strongly connected blocks, the
nonnegative bound and the descending-prime CRT plan); each block residue is
one ``binary_permanent_modp_cuda`` call, with the reference's sentinel and
range checks (perm_binary_cuda.py:103-124).  The module global
``_binary_cuda`` is an object with the reference library's symbol names
bound to ``phip_available`` and ``phip_binary_permanent_modp``.
"""

from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

import perm_binary
from permhip import _native

_U8P = ctypes.POINTER(ctypes.c_uint8)


class _BinaryLibrary:
    """The reference's binary symbols on libpermhip."""

    def __init__(self, lib):
        self._lib = lib

    def cuda_runtime_available(self) -> int:
        return int(_native._touch_device(self._lib))

    def binary_permanent_modp_cuda(self, pointer, order, modulus) -> int:
        _native._check_fork()
        return int(self._lib.phip_binary_permanent_modp(ctypes.cast(pointer, _U8P), int(order), int(modulus)))


_binary_cuda = None
_binary_cuda_load_error: Optional[Exception] = None


def _load_binary_cuda_library() -> None:
    """Bind the library object (perm_binary_cuda.py:42-60)."""
    global _binary_cuda, _binary_cuda_load_error
    try:
        _binary_cuda = _BinaryLibrary(_native.load())
        _binary_cuda_load_error = None
    except (OSError, RuntimeError) as exc:
        _binary_cuda = None
        _binary_cuda_load_error = exc


_load_binary_cuda_library()


def is_binary_cuda_available() -> bool:
    """True when the library is bound and a device answers its probe
    (perm_binary_cuda.py:63-73)."""
    probe = getattr(_binary_cuda, "cuda_runtime_available", None)
    if probe is None:
        return False
    try:
        return bool(probe())
    except (OSError, TypeError, ValueError, _native.ForkedProcessError):
        return False


def _require_binary_cuda() -> None:
    if _binary_cuda is None:
        detail = f" ({_binary_cuda_load_error})" if _binary_cuda_load_error is not None else ""
        raise RuntimeError("binary device library is unavailable; build libpermhip.so" + detail)
    if getattr(_binary_cuda, "cuda_runtime_available", None) is None:
        raise RuntimeError("binary device library is outdated; rebuild libpermhip.so")
    if not is_binary_cuda_available():
        raise RuntimeError("no GPU device is available")


def _validate_cuda_prime(p, order: int) -> int:
    """Odd prime n < p < 2^31 (perm_binary_cuda.py:96-100)."""
    modulus = perm_binary._validate_residue_prime(p, order)
    if modulus >= 1 << 31:
        raise ValueError("the binary device modulus must be less than 2**31")
    return modulus


def _cuda_residue(matrix: Sequence[Sequence[int]], modulus: int) -> int:
    """One block residue through the C ABI with the sentinel and range
    checks (perm_binary_cuda.py:103-124)."""
    _require_binary_cuda()
    block = np.ascontiguousarray(matrix, dtype=np.uint8)
    order = block.shape[0]
    value = int(_binary_cuda.binary_permanent_modp_cuda(block.ctypes.data_as(_U8P), order, modulus))
    if value < 0:
        detail = _native.last_error() if _native.available() else ""
        raise RuntimeError("device binary-permanent residue failed" + (f": {detail}" if detail else ""))
    if value >= modulus:
        raise RuntimeError("device binary-permanent backend returned an invalid residue")
    return value


def _cuda_exact_permanent(matrix: Sequence[Sequence[int]], bound: int) -> int:
    """One block in the certified interval [0, bound] by incremental CRT over
    the descending primes (perm_binary_cuda.py:127-144)."""
    value, product = 0, 1
    for q in perm_binary._crt_primes_above_bound(bound):
        r = _cuda_residue(matrix, q)
        step = (r - value) * pow(product, -1, q) % q
        value += product * step
        product *= q
    if value > bound:
        raise RuntimeError("device binary-permanent residues reconstructed outside the bound")
    return value


def _blocks(normalized):
    feasible, blocks = perm_binary._matching_blocks(perm_binary._row_masks(normalized))
    if not feasible:
        return None
    out = []
    for block in blocks:
        if len(block) == 1:
            continue
        if len(block) > 63:
            raise ValueError("device binary-permanent blocks must have order at most 63")
        out.append(perm_binary._unpack_masks(block))
    return out


def binary_permanent_mod_p_cuda(matrix, p, *, strict: bool = True) -> int:
    """0/1 permanent modulo an odd prime below 2^31 on the device
    (perm_binary_cuda.py:147-179).  Matching and block factoring run on the
    host first, so trivial and infeasible inputs need no device."""
    normalized = perm_binary._normalize_binary_matrix(matrix, strict=strict)
    modulus = _validate_cuda_prime(p, len(normalized))
    if not normalized:
        return 1 % modulus
    if len(normalized) == 1:
        return normalized[0][0] % modulus
    blocks = _blocks(normalized)
    if blocks is None:
        return 0
    result = 1
    for block in blocks:
        result = result * _cuda_residue(block, modulus) % modulus
    return result


def binary_permanent_cuda(matrix, *, strict: bool = True) -> int:
    """Exact nonnegative 0/1 permanent from device residues
    (perm_binary_cuda.py:182-207)."""
    normalized = perm_binary._normalize_binary_matrix(matrix, strict=strict)
    if not normalized:
        return 1
    if len(normalized) == 1:
        return normalized[0][0]
    blocks = _blocks(normalized)
    if blocks is None:
        return 0
    result = 1
    for block in blocks:
        result *= _cuda_exact_permanent(block, perm_binary._normalized_bound(block))
    return result


__all__ = [
    "binary_permanent_cuda",
    "binary_permanent_mod_p_cuda",
    "is_binary_cuda_available",
]
