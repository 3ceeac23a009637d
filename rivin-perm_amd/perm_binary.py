"""Exact permanents of zero-one matrices (perfect matchings) on MI355X.

Drop-in counterpart of the reference's ``perm_binary`` module: the public
functions keep their names, keyword-only options, validation messages and
return values.  Each matching-induced irreducible block (perm_binary.py:709-752)
is sent to the device with all of its CRT primes in one launch; bounds, prime
plans and CRT stay on the host in exact integers.

Backends: 'auto' and 'native' both select the device (``_binary_ext`` is the
GPU object shaped like the reference's native extension).  The reference's
portable Python subset DP and sparse frontier DP are not part of this build
(backend='python' raises), and blocks of order 64 or more are rejected, as
the reference's native backend does.
"""

from __future__ import annotations

import math
from typing import Tuple

import numpy as np

from permhip import _native
from permhip import binary_prep as _prep
from permhip.backends import perm_binary_ext as _binary_ext

HAS_BINARY_EXT = True


def get_binary_backend() -> str:
    """The backend selected by backend='auto'."""
    return "native"


def available_binary_backends() -> Tuple[str, ...]:
    return ("native",)


def _select_backend(backend) -> str:
    if not isinstance(backend, str):
        raise TypeError("backend must be a string")
    backend = backend.lower()
    if backend in ("auto", "native"):
        return "native"
    if backend == "python":
        raise ValueError("the portable Python backend is not part of this build; use backend='auto'")
    raise ValueError("backend must be one of: 'auto', 'native', 'python'")


def _block_array(block_masks) -> np.ndarray:
    return np.array(_prep.unpack_masks(block_masks), dtype=np.uint8)


def _checked_residues(block: np.ndarray, primes) -> list:
    """Device residues with the reference's residue guard (perm_binary.py:817-840)."""
    residues = _native.binary_permanent_residues(block, primes)
    for r, p in zip(residues, primes):
        if isinstance(r, bool) or not isinstance(r, int) or not 0 <= r < p:
            raise RuntimeError("native binary-permanent backend returned an invalid residue")
    return residues


def _exact_block(block_masks) -> int:
    """Nonnegative exact permanent of one irreducible block (perm_binary.py:843-863)."""
    rows = _prep.unpack_masks(block_masks)
    bound = _prep.normalized_bound(rows)
    primes = _prep.crt_primes_above_bound(bound)
    residues = _checked_residues(np.array(rows, dtype=np.uint8), primes)
    result = _prep.crt_nonnegative(residues, primes)
    if result > bound:
        raise RuntimeError("native binary-permanent residues reconstructed outside the bound")
    return result


def binary_permanent_bound(matrix, *, strict: bool = True) -> int:
    """Rigorous upper bound: product over irreducible blocks of
    min(n!, degree products, grouped Bregman-Minc) (perm_binary.py:519-535)."""
    rows = _prep.normalize_binary_matrix(matrix, strict=strict)
    feasible, blocks = _prep.matching_blocks(_prep.row_masks(rows))
    if not feasible:
        return 0
    if not blocks:
        return 1
    return math.prod(_prep.normalized_bound(_prep.unpack_masks(b)) for b in blocks)


def binary_permanent_mod_p(matrix, p, *, backend: str = "auto", parallel: bool = True,
                           num_threads: int = 0, strict: bool = True) -> int:
    """Permanent modulo an odd prime larger than the order (perm_binary.py:893-940)."""
    if not isinstance(parallel, bool):
        raise TypeError("parallel must be a boolean")
    _prep.validate_num_threads(num_threads)
    rows = _prep.normalize_binary_matrix(matrix, strict=strict)
    modulus = _prep.validate_residue_prime(p, len(rows))
    _select_backend(backend)
    if not rows:
        return 1 % modulus
    if len(rows) == 1:
        return rows[0][0] % modulus
    feasible, blocks = _prep.matching_blocks(_prep.row_masks(rows))
    if not feasible:
        return 0
    result = 1
    for block in blocks:
        if len(block) == 1:
            continue
        if len(block) >= 64:
            raise ValueError("native binary-permanent blocks must have order less than 64")
        residue = _checked_residues(_block_array(block), [modulus])[0]
        result = result * residue % modulus
    return result


def binary_permanent(matrix, *, backend: str = "auto", parallel: bool = True,
                     num_threads: int = 0, strict: bool = True) -> int:
    """Exact nonnegative permanent of a square 0/1 matrix (perm_binary.py:943-1005)."""
    if not isinstance(parallel, bool):
        raise TypeError("parallel must be a boolean")
    _prep.validate_num_threads(num_threads)
    rows = _prep.normalize_binary_matrix(matrix, strict=strict)
    _select_backend(backend)
    if not rows:
        return 1
    if len(rows) == 1:
        return rows[0][0]
    feasible, blocks = _prep.matching_blocks(_prep.row_masks(rows))
    if not feasible:
        return 0
    result = 1
    for block in blocks:
        if len(block) == 1:
            continue
        if len(block) >= 64:
            raise ValueError("native binary-permanent blocks must have order less than 64")
        result *= _exact_block(block)
    return result


# Private helper names of the reference module (perm_binary.py:173-227,
# 538-814), used by perm_binary_cuda as they are in the reference.
_crt_primes_above_bound = _prep.crt_primes_above_bound
_is_prime = _prep.is_prime_u64
_matching_blocks = _prep.matching_blocks
_normalize_binary_matrix = _prep.normalize_binary_matrix
_normalized_bound = _prep.normalized_bound
_row_masks = _prep.row_masks
_unpack_masks = _prep.unpack_masks
_validate_residue_prime = _prep.validate_residue_prime

__all__ = [
    "HAS_BINARY_EXT",
    "available_binary_backends",
    "binary_permanent",
    "binary_permanent_bound",
    "binary_permanent_mod_p",
    "get_binary_backend",
]
