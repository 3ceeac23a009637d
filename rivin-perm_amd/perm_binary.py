"""Exact permanents of zero-one matrices (perfect matchings) on MI355X.

Drop-in counterpart of the reference's ``perm_binary`` module: the public
functions keep their names, keyword-only options, validation messages and
return values.  Each matching-induced irreducible block (perm_binary.py:709-752)
is sent to the device with all of its CRT primes in one launch; bounds, prime
plans and CRT stay on the host in exact integers.

Backends: every backend name computes on the device (``_binary_ext`` is the
GPU object shaped like the reference's native extension); the names keep
the reference's block rules (perm_binary.py:893-1005):

* 'auto' first tries the live-column frontier DP on sparse blocks of order
  >= 20 with a narrow frontier (perm_binary.py:475-516, 975-990) under the
  auto state cap, then Glynn for blocks below order 64 and the uncapped DP
  for larger ones;
* 'native' runs Glynn and refuses blocks of order 64 or more;
* 'python' (the reference's portable subset DP, perm_binary.py:230-265)
  runs Glynn below order 64 and the uncapped DP above it.

The DP runs on the device too (phip_binary_frontier_residues), with counts
kept modulo the same descending CRT primes.  For the large blocks it plays
the part of the reference's subset DP, which tracks the set of used
columns; the frontier DP tracks only the live ones and returns the same
count.  It is limited by its 64 live-column slots and by device memory
(``DP_MAX_STATES`` states per layer).
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
    return ("native", "python")


def _select_backend(backend) -> str:
    if not isinstance(backend, str):
        raise TypeError("backend must be a string")
    backend = backend.lower()
    if backend in ("auto", "native"):
        return "native"
    if backend == "python":
        return "python"
    raise ValueError("backend must be one of: 'auto', 'native', 'python'")


# Layer limit of the uncapped DP for blocks of order 64 or more: a layer
# holds 8 + 4k bytes per state for k CRT primes, so 2^27 states stay within
# a few GiB of device memory.
DP_MAX_STATES = 1 << 27


def _large_block_permanent(block) -> int:
    """Exact count of a block of order >= 64 by the frontier DP in the
    library's row order, as the reference's subset DP would give it."""
    order, _, _ = _prep.frontier_order(block)
    try:
        return _frontier_permanent_masks(block, order, DP_MAX_STATES)
    except _FrontierTooWide as exc:
        raise ValueError(f"block of order {len(block)}: {exc}") from exc


class FrontierStateLimitExceeded(RuntimeError):
    """Raised when a frontier layer would exceed its configured state cap
    (perm_binary.py:42-58)."""

    def __init__(self, max_states: int, observed_states: int, processed_rows: int) -> None:
        self.max_states = max_states
        self.observed_states = observed_states
        self.processed_rows = processed_rows
        super().__init__(
            f"frontier state limit {max_states} exceeded while processing row {processed_rows}; "
            f"at least {observed_states} states would be required"
        )


class _FrontierTooWide(ValueError):
    """More columns live during one row than the device state mask holds."""


_AUTO_FRONTIER_MIN_ORDER = _prep.AUTO_FRONTIER_MIN_ORDER
_AUTO_FRONTIER_MAX_AVERAGE_DEGREE = _prep.AUTO_FRONTIER_MAX_AVERAGE_DEGREE
_AUTO_FRONTIER_MAX_WIDTH = _prep.AUTO_FRONTIER_MAX_WIDTH
_AUTO_FRONTIER_MAX_STATES = _prep.AUTO_FRONTIER_MAX_STATES


def _block_array(block_masks) -> np.ndarray:
    return np.array(_prep.unpack_masks(block_masks), dtype=np.uint8)


def _checked_residues(block: np.ndarray, primes) -> list:
    """Device residues with the reference's residue guard (perm_binary.py:817-840)."""
    residues = _native.binary_permanent_residues(block, primes)
    for r, p in zip(residues, primes):
        if isinstance(r, bool) or not isinstance(r, int) or not 0 <= r < p:
            raise RuntimeError("native binary-permanent backend returned an invalid residue")
    return residues


def _native_exact_permanent(block_matrix, bound: int, threads: int) -> int:
    """Nonnegative exact permanent of one irreducible block from device
    residues modulo descending 31-bit primes (perm_binary.py:843-863).
    ``threads`` is accepted for the reference's signature; the device
    decides its own parallelism."""
    primes = _prep.crt_primes_above_bound(bound)
    residues = _checked_residues(np.array(block_matrix, dtype=np.uint8), primes)
    result = _prep.crt_nonnegative(residues, primes)
    if result > bound:
        raise RuntimeError("native binary-permanent residues reconstructed outside the bound")
    return result


def _frontier_permanent_masks(row_masks, row_order, max_states: int) -> int:
    """Exact count of the live-column frontier DP for validated masks and
    row order (perm_binary.py:389-441).  The layers run on the device with
    the counts modulo descending 31-bit primes whose product exceeds the
    binary bound (the same primes as the Glynn path), CRT on the host; a
    layer above max_states raises FrontierStateLimitExceeded with the
    reference's fields.  Raises _FrontierTooWide (a ValueError) when more
    than 64 columns are live during one row."""
    order = len(row_masks)
    if order == 0:
        return 1
    layers = _prep.frontier_layers(row_masks, row_order)
    if layers is None:
        return 0
    row_slots, expiring, slots = layers
    if slots > _prep.FRONTIER_MAX_SLOTS:
        raise _FrontierTooWide(
            f"this row order keeps {slots} columns live at once; the device frontier DP holds at most "
            f"{_prep.FRONTIER_MAX_SLOTS}"
        )
    bound = _prep.normalized_bound_masks(row_masks)
    primes = _prep.crt_primes_above_bound(bound)
    residues = []
    step = _native.FRONTIER_MAX_PRIMES
    for start in range(0, len(primes), step):
        chunk = primes[start:start + step]
        got, exceeded = _native.binary_frontier_residues(row_slots, expiring, slots, chunk, max_states)
        if exceeded is not None:
            raise FrontierStateLimitExceeded(max_states, exceeded[1], exceeded[0])
        for r, p in zip(got, chunk):
            if not 0 <= r < p:
                raise RuntimeError("frontier backend returned an invalid residue")
        residues.extend(got)
    result = _prep.crt_nonnegative(residues, primes)
    if result > bound:
        raise RuntimeError("frontier residues reconstructed outside the bound")
    return result


def binary_permanent_frontier(matrix, *, row_order=None, max_states: int = 500_000,
                              strict: bool = True) -> int:
    """Exact binary permanent by the live-column frontier DP
    (perm_binary.py:444-472).

    ``row_order`` is None for the deterministic greedy order or an exact
    permutation of the row indices.  At most ``max_states`` distinct states
    may be held in any layer; exceeding the cap raises
    FrontierStateLimitExceeded.  Infeasible matrices return zero before the
    capped recurrence.  The result is an arbitrary-size Python integer.
    """
    max_states = _prep.validate_frontier_max_states(max_states)
    rows = _normalize_binary_matrix(matrix, strict=strict)
    order = _prep.validate_frontier_row_order(row_order, len(rows))
    masks = _row_masks(rows)
    if _perfect_matching(masks) is None:
        return 0
    if order is None:
        order, _, _ = _frontier_greedy_order(masks)
    return _frontier_permanent_masks(masks, order, max_states)


def _auto_frontier_plan(row_masks):
    """Orientation and row order for the auto fast path, or None
    (perm_binary.py:475-516)."""
    return _prep.auto_frontier_plan(row_masks)


def binary_permanent_bound(matrix, *, strict: bool = True) -> int:
    """Rigorous upper bound: product over irreducible blocks of
    min(n!, degree products, grouped Bregman-Minc) (perm_binary.py:519-535)."""
    rows = _prep.normalize_binary_matrix(matrix, strict=strict)
    feasible, blocks = _prep.matching_blocks(_prep.row_masks(rows))
    if not feasible:
        return 0
    if not blocks:
        return 1
    return math.prod(_prep.normalized_bound_masks(b) for b in blocks)


def binary_permanent_mod_p(matrix, p, *, backend: str = "auto", parallel: bool = True,
                           num_threads: int = 0, strict: bool = True) -> int:
    """Permanent modulo an odd prime larger than the order (perm_binary.py:893-940)."""
    if not isinstance(parallel, bool):
        raise TypeError("parallel must be a boolean")
    _prep.validate_num_threads(num_threads)
    rows = _prep.normalize_binary_matrix(matrix, strict=strict)
    modulus = _prep.validate_residue_prime(p, len(rows))
    selected = _select_backend(backend)
    explicit_native = selected == "native" and backend.lower() == "native"
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
            if explicit_native:
                raise ValueError("native binary-permanent blocks must have order less than 64")
            residue = _large_block_permanent(block) % modulus
        else:
            residue = _checked_residues(_block_array(block), [modulus])[0]
        result = result * residue % modulus
    return result


def binary_permanent(matrix, *, backend: str = "auto", parallel: bool = True,
                     num_threads: int = 0, strict: bool = True) -> int:
    """Exact nonnegative permanent of a square 0/1 matrix (perm_binary.py:943-1005)."""
    if not isinstance(parallel, bool):
        raise TypeError("parallel must be a boolean")
    num_threads = _prep.validate_num_threads(num_threads)
    threads = num_threads if parallel else 1
    rows = _prep.normalize_binary_matrix(matrix, strict=strict)
    selected = _select_backend(backend)
    automatic = backend.lower() == "auto"
    explicit_native = selected == "native" and not automatic
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
        if automatic:
            plan = _auto_frontier_plan(block)
            if plan is not None:
                try:
                    result *= _frontier_permanent_masks(plan[0], plan[1], _AUTO_FRONTIER_MAX_STATES)
                    continue
                except (FrontierStateLimitExceeded, _FrontierTooWide):
                    pass
        if len(block) >= 64:
            if explicit_native:
                raise ValueError("native binary-permanent blocks must have order less than 64")
            result *= _large_block_permanent(block)
            continue
        block_matrix = _unpack_masks(block)
        result *= _native_exact_permanent(block_matrix, _normalized_bound(block_matrix), threads)
    return result


# Private helper names of the reference module (perm_binary.py:173-227,
# 538-814), used by perm_binary_cuda as they are in the reference.
_crt_primes_above_bound = _prep.crt_primes_above_bound
_frontier_greedy_order = _prep.frontier_order
_is_prime = _prep.is_prime_u64
_matching_blocks = _prep.matching_blocks
_normalize_binary_matrix = _prep.normalize_binary_matrix
_normalized_bound = _prep.normalized_bound
_perfect_matching = _prep.perfect_matching
_row_masks = _prep.row_masks
_transpose_row_masks = _prep.transpose_masks
_unpack_masks = _prep.unpack_masks
_validate_residue_prime = _prep.validate_residue_prime
_validate_frontier_max_states = _prep.validate_frontier_max_states
_validate_frontier_row_order = _prep.validate_frontier_row_order

__all__ = [
    "FrontierStateLimitExceeded",
    "HAS_BINARY_EXT",
    "available_binary_backends",
    "binary_permanent",
    "binary_permanent_bound",
    "binary_permanent_frontier",
    "binary_permanent_mod_p",
    "get_binary_backend",
]
