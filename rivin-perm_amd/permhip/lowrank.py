"""Exact permanents of low-rank integer matrices (host side).

If A has rank r over F_p, its reduced row echelon form R (r x n, identity
at the pivot columns J) spans the rows of A, and every row a of A equals
sum_k a[J_k] R_k.  So A = C R with C = A[:, J], and

    per(C R) = sum_{|alpha| = n} alpha! [x^alpha] prod_i (C_i . x)
                                        [x^alpha] prod_j (R_:j . x),

an identity of polynomials in r variables with no division, valid in every
characteristic (expand each row of C R as a linear form and group the
permutation sum by how many rows pick each variable).  This is the
specialisation the reference calls from ``matrix_permanent``
(perm.py:956-1073, 1158-1169); the factorisation and the coefficient
arithmetic below are this module's own:

* the factorisation is a column-scan Gauss-Jordan elimination mod p on the
  whole matrix with numpy row operations; it stops as soon as the number of
  pivots passes the caller's rank limit, so a probe of a full-rank matrix
  costs (limit + 1) eliminations of O(n^2);
* a homogeneous polynomial of degree d in r variables is stored
  dehomogenised, as a dense (r-1)-dimensional array indexed by the first
  r-1 exponents (the last is d minus their sum), and multiplying by a
  linear form is r shifted multiply-adds over the occupied corner of that
  array;
* all primes of the same rank are processed together along a leading axis,
  so each multiply-add is one numpy operation for every prime.

The limits of an explicit ``method='low_rank'`` request are the reference's
(rank at most 4, at most 50,000 coefficient states C(n + r - 1, r - 1)), with
its exception and message; ``auto`` chooses its rank limit from a cost
model of this module against the device's Glynn throughput
(``rank_limit``).
"""

from __future__ import annotations

import math
from typing import List, Optional, Sequence

import numpy as np

LOW_RANK_MAX_RANK = 4
LOW_RANK_MAX_STATES = 50_000
# Below this order the device answers in well under a millisecond and
# auto never probes.
LOW_RANK_MIN_ORDER = 20


class LowRankLimitExceeded(ValueError):
    """An explicit low-rank request exceeds the rank or coefficient budget."""


def coefficient_states(order: int, rank: int) -> int:
    """Monomials of degree ``order`` in ``rank`` variables."""
    return math.comb(order + rank - 1, rank - 1)


def _residue_matrix(rows: Sequence[Sequence[int]], prime: int) -> np.ndarray:
    n = len(rows)
    dtype = np.int64 if prime < (1 << 31) else object
    return np.array([[int(v) % prime for v in row] for row in rows], dtype=dtype).reshape(n, n)


def _rref_factor(a: np.ndarray, prime: int, max_rank: int, max_states: int):
    """(C, R) with a = C R over F_p, R in reduced row echelon form; raises
    LowRankLimitExceeded once a pivot past the limits is found."""
    n = a.shape[0]
    m = a.copy()
    pivots: List[int] = []
    top = 0
    for col in range(n):
        if top == n:
            break
        nonzero = np.flatnonzero(m[top:, col])
        if nonzero.size == 0:
            continue
        rank = len(pivots) + 1
        states = coefficient_states(n, rank)
        if rank > max_rank or states > max_states:
            raise LowRankLimitExceeded(
                f"rank is at least {rank}; coefficient layer needs {states} "
                f"states (limits: rank {max_rank}, states {max_states})"
            )
        src = top + int(nonzero[0])
        if src != top:
            m[[top, src]] = m[[src, top]]
        inv = pow(int(m[top, col]), -1, prime)
        m[top] = m[top] * inv % prime
        factor = m[:, col].copy()
        factor[top] = 0
        # Entries and factors are below p < 2^31, so each product fits int64.
        m = (m - np.outer(factor, m[top]) % prime) % prime
        pivots.append(col)
        top += 1
    return a[:, pivots], m[: len(pivots)]


def factor_residues(rows, primes: Sequence[int], max_rank: int, max_states: Optional[int] = None) -> list:
    """Factorisations (C, R) for every prime, in order; raises
    LowRankLimitExceeded on the first prime beyond the limits (the state
    limit defaults to LOW_RANK_MAX_STATES)."""
    if max_states is None:
        max_states = LOW_RANK_MAX_STATES
    return [_rref_factor(_residue_matrix(rows, p), p, max_rank, max_states) for p in primes]


def _product_coefficients(forms: np.ndarray, moduli: np.ndarray) -> np.ndarray:
    """Coefficients of prod_t (forms[:, t, :] . x) for every prime.

    forms: (k, n, r) residues; moduli: (k,).  Returns an array of shape
    (k,) + (n + 1,) * (r - 1): entry [q, e_1..e_{r-1}] is the coefficient of
    x_1^e_1 ... x_{r-1}^e_{r-1} x_r^(n - sum e) modulo moduli[q]."""
    k, n, r = forms.shape
    dims = r - 1
    if dims == 0:
        acc = np.ones(k, dtype=forms.dtype)
        for t in range(n):
            acc = acc * forms[:, t, 0] % moduli
        return acc
    shape = (k,) + (n + 1,) * dims
    poly = np.zeros(shape, dtype=forms.dtype)
    poly[(slice(None),) + (0,) * dims] = 1
    expand = (slice(None),) + (None,) * dims
    p = moduli[expand]
    scratch = np.empty(shape, dtype=forms.dtype)
    for t in range(n):
        # Degree t before this factor: exponents of the leading variables
        # lie in [0, t], so only that corner of the array is touched.
        src = (slice(None),) + (slice(0, t + 1),) * dims
        dst = (slice(None),) + (slice(0, t + 2),) * dims
        cur = poly[src]
        out = scratch[dst]
        out[...] = 0
        # Each term is reduced below p < 2^31, so the r terms of one output
        # entry add exactly in int64 and are reduced once.
        np.remainder(cur * forms[:, t, dims][expand], p, out=out[src])
        for axis in range(dims):
            shifted = list(src)
            shifted[axis + 1] = slice(1, t + 2)
            out[tuple(shifted)] += cur * forms[:, t, axis][expand] % p
        np.remainder(out, p, out=poly[dst])
    return poly


def _pairing_weights(n: int, dims: int, prime: int) -> np.ndarray:
    """alpha! = e_1! ... e_dims! (n - sum e)! mod p on the (n+1)^dims grid,
    0 where sum e > n."""
    fact = [1] * (n + 1)
    for i in range(1, n + 1):
        fact[i] = fact[i - 1] * i % prime
    dtype = np.int64 if prime < (1 << 31) else object
    f = np.array(fact, dtype=dtype)
    idx = np.indices((n + 1,) * dims).reshape(dims, -1)
    total = idx.sum(axis=0)
    rest = n - total
    valid = rest >= 0
    w = f[np.where(valid, rest, 0)]
    for axis in range(dims):
        w = w * f[idx[axis]] % prime
    return np.where(valid, w, 0).astype(dtype).reshape((n + 1,) * dims)


def permanents_from_factors(order: int, primes: Sequence[int], factors) -> List[int]:
    """per(C R) mod p for every (prime, (C, R)) pair."""
    out: List[Optional[int]] = [None] * len(primes)
    by_rank = {}
    for i, (p, (c, r)) in enumerate(zip(primes, factors)):
        by_rank.setdefault((r.shape[0], p >= (1 << 31)), []).append(i)
    for (rank, wide), members in by_rank.items():
        if rank == 0:
            for i in members:
                out[i] = int(order == 0) % primes[i]
            continue
        dtype = object if wide else np.int64
        moduli = np.array([primes[i] for i in members], dtype=dtype)
        left = np.stack([factors[i][0] for i in members]).astype(dtype)  # (k, n, r)
        right = np.stack([factors[i][1].T for i in members]).astype(dtype)  # (k, n, r)
        lc = _product_coefficients(left, moduli)
        rc = _product_coefficients(right, moduli)
        dims = rank - 1
        for q, i in enumerate(members):
            p = primes[i]
            if dims == 0:
                out[i] = int(lc[q]) * int(rc[q]) % p * math.factorial(order) % p
                continue
            w = _pairing_weights(order, dims, p)
            terms = lc[q] * rc[q] % p * w % p
            total = terms.sum(dtype=object) if wide else terms.sum(dtype=np.int64)
            out[i] = int(total) % p
    return [int(v) for v in out]


def low_rank_residues(rows, primes, max_rank: int, *, required: bool, max_states: Optional[int] = None):
    """Residues for every prime, or None when some prime exceeds the limits
    and the caller did not require the method.  Every prime is factored
    before any coefficient work: a small rank modulo one prime says nothing
    about the others."""
    try:
        factors = factor_residues(rows, primes, max_rank, max_states)
    except LowRankLimitExceeded:
        if required:
            raise
        return None
    return permanents_from_factors(len(rows), list(primes), factors)


# -- auto-mode dispatch ----------------------------------------------------

# Throughputs of the cost model.  The device figure is the measured Glynn
# rate of the persistent engine on one MI355X with a full prime set
# (reference-equivalent mulmods, n 2^(n-1) per residue); below it the
# kernels' launch and host planning take about a tenth of a millisecond.
# Orders the 31-bit engine does not take (above 46) run on the wide
# kernels at about 2.7e12 mulmods/s.  The host figure is what the numpy
# coefficient updates above sustain per core (element updates per second),
# plus a fixed cost per numpy operation.
DEVICE_MULMODS_PER_S = 5.0e13
WIDE_DEVICE_MULMODS_PER_S = 2.7e12
DEVICE_FLOOR_S = 1.0e-4
HOST_UPDATES_PER_S = 1.0e8
HOST_OP_S = 4.0e-6
DEVICE_MAX_ORDER = 63


def device_seconds(order: int, moduli: int) -> float:
    """Estimated device time of the generic path for one matrix."""
    if order > DEVICE_MAX_ORDER:
        return math.inf
    rate = DEVICE_MULMODS_PER_S if order <= 46 else WIDE_DEVICE_MULMODS_PER_S
    return DEVICE_FLOOR_S + moduli * order * 2.0 ** (order - 1) / rate


def low_rank_seconds(order: int, rank: int, moduli: int) -> float:
    """Estimated host time of the low-rank path at this rank: elimination,
    two coefficient products over the dense grid, the pairing."""
    dims = rank - 1
    elimination = rank * (order * order * moduli / HOST_UPDATES_PER_S + 8 * HOST_OP_S)
    grid = sum((d + 1) ** dims for d in range(order + 1))
    products = 2 * (rank * moduli * grid / HOST_UPDATES_PER_S + order * (2 + 2 * rank) * HOST_OP_S)
    pairing = moduli * ((order + 1) ** dims * 4 / HOST_UPDATES_PER_S + 6 * HOST_OP_S)
    return elimination + products + pairing


def rank_limit(order: int, moduli: int) -> int:
    """Largest rank at which auto tries the low-rank path before the
    device, or 0 to skip even the probe: the largest r <= 4 within the
    coefficient-state cap whose estimated host time beats the device's."""
    if order < LOW_RANK_MIN_ORDER:
        return 0
    generic = device_seconds(order, moduli)
    limit = 0
    for rank in range(1, LOW_RANK_MAX_RANK + 1):
        if coefficient_states(order, rank) > LOW_RANK_MAX_STATES:
            break
        if low_rank_seconds(order, rank, moduli) >= generic:
            break
        limit = rank
    return limit
