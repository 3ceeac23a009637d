"""Exact low-rank permanent identity (host side, explicit ``method='low_rank'``).

For A = U V over F_p with rank r,

    per(A) = sum_alpha alpha! [x^alpha] prod_i (U_i . x) * [x^alpha] prod_j (V_:j . x)

an identity of polynomials with no factorial division, valid in every
characteristic.  This restates the reference's low-rank specialisation
(perm.py:956-1073): an incremental echelon factorisation that aborts once the
rank or the coefficient-state budget is exceeded, the homogeneous coefficient
layers of both products, and the alpha!-weighted pairing.  It runs on the
host in exact Python integers, as in the reference.
"""

from __future__ import annotations

import math
from typing import Dict, List, Sequence, Tuple

LOW_RANK_MAX_RANK = 4  # perm.py:44
LOW_RANK_MAX_STATES = 50_000  # perm.py:45


class LowRankLimitExceeded(ValueError):
    """An explicit low-rank request exceeds the rank or coefficient budget."""


def factor_mod(rows: Sequence[Sequence[int]], prime: int, max_rank: int):
    """A = coordinates @ basis over F_p, basis rows with a unit pivot and zeros
    at every earlier pivot (perm.py:978-1013)."""
    order = len(rows)
    basis: List[Tuple[int, ...]] = []
    pivots: List[int] = []
    coordinates: List[List[int]] = []
    for original in rows:
        row = [int(v) % prime for v in original]
        coeffs = []
        for pivot, vector in zip(pivots, basis):
            c = row[pivot]
            coeffs.append(c)
            if c:
                row = [(x - c * y) % prime for x, y in zip(row, vector)]
        pivot = next((idx for idx, v in enumerate(row) if v), None)
        if pivot is not None:
            rank = len(basis) + 1
            states = math.comb(order + rank - 1, rank - 1)
            if rank > max_rank or states > LOW_RANK_MAX_STATES:
                raise LowRankLimitExceeded(
                    f"rank is at least {rank}; coefficient layer needs {states} "
                    f"states (limits: rank {max_rank}, states {LOW_RANK_MAX_STATES})"
                )
            lead = row[pivot]
            inv = pow(lead, -1, prime)
            basis.append(tuple(v * inv % prime for v in row))
            pivots.append(pivot)
            for previous in coordinates:
                previous.append(0)
            coeffs.append(lead)
        coordinates.append(coeffs)
    return coordinates, basis


def _coefficients(forms, rank: int, prime: int) -> Dict[Tuple[int, ...], int]:
    """Nonzero coefficients of prod over forms of (sum_k f_k x_k) (perm.py:1016-1029)."""
    layer = {(0,) * rank: 1}
    for form in forms:
        terms = [(k, v) for k, v in enumerate(form) if v]
        nxt: Dict[Tuple[int, ...], int] = {}
        for powers, value in layer.items():
            for k, v in terms:
                key = powers[:k] + (powers[k] + 1,) + powers[k + 1:]
                nxt[key] = (nxt.get(key, 0) + value * v) % prime
        layer = {key: v for key, v in nxt.items() if v}
        if not layer:
            break
    return layer


def permanent_from_factors(order: int, prime: int, factors) -> int:
    """sum_alpha alpha! [x^alpha] left * [x^alpha] right mod p (perm.py:1032-1053)."""
    coordinates, basis = factors
    rank = len(basis)
    if rank == 0:
        return int(order == 0)
    left = _coefficients(coordinates, rank, prime)
    right = _coefficients(zip(*basis), rank, prime)
    fact = [1] * (order + 1)
    for i in range(2, order + 1):
        fact[i] = fact[i - 1] * i % prime
    total = 0
    for powers, value in left.items():
        term = value * right.get(powers, 0) % prime
        for e in powers:
            term = term * fact[e] % prime
        total = (total + term) % prime
    return total


def low_rank_residues(rows, primes, max_rank: int, *, required: bool):
    """Residues for every prime, or None when a prime exceeds the budget and
    the caller did not require the method (perm.py:1056-1073)."""
    factors = []
    for p in primes:
        try:
            factors.append(factor_mod(rows, p, max_rank))
        except LowRankLimitExceeded:
            if required:
                raise
            return None
    return [permanent_from_factors(len(rows), p, f) for p, f in zip(primes, factors)]
