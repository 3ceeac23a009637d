"""Exact host-side planning and reconstruction in Python integers.

Everything here runs on the host in arbitrary-precision integers, as the
reference does: bounds, CRT prime plans, the Graham--Lehmer theorem modulus
for Schur matrices, roots of unity, and Chinese remaindering with the signed
lift.  Each function cites the reference function it restates.
"""

from __future__ import annotations

import functools
import math
import threading
from numbers import Integral
from typing import List, Optional, Sequence, Tuple

import numpy as np

INT64_MAX = (1 << 63) - 1
GENERIC_PRIME_FLOOR = 1 << 30  # perm.py:1121
SCHUR_CRT_PRIME_FLOOR = 1 << 30  # perm.py:38


# ---------------------------------------------------------------------------
# Primality and roots of unity
# ---------------------------------------------------------------------------

_MR_WITNESSES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def is_prime(n: int) -> bool:
    """Miller--Rabin with the first twelve prime bases (perm.py:162-195).

    These bases are deterministic below 3.3e24, which covers every modulus
    the planners produce.
    """
    if n < 2:
        return False
    for small in _MR_WITNESSES:
        if n == small:
            return True
        if n % small == 0:
            return False
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for a in _MR_WITNESSES:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def _as_int(value, name: str) -> int:
    if isinstance(value, bool) or not isinstance(value, Integral):
        raise TypeError(f"{name} must be an integer")
    return int(value)


def find_primes_congruent_to_1_mod_n(n: int, count: int, min_prime: int = 0) -> List[int]:
    """The first ``count`` primes p == 1 (mod n) with p >= max(n+1, min_prime)
    (perm.py:198-220)."""
    n = _as_int(n, "n")
    count = _as_int(count, "count")
    min_prime = _as_int(min_prime, "min_prime")
    if n <= 0:
        raise ValueError("n must be positive")
    if count < 0 or min_prime < 0:
        raise ValueError("count and min_prime must be non-negative")
    candidate = max(n + 1, min_prime)
    candidate += (1 - candidate) % n  # next value == 1 (mod n)
    primes = []
    while len(primes) < count:
        if is_prime(candidate):
            primes.append(candidate)
        candidate += n
    return primes


def _distinct_prime_factors(m: int) -> List[int]:
    factors = []
    d = 2
    while d * d <= m:
        if m % d == 0:
            factors.append(d)
            while m % d == 0:
                m //= d
        d += 1 if d == 2 else 2
    if m > 1:
        factors.append(m)
    return factors


def primitive_root_mod_p(p: int) -> int:
    """Smallest generator of (Z/pZ)* (perm.py:223-259)."""
    p = _as_int(p, "p")
    if p <= 1:
        raise ValueError("p must be greater than 1")
    if not is_prime(p):
        raise ValueError(f"p={p} must be prime")
    if p == 2:
        return 1
    phi = p - 1
    factors = _distinct_prime_factors(phi)
    for g in range(2, p):
        if all(pow(g, phi // f, p) != 1 for f in factors):
            return g
    raise ValueError(f"No primitive root found for p={p}")


def nth_root_of_unity_mod_p(n: int, p: int) -> int:
    """omega = g^((p-1)/n) for the smallest generator g (perm.py:262-277)."""
    n = _as_int(n, "n")
    p = _as_int(p, "p")
    if n <= 0:
        raise ValueError("n must be positive")
    if p <= 1:
        raise ValueError("p must be greater than 1")
    if (p - 1) % n != 0:
        raise ValueError(f"p={p} is not ≡ 1 (mod {n})")
    return pow(primitive_root_mod_p(p), (p - 1) // n, p)


# ---------------------------------------------------------------------------
# CRT
# ---------------------------------------------------------------------------

def extended_gcd(a: int, b: int) -> Tuple[int, int, int]:
    """(g, x, y) with a*x + b*y = g (perm.py:534-539), iterative."""
    old_r, r = a, b
    old_x, x = 1, 0
    old_y, y = 0, 1
    while r:
        quotient = old_r // r
        old_r, r = r, old_r - quotient * r
        old_x, x = x, old_x - quotient * x
        old_y, y = y, old_y - quotient * y
    return old_r, old_x, old_y


def chinese_remainder_theorem(residues: Sequence[int], moduli: Sequence[int]) -> int:
    """Unique solution in [0, prod(moduli)) of x == r_i (mod m_i), built
    incrementally (perm.py:542-584)."""
    if len(residues) != len(moduli):
        raise ValueError("residues and moduli must have the same length")
    if len(residues) == 0:
        return 0
    for modulus in moduli:
        if isinstance(modulus, bool) or not isinstance(modulus, Integral):
            raise TypeError("moduli must be integers")
        if modulus <= 1:
            raise ValueError("moduli must be greater than 1")
    moduli = [int(m) for m in moduli]
    residues = [int(r) if isinstance(r, Integral) else r for r in residues]
    result, product = residues[0], moduli[0]
    for r_i, m_i in zip(residues[1:], moduli[1:]):
        g, inverse, _ = extended_gcd(product, m_i)
        if g != 1:
            raise ValueError("Moduli must be pairwise coprime")
        t = ((r_i - result) % m_i) * inverse % m_i
        result += product * t
        product *= m_i
    return result


@functools.lru_cache(maxsize=256)
def _garner_inverses(moduli: Tuple[int, ...]) -> Tuple[int, ...]:
    """(m_0 ... m_{i-1})^-1 mod m_i for i >= 1, the constants of the
    incremental solution above for this modulus sequence."""
    out, product = [], moduli[0]
    for m in moduli[1:]:
        g, inverse, _ = extended_gcd(product, m)
        if g != 1:
            raise ValueError("Moduli must be pairwise coprime")
        out.append(inverse % m)
        product *= m
    return tuple(out)


def crt_known_moduli(residues: Sequence[int], moduli: Tuple[int, ...]) -> int:
    """chinese_remainder_theorem for a tuple of validated moduli (integers
    greater than 1) that recurs, such as a batch's CRT prime lists: the
    inverses are computed once per modulus tuple."""
    if len(residues) != len(moduli):
        raise ValueError("residues and moduli must have the same length")
    if not moduli:
        return 0
    result, product = residues[0], moduli[0]
    for r, m, inverse in zip(residues[1:], moduli[1:], _garner_inverses(moduli)):
        result += product * (((r - result) % m) * inverse % m)
        product *= m
    return result


def signed_lift(value: int, bound: int, modulus: int) -> int:
    """Map a CRT value in [0, modulus) to (-modulus/2, ...] when it exceeds the
    bound (perm.py:1209-1210, 898-899)."""
    return value - modulus if value > bound else value


# ---------------------------------------------------------------------------
# General integer matrices
# ---------------------------------------------------------------------------

def matrix_permanent_bound(rows: Sequence[Sequence[int]]) -> int:
    """prod_i max(1, sum_j |a_ij|) (perm.py:939-953)."""
    if len(rows) == 0:
        return 1
    if isinstance(rows, np.ndarray) and rows.dtype == np.int64 and rows.ndim == 2:
        # Row sums of |a| stay exact in int64 while n * max|a| < 2^63.
        top = int(np.abs(rows.astype(np.float64)).max(initial=0.0))
        if top < (1 << 62) // max(1, rows.shape[1]):
            bound = 1
            for v in np.abs(rows).sum(axis=1).tolist():
                bound *= max(1, int(v))
            return bound
    bound = 1
    for row in rows:
        bound *= max(1, sum(abs(int(x)) for x in row))
    return bound


def matrix_permanent_bounds(arrays: Sequence[np.ndarray]) -> List[int]:
    """matrix_permanent_bound of many square int64 arrays: the arrays of
    each order are stacked and their row sums of |a| taken at once (exact
    while n * max|a| < 2^63, as in the single version; other arrays go
    through it)."""
    out: List[int] = [0] * len(arrays)
    by_order: dict = {}
    for i, a in enumerate(arrays):
        by_order.setdefault(a.shape[0], []).append(i)
    for n, idx in by_order.items():
        if n == 0:
            for i in idx:
                out[i] = 1
            continue
        stack = np.stack([arrays[i] for i in idx])
        limit = (1 << 62) // n
        safe = (stack.max(axis=(1, 2)) < limit) & (stack.min(axis=(1, 2)) > -limit)
        if not safe.all():
            stack[~safe] = 0
        np.abs(stack, out=stack)
        sums = np.maximum(stack.sum(axis=2), 1).tolist()
        for j, i in enumerate(idx):
            out[i] = math.prod(sums[j]) if safe[j] else matrix_permanent_bound(arrays[i])
    return out


_ASCENDING_PRIMES: List[int] = []
_ASCENDING_PRIMES_LOCK = threading.Lock()


def _ascending_prime(index: int) -> int:
    """The index-th prime at or above 2^30 (the sequence is fixed, so it
    is found once per process and reused by every plan)."""
    if index < len(_ASCENDING_PRIMES):
        return _ASCENDING_PRIMES[index]
    with _ASCENDING_PRIMES_LOCK:
        while len(_ASCENDING_PRIMES) <= index:
            candidate = _ASCENDING_PRIMES[-1] + 1 if _ASCENDING_PRIMES else GENERIC_PRIME_FLOOR
            while not is_prime(candidate):
                candidate += 1
            _ASCENDING_PRIMES.append(candidate)
        return _ASCENDING_PRIMES[index]


def generic_crt_primes(bound: int) -> Tuple[List[int], int]:
    """Ascending primes from 2^30 until their product reaches 2*bound+1
    (perm.py:1115-1128).  Returns (primes, product)."""
    target = 2 * bound + 1
    primes, product = [], 1
    while product < target:
        p = _ascending_prime(len(primes))
        primes.append(p)
        product *= p
    return primes, product


# ---------------------------------------------------------------------------
# Schur matrices
# ---------------------------------------------------------------------------

def schur_permanent_bound(n: int) -> int:
    """ceil(sqrt(n^n)) in exact integers (perm.py:587-605)."""
    n = _as_int(n, "n")
    if n < 0:
        raise ValueError("n must be non-negative")
    if n == 0:
        return 1
    power = n ** n
    root = math.isqrt(power)
    return root if root * root == power else root + 1


def graham_lehmer_schur_congruence(n: int) -> Optional[Tuple[int, int]]:
    """Theorem-derived congruence for odd Schur orders (perm.py:280-326).

    Prime n >= 5: perm(S_n) == n! (mod n^3).  Odd composite n (and n = 3):
    for each exact prime power p^a || n, perm(S_n) is divisible by
    p^((n/p^a) * (p^a - 1)/(p - 1)); the product of those divisors is a
    zero-residue modulus.  Even n and n < 3 return None.
    """
    n = _as_int(n, "n")
    if n < 3 or n % 2 == 0:
        return None
    if n >= 5 and is_prime(n):
        modulus = n ** 3
        return math.factorial(n) % modulus, modulus
    remaining, modulus, prime = n, 1, 3
    while prime * prime <= remaining:
        if remaining % prime == 0:
            prime_power = 1
            while remaining % prime == 0:
                remaining //= prime
                prime_power *= prime
            modulus *= prime ** ((n // prime_power) * ((prime_power - 1) // (prime - 1)))
        prime += 2
    if remaining > 1:
        modulus *= remaining ** (n // remaining)
    return 0, modulus


def plan_schur_crt_primes(n: int, target: int, min_prime: int) -> Tuple[List[int], int]:
    """Primes == 1 (mod n) from min_prime upward until the product reaches
    target (perm.py:769-784)."""
    primes, product = [], 1
    while product < target:
        for p in find_primes_congruent_to_1_mod_n(n, 10, min_prime):
            if product >= target:
                break
            primes.append(p)
            product *= p
            min_prime = p + 1
    return primes, product


def plan_schur_exact_crt(n: int, target: int, min_prime: int):
    """Kernel primes plus the optional theorem congruence (perm.py:787-803).

    Returns (primes, prime_product, congruence-or-None)."""
    congruence = graham_lehmer_schur_congruence(n)
    if congruence is None:
        primes, product = plan_schur_crt_primes(n, target, min_prime)
        return primes, product, None
    theorem_modulus = congruence[1]
    remaining = (target + theorem_modulus - 1) // theorem_modulus
    primes, product = plan_schur_crt_primes(n, remaining, min_prime)
    return primes, product, congruence


def schur_minor_mod_p(n: int, omega: int, p: int) -> List[List[int]]:
    """The (0,0) minor of S_n: entries omega^((r+1)(c+1)) mod p for
    0 <= r, c < n-1 (perm.py:501-510)."""
    m = n - 1
    rows = []
    omega_j = omega % p
    for _ in range(m):
        row, power = [], omega_j
        for _ in range(m):
            row.append(power)
            power = power * omega_j % p
        rows.append(row)
        omega_j = omega_j * omega % p
    return rows


def build_schur_matrix_mod_p(n: int, omega: int, p: int) -> List[List[int]]:
    """S_n entries omega^(jk) mod p, 0 <= j, k < n (perm.py:420-442)."""
    omega %= p
    matrix, omega_j = [], 1 % p
    for _ in range(n):
        row, power = [], 1 % p
        for _ in range(n):
            row.append(power)
            power = power * omega_j % p
        matrix.append(row)
        omega_j = omega_j * omega % p
    return matrix
