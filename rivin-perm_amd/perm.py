"""Exact permanents of integer and Schur matrices on MI355X.

Drop-in counterpart of the reference's ``perm`` module (perm.py): the public
functions keep their names, signatures, validation and exceptions, and
return the same Python integers.  Residues are computed by the HIP kernels
of ``libpermhip.so``; every CRT prime of one permanent goes to the device in
a single launch.  Bounds, prime plans and Chinese remaindering stay on the
host in exact integers.

``perm_c_ext`` is the GPU object shaped like the reference's C extension
(``permhip.backends.HipPermanentExtension``) and ``HAS_C_EXT`` is True, so
code written against the reference's dispatch seam (perm.py:908-930) keeps
working.  Without a HIP device the residue functions raise RuntimeError:
there is no CPU fallback.
"""

from __future__ import annotations

import math
from numbers import Integral
from typing import List, Optional, Tuple

import numpy as np

from permhip import _native
from permhip import planner as _planner
from permhip import lowrank as _lowrank
from permhip.backends import perm_c_ext

HAS_C_EXT = True
HAS_NUMBA = False

_INT64_MAX = (1 << 63) - 1
_OPENMP_MATRIX_MIN_ORDER = 17
_SCHUR_CRT_PRIME_FLOOR = _planner.SCHUR_CRT_PRIME_FLOOR

LowRankLimitExceeded = _lowrank.LowRankLimitExceeded

is_prime = _planner.is_prime
find_primes_congruent_to_1_mod_n = _planner.find_primes_congruent_to_1_mod_n
primitive_root_mod_p = _planner.primitive_root_mod_p
nth_root_of_unity_mod_p = _planner.nth_root_of_unity_mod_p
extended_gcd = _planner.extended_gcd
chinese_remainder_theorem = _planner.chinese_remainder_theorem
schur_permanent_bound = _planner.schur_permanent_bound
_graham_lehmer_schur_congruence = _planner.graham_lehmer_schur_congruence
_plan_schur_crt_primes = _planner.plan_schur_crt_primes
_plan_schur_exact_crt = _planner.plan_schur_exact_crt


# ---------------------------------------------------------------------------
# Validation (perm.py:52-122)
# ---------------------------------------------------------------------------

def _validate_nonnegative_int(value, name: str) -> int:
    if isinstance(value, bool) or not isinstance(value, Integral):
        raise TypeError(f"{name} must be an integer")
    value = int(value)
    if value < 0:
        raise ValueError(f"{name} must be non-negative")
    return value


def _validate_positive_int(value, name: str) -> int:
    value = _validate_nonnegative_int(value, name)
    if value == 0:
        raise ValueError(f"{name} must be positive")
    return value


def _validate_modulus(p) -> int:
    if isinstance(p, bool) or not isinstance(p, Integral):
        raise TypeError("p must be an integer")
    p = int(p)
    if p <= 1:
        raise ValueError("p must be greater than 1")
    return p


def _validate_num_threads(num_threads) -> int:
    if isinstance(num_threads, bool) or not isinstance(num_threads, Integral):
        raise TypeError("num_threads must be an integer")
    num_threads = int(num_threads)
    if num_threads < 0:
        raise ValueError("num_threads must be non-negative")
    return num_threads


def _validate_square_matrix(matrix) -> int:
    shape = getattr(matrix, "shape", None)
    if shape is not None:
        if len(shape) != 2:
            raise ValueError("Matrix must be two-dimensional")
        if shape[0] != shape[1]:
            raise ValueError("Matrix must be square")
    try:
        n = len(matrix)
    except TypeError as exc:
        raise TypeError("Matrix must be a sequence of rows") from exc
    for row in matrix:
        try:
            row_length = len(row)
        except TypeError as exc:
            raise TypeError("Matrix rows must be sequences") from exc
        if row_length != n:
            raise ValueError("Matrix must be square")
    return n


def _validate_integer_matrix(matrix) -> int:
    n = _validate_square_matrix(matrix)
    for row in matrix:
        for entry in row:
            if not isinstance(entry, Integral):
                raise TypeError("Matrix entries must be integers")
    return n


def get_backend() -> str:
    """The active residue backend: always 'hip' (no CPU tiers)."""
    return "hip"


def gray_code_next(i: int) -> int:
    """Bit that changes between Gray codes i and i+1 (perm.py:329-334)."""
    return (i + 1) & -(i + 1)


def gray_code_bit_position(i: int) -> int:
    """Index of the lowest set bit of i, 0 for i == 0 (perm.py:337-345)."""
    if i == 0:
        return 0
    return (i & -i).bit_length() - 1


# ---------------------------------------------------------------------------
# Device residues
# ---------------------------------------------------------------------------

def _as_int64_matrix(rows) -> Optional[np.ndarray]:
    """int64 copy of integer rows when every entry fits, else None."""
    n = len(rows)
    flat = [int(x) for row in rows for x in row]
    if any(v < -(1 << 63) or v > _INT64_MAX for v in flat):
        return None
    return np.array(flat, dtype=np.int64).reshape(n, n)


def _device_residues(rows, moduli: List[int]) -> List[int]:
    """perm(rows) mod each modulus (every modulus < 2^63), one device pass."""
    n = len(rows)
    if n == 0:
        return [1 % m for m in moduli]
    a = _as_int64_matrix(rows)
    if a is None:
        # Entries beyond int64: reduce per modulus in exact integers first.
        a = np.array(
            [[[int(x) % m for x in row] for row in rows] for m in moduli], dtype=np.int64
        )
    return _native.permanent_residues(a, moduli)


def ryser_permanent_mod_p(matrix, p: int) -> int:
    """perm(matrix) mod p (perm.py:348-417); computed on the device."""
    n = _validate_integer_matrix(matrix)
    p = _validate_modulus(p)
    matrix = [[int(e) for e in row] for row in matrix]
    if n == 0:
        return 1
    if n == 1:
        return matrix[0][0] % p
    if p > _INT64_MAX:
        return _big_modulus_residue(matrix, p)
    return _device_residues(matrix, [p])[0]


ryser_permanent_mod_p_fast = ryser_permanent_mod_p


def _big_modulus_residue(matrix, p: int) -> int:
    """Residue for a modulus beyond 63 bits: exact permanent, then reduce."""
    return matrix_permanent(matrix, parallel=True) % p


def build_schur_matrix_mod_p(n: int, omega: int, p: int) -> List[List[int]]:
    """S_n with entries omega^(jk) mod p (perm.py:420-442)."""
    n = _validate_nonnegative_int(n, "n")
    p = _validate_modulus(p)
    if isinstance(omega, Integral):
        omega = int(omega)
    return _planner.build_schur_matrix_mod_p(n, omega, p)


def schur_permanent_mod_p(n: int, p: int) -> int:
    """perm(S_n) mod p for p == 1 (mod n) (perm.py:515-531)."""
    n = _validate_nonnegative_int(n, "n")
    p = _validate_modulus(p)
    if n <= 1:
        return 1
    if (p - 1) % n != 0:
        raise ValueError(f"p={p} is not ≡ 1 (mod {n})")
    if n % 2 == 0:
        return 0
    omega = nth_root_of_unity_mod_p(n, p)
    return perm_c_ext.schur_permanent_mod_p(n, omega, p)


def schur_permanent_mod_p_fast(n: int, p: int, num_threads: int = 0) -> int:
    """perm(S_n) mod p (perm.py:729-759)."""
    n = _validate_nonnegative_int(n, "n")
    p = _validate_modulus(p)
    _validate_num_threads(num_threads)
    if n <= 1:
        return 1
    if (p - 1) % n != 0:
        raise ValueError(f"p={p} is not ≡ 1 (mod {n})")
    if n % 2 == 0:
        return 0
    omega = nth_root_of_unity_mod_p(n, p)
    return perm_c_ext.schur_permanent_mod_p(n, omega, p)


def _schur_residues(n: int, primes: List[int]) -> List[int]:
    """perm(S_n) mod every prime in one device launch.

    For odd n and p == 1 (mod n), expansion along the all-ones first row
    gives n equal permanental cofactors (perm_c_omp.c:964-975), so each
    residue is n * per(minor) with minor entries omega^((r+1)(c+1)).
    """
    if not primes:
        return []
    if (_planner.is_prime(n) and _native.get_tuning(_native.TUNE_NECKLACE_MIN_ORDER) <= n <= 47
            and all(p < (1 << 32) for p in primes)):
        # Prime orders: the minor is circulant after reindexing by a primitive
        # root, and the device sums Glynn terms over necklace orbits
        # (perm_c_omp.c:979-1006), about 2(n-1) times fewer terms.
        return [_native.schur_permanent_modp(n, nth_root_of_unity_mod_p(n, p), p) for p in primes]
    minors = []
    for p in primes:
        omega = nth_root_of_unity_mod_p(n, p)
        minors.append(_planner.schur_minor_mod_p(n, omega, p))
    stack = np.array(minors, dtype=np.int64)
    cofactors = _native.permanent_residues(stack, primes)
    return [(n % p) * c % p for p, c in zip(primes, cofactors)]


def _schur_exact(n: int, min_prime: int, verbose: bool) -> int:
    bound = schur_permanent_bound(n)
    target = 2 * bound + 1
    primes, prime_product, congruence = _plan_schur_exact_crt(n, target, min_prime)
    theorem_modulus = congruence[1] if congruence is not None else 1
    reconstruction_modulus = prime_product * theorem_modulus
    if verbose:
        print(f"Using {len(primes)} primes: {primes[:5]}{'...' if len(primes) > 5 else ''}")
        if congruence is not None:
            print(f"Graham-Lehmer CRT modulus: {theorem_modulus}")
        print(f"CRT modulus product: {reconstruction_modulus:.2e}, bound: {bound:.2e}")
    residues = _schur_residues(n, primes)
    if verbose:
        for p, r in zip(primes, residues):
            print(f"  perm(S_{n}) mod {p} = {r}")
    moduli = list(primes)
    if congruence is not None:
        residues.append(congruence[0])
        moduli.append(congruence[1])
    result = chinese_remainder_theorem(residues, moduli)
    return _planner.signed_lift(result, bound, reconstruction_modulus)


def schur_permanent(n: int, verbose: bool = False) -> int:
    """Exact perm(S_n) with primes from n+1 upward (perm.py:608-668)."""
    n = _validate_nonnegative_int(n, "n")
    if n <= 1:
        return 1
    if n % 2 == 0:
        return 0
    return _schur_exact(n, n + 1, verbose)


def schur_permanent_fast(n: int, verbose: bool = False, parallel: bool = True,
                         num_threads: int = 0) -> int:
    """Exact perm(S_n) with primes from 2^30 upward (perm.py:806-901)."""
    n = _validate_nonnegative_int(n, "n")
    _validate_num_threads(num_threads)
    if n <= 1:
        return 1
    if n % 2 == 0:
        return 0
    return _schur_exact(n, max(n + 1, _SCHUR_CRT_PRIME_FLOOR), verbose)


def schur_permanent_of_size(n: int, verbose: bool = False, parallel: bool = True,
                            num_threads: int = 0) -> int:
    """The permanent of the n x n Schur matrix (perm.py:1219-1240)."""
    return schur_permanent_fast(n, verbose, parallel, num_threads)


permanent = schur_permanent_of_size


# ---------------------------------------------------------------------------
# General integer matrices
# ---------------------------------------------------------------------------

def matrix_permanent_bound(matrix) -> int:
    """prod_i max(1, sum_j |a_ij|) (perm.py:939-953)."""
    n = _validate_integer_matrix(matrix)
    if n == 0:
        return 1
    return _planner.matrix_permanent_bound(matrix)


def matrix_permanent(
    matrix,
    verbose: bool = False,
    parallel: bool = True,
    num_threads: int = 0,
    *,
    method: str = "auto",
) -> int:
    """Exact permanent of an integer matrix (perm.py:1076-1212).

    All CRT residues of the generic path are computed in one device launch.
    ``method='low_rank'`` runs the exact low-rank coefficient identity on
    the host and raises LowRankLimitExceeded beyond rank 4 or 50,000
    coefficient states.  ``method='auto'`` first probes the rank modulo
    every CRT prime, up to the largest rank whose host cost beats the
    device's estimated Glynn time (``permhip.lowrank.rank_limit``; orders
    below 20 are never probed), as the reference's auto mode does
    (perm.py:1158-1169); a matrix over that rank at any prime goes to the
    device.  'generic' always uses the device.  ``parallel`` and
    ``num_threads`` are validated and otherwise ignored: the device is
    always used whole.
    """
    n = _validate_integer_matrix(matrix)
    _validate_num_threads(num_threads)
    if not isinstance(method, str):
        raise TypeError("method must be a string")
    method = method.lower()
    if method not in ("auto", "generic", "low_rank"):
        raise ValueError("method must be one of: 'auto', 'generic', 'low_rank'")
    if n == 0:
        return 1
    if n == 1:
        return int(matrix[0][0])
    rows = [[int(x) for x in row] for row in matrix]
    bound = _planner.matrix_permanent_bound(rows)
    primes, prime_product = _planner.generic_crt_primes(bound)
    if verbose:
        print(f"Using {len(primes)} primes")
        print(f"Prime product: {prime_product:.2e}, bound: {bound:.2e}")
    residues = None
    if method == "low_rank":
        residues = _lowrank.low_rank_residues(rows, primes, _lowrank.LOW_RANK_MAX_RANK, required=True)
    elif method == "auto":
        limit = _lowrank.rank_limit(n, len(primes))
        if limit:
            residues = _lowrank.low_rank_residues(rows, primes, limit, required=False)
    if residues is not None:
        if verbose:
            print("Using exact low-rank coefficient computation (host residues)")
    else:
        if n >= 64:
            raise ValueError("matrix dimension must be less than 64")
        residues = _device_residues(rows, primes)
    result = chinese_remainder_theorem(residues, primes)
    return _planner.signed_lift(result, bound, prime_product)


if __name__ == "__main__":
    import sys

    size = int(sys.argv[1]) if len(sys.argv) > 1 else 5
    print(f"perm(S_{size}) = {permanent(size, '-v' in sys.argv)}")
