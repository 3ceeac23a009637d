"""Permanents modulo 3 on MI355X (bit-sliced F_3 kernel).

Drop-in counterpart of the reference's ``perm_mod3`` module.  The encoding
helpers (perm_mod3.py:62-140) are kept as host utilities; every permanent is
computed by the device kernel (``perm_mod3_omp`` is the GPU object shaped
like the reference's OpenMP extension).  ``num_threads`` is validated and
otherwise ignored.
"""

from __future__ import annotations

from numbers import Integral
from typing import Tuple

import numpy as np

from permhip import _native
from permhip.backends import perm_mod3_omp

HAS_MOD3_OMP = True
HAS_NUMBA = False


def _validate_num_threads(num_threads):
    if num_threads is None:
        return None
    if isinstance(num_threads, bool) or not isinstance(num_threads, Integral):
        raise TypeError("num_threads must be a positive integer or None")
    num_threads = int(num_threads)
    if num_threads <= 0:
        raise ValueError("num_threads must be positive")
    return num_threads


def _as_square_integer_matrix(matrix) -> np.ndarray:
    try:
        matrix = np.asarray(matrix)
    except ValueError as exc:
        raise ValueError("Matrix must be a rectangular two-dimensional array") from exc
    if matrix.ndim == 1 and matrix.size == 0:
        matrix = np.empty((0, 0), dtype=np.int64)
    if matrix.ndim != 2:
        raise ValueError("Matrix must be two-dimensional")
    if matrix.shape[0] != matrix.shape[1]:
        raise ValueError("Matrix must be square")
    if not np.issubdtype(matrix.dtype, np.integer):
        raise TypeError("Matrix entries must be integers")
    return matrix


def encode_f3_element(x: int) -> Tuple[int, int]:
    """0 -> (0,0), 1 -> (1,0), -1 == 2 -> (1,1)."""
    x = x % 3
    return (0, 0) if x == 0 else (1, 0) if x == 1 else (1, 1)


def decode_f3_element(mag: int, sgn: int) -> int:
    return 0 if mag == 0 else 1 if sgn == 0 else 2


def encode_matrix_f3(matrix) -> Tuple[np.ndarray, np.ndarray]:
    """Column words: bit i of mags[j] is a_ij != 0, of sgns[j] is a_ij == -1."""
    matrix = _as_square_integer_matrix(matrix)
    n = matrix.shape[0]
    if n >= 64:
        raise ValueError("Matrix size must be less than 64 for bipedal encoding")
    reduced = np.asarray(matrix % 3, dtype=np.int64)
    mags = np.zeros(n, dtype=np.uint64)
    sgns = np.zeros(n, dtype=np.uint64)
    for j in range(n):
        m = s = 0
        for i in range(n):
            v = int(reduced[i, j])
            if v:
                m |= 1 << i
                if v == 2:
                    s |= 1 << i
        mags[j] = m
        sgns[j] = s
    return mags, sgns


def bipedal_add(m1, s1, m2, s2):
    """(m1,s1) + (m2,s2) (perm_mod3.py:111-123)."""
    t = m1 ^ s1 ^ s2
    return (m2 & t) | (m1 ^ m2), (m2 & t) ^ s1


def bipedal_sub(m1, s1, m2, s2):
    """(m1,s1) - (m2,s2) (perm_mod3.py:126-138)."""
    t = s1 ^ s2
    return (m1 & t) | (m1 ^ m2), (m1 & t) ^ (m2 ^ s2)


def permanent_mod3_fast(matrix, num_threads=None) -> int:
    """perm(matrix) mod 3 in {0, 1, 2} (perm_mod3.py:355-383)."""
    matrix = _as_square_integer_matrix(matrix)
    _validate_num_threads(num_threads)
    n = matrix.shape[0]
    if n >= 64:
        raise ValueError("Matrix size must be less than 64")
    if n == 0:
        return 1
    if n == 1:
        return int(matrix[0, 0]) % 3
    return _native.permanent_mod3(np.asarray(matrix % 3, dtype=np.int64))


def permanent_mod3_ryser(matrix) -> int:
    """Same residue as permanent_mod3_fast (the algorithm is internal)."""
    matrix = _as_square_integer_matrix(matrix)
    if matrix.shape[0] >= 64:
        raise ValueError("Matrix size must be less than 64 for bipedal encoding")
    return permanent_mod3_fast(matrix)


def permanent_mod3_naive(matrix) -> int:
    """Same residue as permanent_mod3_fast."""
    matrix = _as_square_integer_matrix(matrix)
    if matrix.shape[0] >= 64:
        raise ValueError("Matrix size must be less than 64")
    return permanent_mod3_fast(matrix)


def _general_residue_mod3(matrix) -> int:
    """perm(matrix) mod 3 through the general residue kernel (Glynn over
    the integers modulo 3), a device path independent of the bit-sliced
    mod-3 kernel."""
    matrix = _as_square_integer_matrix(matrix)
    n = matrix.shape[0]
    if n == 0:
        return 1
    return _native.permanent_residues(np.asarray(matrix % 3, dtype=np.int64), [3])[0]


def benchmark_mod3(n: int = 20, num_trials: int = 10):
    """Time the mod-3 kernel against the general residue kernel modulo 3 on
    seeded random matrices, after checking that they agree (the demo of
    perm_mod3.py:386-439, with the two device paths in place of the
    reference's naive and bipedal CPU loops)."""
    import time

    print(f"Benchmarking permanent mod 3 for {n}x{n} matrices")
    print("=" * 50)
    np.random.seed(42)
    matrices = [np.random.randint(0, 3, (n, n)) for _ in range(num_trials)]
    print("Verifying correctness...")
    for M in matrices[:3]:
        general = _general_residue_mod3(M)
        fast = permanent_mod3_fast(M)
        if general != fast:
            print(f"MISMATCH: general={general}, fast={fast}")
            return
    print("Correctness verified.")
    print()
    start = time.perf_counter()
    for M in matrices:
        _general_residue_mod3(M)
    general_time = time.perf_counter() - start
    start = time.perf_counter()
    for M in matrices:
        permanent_mod3_fast(M)
    fast_time = time.perf_counter() - start
    print(f"General residue kernel: {general_time:.4f}s ({general_time / num_trials * 1000:.2f}ms per matrix)")
    print(f"Mod-3 kernel:           {fast_time:.4f}s ({fast_time / num_trials * 1000:.2f}ms per matrix)")
    print(f"Speedup:                {general_time / fast_time:.1f}x")


def test_random_matrices():
    """Random integer matrices mod 3, seeded as perm_mod3.py:442-457: the
    mod-3 kernel against the general residue kernel modulo 3."""
    print("Testing random integer matrices mod 3")
    print("=" * 50)
    np.random.seed(123)
    for n in [5, 10, 15, 20]:
        M = np.random.randint(-10, 11, (n, n))
        general = _general_residue_mod3(M)
        fast = permanent_mod3_fast(M)
        print(f"n={n:2d}: general={general}, fast={fast}, match={general == fast}")
