"""Backend objects shaped like the reference's native extensions.

The reference selects its residue backend through module globals
(``perm.HAS_C_EXT`` / ``perm.perm_c_ext``, perm.py:130-136;
``perm_binary._binary_ext``, perm_binary.py:23-32;
``perm_mod3.perm_mod3_omp``, perm_mod3.py:24-30).  The classes below expose
the same method names, argument meaning, validation order and exception
types as those CPython extensions, but compute every residue on the GPU
through ``libpermhip.so``.  They can be assigned to the reference's globals
as a drop-in, and the mirror modules in this package use them directly.

There is no CPU fallback: without a HIP device every residue call raises.
"""

from __future__ import annotations

import operator
from numbers import Integral

import numpy as np

from . import _native

_INT64_MIN = -(1 << 63)
_INT64_MAX = (1 << 63) - 1
_INT32_MAX = (1 << 31) - 1


def _parse_long_long(value, what: str = "argument") -> int:
    """CPython 'L' conversion: __index__ required, int64 range enforced."""
    if isinstance(value, float):
        raise TypeError(f"{what} must be an integer, not float")
    try:
        v = operator.index(value)
    except TypeError as exc:
        raise TypeError(f"{what} must be an integer") from exc
    if not _INT64_MIN <= v <= _INT64_MAX:
        raise OverflowError("Python int too large to convert to C long long")
    return int(v)


def _parse_int(value, what: str = "argument") -> int:
    """CPython 'i' conversion (32-bit signed)."""
    if isinstance(value, float):
        raise TypeError(f"{what} must be an integer, not float")
    try:
        v = operator.index(value)
    except TypeError as exc:
        raise TypeError(f"{what} must be an integer") from exc
    if v > _INT32_MAX:
        raise OverflowError("signed integer is greater than maximum")
    if v < -_INT32_MAX - 1:
        raise OverflowError("signed integer is less than minimum")
    return int(v)


class HipPermanentExtension:
    """GPU counterpart of ``perm_c_ext`` (perm_c_omp.c:1103-1122)."""

    __name__ = "permhip.perm_c_ext"

    def _matrix(self, matrix, p: int) -> np.ndarray:
        # Mirrors py_matrix_permanent_mod_p's checks (perm_c_omp.c:858-916).
        if not isinstance(matrix, list):
            raise TypeError("matrix must be a list of lists")
        if p <= 0:
            raise ValueError("p must be a positive signed 64-bit integer")
        n = len(matrix)
        if n >= 64:
            raise ValueError("matrix dimension must be less than 64")
        out = np.empty((n, n), dtype=np.int64)
        for i, row in enumerate(matrix):
            if not isinstance(row, list) or len(row) != n:
                raise ValueError("matrix must be square")
            for j, item in enumerate(row):
                out[i, j] = _parse_long_long(item, "matrix entry")
        return out

    def _residue(self, args, what: str) -> int:
        if len(args) < 2 or len(args) > 3:
            raise TypeError(f"{what}() takes 2 or 3 arguments")
        matrix = args[0]
        p = _parse_long_long(args[1], "p")
        if len(args) == 3:
            _parse_int(args[2], "num_threads")
        a = self._matrix(matrix, p)
        if a.shape[0] == 0:
            return 1 % p
        return _native.permanent_residues(a, [p])[0]

    def permanent_mod_p(self, *args) -> int:
        """perm(matrix) mod p for 0 < p < 2^63; num_threads is accepted and
        ignored (the GPU uses the whole device)."""
        return self._residue(args, "permanent_mod_p")

    def ryser_permanent_mod_p(self, *args) -> int:
        """Same residue as permanent_mod_p (the algorithm is internal)."""
        return self._residue(args, "ryser_permanent_mod_p")

    def schur_permanent_mod_p(self, *args) -> int:
        """perm(S_n) mod p for S_n[j][k] = omega^(jk) (perm_c_omp.c:942-1060)."""
        if len(args) < 3 or len(args) > 4:
            raise TypeError("schur_permanent_mod_p() takes 3 or 4 arguments")
        n = _parse_int(args[0], "n")
        omega = _parse_long_long(args[1], "omega")
        p = _parse_long_long(args[2], "p")
        if len(args) == 4:
            _parse_int(args[3], "num_threads")
        if n < 0 or n >= 64:
            raise ValueError("n must satisfy 0 <= n < 64")
        if p <= 0:
            raise ValueError("p must be a positive signed 64-bit integer")
        if n <= 1:
            return 1 % p
        if p == 1:
            return 0
        return _native.schur_permanent_modp(n, omega, p)

    def schur_necklace_orbit_stats(self, *args):
        """(representatives, words) of the necklace walk the Schur path uses
        for prime order n (perm_c_omp.c:1062-1080): the device runs the walk
        with no terms formed.  Orders above 47 are refused by the device."""
        if len(args) != 1:
            raise TypeError("schur_necklace_orbit_stats() takes exactly one argument")
        n = _parse_int(args[0], "n")
        if n < 3 or n >= 64 or n % 2 == 0 or not _is_prime_u64(n):
            raise ValueError("n must be an odd prime satisfying 3 <= n < 64")
        return _native.schur_necklace_orbit_stats(n)

    def get_max_threads(self) -> int:
        """Parallel width reported to thread-count heuristics: the device's
        compute units, so callers never fan residues out over processes."""
        info = _native.device_info()
        for token in info.split():
            if token.startswith("CUs="):
                return max(1, int(token[4:]))
        return 1

    def set_num_threads(self, num_threads) -> None:
        n = _parse_int(num_threads, "num_threads")
        if n <= 0:
            raise ValueError("num_threads must be positive")


def _index_long_long(obj, name: str) -> int:
    # parse_index_long_long (perm_binary_omp.c:367-384).
    if isinstance(obj, (bool, np.bool_)):
        raise TypeError(f"{name} must be an integer")
    try:
        value = operator.index(obj)
    except TypeError as exc:
        raise TypeError(f"{name} must be an integer") from exc
    if not _INT64_MIN <= value <= _INT64_MAX:
        raise OverflowError("Python int too large to convert to C long long")
    return int(value)


class HipBinaryExtension:
    """GPU counterpart of ``perm_binary_ext`` (perm_binary_omp.c:386-534)."""

    __name__ = "permhip.perm_binary_ext"

    def binary_permanent_mod_p(self, matrix, p, num_threads=None) -> int:
        p = _index_long_long(p, "p")
        if num_threads is not None:
            threads = _index_long_long(num_threads, "num_threads")
            if threads > _INT32_MAX:
                raise ValueError("num_threads is too large")
            if threads < 0:
                raise ValueError("num_threads must be non-negative")
        try:
            rows = list(matrix)
        except TypeError as exc:
            raise TypeError("matrix must be a sequence of rows") from exc
        n = len(rows)
        if n > 63:
            raise ValueError("native binary backend supports order at most 63")
        if p <= 1:
            raise ValueError("p must be greater than 1")
        if p == 2:
            raise ValueError("p must be odd")
        if p <= n:
            raise ValueError("p must be greater than the matrix order")
        if not _is_prime_u64(p):
            raise ValueError("p must be prime")
        a = np.zeros((n, n), dtype=np.uint8)
        for i, row in enumerate(rows):
            try:
                items = list(row)
            except TypeError as exc:
                raise TypeError("matrix rows must be sequences") from exc
            if len(items) != n:
                raise ValueError("matrix must be square")
            for j, entry in enumerate(items):
                value = _parse_long_long(entry, "matrix entry")
                if value not in (0, 1):
                    raise ValueError("matrix entries must be 0 or 1")
                a[i, j] = value
        if n == 0:
            return 1
        if n == 1:
            return int(a[0, 0])
        if not a.any(axis=0).all() or not a.any(axis=1).all():
            return 0
        return _native.binary_permanent_residues(a, [p])[0]

    def get_max_threads(self) -> int:
        return HipPermanentExtension().get_max_threads()


def _is_prime_u64(value: int) -> bool:
    """Deterministic Miller--Rabin for 64-bit values (bases of
    perm_binary.py:764-793)."""
    if value < 2:
        return False
    for small in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if value == small:
            return True
        if value % small == 0:
            return False
    d, r = value - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for base in (2, 325, 9375, 28178, 450775, 9780504, 1795265022):
        base %= value
        if base == 0:
            continue
        x = pow(base, d, value)
        if x in (1, value - 1):
            continue
        for _ in range(r - 1):
            x = x * x % value
            if x == value - 1:
                break
        else:
            return False
    return True


class HipMod3Extension:
    """GPU counterpart of ``perm_mod3_omp`` (perm_mod3_omp.c:220-313)."""

    __name__ = "permhip.perm_mod3_omp"

    def permanent_mod3(self, matrix, num_threads=-1) -> int:
        if not isinstance(matrix, np.ndarray):
            raise TypeError("argument 1 must be numpy.ndarray")
        _parse_int(num_threads, "num_threads")
        if matrix.ndim != 2:
            raise ValueError("Matrix must be 2-dimensional")
        if matrix.shape[0] != matrix.shape[1]:
            raise ValueError("Matrix must be square")
        if matrix.shape[0] >= 64:
            raise ValueError("Matrix size must be less than 64")
        n = int(matrix.shape[0])
        if n == 0:
            return 1
        reduced = np.asarray(matrix, dtype=np.int64) % 3
        return _native.permanent_mod3(reduced)


def _poly_matrix(obj, name: str, expected_n: int = -1):
    # parse_matrix (perm_polynomial_omp.c:165-240): a sequence of square rows
    # of int64 integers, order below 64.
    if isinstance(obj, (str, bytes)) or not hasattr(obj, "__len__") or not hasattr(obj, "__getitem__"):
        raise TypeError(f"{name} must be a sequence of rows")
    rows = list(obj)
    n = len(rows)
    if n >= 64:
        raise ValueError(f"{name} order must be less than 64")
    if expected_n >= 0 and n != expected_n:
        raise ValueError("constant_matrix and linear_matrix must have the same order")
    out = np.zeros((n, n), dtype=np.int64)
    for i, row in enumerate(rows):
        if isinstance(row, (str, bytes)) or not hasattr(row, "__len__") or not hasattr(row, "__getitem__"):
            raise TypeError("matrix rows must be sequences")
        items = list(row)
        if len(items) != n:
            raise ValueError(f"{name} must be square")
        for j, entry in enumerate(items):
            try:
                value = operator.index(entry)
            except TypeError as exc:
                raise TypeError("matrix entries must be integers") from exc
            if not _INT64_MIN <= value <= _INT64_MAX:
                raise OverflowError("Python int too large to convert to C long long")
            out[i, j] = value
    return out


def _poly_threads(num_threads) -> None:
    if num_threads is None:
        return
    if isinstance(num_threads, (bool, np.bool_)):
        raise TypeError("num_threads must be an integer")
    value = _index_long_long(num_threads, "num_threads")
    if value < 0:
        raise ValueError("num_threads must be non-negative")
    if value > _INT32_MAX:
        raise ValueError("num_threads is too large")


_ROOK_MAX_MODULI = 16
_WEIGHTED_ROOK_MAX_ENTRY = (1 << 16) - 1


class HipPolynomialExtension:
    """GPU counterpart of ``perm_polynomial_ext`` (perm_polynomial_omp.c:
    1207-1260, 1706-1836): coefficient vectors of per(C + x L) modulo odd
    moduli, and rook / weighted-rook residues, from the coefficient-valued
    Glynn kernel (phip_affine_permanent_residues)."""

    __name__ = "permhip.perm_polynomial_ext"

    def affine_permanent_mod(self, constant_matrix, linear_matrix, modulus, num_threads=None):
        m = _index_long_long(modulus, "modulus")
        if isinstance(modulus, (bool, np.bool_)):
            raise TypeError("modulus must be an integer")
        if m <= 1 or m % 2 == 0:
            raise ValueError("modulus must be an odd signed 64-bit integer greater than 1")
        _poly_threads(num_threads)
        C = _poly_matrix(constant_matrix, "constant_matrix")
        L = _poly_matrix(linear_matrix, "linear_matrix", C.shape[0])
        n = C.shape[0]
        if n == 0:
            return (1 % m,)
        if n == 1:
            return (int(C[0, 0]) % m, int(L[0, 0]) % m)
        import perm_polynomial

        if not perm_polynomial._device_takes(n, m):
            # Beyond the kernels' orders for this modulus: exact
            # coefficients (30-bit primes, CRT), reduced.
            exact = perm_polynomial._affine_permanent_validated(tuple(map(tuple, C.tolist())),
                                                                tuple(map(tuple, L.tolist())))
            return tuple(v % m for v in exact)
        return _native.affine_permanent_residues(C, L, [m])[0]

    def _rook_many(self, matrix, moduli, num_threads, binary: bool):
        _poly_threads(num_threads)
        A = _poly_matrix(matrix, "matrix")
        n = A.shape[0]
        if binary and not np.isin(A, (0, 1)).all():
            raise ValueError("matrix entries must be 0 or 1")
        if not binary and ((A < 0).any() or (A > _WEIGHTED_ROOK_MAX_ENTRY).any()):
            raise ValueError(f"matrix entries must be between 0 and {_WEIGHTED_ROOK_MAX_ENTRY}")
        try:
            mods = [_index_long_long(q, "modulus") for q in moduli]
        except TypeError as exc:
            raise TypeError("moduli must be a sequence of integers") from exc
        if not 1 <= len(mods) <= _ROOK_MAX_MODULI:
            raise ValueError(f"moduli must hold between 1 and {_ROOK_MAX_MODULI} primes")
        for q in mods:
            if q <= max(n, 2) or q % 2 == 0 or not _is_prime_u64(q):
                raise ValueError("moduli must be odd primes greater than the matrix order")
        if n == 0:
            return tuple((1,) for _ in mods)
        if n == 1:
            return tuple((1, int(A[0, 0]) % q) for q in mods)
        # r_k = [x^k] per(J + x A) / (n - k)!
        J = np.ones((n, n), dtype=np.int64)
        import perm_polynomial

        if all(perm_polynomial._device_takes(n, q) for q in mods):
            rows = _native.affine_permanent_residues(J, A, mods)
        else:
            # Orders the kernels take only below 2^30 for these primes: the
            # exact coefficients (30-bit primes, CRT), reduced per prime.
            exact = perm_polynomial._affine_permanent_validated(tuple(map(tuple, J.tolist())),
                                                                tuple(map(tuple, A.tolist())))
            rows = [[v % q for v in exact] for q in mods]
        out = []
        for q, row in zip(mods, rows):
            fact = [1] * (n + 1)
            for i in range(1, n + 1):
                fact[i] = fact[i - 1] * i % q
            out.append(tuple(c * pow(fact[n - k], -1, q) % q for k, c in enumerate(row)))
        return tuple(out)

    def rook_polynomial_mod_many(self, matrix, moduli, num_threads=None):
        return self._rook_many(matrix, moduli, num_threads, True)

    def weighted_rook_polynomial_mod_many(self, matrix, moduli, num_threads=None):
        return self._rook_many(matrix, moduli, num_threads, False)


perm_c_ext = HipPermanentExtension()
perm_binary_ext = HipBinaryExtension()
perm_mod3_omp = HipMod3Extension()
perm_polynomial_ext = HipPolynomialExtension()
