"""Exact and modular permanents of affine polynomial matrices on MI355X.

Drop-in counterpart of the reference's ``perm_polynomial`` module
(perm_polynomial.py): the public functions keep their names, keyword-only
options, validation messages and return values.

* ``affine_permanent_mod(C, L, m)``: the coefficients of per(C + x L) modulo
  an odd m, ascending degree, ``order + 1`` entries.
* ``affine_permanent(C, L)``: the exact integer coefficients, from residues
  modulo descending 30-bit primes (all of them in one device pass) and CRT,
  certified by ``affine_permanent_bounds``.
* ``principal_permanent_polynomial[_mod](A)``: per(I + x A).
* ``rook_polynomial(A)`` / ``weighted_rook_polynomial(W)``: matching counts
  and weighted matching sums, the coefficients of per(J + x A) divided by
  (n - k)!.

Every residue comes from the coefficient-valued Glynn kernel on the device
(``_polynomial_ext`` is the GPU object shaped like the reference's native
extension).  Bounds, prime plans and CRT stay on the host in exact
integers.  backend='python' (the reference's portable backend) is accepted
and computes on the device like 'native': the coefficients are the same
integers, and this build has no host arithmetic path.  The device kernel
takes every order the reference accepts (below 64) for moduli below 2^30,
up to order 40 below 2^32 and up to 32 above; beyond those orders a
modulus gets the exact coefficients (30-bit primes and CRT) reduced.  Orders
above 40 walk 2^40 and more Gray states of degree-n products per modulus.
"""

from __future__ import annotations

import operator
import threading
from math import factorial
from typing import Sequence

import numpy as np

from permhip import _native
from permhip import binary_prep as _prep
from permhip.backends import perm_polynomial_ext as _polynomial_ext

_INT64_MAX = (1 << 63) - 1
_INT32_MAX = (1 << 31) - 1
_CRT_PRIME_START = (1 << 30) - 1
_ROOK_NATIVE_MAX_MODULI = 16

HAS_POLYNOMIAL_EXT = True

_POLYNOMIAL_NATIVE_HOOK_NAMES = (
    "affine_permanent_mod",
    "rook_polynomial_mod_many",
    "weighted_rook_polynomial_mod_many",
)


def polynomial_backend_capabilities() -> dict:
    """Available backends and native hooks (perm_polynomial.py:70-95)."""
    hooks = tuple(name for name in _POLYNOMIAL_NATIVE_HOOK_NAMES if hasattr(_polynomial_ext, name))
    return {
        "available_backends": ("native", "python"),
        "default_backend": "native",
        "native_extension_available": True,
        "native_hooks": hooks,
    }


# ---------------------------------------------------------------------------
# Validation (messages of perm_polynomial.py:98-223).


def _validate_modulus(modulus) -> int:
    if isinstance(modulus, bool):
        raise TypeError("modulus must be an integer")
    try:
        value = int(operator.index(modulus))
    except TypeError as exc:
        raise TypeError("modulus must be an integer") from exc
    if value <= 1 or value % 2 == 0:
        raise ValueError("modulus must be an odd integer greater than 1")
    if value > _INT64_MAX:
        raise ValueError("modulus must fit in a positive signed 64-bit integer")
    return value


def _validate_num_threads(num_threads) -> int:
    if isinstance(num_threads, bool):
        raise TypeError("num_threads must be an integer")
    try:
        value = int(operator.index(num_threads))
    except TypeError as exc:
        raise TypeError("num_threads must be an integer") from exc
    if value < 0:
        raise ValueError("num_threads must be non-negative")
    if value > _INT32_MAX:
        raise ValueError("num_threads is too large")
    return value


def _validate_parallel(parallel) -> None:
    if not isinstance(parallel, bool):
        raise TypeError("parallel must be a boolean")


def _select_backend(backend) -> str:
    if not isinstance(backend, str):
        raise TypeError("backend must be a string")
    chosen = backend.lower()
    if chosen in ("auto", "native", "python"):
        return "native"
    raise ValueError("backend must be one of: 'auto', 'native', 'python'")


def _snapshot_integer_matrix(matrix, name: str):
    """Square matrix of arbitrary Python integers as a tuple of tuples."""
    shape = getattr(matrix, "shape", None)
    if shape is not None:
        try:
            shape = tuple(shape)
        except TypeError as exc:
            raise ValueError(f"{name} must have a two-dimensional square shape") from exc
        if len(shape) != 2:
            raise ValueError(f"{name} must be two-dimensional")
        if shape[0] != shape[1]:
            raise ValueError(f"{name} must be square")
    try:
        order = len(matrix)
    except TypeError as exc:
        raise TypeError(f"{name} must be a sequence of rows") from exc
    if order >= 64:
        raise ValueError("matrix order must be less than 64")
    rows = []
    for r, row in enumerate(matrix):
        try:
            width = len(row)
        except TypeError as exc:
            raise TypeError(f"{name} rows must be sequences") from exc
        if width != order:
            raise ValueError(f"{name} must be square")
        values = []
        for c, entry in enumerate(row):
            try:
                values.append(int(operator.index(entry)))
            except TypeError as exc:
                raise TypeError(
                    f"{name} entries must be integers; entry ({r}, {c}) is {type(entry).__name__}"
                ) from exc
        rows.append(tuple(values))
    if len(rows) != order:
        raise ValueError(f"{name} row count changed during validation")
    return tuple(rows)


def _snapshot_binary_matrix(matrix):
    rows = _snapshot_integer_matrix(matrix, "matrix")
    for r, row in enumerate(rows):
        for c, value in enumerate(row):
            if value not in (0, 1):
                raise ValueError(f"matrix entries must be 0 or 1; entry ({r}, {c}) is {value}")
    return rows


def _snapshot_nonnegative_integer_matrix(matrix):
    rows = _snapshot_integer_matrix(matrix, "matrix")
    for r, row in enumerate(rows):
        for c, value in enumerate(row):
            if value < 0:
                raise ValueError(f"matrix entries must be nonnegative; entry ({r}, {c}) is {value}")
    return rows


def _validate_affine_matrices(constant_matrix, linear_matrix):
    constant = _snapshot_integer_matrix(constant_matrix, "constant_matrix")
    linear = _snapshot_integer_matrix(linear_matrix, "linear_matrix")
    if len(constant) != len(linear):
        raise ValueError("constant_matrix and linear_matrix must have the same shape")
    return constant, linear


# ---------------------------------------------------------------------------
# Bounds and prime plans.


def _affine_product(constants: Sequence[int], linears: Sequence[int]) -> list:
    """Coefficients of prod_i (constants[i] + x linears[i]) over the integers."""
    poly = [1]
    for a, b in zip(constants, linears):
        nxt = [0] * (len(poly) + 1)
        for d, v in enumerate(poly):
            nxt[d] += v * a
            nxt[d + 1] += v * b
        poly = nxt
    return poly


def _affine_bounds(constant, linear) -> tuple:
    """Coefficientwise absolute bounds of per(C + x L) (perm_polynomial.py:
    782-860): the smallest of four relaxations -- the product over rows of
    (sum_j |C_ij| + x sum_j |L_ij|), the same over columns, and n! times the
    products of the row and of the column maxima."""
    n = len(constant)
    if n == 0:
        return (1,)
    cols_c = [[abs(constant[i][j]) for i in range(n)] for j in range(n)]
    cols_l = [[abs(linear[i][j]) for i in range(n)] for j in range(n)]
    rows_c = [[abs(v) for v in row] for row in constant]
    rows_l = [[abs(v) for v in row] for row in linear]
    fact = factorial(n)
    candidates = [
        _affine_product([sum(r) for r in rows_c], [sum(r) for r in rows_l]),
        _affine_product([sum(c) for c in cols_c], [sum(c) for c in cols_l]),
        [fact * v for v in _affine_product([max(r) for r in rows_c], [max(r) for r in rows_l])],
        [fact * v for v in _affine_product([max(c) for c in cols_c], [max(c) for c in cols_l])],
    ]
    return tuple(min(values) for values in zip(*candidates))


def affine_permanent_bounds(constant_matrix, linear_matrix) -> tuple:
    """Certified absolute bounds for the coefficients of per(C + x L), in
    ascending degree (perm_polynomial.py:883-895)."""
    constant, linear = _validate_affine_matrices(constant_matrix, linear_matrix)
    return _affine_bounds(constant, linear)


_PRIMES: list = []
_PRIMES_LOCK = threading.Lock()


def _prime(index: int) -> int:
    """The index-th prime counting down from 2^30 - 1 (a fixed sequence,
    found once per process)."""
    if index < len(_PRIMES):
        return _PRIMES[index]
    with _PRIMES_LOCK:
        while len(_PRIMES) <= index:
            c = _PRIMES[-1] - 2 if _PRIMES else _CRT_PRIME_START
            while not _prep.is_prime_u64(c):
                c -= 2
            _PRIMES.append(c)
        return _PRIMES[index]


def _descending_primes(target: int, minimum: int = 3) -> list:
    """Descending primes until their product exceeds target, all above
    `minimum` (the plan of perm_polynomial.py:929-945, started below 2^30 so
    that every residue runs on the device's signed lazy Montgomery kernel;
    the exact coefficients do not depend on the primes)."""
    primes, product = [], 1
    while product <= target:
        p = _prime(len(primes))
        if p <= minimum:
            raise RuntimeError("exhausted the deterministic CRT prime range")
        primes.append(p)
        product *= p
    return primes


def _crt_vectors(rows: Sequence[Sequence[int]], primes: Sequence[int]) -> tuple:
    """Coefficientwise CRT (Garner) of residue vectors; returns (values in
    [0, M), M)."""
    values = list(rows[0])
    modulus = primes[0]
    for residues, p in zip(rows[1:], primes[1:]):
        inv = pow(modulus % p, -1, p)
        for d, r in enumerate(residues):
            values[d] += modulus * ((r - values[d]) * inv % p)
        modulus *= p
    return values, modulus


# ---------------------------------------------------------------------------
# Device residues.


# Largest orders of the device kernel: every order below 64 for moduli below
# 2^30, 40 for moduli below 2^32, 32 for larger ones.
DEVICE_MAX_ORDER = 63
DEVICE_MAX_ORDER_32 = 40
DEVICE_MAX_ORDER_WIDE = 32


def _device_takes(n: int, modulus: int) -> bool:
    """Whether the device kernel computes order n modulo `modulus` itself
    (otherwise the exact coefficients are reduced)."""
    if modulus < (1 << 30):
        return n <= DEVICE_MAX_ORDER
    if modulus < (1 << 32):
        return n <= DEVICE_MAX_ORDER_32
    return n <= DEVICE_MAX_ORDER_WIDE


def _check_device_order(n: int) -> None:
    if n > DEVICE_MAX_ORDER:
        raise ValueError("matrix order must be less than 64")


def _device_affine_residues(constant, linear, primes) -> list:
    """Coefficient vectors modulo each prime, one device pass.  Entries that
    do not fit int64 are reduced per prime on the host first."""
    n = len(constant)
    _check_device_order(n)
    flat = [v for m in (constant, linear) for row in m for v in row]
    if all(-(1 << 63) <= v <= _INT64_MAX for v in flat):
        C = np.array(constant, dtype=np.int64).reshape(n, n)
        L = np.array(linear, dtype=np.int64).reshape(n, n)
    else:
        C = np.array([[[v % p for v in row] for row in constant] for p in primes], dtype=np.int64)
        L = np.array([[[v % p for v in row] for row in linear] for p in primes], dtype=np.int64)
    rows = _native.affine_permanent_residues(C, L, primes)
    for row, p in zip(rows, primes):
        if len(row) != n + 1 or not all(isinstance(v, int) and 0 <= v < p for v in row):
            raise RuntimeError("native polynomial-permanent backend returned invalid coefficients")
    return rows


def _affine_permanent_mod_validated(constant, linear, modulus: int) -> tuple:
    n = len(constant)
    _check_device_order(n)
    if not _device_takes(n, modulus):
        # Beyond the kernels' orders for this modulus: the exact
        # coefficients (30-bit primes) reduced.
        return tuple(v % modulus for v in _affine_permanent_validated(constant, linear))
    reduced_c = tuple(tuple(v % modulus for v in row) for row in constant)
    reduced_l = tuple(tuple(v % modulus for v in row) for row in linear)
    if not reduced_c:
        return (1 % modulus,)
    if len(reduced_c) == 1:
        return (reduced_c[0][0], reduced_l[0][0])
    return tuple(_polynomial_ext.affine_permanent_mod(reduced_c, reduced_l, modulus))


def affine_permanent_mod(constant_matrix, linear_matrix, modulus, *, backend: str = "auto",
                         parallel: bool = True, num_threads: int = 0) -> tuple:
    """Coefficients of per(constant_matrix + x linear_matrix) modulo an odd
    modulus (not necessarily prime), ascending degree, ``order + 1`` entries
    (perm_polynomial.py:1086-1114)."""
    modulus = _validate_modulus(modulus)
    _validate_parallel(parallel)
    _validate_num_threads(num_threads)
    _select_backend(backend)
    constant, linear = _validate_affine_matrices(constant_matrix, linear_matrix)
    return _affine_permanent_mod_validated(constant, linear, modulus)


def _affine_permanent_validated(constant, linear) -> tuple:
    """Exact coefficients (perm_polynomial.py:1117-1180)."""
    n = len(constant)
    if n == 0:
        return (1,)
    if n == 1:
        return (constant[0][0], linear[0][0])
    bounds = _affine_bounds(constant, linear)
    top = max(bounds)
    if top == 0:
        return (0,) * (n + 1)
    nonnegative = all(v >= 0 for m in (constant, linear) for row in m for v in row)
    primes = _descending_primes(top if nonnegative else 2 * top)
    values, modulus = _crt_vectors(_device_affine_residues(constant, linear, primes), primes)
    if nonnegative:
        exact = tuple(values)
        valid = all(0 <= v <= b for v, b in zip(exact, bounds))
    else:
        half = modulus // 2
        exact = tuple(v - modulus if v > half else v for v in values)
        valid = all(abs(v) <= b for v, b in zip(exact, bounds))
    if not valid:
        raise RuntimeError("CRT reconstruction lies outside its certified coefficient bound")
    return exact


def affine_permanent(constant_matrix, linear_matrix, *, backend: str = "auto", parallel: bool = True,
                     num_threads: int = 0) -> tuple:
    """Exact integer coefficients of per(C + x L), ascending degree,
    including trailing zeroes; the empty permanent is (1,)
    (perm_polynomial.py:1183-1214)."""
    _validate_parallel(parallel)
    _validate_num_threads(num_threads)
    _select_backend(backend)
    constant, linear = _validate_affine_matrices(constant_matrix, linear_matrix)
    return _affine_permanent_validated(constant, linear)


def _identity(order: int):
    return tuple(tuple(int(r == c) for c in range(order)) for r in range(order))


def principal_permanent_polynomial_mod(matrix, modulus, *, backend: str = "auto", parallel: bool = True,
                                       num_threads: int = 0) -> tuple:
    """Entry k: the sum of per(matrix[S, S]) over k-subsets S, modulo an odd
    modulus, i.e. the coefficients of per(I + x matrix)
    (perm_polynomial.py:1217-1250)."""
    modulus = _validate_modulus(modulus)
    _validate_parallel(parallel)
    _validate_num_threads(num_threads)
    _select_backend(backend)
    rows = _snapshot_integer_matrix(matrix, "matrix")
    return _affine_permanent_mod_validated(_identity(len(rows)), rows, modulus)


def principal_permanent_polynomial(matrix, *, backend: str = "auto", parallel: bool = True,
                                   num_threads: int = 0) -> tuple:
    """Exact principal-permanent polynomial, the coefficients of
    per(I + x matrix) (perm_polynomial.py:1253-1285)."""
    _validate_parallel(parallel)
    _validate_num_threads(num_threads)
    _select_backend(backend)
    rows = _snapshot_integer_matrix(matrix, "matrix")
    return _affine_permanent_validated(_identity(len(rows)), rows)


# ---------------------------------------------------------------------------
# Rook polynomials: [x^k] per(J + x A) = (n - k)! r_k.


def _rook_bounds(rows) -> tuple:
    """Bounds of r_k: the affine bounds of (J, A) divided by (n - k)!
    (the reference's weighted bound, perm_polynomial.py:862-880)."""
    n = len(rows)
    ones = ((1,) * n,) * n
    return tuple(b // factorial(n - k) for k, b in enumerate(_affine_bounds(ones, rows)))


# Fixed-point scale (bits) of the partial-matching bounds below.
_ROOK_SCALE_BITS = 32


def _ceil_scaled_power_root(base: int, power: int, degree: int, shift: int) -> int:
    """ceil(2^shift * base^(power/degree)), exactly (an integer root of an
    exact integer)."""
    return _prep.ceil_nth_root(base ** power << (shift * degree), degree)


def _matching_bounds(rows) -> tuple:
    """Upper bounds of the k-matching counts r_k of a 0/1 matrix from
    Bregman-Minc (the reference's partial-matching bound,
    perm_polynomial.py:643-692, restated).  For a set S of k rows, append
    n - k all-one rows: every k-matching of S extends to exactly (n - k)!
    permutations of that n x n matrix, whose permanent is at most
    prod_{i in S} (d_i!)^(1/d_i) (n!)^((n-k)/n).  Summing over S gives
    r_k <= e_k(w) (n!)^((n-k)/n) / (n-k)! with w_i = (d_i!)^(1/d_i) and e_k
    the elementary symmetric sum; columns give a second bound.  Every
    factor is rounded up at a 2^32 fixed-point scale, so the floors are
    rigorous."""
    n = len(rows)
    shift = _ROOK_SCALE_BITS
    roots = [0] + [_ceil_scaled_power_root(factorial(d), 1, d, shift) for d in range(1, n + 1)]

    def symmetric(weights):
        e = [1] + [0] * n
        for count, w in enumerate(weights, start=1):
            for k in range(count, 0, -1):
                e[k] += e[k - 1] * w
        return e

    e_rows = symmetric([roots[sum(row)] for row in rows])
    e_cols = symmetric([roots[sum(rows[i][j] for i in range(n))] for j in range(n)])
    n_fact = factorial(n)
    out = [1]
    for k in range(1, n + 1):
        dummy = -(-_ceil_scaled_power_root(n_fact, n - k, n, shift) // factorial(n - k))
        out.append(min(e_rows[k], e_cols[k]) * dummy >> (shift * (k + 1)))
    return tuple(out)


def _rook_parities(rows) -> tuple:
    """r_k mod 2 for every k, without a Glynn walk: over GF(2),
    det(I + x A A^T) = sum_k x^k sum_{|R| = |C| = k} det(A[R, C])^2 (Cauchy-
    Binet), and det = per mod 2, so its coefficients are the matching
    counts' parities (the reference's certificate, perm_polynomial.py:
    700-760, by another route: Hessenberg reduction of A A^T mod 2 by
    similarity, then the characteristic-polynomial recurrence of a
    Hessenberg matrix, polynomials as bit masks).  For a weighted matrix
    pass the entries mod 2."""
    n = len(rows)
    masks = [sum(1 << j for j, v in enumerate(row) if v & 1) for row in rows]
    H = [[bin(masks[i] & masks[j]).count("1") & 1 for j in range(n)] for i in range(n)]
    for k in range(n - 2):
        pivot = next((i for i in range(k + 1, n) if H[i][k]), None)
        if pivot is None:
            continue
        if pivot != k + 1:
            H[k + 1], H[pivot] = H[pivot], H[k + 1]
            for row in H:
                row[k + 1], row[pivot] = row[pivot], row[k + 1]
        for r in range(k + 2, n):
            if H[r][k]:
                H[r] = [a ^ b for a, b in zip(H[r], H[k + 1])]
                for row in H:
                    row[k + 1] ^= row[r]
    # p_m: characteristic polynomial of the leading m x m block (bit d is
    # the coefficient of t^d); over GF(2) every sign is +.
    p = [1]
    for m in range(1, n + 1):
        acc = (p[m - 1] << 1) ^ (p[m - 1] if H[m - 1][m - 1] else 0)
        chain = 1
        for i in range(m - 1, 0, -1):
            chain &= H[i][i - 1]
            if not chain:
                break
            if H[i - 1][m - 1]:
                acc ^= p[i - 1]
        p.append(acc)
    char = p[n]
    # det(I + x G) = x^n char(1/x): the coefficient of x^k is that of t^(n-k).
    return tuple((char >> (n - k)) & 1 for k in range(n + 1))


_WEIGHTED_ROOK_NATIVE_MAX_ENTRY = (1 << 16) - 1


def _rook_residue_rows(rows, primes, weighted: bool) -> list:
    """r_k modulo each prime: through the rook hooks, or -- for weights
    beyond the weighted hook's entry range -- from per(J + x W) residues
    divided by (n - k)! (perm_polynomial.py:1607-1719 falls back the same
    way)."""
    n = len(rows)
    if weighted and max(max(row) for row in rows) > _WEIGHTED_ROOK_NATIVE_MAX_ENTRY:
        ones = ((1,) * n,) * n
        out = []
        for row, p in zip(_device_affine_residues(ones, rows, primes), primes):
            out.append(tuple(c * pow(factorial(n - k) % p, -1, p) % p for k, c in enumerate(row)))
        return out
    hook = "weighted_rook_polynomial_mod_many" if weighted else "rook_polynomial_mod_many"
    out = []
    for start in range(0, len(primes), _ROOK_NATIVE_MAX_MODULI):
        chunk = primes[start:start + _ROOK_NATIVE_MAX_MODULI]
        got = getattr(_polynomial_ext, hook)(rows, chunk)
        for row, p in zip(got, chunk):
            if len(row) != n + 1 or not all(isinstance(v, int) and 0 <= v < p for v in row):
                raise RuntimeError(
                    "native polynomial-permanent backend returned invalid "
                    + ("weighted-rook" if weighted else "rook") + " residues"
                )
        out.extend(got)
    return out


def _rook_exact(rows, weighted: bool) -> tuple:
    """Exact (weighted) matching counts: bounds, primes whose product
    exceeds half the largest bound, CRT, and the parity certificate to pick
    between the residue and its one possible lift (perm_polynomial.py:
    1324-1409)."""
    n = len(rows)
    bounds = _rook_bounds(rows)
    if not weighted:
        bounds = tuple(min(a, b) for a, b in zip(bounds, _matching_bounds(rows)))
    parities = _rook_parities(rows)
    primes = _descending_primes(max(bounds) // 2, minimum=n)
    values, modulus = _crt_vectors(_rook_residue_rows(rows, primes, weighted), primes)
    # The modulus is odd and above half of every bound: a count is its
    # residue or the residue plus the modulus, whichever has its parity.
    exact = tuple(v if v % 2 == parity else v + modulus for v, parity in zip(values, parities))
    if not all(0 <= v <= b for v, b in zip(exact, bounds)):
        what = "weighted-rook" if weighted else "rook"
        raise RuntimeError(f"CRT reconstruction lies outside its certified {what} bound")
    return exact


def rook_polynomial(matrix, *, backend: str = "auto", parallel: bool = True, num_threads: int = 0) -> tuple:
    """Exact matching counts r_0..r_n of a square zero-one matrix
    (perm_polynomial.py:1288-1409)."""
    _validate_parallel(parallel)
    _validate_num_threads(num_threads)
    _select_backend(backend)
    rows = _snapshot_binary_matrix(matrix)
    n = len(rows)
    if n == 0:
        return (1,)
    if n == 1:
        return (1, rows[0][0])
    if not any(any(row) for row in rows):
        return (1,) + (0,) * n
    return _rook_exact(rows, weighted=False)


def _support_components(rows):
    """Connected components of the bipartite support graph as (row list,
    column list), rows and columns in ascending order."""
    n = len(rows)
    parent = list(range(2 * n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for i in range(n):
        for j in range(n):
            if rows[i][j]:
                a, b = find(i), find(n + j)
                if a != b:
                    parent[a] = b
    groups = {}
    for i in range(n):
        if any(rows[i]):
            groups.setdefault(find(i), ([], []))[0].append(i)
    for j in range(n):
        if any(rows[i][j] for i in range(n)):
            groups.setdefault(find(n + j), ([], []))[1].append(j)
    return list(groups.values())


def _poly_mul(a, b):
    out = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        if x:
            for j, y in enumerate(b):
                out[i + j] += x * y
    return out


def _weighted_component(rows, r_idx, c_idx) -> list:
    """Weighted matching polynomial of one support component, padded with
    zero rows or columns to a square (padding adds no matchings)."""
    if len(r_idx) == 1 or len(c_idx) == 1:
        # One row or one column: a matching takes at most one entry.
        return [1, sum(rows[i][j] for i in r_idx for j in c_idx)]
    size = max(len(r_idx), len(c_idx))
    sub = [[0] * size for _ in range(size)]
    for a, i in enumerate(r_idx):
        for b, j in enumerate(c_idx):
            sub[a][b] = rows[i][j]
    sub = tuple(tuple(r) for r in sub)
    return list(_rook_exact(sub, weighted=True))[: min(len(r_idx), len(c_idx)) + 1]


def weighted_rook_polynomial(matrix, *, backend: str = "auto", parallel: bool = True,
                             num_threads: int = 0) -> tuple:
    """Exact weighted matching sums of a square nonnegative integer matrix:
    entry k sums, over k-matchings, the product of their entries
    (perm_polynomial.py:1512-1606).  Disconnected support components are
    evaluated separately and their polynomials multiplied."""
    _validate_parallel(parallel)
    _validate_num_threads(num_threads)
    _select_backend(backend)
    rows = _snapshot_nonnegative_integer_matrix(matrix)
    n = len(rows)
    if n == 0:
        return (1,)
    if n == 1:
        return (1, rows[0][0])
    result = [1]
    for r_idx, c_idx in _support_components(rows):
        result = _poly_mul(result, _weighted_component(rows, r_idx, c_idx))
    result = result[: n + 1]
    return tuple(result + [0] * (n + 1 - len(result)))
