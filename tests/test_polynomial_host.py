"""Host logic of the polynomial-permanent mirror on the CPU: validation,
bounds, prime plans, CRT and the rook / component reductions, with the
device's affine residues replaced by the oracle's coefficient-valued Glynn
(the reference tests its native seams with fakes the same way,
tests/test_polynomial_native.py).  Known answers are the reference's own
(tests/test_polynomial_permanent.py:159-412)."""

import itertools
import math
import random

import numpy as np
import pytest

import oracle


def _fake_affine(constant, linear, moduli):
    c = np.asarray(constant, dtype=np.int64)
    l = np.asarray(linear, dtype=np.int64)
    out = []
    for i, m in enumerate(moduli):
        ci = c if c.ndim == 2 else c[i]
        li = l if l.ndim == 2 else l[i]
        out.append(oracle.C.affine_mod_p(ci.tolist(), li.tolist(), m, 2))
    return out


@pytest.fixture
def poly(monkeypatch):
    from permhip import _native

    monkeypatch.setattr(_native, "require_device", lambda: None)
    monkeypatch.setattr(_native, "affine_permanent_residues", _fake_affine)
    import perm_polynomial

    return perm_polynomial


def _poly_mul(a, b):
    out = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


def definition_affine(constant, linear):
    """Sum over permutations of prod_i (C[i][s(i)] + x L[i][s(i)])."""
    n = len(constant)
    total = [0] * (n + 1)
    for perm in itertools.permutations(range(n)):
        term = [1]
        for i, j in enumerate(perm):
            term = _poly_mul(term, [constant[i][j], linear[i][j]])
        for d, v in enumerate(term):
            total[d] += v
    return tuple(total)


def test_oracle_affine_matches_definition():
    """The oracle's coefficient Glynn against the permutation expansion,
    including composite and wide odd moduli."""
    rng = random.Random(11)
    for n in range(0, 7):
        C = [[rng.randrange(-9, 10) for _ in range(n)] for _ in range(n)]
        L = [[rng.randrange(-9, 10) for _ in range(n)] for _ in range(n)]
        exact = definition_affine(C, L) if n else (1,)
        for m in (3, 9, 101, 2_147_483_647, (1 << 61) - 1, (1 << 63) - 25):
            assert oracle.C.affine_mod_p(C, L, m, 2) == tuple(v % m for v in exact), (n, m)


def test_oracle_affine_evaluates_to_plain_permanents():
    """sum_d coeff_d t^d = per(C + t L) for several t: a dropped or shifted
    coefficient fails at one of the points."""
    rng = random.Random(12)
    n, p = 9, 1_000_000_007
    C = [[rng.randrange(-50, 51) for _ in range(n)] for _ in range(n)]
    L = [[rng.randrange(-50, 51) for _ in range(n)] for _ in range(n)]
    coeffs = oracle.C.affine_mod_p(C, L, p, 2)
    for t in (0, 1, 2, -3, 17, 123456):
        M = [[(C[i][j] + t * L[i][j]) % p for j in range(n)] for i in range(n)]
        assert sum(c * pow(t, d, p) for d, c in enumerate(coeffs)) % p == oracle.C.permanent_mod_p(M, p, 2), t


def test_known_cases(poly):
    assert poly.affine_permanent([[3, 4], [5, 6]], [[0, 0], [0, 0]]) == (38, 0, 0)
    assert poly.principal_permanent_polynomial([]) == (1,)
    assert poly.principal_permanent_polynomial([[7]]) == (1, 7)
    assert poly.principal_permanent_polynomial([[2, -3], [5, 7]]) == (1, 9, -1)
    assert poly.principal_permanent_polynomial([[2, 0, 0], [0, 3, 0], [0, 0, 5]]) == (1, 10, 31, 30)
    assert poly.principal_permanent_polynomial_mod([], 9) == (1,)
    assert poly.principal_permanent_polynomial_mod([[14]], 9) == (1, 5)
    assert poly.principal_permanent_polynomial_mod([[2, -3], [5, 7]], 101) == (1, 9, 100)
    assert poly.principal_permanent_polynomial_mod([[2, 0, 0], [0, 3, 0], [0, 0, 5]], 25) == (1, 10, 6, 5)
    assert poly.rook_polynomial([]) == (1,)
    assert poly.rook_polynomial([[0, 0, 0]] * 3) == (1, 0, 0, 0)
    assert poly.rook_polynomial([[1, 0, 0], [0, 1, 0], [0, 0, 1]]) == (1, 3, 3, 1)
    assert poly.rook_polynomial([[1, 1, 1, 1]] * 4) == (1, 16, 72, 96, 24)
    assert poly.rook_polynomial([[1, 1, 1]] * 3) == (1, 9, 18, 6)
    assert poly.weighted_rook_polynomial([]) == (1,)
    assert poly.weighted_rook_polynomial([[0, 0, 0]] * 3) == (1, 0, 0, 0)
    assert poly.weighted_rook_polynomial([[2, 0, 0], [0, 3, 0], [0, 0, 5]]) == (1, 10, 31, 30)
    assert poly.weighted_rook_polynomial([[2, 2, 2]] * 3) == (1, 18, 72, 48)
    huge = 10 ** 100 + 7
    assert poly.weighted_rook_polynomial([[huge]]) == (1, huge)


def test_zero_matrices_short_circuit(poly, monkeypatch):
    from permhip import _native

    def fail(*args, **kwargs):
        raise AssertionError("zero matrix should not reach the device")

    monkeypatch.setattr(_native, "affine_permanent_residues", fail)
    assert poly.rook_polynomial([[0] * 63 for _ in range(63)]) == (1,) + (0,) * 63
    assert poly.weighted_rook_polynomial([[0] * 63 for _ in range(63)]) == (1,) + (0,) * 63


def test_exact_affine_against_definition(poly):
    rng = random.Random(13)
    for n in range(2, 7):
        for lo in (-(10 ** 12), 0):
            C = [[rng.randrange(lo, 10 ** 12) for _ in range(n)] for _ in range(n)]
            L = [[rng.randrange(lo, 10 ** 12) for _ in range(n)] for _ in range(n)]
            assert poly.affine_permanent(C, L) == definition_affine(C, L), (n, lo)
    # Entries beyond int64 are reduced per prime on the host.
    C = [[(1 << 80) + 3, 2], [5, -(1 << 70)]]
    L = [[1, 1 << 65], [7, 3]]
    assert poly.affine_permanent(C, L) == definition_affine(C, L)


def test_rook_and_principal_closed_forms(poly):
    for n in (2, 5, 7):
        full = [[1] * n for _ in range(n)]
        assert poly.rook_polynomial(full) == tuple(math.comb(n, k) ** 2 * math.factorial(k) for k in range(n + 1))
        assert poly.principal_permanent_polynomial(full) == tuple(
            math.comb(n, k) * math.factorial(k) for k in range(n + 1))
        eye = [[int(i == j) for j in range(n)] for i in range(n)]
        assert poly.rook_polynomial(eye) == tuple(math.comb(n, k) for k in range(n + 1))


def brute_rook(W):
    n = len(W)
    out = [0] * (n + 1)
    for k in range(n + 1):
        for rows in itertools.combinations(range(n), k):
            for cols in itertools.permutations(range(n), k):
                prod = 1
                for r, c in zip(rows, cols):
                    prod *= W[r][c]
                out[k] += prod
    return tuple(out)


def test_weighted_rook_components_and_wide_weights(poly):
    rng = random.Random(14)
    # Block-diagonal support (two components plus an isolated edge), weights
    # past the weighted hook's 16-bit range in one block.
    W = [[0] * 7 for _ in range(7)]
    for i in range(3):
        for j in range(3):
            W[i][j] = rng.randrange(0, 100)
    for i in range(3, 6):
        for j in range(3, 6):
            W[i][j] = rng.randrange(0, 10 ** 9)
    W[6][6] = 12345
    assert poly.weighted_rook_polynomial(W) == brute_rook(W)
    # A permuted diagonal: every component is one edge.
    P = [[0] * 6 for _ in range(6)]
    weights = [3, 1, 4, 1, 5, 9]
    for i, w in enumerate(weights):
        P[i][(i * 5 + 2) % 6] = w
    expected = [1]
    for w in weights:
        expected = _poly_mul(expected, [1, w])
    assert poly.weighted_rook_polynomial(P) == tuple(expected)
    # Rectangular components (rows share columns unevenly).
    R = [[1, 2, 0, 0], [0, 3, 0, 0], [4, 0, 0, 0], [0, 0, 5, 6]]
    assert poly.weighted_rook_polynomial(R) == brute_rook(R)


def test_bounds(poly):
    assert poly.affine_permanent_bounds([], []) == (1,)
    C = [[1, -2], [3, 4]]
    L = [[0, 5], [-6, 0]]
    b = poly.affine_permanent_bounds(C, L)
    exact = definition_affine(C, L)
    assert len(b) == 3 and all(abs(v) <= w for v, w in zip(exact, b))
    # Row relaxation (1+2 + 5x)(3+4 + 6x) = 21 + 53x + 30x^2; column (4 + 6x)(6 + 5x) = 24 + 56x + 30x^2;
    # 2! times the products of maxima: 2 (2 + 5x)(4 + 6x) = 16 + 64x + 60x^2, columns 2 (3 + 6x)(4 + 5x).
    assert b == (16, 53, 30)


@pytest.mark.parametrize("bad,exc,msg", [
    (2, ValueError, "odd integer greater than 1"),
    (1, ValueError, "odd integer greater than 1"),
    ((1 << 63) + 1, ValueError, "positive signed 64-bit"),
    (True, TypeError, "modulus must be an integer"),
    (3.0, TypeError, "modulus must be an integer"),
])
def test_modulus_validation(poly, bad, exc, msg):
    with pytest.raises(exc, match=msg):
        poly.affine_permanent_mod([[1]], [[1]], bad)


def test_matrix_and_option_validation(poly):
    with pytest.raises(ValueError, match="same shape"):
        poly.affine_permanent([[1]], [[1, 2], [3, 4]])
    with pytest.raises(ValueError, match="must be square"):
        poly.affine_permanent([[1, 2]], [[1, 2]])
    with pytest.raises(TypeError, match="entries must be integers"):
        poly.affine_permanent([[1.5]], [[1]])
    with pytest.raises(ValueError, match="less than 64"):
        poly.rook_polynomial([[0] * 64 for _ in range(64)])
    with pytest.raises(ValueError, match="0 or 1"):
        poly.rook_polynomial([[2]])
    with pytest.raises(ValueError, match="nonnegative"):
        poly.weighted_rook_polynomial([[-1]])
    with pytest.raises(TypeError, match="parallel must be a boolean"):
        poly.rook_polynomial([[1]], parallel=1)
    with pytest.raises(ValueError, match="num_threads must be non-negative"):
        poly.rook_polynomial([[1]], num_threads=-1)
    with pytest.raises(ValueError, match="backend must be one of"):
        poly.rook_polynomial([[1]], backend="gpu")
    # The portable backend's name is accepted and computes like the native
    # one (a one-cell board's rook polynomial is closed form on the host).
    assert poly.rook_polynomial([[1]], backend="python") == poly.rook_polynomial([[1]]) == (1, 1)
    caps = poly.polynomial_backend_capabilities()
    assert caps["default_backend"] == "native" and "affine_permanent_mod" in caps["native_hooks"]
    assert caps["available_backends"] == ("native", "python")


def test_native_hooks_validate_like_the_extension(poly):
    ext = poly._polynomial_ext
    with pytest.raises(ValueError, match="odd signed 64-bit integer greater than 1"):
        ext.affine_permanent_mod([[1]], [[1]], 4)
    with pytest.raises(ValueError, match="same order"):
        ext.affine_permanent_mod([[1]], [[1, 2], [3, 4]], 7)
    with pytest.raises(ValueError, match="odd primes greater than the matrix order"):
        ext.rook_polynomial_mod_many([[1, 0], [0, 1]], [9])
    with pytest.raises(ValueError, match="between 0 and 65535"):
        ext.weighted_rook_polynomial_mod_many([[70000]], [101])
    assert ext.rook_polynomial_mod_many([[1, 1], [1, 1]], [101, 103]) == ((1, 4, 2), (1, 4, 2))


def test_device_order_routing(poly):
    """Which (order, modulus) pairs the device kernel takes itself: every
    order below 64 for moduli below 2^30, up to 40 below 2^32, up to 32
    above; other pairs reduce the exact coefficients."""
    takes = poly._device_takes
    assert takes(63, (1 << 30) - 35) and takes(41, 101)
    assert takes(40, (1 << 31) - 1) and not takes(41, (1 << 31) - 1)
    assert takes(40, (1 << 32) - 5) and not takes(41, (1 << 32) - 5)
    assert takes(32, (1 << 61) - 1) and not takes(33, (1 << 61) - 1)


def _matching_sums(rows):
    """Weighted k-matching sums by a row-by-row DP over used-column masks
    (plain counts for a 0/1 matrix): independent of the bounds, the
    GF(2) determinant and the device."""
    from collections import defaultdict

    n = len(rows)
    states = defaultdict(int)
    states[(0, 0)] = 1
    for i in range(n):
        nxt = defaultdict(int)
        for (mask, k), c in states.items():
            nxt[(mask, k)] += c
            for j in range(n):
                if rows[i][j] and not (mask >> j) & 1:
                    nxt[(mask | 1 << j, k + 1)] += c * rows[i][j]
        states = nxt
    out = [0] * (n + 1)
    for (_, k), c in states.items():
        out[k] += c
    return out


def test_rook_parity_certificate_and_matching_bounds(poly):
    """The GF(2) determinant gives every matching count's parity, and the
    partial-matching (Bregman-Minc) bounds hold, on random boards of orders
    1-9 (weighted boards through their entries mod 2); the closed forms of
    J_n (C(n,k)^2 k!) and I_n (C(n,k)) at order 24."""
    import math
    import random

    rng = random.Random(4401)
    for _ in range(200):
        n = rng.randint(1, 9)
        density = rng.random()
        A = tuple(tuple(int(rng.random() < density) for _ in range(n)) for _ in range(n))
        counts = _matching_sums(A)
        assert poly._rook_parities(A) == tuple(c % 2 for c in counts), A
        assert all(b >= c for b, c in zip(poly._matching_bounds(A), counts)), A
        W = tuple(tuple(rng.randrange(0, 5) for _ in range(n)) for _ in range(n))
        assert poly._rook_parities(W) == tuple(c % 2 for c in _matching_sums(W)), W
    n = 24
    J = tuple((1,) * n for _ in range(n))
    I = tuple(tuple(int(i == j) for j in range(n)) for i in range(n))
    rJ = [math.comb(n, k) ** 2 * math.factorial(k) for k in range(n + 1)]
    rI = [math.comb(n, k) for k in range(n + 1)]
    assert poly._rook_parities(J) == tuple(v % 2 for v in rJ)
    assert poly._rook_parities(I) == tuple(v % 2 for v in rI)
    assert all(b >= v for b, v in zip(poly._matching_bounds(J), rJ))
    assert all(b >= v for b, v in zip(poly._matching_bounds(I), rI))
