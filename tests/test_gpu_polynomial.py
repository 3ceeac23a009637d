"""Coefficient-valued Glynn on the device (phip_affine_permanent_residues)
against the oracle, the reference's own compiled extension (when built into
oracle/_ref) and, at orders the oracle cannot reach in seconds, against the
device's plain permanent kernels through evaluation: sum_d c_d t^d must be
per(C + t L).  Every case names its seed, order and modulus."""

import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# One modulus per device arithmetic: composite and prime below 2^31 (lazy
# 32-bit Montgomery), 32-bit, Mersenne 2^61 - 1, pseudo-Mersenne, 64-bit.
MODULI = [3, 9, 3 * 5 * 7 * 11 * 13, 1_000_000_007, 2_147_483_647, 4_294_967_291,
          (1 << 61) - 1, (1 << 61) - 31, (1 << 63) - 25]


def _pair(rng, n, lo, hi):
    C = [[rng.randrange(lo, hi + 1) for _ in range(n)] for _ in range(n)]
    L = [[rng.randrange(lo, hi + 1) for _ in range(n)] for _ in range(n)]
    return C, L


@pytest.mark.parametrize("tier", [8, 12, 16, 20, 24, 28, 32, 36, 40, 48, 56, 64])
def test_affine_register_tiers(native, oracle_c, tier):
    """Every register tier against the oracle at small orders (rows beyond
    the order are identity factors).  Tiers 48, 56 and 64 are taken by the
    moduli below 2^30; the others run their own largest tier."""
    old = native.get_tuning(native.TUNE_AFFINE_MIN_TIER)
    try:
        native.set_tuning(native.TUNE_AFFINE_MIN_TIER, tier)
        for n in (0, 1, 2, 3, 5, 8, 11):
            rng = random.Random(7000 + tier * 31 + n)
            C, L = _pair(rng, n, -(10 ** 15), 10 ** 15)
            got = native.affine_permanent_residues(np.array(C, dtype=np.int64).reshape(n, n),
                                                   np.array(L, dtype=np.int64).reshape(n, n), MODULI)
            for m, row in zip(MODULI, got):
                assert row == oracle_c.affine_mod_p(C, L, m, 4), (tier, n, m)
    finally:
        native.set_tuning(native.TUNE_AFFINE_MIN_TIER, old)


@pytest.mark.parametrize("n", [13, 17, 20])
def test_affine_against_oracle(native, oracle_c, n):
    rng = random.Random(7100 + n)
    C, L = _pair(rng, n, -50, 50)
    got = native.affine_permanent_residues(np.array(C, dtype=np.int64), np.array(L, dtype=np.int64), MODULI)
    for m, row in zip(MODULI, got):
        assert row == oracle_c.affine_mod_p(C, L, m, 16), (n, m)


@pytest.mark.parametrize("n", [24, 28])
def test_affine_evaluates_to_plain_permanents(native, n):
    """At larger orders: the coefficient vector evaluated at t equals the
    plain permanent of C + t L from the integer kernels."""
    rng = random.Random(7200 + n)
    C, L = _pair(rng, n, -3, 3)
    primes = [2_147_483_647, 4_294_967_291, (1 << 61) - 1]
    rows = native.affine_permanent_residues(np.array(C, dtype=np.int64), np.array(L, dtype=np.int64), primes)
    for t in (1, -2, 5):
        M = np.array([[C[i][j] + t * L[i][j] for j in range(n)] for i in range(n)], dtype=np.int64)
        plain = native.permanent_residues(M, primes)
        for p, row, want in zip(primes, rows, plain):
            assert sum(c * pow(t, d, p) for d, c in enumerate(row)) % p == want, (n, t, p)


@pytest.mark.parametrize("n,moduli", [(33, [1_000_000_007, 2_147_483_647, 4_294_967_291]),
                                      (36, [1_073_741_789])])
def test_affine_above_order_32_evaluates_to_plain_permanents(native, n, moduli):
    """Orders 33-40 run the register tiers 36 and 40 (32-bit residues): the
    coefficient vector at t equals the plain permanent of C + t L from the
    integer kernels, for several t."""
    rng = random.Random(7250 + n)
    C, L = _pair(rng, n, -2, 2)
    rows = native.affine_permanent_residues(np.array(C, dtype=np.int64), np.array(L, dtype=np.int64), moduli)
    for t in (1, -1, 3):
        M = np.array([[C[i][j] + t * L[i][j] for j in range(n)] for i in range(n)], dtype=np.int64)
        plain = native.permanent_residues(M, moduli)
        for p, row, want in zip(moduli, rows, plain):
            assert len(row) == n + 1
            assert sum(c * pow(t, d, p) for d, c in enumerate(row)) % p == want, (n, t, p)


def test_affine_wide_moduli_above_order_32(native):
    """Moduli of 2^32 and above stop at order 32 on the device; above it
    the mirror and the seam object reduce the exact coefficients."""
    import perm_polynomial as pp
    from permhip.backends import perm_polynomial_ext

    n = 33
    rng = random.Random(7260)
    A = [[rng.randrange(2) for _ in range(n)] for _ in range(n)]
    J = [[1] * n for _ in range(n)]
    m = (1 << 61) - 1
    exact = pp.affine_permanent(J, A)
    assert pp.affine_permanent_mod(J, A, m) == tuple(v % m for v in exact)
    assert tuple(perm_polynomial_ext.affine_permanent_mod(J, A, m)) == tuple(v % m for v in exact)


def test_affine_largest_tier_at_order_24(native):
    """The tier-64 kernel (its excess values in AGPRs) at order 24 gives
    the default tier's coefficient vectors, for moduli below 2^30."""
    rng = random.Random(7290)
    n = 24
    C, L = _pair(rng, n, -5, 5)
    C = np.array(C, dtype=np.int64)
    L = np.array(L, dtype=np.int64)
    moduli = [1_000_003, 536_870_909, 1_073_741_789]
    want = native.affine_permanent_residues(C, L, moduli)
    old = native.get_tuning(native.TUNE_AFFINE_MIN_TIER)
    try:
        native.set_tuning(native.TUNE_AFFINE_MIN_TIER, 64)
        print("affine tier 64 n=24", flush=True)
        assert native.affine_permanent_residues(C, L, moduli) == want
    finally:
        native.set_tuning(native.TUNE_AFFINE_MIN_TIER, old)


def test_affine_per_modulus_stride(native, oracle_c):
    """prime_stride = n*n: one pre-reduced pair per modulus."""
    rng = random.Random(7300)
    n = 9
    moduli = [101, 1_000_000_007, (1 << 61) - 1]
    pairs = [_pair(rng, n, 0, m - 1) for m in moduli]
    C = np.array([p[0] for p in pairs], dtype=np.int64)
    L = np.array([p[1] for p in pairs], dtype=np.int64)
    got = native.affine_permanent_residues(C, L, moduli)
    for (c, l), m, row in zip(pairs, moduli, got):
        assert row == oracle_c.affine_mod_p(c, l, m, 4), m


def test_affine_rejections(native):
    A = np.eye(3, dtype=np.int64)
    with pytest.raises(RuntimeError, match="odd"):
        native.affine_permanent_residues(A, A, [4])
    with pytest.raises(RuntimeError, match="2\\^30"):
        big = np.zeros((41, 41), dtype=np.int64)
        native.affine_permanent_residues(big, big, [2_147_483_647])
    with pytest.raises(RuntimeError, match="63"):
        big = np.zeros((64, 64), dtype=np.int64)
        native.affine_permanent_residues(big, big, [101])
    with pytest.raises(RuntimeError, match="2\\^32"):
        big = np.zeros((33, 33), dtype=np.int64)
        native.affine_permanent_residues(big, big, [(1 << 61) - 1])
    assert native.affine_permanent_residues(A, A, [101]) == [(1, 3, 3, 1)]


def test_against_reference_extension(native, oracle_c):
    """The reference's own perm_polynomial_ext, compiled from its source into
    oracle/_ref, gives the same coefficient vectors."""
    import oracle

    ext = oracle.reference_extensions().get("perm_polynomial_ext")
    if ext is None:
        pytest.skip("reference polynomial extension not built (oracle/_ref)")
    rng = random.Random(7400)
    for n in (2, 4, 7, 10, 14):
        C, L = _pair(rng, n, -(10 ** 9), 10 ** 9)
        got = native.affine_permanent_residues(np.array(C, dtype=np.int64), np.array(L, dtype=np.int64), MODULI)
        for m, row in zip(MODULI, got):
            assert row == tuple(ext.affine_permanent_mod(C, L, m, 4)), (n, m)


def test_exact_public_functions(native, oracle_c):
    import math

    import oracle
    import perm_polynomial as pp

    rng = random.Random(7500)
    n = 12
    A = [[rng.randrange(2) for _ in range(n)] for _ in range(n)]
    full = [[1] * n for _ in range(n)]
    assert pp.rook_polynomial(full) == tuple(math.comb(n, k) ** 2 * math.factorial(k) for k in range(n + 1))
    rook = pp.rook_polynomial(A)
    # r_n is the permanent of A; r_1 the number of ones.
    assert rook[n] == oracle.subset_permanent(A)
    assert rook[1] == sum(map(sum, A)) and rook[0] == 1
    # Principal polynomial: constant 1, linear coefficient the trace, top the permanent.
    M = [[rng.randrange(-20, 21) for _ in range(n)] for _ in range(n)]
    pr = pp.principal_permanent_polynomial(M)
    assert pr[0] == 1 and pr[1] == sum(M[i][i] for i in range(n))
    assert pr[n] == oracle.subset_permanent(M)
    # affine_permanent at x = 0 and the top degree: per(C) and per(L).
    C, L = _pair(rng, n, -30, 30)
    coeffs = pp.affine_permanent(C, L)
    assert coeffs[0] == oracle.subset_permanent(C) and coeffs[n] == oracle.subset_permanent(L)
    assert all(abs(v) <= b for v, b in zip(coeffs, pp.affine_permanent_bounds(C, L)))
    W = [[rng.randrange(0, 70000) for _ in range(n)] for _ in range(n)]
    wr = pp.weighted_rook_polynomial(W)
    assert wr[1] == sum(map(sum, W)) and wr[n] == oracle.subset_permanent(W)


def test_python_backend_name_runs_on_the_device():
    """backend='python' (the reference's portable backend) is accepted and
    gives the native coefficients: the rook polynomial of J_9 is
    C(9, k)^2 k!, and a random board's equals the 'auto' result."""
    import math

    import numpy as np
    import perm_polynomial as poly

    J = [[1] * 9 for _ in range(9)]
    want = tuple(math.comb(9, k) ** 2 * math.factorial(k) for k in range(10))
    assert poly.rook_polynomial(J, backend="python") == want
    board = np.random.default_rng(12).integers(0, 2, (11, 11)).tolist()
    assert poly.rook_polynomial(board, backend="python") == poly.rook_polynomial(board, backend="auto")
