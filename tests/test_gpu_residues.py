"""Device residues through the C ABI against the oracle (bit-exact).

Every case names its seed, order and moduli so a failure identifies itself.
"""

import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INT64_MIN = -(1 << 63)
INT64_MAX = (1 << 63) - 1

PRIMES_31 = [1073741827, 1073741831, 1073741833, 2147483647, 2147483629]
ODD_SMALL = [3, 5, 7, 101, 65537]
ODD_COMPOSITE = [9, 15, 45, 3 * 5 * 7 * 11 * 13 * 17]
EVEN = [2, 4, 6, 12, 1 << 20, (1 << 40) * 3, 1 << 62]
WIDE = [(1 << 61) - 1, 3_037_000_499, 3_037_000_501, INT64_MAX - 24, INT64_MAX]
ALL_MODULI = PRIMES_31 + ODD_SMALL + ODD_COMPOSITE + EVEN + WIDE + [1]


def _random_matrix(rng, n, lo, hi):
    return [[rng.randrange(lo, hi + 1) for _ in range(n)] for _ in range(n)]


@pytest.fixture(params=["default", "modular"])
def engine(request, native):
    """Runs a test twice: with the default planning (tables with enough
    primes take the exact-accumulation walk) and with the exact variant off,
    so the modular prime-set kernels stay covered."""
    if request.param == "default":
        yield request.param
        return
    old = native.get_tuning(native.TUNE_EXACT_MIN_PRIMES)
    native.set_tuning(native.TUNE_EXACT_MIN_PRIMES, 0)
    try:
        yield request.param
    finally:
        native.set_tuning(native.TUNE_EXACT_MIN_PRIMES, old)


@pytest.mark.parametrize("n", list(range(0, 15)))
def test_all_moduli_small_orders(native, oracle_c, n):
    rng = random.Random(20261203 + n)
    A = _random_matrix(rng, n, -10, 10)
    got = native.permanent_residues(np.array(A, dtype=np.int64).reshape(n, n), ALL_MODULI)
    for m, r in zip(ALL_MODULI, got):
        assert r == oracle_c.permanent_mod_p(A, m, 4), (n, m)


@pytest.mark.parametrize(
    "n,lo,hi",
    [
        (16, -1, 1),      # small row bounds: groups of 6
        (17, -10, 10),    # groups of 4
        (18, -100, 100),  # groups of 3
        (14, -1000, 1000),  # groups of 2
        (23, -3, 3),      # groups of 6, odd order (last group partial)
        (19, -(10 ** 6), 10 ** 6),  # wide rows
        (20, -10, 10),
        (21, 0, 1),
        (22, -50, 50),
    ],
)
def test_row_sum_representations(native, oracle_c, n, lo, hi):
    rng = random.Random(1000 + n)
    A = _random_matrix(rng, n, lo, hi)
    moduli = PRIMES_31 + [101, 9]
    got = native.permanent_residues(np.array(A, dtype=np.int64), moduli)
    for m, r in zip(moduli, got):
        assert r == oracle_c.permanent_mod_p(A, m, 8), (n, lo, hi, m)


@pytest.mark.parametrize("n", [23, 24, 25])
def test_larger_orders_against_openmp_oracle(native, oracle_c, n):
    rng = random.Random(77 + n)
    A = _random_matrix(rng, n, -10, 10)
    moduli = [1073741827, 2147483647]
    got = native.permanent_residues(np.array(A, dtype=np.int64), moduli)
    for m, r in zip(moduli, got):
        assert r == oracle_c.permanent_mod_p(A, m, 8), (n, m)


def _primes_from(start, count):
    from permhip import planner

    out, q = [], start | 1
    while len(out) < count:
        if planner.is_prime(q):
            out.append(q)
        q += 2
    return out


@pytest.mark.parametrize("k", [2, 3, 4, 5, 6, 7, 8, 9, 12])
@pytest.mark.parametrize(
    "n,lo,hi",
    [
        (14, -10, 10),    # groups of 4
        (17, -3, 3),      # groups of 6, partial last group
        (16, 0, 1),       # groups of 6
        (13, -100, 100),  # groups of 3
        (12, -1000, 1000),  # groups of 2
        (5, -2, 2),       # a single group
        (7, -10, 10),
    ],
)
def test_prime_sets_against_oracle(native, oracle_c, engine, n, lo, hi, k):
    """Primes whose centred tables coincide share one Gray walk (sets of 2, 4
    or 8, padded); every residue must equal the oracle's."""
    rng = random.Random(4000 + 31 * n + k)
    A = _random_matrix(rng, n, lo, hi)
    primes = _primes_from(1 << 30, k)
    got = native.permanent_residues(np.array(A, dtype=np.int64), primes)
    for p, r in zip(primes, got):
        assert r == oracle_c.permanent_mod_p(A, p, 8), (n, lo, hi, k, p)


def test_prime_sets_mixed_tables(native, oracle_c, engine):
    """Moduli below twice the largest entry get their own tables; the rest
    share one.  Order of the answers follows the caller's modulus order."""
    rng = random.Random(77)
    n = 15
    A = _random_matrix(rng, n, -60, 60)
    big = _primes_from(1 << 30, 6)
    moduli = [big[0], 101, big[1], 103, big[2], 9, big[3], 2147483647, big[4], 65537, big[5]]
    got = native.permanent_residues(np.array(A, dtype=np.int64), moduli)
    for m, r in zip(moduli, got):
        assert r == oracle_c.permanent_mod_p(A, m, 8), m


def test_int64_extremes(native, oracle_c):
    A = [[INT64_MIN, INT64_MAX, -1, 5, 7], [INT64_MAX, INT64_MIN, 1, 0, 3],
         [0, -1, INT64_MAX, 2, 2], [1, 1, 1, 1, 1], [INT64_MIN, 0, 0, 0, 9]]
    got = native.permanent_residues(np.array(A, dtype=np.int64), ALL_MODULI)
    for m, r in zip(ALL_MODULI, got):
        assert r == oracle_c.permanent_mod_p(A, m, 2), m


def test_per_modulus_matrices_stride(native, oracle_c):
    """k x n x n input: each modulus gets its own (already reduced) matrix."""
    rng = random.Random(5)
    n = 12
    moduli = [1073741827, 101, 2147483647]
    mats = [_random_matrix(rng, n, 0, m - 1) for m in moduli]
    got = native.permanent_residues(np.array(mats, dtype=np.int64), moduli)
    for M, m, r in zip(mats, moduli, got):
        assert r == oracle_c.permanent_mod_p(M, m, 4)


def test_schur_entry_point(native, oracle_c):
    from permhip import planner

    for n, p in [(3, 7), (5, 11), (7, 29), (11, 23), (13, 53), (9, 19), (15, 31), (17, 103)]:
        omega = planner.nth_root_of_unity_mod_p(n, p)
        S = planner.build_schur_matrix_mod_p(n, omega, p)
        assert native.schur_permanent_modp(n, omega, p) == oracle_c.permanent_mod_p(S, p, 2), (n, p)
    # omega that is not an n-th root of unity takes the full-matrix path.
    S = planner.build_schur_matrix_mod_p(6, 5, 101)
    assert native.schur_permanent_modp(6, 5, 101) == oracle_c.permanent_mod_p(S, 101, 2)


@pytest.mark.parametrize("n", [2, 3, 5, 8, 13, 20, 24])
def test_binary_residues(native, oracle_c, n):
    rng = np.random.default_rng(20260815 + n)
    B = rng.integers(0, 2, size=(n, n), dtype=np.uint8)
    primes = [101, 2147483647, 2147483629, 2147483659, 9223372036854775783]
    primes = [p for p in primes if p > n]
    got = native.binary_permanent_residues(B, primes)
    for p, r in zip(primes, got):
        assert r == oracle_c.binary_mod_p(B, p, 8), (n, p)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 16, 21, 24, 31, 32, 33, 34])
def test_mod3(native, oracle_c, n):
    """Orders below 5 take the per-state loop, the others the 16-state
    unrolled one (with one and two 32-bit words per row vector)."""
    for seed in range(3):
        rng = np.random.default_rng(100 * seed + 3 + n)
        A = rng.integers(-5, 6, size=(n, n), dtype=np.int64)
        if seed == 2:
            A = (rng.random((n, n)) < 0.5).astype(np.int64)
        assert native.permanent_mod3(A % 3) == oracle_c.mod3(A, 8), (n, seed)


def test_abi_rejections_do_not_poison_state(native):
    import ctypes

    lib = native.load()
    eye = np.eye(5, dtype=np.uint8)
    ptr = eye.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    null = ctypes.POINTER(ctypes.c_uint8)()
    assert lib.phip_binary_permanent_modp(ptr, 5, 9) == -1
    assert lib.phip_binary_permanent_modp(ptr, 5, 5) == -1
    assert lib.phip_binary_permanent_modp(ptr, -1, 101) == -1
    assert lib.phip_binary_permanent_modp(ptr, 64, 101) == -1
    assert lib.phip_binary_permanent_modp(null, 5, 101) == -1
    bad = eye.copy()
    bad[0, 0] = 2
    assert lib.phip_binary_permanent_modp(bad.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), 5, 101) == -1
    assert lib.phip_binary_permanent_modp(null, 0, 3) == 1
    assert lib.phip_binary_permanent_modp(ptr, 5, 101) == 1
    a = np.eye(6, dtype=np.int64)
    iptr = a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    assert lib.phip_permanent_modp(iptr, 6, 0) == -1
    assert lib.phip_permanent_modp(iptr, 64, 101) == -1
    assert lib.phip_permanent_modp(iptr, 6, 101) == 1


# One modulus of each wide arithmetic class and edge: 32-bit Montgomery (the
# reference's 32-bit device primes among them), 61-bit pseudo-Mersenne
# (Mersenne 2^61 - 1 and the largest complement taken) and 64-bit
# Montgomery.
WIDE_CLASSES = [
    2_147_483_659, 4_294_967_291, 4_294_966_153, 4_294_967_161,
    (1 << 61) - 1, (1 << 61) - 31, (1 << 61) - ((1 << 28) - 1),
    (1 << 61) - ((1 << 28) + 1), (1 << 61) + 1, INT64_MAX - 24, INT64_MAX,
]


@pytest.mark.parametrize("tier", [16, 32, 48, 64])
def test_wide_register_tiers(native, oracle_c, tier):
    """Every row-register tier of the wide kernels, forced at small orders."""
    old = native.get_tuning(native.TUNE_WIDE_MIN_TIER)
    try:
        native.set_tuning(native.TUNE_WIDE_MIN_TIER, tier)
        for n in (2, 3, 4, 5, 9, 13, 16):
            rng = random.Random(4000 + 17 * tier + n)
            A = _random_matrix(rng, n, -(10 ** 12), 10 ** 12)
            got = native.permanent_residues(np.array(A, dtype=np.int64), WIDE_CLASSES)
            for m, r in zip(WIDE_CLASSES, got):
                assert r == oracle_c.permanent_mod_p(A, m, 4), (tier, n, m)
    finally:
        native.set_tuning(native.TUNE_WIDE_MIN_TIER, old)


@pytest.mark.parametrize("n", [17, 22])
def test_wide_moduli_against_oracle(native, oracle_c, n):
    rng = random.Random(5000 + n)
    A = _random_matrix(rng, n, -(10 ** 15), 10 ** 15)
    got = native.permanent_residues(np.array(A, dtype=np.int64), WIDE_CLASSES)
    for m, r in zip(WIDE_CLASSES, got):
        assert r == oracle_c.permanent_mod_p(A, m, 16), (n, m)


@pytest.mark.parametrize("n", [27, 33])
def test_wide_moduli_agree_with_exact_permanent(native, n):
    """At orders the oracle cannot reach in seconds: every wide residue is
    the exact permanent (from the 31-bit engine and CRT) reduced."""
    import perm

    rng = random.Random(6000 + n)
    A = _random_matrix(rng, n, -10, 10)
    exact = perm.matrix_permanent(A)
    moduli = [4_294_966_153, (1 << 61) - 1, (1 << 61) - 31, INT64_MAX - 24]
    got = native.permanent_residues(np.array(A, dtype=np.int64), moduli)
    assert got == [exact % m for m in moduli], n


PAIRED_LIMIT = 1_342_177_280  # 1.25 * 2^30: sets below it pair consecutive states


def _primes_below(limit, count):
    from permhip import planner

    out, c = [], (limit - 1) | 1
    while len(out) < count:
        if c < limit and planner.is_prime(c):
            out.append(c)
        c -= 2
    return out


@pytest.mark.parametrize("n,lo,hi", [(16, -1, 1), (20, -10, 10), (23, -100, 100), (21, -3, 3), (25, -10, 10)])
def test_paired_state_sets_against_oracle(native, oracle_c, engine, n, lo, hi):
    """Sets of 8 and 4 primes just below the paired bound (paired kernels)
    and just above it (per-state kernels) agree with the oracle; groups of
    every row-sum representation and odd group counts are covered."""
    rng = random.Random(8000 + n)
    A = _random_matrix(rng, n, lo, hi)
    below = _primes_below(PAIRED_LIMIT, 8)
    above = _primes_from(PAIRED_LIMIT, 8)
    for moduli in (below, above, below[:4], above[:4], below[:4] + above[:4]):
        got = native.permanent_residues(np.array(A, dtype=np.int64), moduli)
        for m, r in zip(moduli, got):
            assert r == oracle_c.permanent_mod_p(A, m, 16), (n, lo, hi, m)


def test_paired_classes_are_planned(native):
    """With the exact variant off, the plan routes a set of 8 primes from
    2^30 to a paired class and a set from 2^31 - 1 down to a per-state
    class; with it on, both run as one exact walk."""
    rng = np.random.default_rng(3)
    A = rng.integers(-10, 11, (20, 20)).astype(np.int64)
    assert set(native.batch_plan_classes([A], [_primes_from(1 << 30, 8)])) == {(4, 8, 6)}
    old = native.get_tuning(native.TUNE_EXACT_MIN_PRIMES)
    native.set_tuning(native.TUNE_EXACT_MIN_PRIMES, 0)
    try:
        _paired_classes(native)
    finally:
        native.set_tuning(native.TUNE_EXACT_MIN_PRIMES, old)


def _paired_classes(native):
    rng = np.random.default_rng(3)
    A = rng.integers(-10, 11, (20, 20)).astype(np.int64)
    low = native.Batch([A], [_primes_from(1 << 30, 8)])
    high = native.Batch([A], [_primes_below((1 << 31) - 1, 8)])
    try:
        assert low.stats()["variant"] == 1.0 and high.stats()["variant"] == 0.0
    finally:
        low.close()
        high.close()
