"""The exact-accumulation kernels (ExactSink: each Gray state's product of
the exact group products formed in radix-2^31 limbs, one walk for every
prime of a table) against the oracle, bit for bit, and against the modular
kernels they replace (PHIP_TUNE_EXACT_MIN_PRIMES = 0): tables of at most
eight groups (full products in 8 limbs; groups of 3 or 4 rows whose
four-group bound products are below 2^122 in signed limbs, variant 6, else
in magnitudes and a sign, variant 3) and of 9..16 groups (products kept
modulo 2^(31 L), L = 10 or 12, exact while the row-bound product is below
2^(31 L)).  Every case names its seed, order, entry range and prime
count."""

import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRIMES = [1073741827, 1073741831, 1073741833, 1073741839, 1073741843, 1073741857, 1073741891, 1073741909,
          1073741939, 1073741953, 1073741969, 1073741971, 2147483647, 2147483629, 2147483587, 2147483579]


def _with_tuning(native, key, value, fn):
    old = native.get_tuning(key)
    native.set_tuning(key, value)
    try:
        return fn()
    finally:
        native.set_tuning(key, old)


def _classes(native, A, primes):
    return set(native.batch_plan_classes([A], [primes]))


# (order, entry bound): row bounds that make the planner pick groups of 6, 4,
# 3 and 2 rows, with at most eight groups (the exact variant's limit).
SHAPES = [(5, 1), (8, 3), (12, 1), (13, 40), (16, 100), (18, 4), (20, 9), (23, 12), (24, 1), (27, 2),
          (28, 6), (30, 1), (31, 3), (32, 10)]


@pytest.mark.parametrize("n,bound", SHAPES)
def test_exact_matches_oracle_and_modular(native, oracle_c, n, bound):
    rng = np.random.default_rng(900 + 31 * n + bound)
    A = rng.integers(-bound, bound + 1, (n, n)).astype(np.int64)
    for k in (6, 8, 10, 16):
        primes = PRIMES[:k]
        assert any(c[2] in (3, 6) for c in _classes(native, A, primes)), (n, bound, k)
        got = native.permanent_residues(A, primes)
        modular = _with_tuning(native, native.TUNE_EXACT_MIN_PRIMES, 0, lambda: native.permanent_residues(A, primes))
        magnitudes = _with_tuning(native, native.TUNE_EXACT_SIGNED, 0, lambda: native.permanent_residues(A, primes))
        assert got == modular == magnitudes, (n, bound, k)
        if n <= 20:
            assert got == [oracle_c.permanent_mod_p(A.tolist(), p, 8) for p in primes], (n, bound, k)


def test_exact_extreme_group_products(native, oracle_c):
    """Rows whose sums stay at their bounds: every group product at the
    largest magnitude the planner admits, with both signs."""
    for seed, n in ((1, 16), (2, 20), (3, 24)):
        rng = random.Random(seed)
        # Rows of equal entries: every state's row sums are +-(n-2)*a ... +-n*a.
        A = np.array([[rng.choice((-1, 1)) * 35 for _ in range(n)] for _ in range(n)], dtype=np.int64)
        primes = PRIMES[:8]
        assert any(c[2] in (3, 6) for c in _classes(native, A, primes))
        got = native.permanent_residues(A, primes)
        if n <= 20:
            assert got == [oracle_c.permanent_mod_p(A.tolist(), p, 8) for p in primes]
        assert got == _with_tuning(native, native.TUNE_EXACT_MIN_PRIMES, 0,
                                   lambda: native.permanent_residues(A, primes))


def test_exact_zero_rows_and_identity(native):
    n = 24
    Z = np.zeros((n, n), dtype=np.int64)
    Z[3] = 0
    Z[np.arange(n), np.arange(n)] = 1
    assert native.permanent_residues(Z, PRIMES[:8]) == [1] * 8
    Z[5] = 0
    assert native.permanent_residues(Z, PRIMES[:8]) == [0] * 8


def test_exact_long_chunks(native, oracle_c):
    """A large minimum chunk size (the whole 2^17-state walk in one lane's
    chunk when 2^30 states per lane are requested): the lane columns grow to
    their largest; results are unchanged."""
    rng = np.random.default_rng(77)
    A = rng.integers(-3, 4, (18, 18)).astype(np.int64)
    primes = PRIMES[:8]
    want = [oracle_c.permanent_mod_p(A.tolist(), p, 8) for p in primes]
    got = _with_tuning(native, native.TUNE_MIN_CBITS, 30, lambda: native.permanent_residues(A, primes))
    assert got == want


def test_exact_in_batches_with_modular_jobs(native, oracle_c):
    """One plan mixing exact jobs (many primes) with modular ones (few
    primes, more than eight groups, 0/1 tables) in every class."""
    rng = np.random.default_rng(5)
    mats, moduli = [], []
    for i in range(40):
        n = int(rng.integers(5, 21))
        bound = int(rng.choice((1, 5, 60, 3000)))
        mats.append(rng.integers(-bound, bound + 1, (n, n)).astype(np.int64))
        moduli.append(PRIMES[: int(rng.integers(1, 13))])
    plan = native.Batch(mats, moduli)
    try:
        flat = plan.run()
    finally:
        plan.close()
    classes = native.batch_plan_classes(mats, moduli)
    assert any(c[2] in (3, 6) for c in classes) and any(c[2] not in (3, 6) for c in classes)
    pos = 0
    for A, ps in zip(mats, moduli):
        want = [oracle_c.permanent_mod_p(A.tolist(), p, 8) for p in ps]
        assert flat[pos:pos + len(ps)] == want
        pos += len(ps)


def test_headline_goes_exact(native):
    """The 32 x 32 headline matrix with its 8 primes runs on one exact walk."""
    from permhip import planner, workloads

    A = workloads.headline()
    primes, _ = planner.generic_crt_primes(planner.matrix_permanent_bound(A.tolist()))
    assert _classes(native, A, primes) == {(4, 8, 6)}
    got = native.permanent_residues(A, primes)
    assert got == _with_tuning(native, native.TUNE_EXACT_MIN_PRIMES, 0, lambda: native.permanent_residues(A, primes))
    assert got == _with_tuning(native, native.TUNE_EXACT_SIGNED, 0, lambda: native.permanent_residues(A, primes))


@pytest.mark.parametrize("n,entry,variant", [(32, 6, 6), (24, 47, 6), (24, 48, 3)])
def test_signed_limbs_at_the_quad_limit(native, n, entry, variant):
    """Constant matrices c J_n with rows negated at random (perm = +-c^n n!),
    whose first Gray state puts every row sum at its bound.  Order 32 in
    groups of 4: 16 rows of bound 192 give four-group products up to
    2^121.4, signed limbs.  Order 24 in groups of 3: 12 rows of bound
    24 * 47 = 1128 give 2^121.7 (signed limbs), of bound 1152 give 2^122.04
    (the magnitude class)."""
    import math

    primes = PRIMES[:8]
    rng = np.random.default_rng(entry)
    signs = rng.choice((-1, 1), n)
    for flip in (False, True):
        M = np.full((n, n), entry, dtype=np.int64)
        if flip:
            M = M * signs[:, None]
        print(f"signed limit n={n} entry={entry} flip={flip}", flush=True)
        assert _classes(native, M, primes) == {(4 if n == 32 else 3, 8, variant)}
        exact = (int(np.prod(signs)) if flip else 1) * entry ** n * math.factorial(n)
        assert native.permanent_residues(M, primes) == [exact % p for p in primes], (n, entry, flip)


def test_signed_limbs_decline_large_quads(native):
    """Order 32 in groups of 4 with row bounds 208 (sixteen entries of
    magnitude 6 and sixteen of 7): groups below 2^31 but four-group bound
    products of 2^123.2, so the magnitude class; its residues equal the
    modular kernels'."""
    rng = np.random.default_rng(208)
    mags = np.array([6] * 16 + [7] * 16)
    A = np.array([rng.permutation(mags) * rng.choice((-1, 1), 32) for _ in range(32)], dtype=np.int64)
    primes = PRIMES[:8]
    assert _classes(native, A, primes) == {(4, 8, 3)}
    assert native.permanent_residues(A, primes) == _with_tuning(
        native, native.TUNE_EXACT_MIN_PRIMES, 0, lambda: native.permanent_residues(A, primes))


def _row_bound_bits(A):
    prod = 1
    for v in np.abs(A).sum(axis=1).tolist():
        prod *= int(v)
    return prod.bit_length()


# (order, entry bound, group size, limb columns, variant): 9..16 groups of
# 2, 3 or 4 rows, on both sides of the 2^310 limit of the 10-column class;
# groups of 2 rows whose partial bound products fit 22-bit values take the
# compact classes (variant 4).
WIDE_SHAPES = [(17, 1500, 2, 10, 3), (20, 1000, 2, 10, 3), (26, 100, 3, 10, 3), (28, 100, 2, 10, 4),
               (19, 250, 2, 10, 4), (25, 120, 2, 10, 4), (30, 300, 2, 12, 3), (32, 20, 3, 10, 3), (33, 6, 4, 10, 3),
               (34, 50, 3, 12, 3)]


def _exact(c):
    return c[2] in (3, 4)


@pytest.mark.parametrize("n,bound,group,columns,variant", WIDE_SHAPES)
def test_wide_exact_matches_oracle_and_modular(native, oracle_c, n, bound, group, columns, variant):
    rng = np.random.default_rng(1900 + 31 * n + bound)
    A = rng.integers(-bound, bound + 1, (n, n)).astype(np.int64)
    for k in (8, 10, 16):
        primes = PRIMES[:k]
        assert _classes(native, A, primes) == {(group, columns, variant)}, (n, bound, k, _row_bound_bits(A))
        got = native.permanent_residues(A, primes)
        modular = _with_tuning(native, native.TUNE_EXACT_MIN_PRIMES, 0, lambda: native.permanent_residues(A, primes))
        assert got == modular, (n, bound, k)
        if n <= 20:
            assert got == [oracle_c.permanent_mod_p(A.tolist(), p, 8) for p in primes], (n, bound, k)


@pytest.mark.parametrize("entry,columns,variant", [(76, 10, 4), (77, 12, 3), (350, 12, 3)])
def test_wide_exact_at_the_truncation_limit(native, entry, columns, variant):
    """28 x 28 tables of entries +-entry: every row bound is 28 * entry and
    the largest state product equals the row-bound product, which sits just
    below 2^310 (76), just above it (77: 12 columns) and just below 2^372
    (350).  The product kept modulo 2^(31 L) must still be exact.  At 76
    the group values reach 2^22.1, and the compact tree's nodes of 4, 6 and
    8 groups are within 5, 22 and 9 bits of their 3, 5 and 6 limbs."""
    rng = np.random.default_rng(entry)
    A = (rng.choice((-1, 1), (28, 28)) * entry).astype(np.int64)
    bits = _row_bound_bits(A)
    assert bits <= 31 * columns and (columns == 10 or bits > 310)
    primes = PRIMES[:10]
    assert _classes(native, A, primes) == {(2, columns, variant)}
    want = _with_tuning(native, native.TUNE_EXACT_MIN_PRIMES, 0, lambda: native.permanent_residues(A, primes))
    assert native.permanent_residues(A, primes) == want
    if variant == 4:
        assert _with_tuning(native, native.TUNE_EXACT_COMPACT, 0,
                            lambda: _classes(native, A, primes)) == {(2, columns, 3)}
        assert _with_tuning(native, native.TUNE_EXACT_COMPACT, 0, lambda: native.permanent_residues(A, primes)) == want


def test_wide_exact_declines_large_bounds_and_few_primes(native):
    """Row-bound products above 2^372 and tables with fewer primes than
    PHIP_TUNE_EXACT_WIDE_MIN_PRIMES stay on the modular kernels (host
    planning only)."""
    rng = np.random.default_rng(3)
    big = rng.integers(-1000, 1001, (32, 32)).astype(np.int64)
    assert _row_bound_bits(big) > 372
    assert not any(_exact(c) for c in _classes(native, big, PRIMES[:16]))
    A = rng.integers(-100, 101, (28, 28)).astype(np.int64)
    assert not any(_exact(c) for c in _classes(native, A, PRIMES[:5]))
    assert _classes(native, A, PRIMES[:6]) == {(2, 10, 4)}
    assert _with_tuning(native, native.TUNE_EXACT_WIDE_MIN_PRIMES, 0,
                        lambda: not any(_exact(c) for c in _classes(native, A, PRIMES[:10])))


def test_wide_exact_long_chunks(native):
    """Chunks of 2^24 states per lane (the exact variant's cap) on a 14-group
    table: the lane columns grow to their largest."""
    rng = np.random.default_rng(78)
    A = rng.integers(-100, 101, (26, 26)).astype(np.int64)
    primes = PRIMES[:10]
    want = _with_tuning(native, native.TUNE_EXACT_MIN_PRIMES, 0, lambda: native.permanent_residues(A, primes))
    got = _with_tuning(native, native.TUNE_MIN_CBITS, 30, lambda: native.permanent_residues(A, primes))
    assert got == want


@pytest.mark.parametrize("entry,variant", [(157, 4), (158, 3)])
def test_compact_exact_at_its_node_limits(native, oracle_c, entry, variant):
    """20 x 20 tables whose rows are constant (+-entry, random signs): the
    first Gray state attains every row bound at once, so its group products
    (2 rows, 23.2 bits) and the tree's nodes of 4 and 8 groups reach their
    bounds.  At 157 those nodes need exactly 93 and 186 bits, the most that
    3 and 6 limbs hold (compact class); at 158 they need 94 and 187 and the
    planner keeps the 31-bit exact class.  Both agree with the 31-bit exact
    class, the modular kernels and the oracle."""
    rng = np.random.default_rng(entry)
    A = np.outer(rng.choice((-1, 1), 20), np.ones(20, dtype=np.int64)) * entry
    A = A.astype(np.int64)
    primes = PRIMES[:10]
    assert _classes(native, A, primes) == {(2, 10, variant)}
    got = native.permanent_residues(A, primes)
    assert got == _with_tuning(native, native.TUNE_EXACT_COMPACT, 0, lambda: native.permanent_residues(A, primes))
    assert got == _with_tuning(native, native.TUNE_EXACT_MIN_PRIMES, 0, lambda: native.permanent_residues(A, primes))
    assert got == [oracle_c.permanent_mod_p(A.tolist(), p, 8) for p in primes]


def test_compact_exact_random_tables(native, oracle_c):
    """Random tables of 2-row groups on the compact class at orders 18..28
    (entries up to a few hundred), against the 31-bit exact class and, up
    to order 20, the oracle."""
    rng = np.random.default_rng(4040)
    seen = 0
    for _ in range(24):
        n = int(rng.integers(18, 29))
        bound = int(rng.integers(90, 400))
        A = rng.integers(-bound, bound + 1, (n, n)).astype(np.int64)
        primes = PRIMES[:int(rng.integers(6, 17))]
        if _classes(native, A, primes) != {(2, 10, 4)}:
            continue
        seen += 1
        print(f"compact n={n} bound={bound} k={len(primes)}", flush=True)
        got = native.permanent_residues(A, primes)
        assert got == _with_tuning(native, native.TUNE_EXACT_COMPACT, 0,
                                   lambda: native.permanent_residues(A, primes)), (n, bound)
        if n <= 20:
            assert got == [oracle_c.permanent_mod_p(A.tolist(), p, 8) for p in primes], (n, bound)
    assert seen >= 8
