"""Zero-skipping Glynn kernels (0/1 tables whose Gray states are mostly
zero terms): the planner picks them, and their residues equal the plain
kernels' and the oracle's, bit for bit, including chunks with inactive
lanes, padded single primes, zero rows and mixed plans.  Both walks are
covered: the block walk (32 states per row lookup, the default from order
8 on) and the per-state zero test; the block walk in both its forms, with
a Montgomery tree per prime and with exact accumulation (the default for
five primes or more; the block_exact fixture runs the tests that take it
once with it from three primes on and once with it off)."""

import math
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from permhip import binary_prep as _prep

# The binary planner's primes (descending from 2^31 - 1) and the integer
# planner's (ascending from 2^30, which pair consecutive states).
BINARY_PRIMES = [_prep.descending_prime(i) for i in range(8)]
LOW_PRIMES = [q for q in range((1 << 30) + 1, (1 << 30) + 2000, 2) if _prep.is_prime_u64(q)][:4]


def _with_skip(native, permille, fn):
    old = native.get_tuning(native.TUNE_SKIP_ZEROS_PERMILLE)
    native.set_tuning(native.TUNE_SKIP_ZEROS_PERMILLE, permille)
    try:
        return fn()
    finally:
        native.set_tuning(native.TUNE_SKIP_ZEROS_PERMILLE, old)


# Plan variants of the zero-skipping walks: per prime (2) and the block walk
# with exact accumulation (5).
SKIP_VARIANTS = (2.0, 5.0)


@pytest.fixture(params=[3, 0], ids=["exact-blocks", "modular-blocks"])
def block_exact(request, native):
    """Runs a test with the exact block walk from three primes on (the
    default is five) and switched off (every block table on the per-prime
    walk)."""
    old = native.get_tuning(native.TUNE_SKIP_EXACT_MIN_PRIMES)
    native.set_tuning(native.TUNE_SKIP_EXACT_MIN_PRIMES, request.param)
    try:
        yield request.param
    finally:
        native.set_tuning(native.TUNE_SKIP_EXACT_MIN_PRIMES, old)


def _with_blocks(native, on, fn):
    old = native.get_tuning(native.TUNE_SKIP_BLOCKS)
    native.set_tuning(native.TUNE_SKIP_BLOCKS, 1 if on else 0)
    try:
        return fn()
    finally:
        native.set_tuning(native.TUNE_SKIP_BLOCKS, old)


def _even_degree_matrix(rng, n, density):
    """A 0/1 matrix whose rows mostly have even degree (many zero terms)."""
    A = (rng.random((n, n)) < density).astype(np.int64)
    for i in range(n):
        if A[i].sum() % 2:
            j = int(rng.integers(n))
            A[i, j] ^= 1
    return A


def _variant(native, A, primes):
    plan = native.Batch([A], [primes])
    try:
        return plan.stats()["variant"]
    finally:
        plan.close()


@pytest.mark.parametrize("n", [5, 6, 8, 11, 14, 17, 20, 23])
def test_skip_matches_plain_and_oracle_small(native, oracle_c, n):
    """Small orders (jobs of fewer than 64 lanes leave lanes inactive)."""
    rng = np.random.default_rng(100 + n)
    for density in (0.35, 0.6):
        A = _even_degree_matrix(rng, n, density)
        for primes in (BINARY_PRIMES[:2], BINARY_PRIMES[:4], BINARY_PRIMES[:8], BINARY_PRIMES[:1]):
            print(f"skip n={n} density={density} k={len(primes)}", flush=True)
            skip = native.permanent_residues(A, primes)
            per_state = _with_blocks(native, False, lambda: native.permanent_residues(A, primes))
            plain = _with_skip(native, 0, lambda: native.permanent_residues(A, primes))
            assert skip == plain == per_state, (n, density, primes)
            for p, r in zip(primes, skip):
                assert r == oracle_c.binary_mod_p(A.astype(np.uint8), p, 16), (n, p)


@pytest.mark.parametrize("n", [28, 32])
def test_skip_matches_plain_large(native, block_exact, n):
    rng = np.random.default_rng(7 + n)
    A = _even_degree_matrix(rng, n, 0.5)
    assert _variant(native, A, BINARY_PRIMES[:4]) == (5.0 if block_exact else 2.0)
    print(f"skip large n={n}", flush=True)
    skip = native.permanent_residues(A, BINARY_PRIMES[:4])
    per_state = _with_blocks(native, False, lambda: native.permanent_residues(A, BINARY_PRIMES[:4]))
    plain = _with_skip(native, 0, lambda: native.permanent_residues(A, BINARY_PRIMES[:4]))
    assert skip == plain == per_state


def test_planner_picks_the_variant(native):
    rng = np.random.default_rng(11)
    A = _even_degree_matrix(rng, 20, 0.5)
    assert _variant(native, A, BINARY_PRIMES[:5]) == 5.0
    assert _variant(native, A, BINARY_PRIMES[:4]) == 2.0
    assert _with_blocks(native, False, lambda: _variant(native, A, BINARY_PRIMES[:4])) == 2.0
    assert _with_skip(native, 0, lambda: _variant(native, A, BINARY_PRIMES[:4])) == 0.0
    # Odd degrees everywhere: no row ever sums to zero, so no skipping.
    odd = np.zeros((20, 20), dtype=np.int64)
    for i in range(20):
        for j in range(i, i + 5):
            odd[i, j % 20] = 1
    assert _variant(native, odd, BINARY_PRIMES[:4]) == 0.0
    # Integer tables never skip.
    B = rng.integers(-3, 4, (20, 20)).astype(np.int64)
    assert _variant(native, B, BINARY_PRIMES[:4]) not in SKIP_VARIANTS


def _block_ones(sizes, rng):
    """A block-diagonal 0/1 matrix of all-ones blocks (permanent the
    product of the blocks' factorials), rows and columns shuffled."""
    n = sum(sizes)
    A = np.zeros((n, n), dtype=np.int64)
    at = 0
    for k in sizes:
        A[at:at + k, at:at + k] = 1
        at += k
    return A[rng.permutation(n)][:, rng.permutation(n)]


@pytest.mark.parametrize("sizes,half", [((4,) * 11, False), ((3,) * 9 + (4,) * 4 + (2,), True)])
def test_skip_largest_orders_closed_form(native, block_exact, sizes, half):
    """Orders 44 and 45 (near the byte walk's largest, 46): eleven J_4
    blocks (every row of even degree, full-width test) and nine J_3, four
    J_4 and one J_2 block (18 even rows of 45, half-width test; an order
    that leaves padding bytes).  2^43 and 2^44 Gray states; the residues
    equal the product of the blocks' factorials."""
    rng = np.random.default_rng(len(sizes))
    A = _block_ones(sizes, rng)
    n = A.shape[0]
    primes = BINARY_PRIMES[:4]
    exact = math.prod(math.factorial(k) for k in sizes)

    def per_state():
        plan = native.Batch([A], [primes])
        try:
            st = plan.stats()
        finally:
            plan.close()
        words = (n + 3) // 4
        assert st["variant"] == 2.0 and st["skip_block_rows"] == -1.0
        assert st["skip_test_words"] == ((words + 1) // 2 if half else words)
        print(f"skip largest n={n} half={half} per state", flush=True)
        return native.permanent_residues(A, primes)

    if block_exact:  # the per-state walk does not depend on the block walk's form
        assert _with_blocks(native, False, per_state) == [exact % p for p in primes]
    print(f"skip largest n={n} half={half} blocks", flush=True)
    assert native.permanent_residues(A, primes) == [exact % p for p in primes]


def test_skip_known_counts(native, block_exact):
    """Closed forms through the zero-skipping kernels: all-ones J_n
    (n!, rows of even degree for even n), J_n - I (derangements), a zero
    row, and block-diagonal ones."""
    import perm_binary

    for n in (10, 12, 14):
        J = np.ones((n, n), dtype=np.int64)
        assert native.permanent_residues(J, BINARY_PRIMES[:2]) == [math.factorial(n) % p for p in BINARY_PRIMES[:2]]
    der = [1, 0]
    for n in range(2, 16):
        der.append((n - 1) * (der[-1] + der[-2]))
    for n in (11, 13, 15):  # J - I has rows of degree n - 1, even for odd n
        D = np.ones((n, n), dtype=np.int64) - np.eye(n, dtype=np.int64)
        assert _variant(native, D, BINARY_PRIMES[:2]) == 2.0
        assert native.permanent_residues(D, BINARY_PRIMES[:2]) == [der[n] % p for p in BINARY_PRIMES[:2]]
    Z = _even_degree_matrix(np.random.default_rng(5), 12, 0.5)
    Z[3] = 0
    assert native.permanent_residues(Z, BINARY_PRIMES[:4]) == [0, 0, 0, 0]
    B = [[1 if i // 4 == j // 4 else 0 for j in range(24)] for i in range(24)]
    assert perm_binary.binary_permanent(B) == math.factorial(4) ** 6


def test_skip_batch_mixes_variants(native, oracle_c, block_exact):
    """One plan with zero-skipping, plain and paired jobs: every residue
    equals the oracle's."""
    rng = np.random.default_rng(21)
    mats, mods = [], []
    for n in (12, 15, 18):
        mats.append(_even_degree_matrix(rng, n, 0.5))
        mods.append(BINARY_PRIMES[:4])
        mats.append(rng.integers(-5, 6, (n, n)).astype(np.int64))
        mods.append(BINARY_PRIMES[:4])
        mats.append(rng.integers(-5, 6, (n, n)).astype(np.int64))
        mods.append(LOW_PRIMES)
    plan = native.Batch(mats, mods)
    try:
        flat = plan.run()
    finally:
        plan.close()
    at = 0
    for A, ps in zip(mats, mods):
        for p in ps:
            assert flat[at] == oracle_c.permanent_mod_p(A.tolist(), p, 16), (A.shape[0], p)
            at += 1


def test_grid_and_binary36_through_skip(native):
    """The 40x40 periodic grid (exact count) and the binary36 golden
    residues run on the zero-skipping kernels."""
    import perm_binary
    from permhip import grids

    G = np.array(grids.grid_biadjacency((4, 4, 5), periodic_axes=(0, 1)), dtype=np.int64)
    assert _variant(native, G, [2147483647, 2147483629]) == 2.0
    print("skip grid 40", flush=True)
    assert perm_binary.binary_permanent(G.tolist()) == 2551276235535120


def _odd_heavy_matrix(rng, n, density, even_rows):
    """A 0/1 matrix with exactly `even_rows` rows of even degree (the rest
    odd), rows shuffled, so the zero test may read only half the words.  Its
    columns have odd degree (all but the last when the parities require it),
    so the planner keeps this orientation rather than walk the transpose."""
    A = (rng.random((n, n)) < density).astype(np.int64)
    row_odd = [i >= even_rows for i in range(n)]
    col_odd = [True] * n
    if sum(row_odd) % 2 != n % 2:
        col_odd[n - 1] = False
    for i in range(n - 1):
        if (A[i].sum() % 2 == 1) != row_odd[i]:
            A[i, n - 1] ^= 1
    for j in range(n - 1):
        if (A[:, j].sum() % 2 == 1) != col_odd[j]:
            A[n - 1, j] ^= 1
    if (A[n - 1].sum() % 2 == 1) != row_odd[n - 1]:
        A[n - 1, n - 1] ^= 1
    assert [d % 2 == 1 for d in A.sum(axis=1)] == row_odd
    return A[rng.permutation(n)]


@pytest.mark.parametrize("n", [20, 23, 26, 29, 33])
def test_half_width_zero_test(native, oracle_c, n):
    _with_blocks(native, False, lambda: _half_width_case(native, oracle_c, n))


def _half_width_case(native, oracle_c, n):
    """Tables whose rows of even degree fit in the leading half of the byte
    words run the half-width test; residues equal the plain kernels' and,
    up to order 26, the oracle's."""
    rng = np.random.default_rng(300 + n)
    words = (n + 3) // 4
    half = (words + 1) // 2
    for density, even_rows in ((0.5, 4 * half), (0.35, 4 * half - 3), (0.65, 3)):
        A = _odd_heavy_matrix(rng, n, density, min(even_rows, n))
        primes = BINARY_PRIMES[:4]
        plan = native.Batch([A], [primes])
        try:
            st = plan.stats()
        finally:
            plan.close()
        print(f"half test n={n} density={density} even={even_rows} stats={st['variant']},{st['skip_test_words']}",
              flush=True)
        if st["variant"] != 2.0:
            continue
        assert st["skip_test_words"] == half
        skip = native.permanent_residues(A, primes)
        plain = _with_skip(native, 0, lambda: native.permanent_residues(A, primes))
        assert skip == plain, (n, density)
        if n <= 26:
            for p, r in zip(primes, skip):
                assert r == oracle_c.binary_mod_p(A.astype(np.uint8), p, 16), (n, p)


def test_full_width_when_even_rows_spill(native):
    rng = np.random.default_rng(17)
    n = 30
    A = _even_degree_matrix(rng, n, 0.5)

    def stats():
        plan = native.Batch([A], [BINARY_PRIMES[:4]])
        try:
            return plan.stats()
        finally:
            plan.close()

    st = _with_blocks(native, False, stats)
    assert st["variant"] == 2.0 and st["skip_test_words"] == (n + 3) // 4


@pytest.mark.parametrize("k", [3, 5, 6, 7])
def test_one_walk_for_up_to_eight_primes(native, block_exact, k):
    """A 0/1 table's primes share one zero-skipping walk (a padded set of 4
    or 8, or the exact block walk's eight limb columns), with the plain
    kernels' residues."""
    rng = np.random.default_rng(40 + k)
    A = _even_degree_matrix(rng, 24, 0.5)
    primes = BINARY_PRIMES[:k]
    plan = native.Batch([A], [primes])
    try:
        st = plan.stats()
    finally:
        plan.close()
    if block_exact:
        assert st["variant"] == 5.0 and st["jobs"] == 1.0 and st["set"] == 8.0
    else:
        assert st["variant"] == 2.0 and st["jobs"] == 1.0 and st["set"] == (4.0 if k <= 4 else 8.0)
    print(f"one walk k={k}", flush=True)
    skip = native.permanent_residues(A, primes)
    per_state = _with_blocks(native, False, lambda: native.permanent_residues(A, primes))
    plain = _with_skip(native, 0, lambda: native.permanent_residues(A, primes))
    assert skip == plain == per_state


def _block_rows(n, even):
    words = (n + 3) // 4
    return 4 * min(words, 2 * ((even + 7) // 8))


@pytest.mark.parametrize("n,even_rows", [(8, 8), (13, 5), (24, 24), (31, 9), (38, 20), (38, 38)])
def test_block_walk_stats_and_parity(native, oracle_c, block_exact, n, even_rows):
    """The block walk tests whole pairs of row words covering the rows of
    even degree; its residues equal the per-state walk's, and up to order
    24 the oracle's."""
    rng = np.random.default_rng(500 + n + even_rows)
    A = _odd_heavy_matrix(rng, n, 0.5, even_rows)
    even = int(sum(1 for d in A.sum(axis=1) if d % 2 == 0))
    primes = BINARY_PRIMES[:4]
    plan = native.Batch([A], [primes])
    try:
        st = plan.stats()
    finally:
        plan.close()
    print(f"block n={n} even={even} stats={st['variant']},{st['skip_block_rows']},{st['chunk_bits_min']}", flush=True)
    if st["variant"] not in SKIP_VARIANTS:
        assert even < 4
        return
    assert st["variant"] == (5.0 if block_exact else 2.0)
    assert st["skip_block_rows"] == _block_rows(n, even)
    assert st["chunk_bits_min"] >= 6
    blocks = native.permanent_residues(A, primes)
    per_state = _with_blocks(native, False, lambda: native.permanent_residues(A, primes))
    assert blocks == per_state, (n, even)
    if n <= 24:
        for p, r in zip(primes, blocks):
            assert r == oracle_c.binary_mod_p(A.astype(np.uint8), p, 16), (n, p)


def test_block_walk_many_jobs_one_class(native, oracle_c, block_exact):
    """One plan of 0/1 matrices of orders 8..22 on one block class: the
    wavefronts move between jobs of different table sizes (the LDS region
    is sized for the largest; tables reload on a job change), and single
    primes, sets of 2, 3, 4 and 5 mix (with the exact block walk on, the
    sets of 3 or more on its class, the rest on the per-prime walk)."""
    rng = np.random.default_rng(61)
    mats, mods = [], []
    for i, n in enumerate(list(range(8, 23)) * 2):
        mats.append(_even_degree_matrix(rng, n, 0.4 + 0.02 * (i % 10)))
        mods.append(BINARY_PRIMES[: (1, 2, 4, 3, 5)[i % 5]])
    plan = native.Batch(mats, mods)
    try:
        flat = plan.run()
        flat_again = plan.run()
    finally:
        plan.close()
    assert flat == flat_again
    at = 0
    for A, ps in zip(mats, mods):
        for p in ps:
            assert flat[at] == oracle_c.binary_mod_p(A.astype(np.uint8), p, 16), (A.shape[0], p)
            at += 1


@pytest.mark.parametrize("n", [40, 42])
def test_block_walk_large_closed_forms(native, block_exact, n):
    """Orders 40 and 42 on the block walk: block-diagonal all-ones blocks of
    sizes 2..5 (rows of both parities), residues the product of the blocks'
    factorials."""
    rng = np.random.default_rng(n)
    sizes = []
    while sum(sizes) < n:
        sizes.append(min(int(rng.integers(2, 6)), n - sum(sizes)))
    if sizes[-1] == 1:
        sizes[-2] -= 1
        sizes[-1] = 2
    A = _block_ones(sizes, rng)
    primes = BINARY_PRIMES[:3]
    plan = native.Batch([A], [primes])
    try:
        st = plan.stats()
    finally:
        plan.close()
    assert st["variant"] == (5.0 if block_exact else 2.0) and st["skip_block_rows"] > 0
    exact = math.prod(math.factorial(k) for k in sizes)
    print(f"block closed form n={n} sizes={sizes}", flush=True)
    assert native.permanent_residues(A, primes) == [exact % p for p in primes]


def _with_mod3_blocks(native, on, fn):
    old = native.get_tuning(native.TUNE_MOD3_BLOCKS)
    native.set_tuning(native.TUNE_MOD3_BLOCKS, 1 if on else 0)
    try:
        return fn()
    finally:
        native.set_tuning(native.TUNE_MOD3_BLOCKS, old)


@pytest.mark.parametrize("n", [8, 9, 13, 17, 22, 27, 31, 36])
def test_mod3_block_walk_matches_bit_sliced_and_oracle(native, oracle_c, n):
    """Permanents modulo 3 on the modulo-3 block walk (orders 8..46) equal
    the bit-sliced kernel's and, up to order 27, the oracle's: integer
    matrices, 0/1 matrices, matrices of -1/0/1 entries heavy in +-1, and a
    matrix with a row of multiples of 3."""
    rng = np.random.default_rng(600 + n)
    cases = [rng.integers(-7, 8, (n, n)).astype(np.int64),
             (rng.random((n, n)) < 0.5).astype(np.int64),
             rng.choice([-1, 1, 1, 0], size=(n, n)).astype(np.int64)]
    Z = rng.integers(-4, 5, (n, n)).astype(np.int64)
    Z[n // 2] = 3 * rng.integers(-2, 3, n)
    cases.append(Z)
    for idx, A in enumerate(cases):
        print(f"mod3 blocks n={n} case={idx}", flush=True)
        blocks = native.permanent_mod3(A % 3)
        sliced = _with_mod3_blocks(native, False, lambda: native.permanent_mod3(A % 3))
        assert blocks == sliced, (n, idx)
        if n <= 27:
            assert blocks == oracle_c.mod3(A, 8), (n, idx)
        if idx == 3:
            assert blocks == 0


@pytest.mark.parametrize("n", [40, 44])
def test_mod3_block_walk_large_orders(native, n):
    """Orders 40 and 44 (2^39 and 2^43 Gray states): the modulo-3 block walk
    against the bit-sliced kernel, and a block-diagonal matrix against the
    product of its blocks' permanents modulo 3."""
    rng = np.random.default_rng(n)
    A = rng.integers(-2, 3, (n, n)).astype(np.int64)
    print(f"mod3 blocks large n={n}", flush=True)
    blocks = native.permanent_mod3(A % 3)
    if n == 40:
        sliced = _with_mod3_blocks(native, False, lambda: native.permanent_mod3(A % 3))
        assert blocks == sliced
    sizes = [5] * (n // 5)
    sizes[-1] += n - sum(sizes)
    D = np.zeros((n, n), dtype=np.int64)
    at = 0
    want = 1
    import itertools

    for k in sizes:
        blk = rng.integers(-1, 2, (k, k)).astype(np.int64)
        D[at:at + k, at:at + k] = blk
        want = want * sum(math.prod(int(blk[i, s[i]]) for i in range(k)) for s in itertools.permutations(range(k)))
        at += k
    perm_r = rng.permutation(n)
    D = D[perm_r][:, rng.permutation(n)]
    assert native.permanent_mod3(D % 3) == want % 3
