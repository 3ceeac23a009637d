"""Live-column frontier DP on the device (perm_binary.py:389-516) against
the oracle's frontier_count, the permutation definition, closed forms and
the Glynn path; layer sizes through the state cap; the reference's own
frontier cases (tests/test_binary_frontier.py)."""

import math
import random

import pytest

import oracle

pytestmark = pytest.mark.gpu


def _random_masks(rng, n, density, diagonal=True):
    masks = [sum(1 << c for c in range(n) if rng.random() < density) for _ in range(n)]
    if diagonal:
        perm = list(range(n))
        rng.shuffle(perm)
        masks = [m | (1 << perm[r]) for r, m in enumerate(masks)]
    return tuple(masks)


def _cycle_adjacency(order):
    matrix = [[0] * order for _ in range(order)]
    for v in range(order):
        u = (v + 1) % order
        matrix[v][u] = matrix[u][v] = 1
    return matrix


def _block_diagonal_ones(size, count):
    n = size * count
    return [[int(r // size == c // size) for c in range(n)] for r in range(n)]


def _tridiagonal(n):
    return [[int(abs(r - c) <= 1) for c in range(n)] for r in range(n)]


def _fibonacci(k):
    a, b = 0, 1
    for _ in range(k):
        a, b = b, a + b
    return a


def test_frontier_matches_definition_on_random_small_matrices(native):
    import perm_binary
    from permhip import binary_prep as prep

    rng = random.Random(20261203)
    for n in range(1, 8):
        for _ in range(6):
            masks = _random_masks(rng, n, 0.5, diagonal=False)
            matrix = prep.unpack_masks(masks)
            expected = oracle.definition_permanent(matrix)
            print(f"frontier random n={n}", flush=True)
            assert perm_binary.binary_permanent_frontier(matrix) == expected
            order = list(range(n))
            rng.shuffle(order)
            assert perm_binary.binary_permanent_frontier(matrix, row_order=order) == expected


@pytest.mark.parametrize("n,density", [(20, 0.12), (30, 0.08), (40, 0.06), (60, 0.04)])
def test_frontier_matches_oracle_and_glynn_on_sparse_blocks(native, n, density):
    import perm_binary
    from permhip import binary_prep as prep

    rng = random.Random(n * 7919)
    checked = 0
    for _ in range(4):
        masks = _random_masks(rng, n, density)
        order, _, _ = prep.frontier_order(masks)
        try:
            expected = oracle.frontier_count(masks, order, max_states=200_000)
        except oracle.FrontierCapExceeded:
            continue
        matrix = prep.unpack_masks(masks)
        print(f"frontier sparse n={n}", flush=True)
        got = perm_binary.binary_permanent_frontier(matrix, max_states=200_000)
        assert got == expected
        if n <= 40:
            assert perm_binary.binary_permanent(matrix, backend="native") == expected
        checked += 1
    assert checked


def test_frontier_layer_sizes_match_the_oracle_through_the_cap(native):
    """The device's layer sizes equal the oracle's: a cap equal to the
    largest layer passes, one below it stops at that layer's row."""
    import perm_binary
    from permhip import binary_prep as prep

    rng = random.Random(4242)
    for n in (12, 24, 36):
        masks = _random_masks(rng, n, 0.12)
        order, _, _ = prep.frontier_order(masks)
        sizes = []
        expected = oracle.frontier_count(masks, order, layer_sizes=sizes)
        matrix = prep.unpack_masks(masks)
        widest = max(sizes)
        print(f"frontier cap n={n} widest={widest}", flush=True)
        assert perm_binary.binary_permanent_frontier(matrix, row_order=order, max_states=widest) == expected
        if widest > 1:
            first = next(i for i, s in enumerate(sizes) if s > widest - 1)
            with pytest.raises(perm_binary.FrontierStateLimitExceeded) as caught:
                perm_binary.binary_permanent_frontier(matrix, row_order=order, max_states=widest - 1)
            assert caught.value.processed_rows == first + 1
            assert caught.value.observed_states == widest


def test_frontier_uses_arbitrary_size_python_integers(native):
    import perm_binary

    result = perm_binary.binary_permanent_frontier(_block_diagonal_ones(10, 5), max_states=1_000)
    assert result == math.factorial(10) ** 5
    assert result.bit_length() > 64 and isinstance(result, int)


def test_frontier_chunks_many_primes(native, monkeypatch):
    """A bound above 64 primes is evaluated in chunks; forcing chunks of
    two primes must give the same integer."""
    import perm_binary
    from permhip import _native

    matrix = _block_diagonal_ones(8, 6)
    expected = math.factorial(8) ** 6
    monkeypatch.setattr(_native, "FRONTIER_MAX_PRIMES", 2)
    assert perm_binary.binary_permanent_frontier(matrix) == expected


def test_frontier_accepts_64_moduli_in_one_call(native):
    from permhip import binary_prep as prep

    masks = prep.row_masks(_tridiagonal(70))
    row_slots, expiring, slots = prep.frontier_layers(masks, range(70))
    primes = [prep.descending_prime(i) for i in range(64)]
    got, exceeded = native.binary_frontier_residues(row_slots, expiring, slots, primes, 1000)
    assert exceeded is None
    assert got == [_fibonacci(71) % p for p in primes]
    with pytest.raises(RuntimeError, match="frontier"):
        native.binary_frontier_residues(row_slots, expiring, slots, primes + [3], 1000)


def test_frontier_rejects_masks_outside_the_slots(native):
    with pytest.raises(RuntimeError, match="slots"):
        native.binary_frontier_residues([0b100], [0b100], 2, [2147483647], 10)
    with pytest.raises(RuntimeError, match="slots"):
        native.binary_frontier_residues([0b01], [0b10], 2, [2147483647], 10)


def test_frontier_state_cap_raises_typed_exception_immediately(native):
    import perm_binary

    matrix = [[1] * 8 for _ in range(8)]
    with pytest.raises(perm_binary.FrontierStateLimitExceeded, match="state limit 5 exceeded") as caught:
        perm_binary.binary_permanent_frontier(matrix, row_order=range(8), max_states=5)
    assert caught.value.max_states == 5
    assert caught.value.observed_states == 6
    assert caught.value.processed_rows == 1


def test_frontier_conventions_on_device(native):
    import perm_binary

    assert perm_binary.binary_permanent_frontier([[1]]) == 1
    assert perm_binary.binary_permanent_frontier([[1, 1], [1, 1]], row_order=(1, 0)) == 2
    assert perm_binary.binary_permanent_frontier([[3, 0], [1, 7]], strict=False) == 1
    assert perm_binary.binary_permanent_frontier(_cycle_adjacency(21)) == 2


def test_auto_uses_frontier_for_eligible_sparse_block(native, monkeypatch):
    import perm_binary

    def fail_native(*args, **kwargs):
        raise AssertionError("eligible sparse auto input should use frontier")

    monkeypatch.setattr(perm_binary, "_native_exact_permanent", fail_native)
    assert perm_binary.binary_permanent(_cycle_adjacency(21), backend="auto") == 2


def test_auto_frontier_cap_falls_back_to_native(native, monkeypatch):
    import perm_binary

    calls = []

    def fake_native(argument, bound, threads):
        calls.append((argument, bound, threads))
        return 2

    monkeypatch.setattr(perm_binary, "_AUTO_FRONTIER_MAX_STATES", 1)
    monkeypatch.setattr(perm_binary, "_native_exact_permanent", fake_native)
    assert perm_binary.binary_permanent(_cycle_adjacency(21), backend="auto") == 2
    assert len(calls) == 1


def test_auto_frontier_cap_fallback_is_exact(native, monkeypatch):
    import perm_binary

    monkeypatch.setattr(perm_binary, "_AUTO_FRONTIER_MAX_STATES", 1)
    assert perm_binary.binary_permanent(_cycle_adjacency(21), backend="auto") == 2


def test_explicit_native_backend_bypasses_frontier(native, monkeypatch):
    import perm_binary

    calls = []

    def fail_plan(*args, **kwargs):
        raise AssertionError("explicit native must not plan a frontier")

    def fake_native(argument, bound, threads):
        calls.append((argument, bound, threads))
        return 2

    monkeypatch.setattr(perm_binary, "_auto_frontier_plan", fail_plan)
    monkeypatch.setattr(perm_binary, "_native_exact_permanent", fake_native)
    assert perm_binary.binary_permanent(_cycle_adjacency(21), backend="native") == 2
    assert len(calls) == 1


def test_auto_handles_sparse_blocks_of_order_64_and_more(native):
    """A tridiagonal band of order 120 is one irreducible block with
    permanent F(121); auto takes it through the frontier, native rejects it
    as the reference's native backend does."""
    import perm_binary

    band = _tridiagonal(120)
    assert perm_binary.binary_permanent(band) == _fibonacci(121)
    with pytest.raises(ValueError, match="less than 64"):
        perm_binary.binary_permanent(band, backend="native")
    assert perm_binary.binary_permanent_frontier(band) == _fibonacci(121)


def test_large_order_with_small_blocks(native):
    """Order 100 in ten all-ones blocks of 10: the block split runs on the
    host restatement (masks wider than 64 bits), each block on Glynn."""
    import perm_binary

    matrix = _block_diagonal_ones(10, 10)
    assert perm_binary.binary_permanent(matrix) == math.factorial(10) ** 10
    assert perm_binary.binary_permanent_bound(matrix) >= math.factorial(10) ** 10


def _kasteleyn(m, n):
    """Domino tilings of an m x n board (Kasteleyn's product formula)."""
    value = 1.0
    for j in range(1, (m + 1) // 2 + 1):
        for k in range(1, (n + 1) // 2 + 1):
            value *= 4 * math.cos(math.pi * j / (m + 1)) ** 2 + 4 * math.cos(math.pi * k / (n + 1)) ** 2
    return round(value)


@pytest.mark.parametrize("shape", [(6, 8), (8, 10), (4, 30)])
def test_grid_domino_tilings_through_the_frontier(native, shape):
    """Perfect matchings of an open 2D grid (domino tilings, Kasteleyn's
    closed form): the biadjacency block is sparse (average degree < 4) with
    a frontier as wide as the short side, so auto plans a frontier."""
    import perm_binary
    from permhip import binary_prep as prep
    from permhip import grids

    matrix = grids.grid_biadjacency(shape)
    feasible, blocks = prep.matching_blocks(prep.row_masks(matrix))
    assert feasible and len(blocks) == 1
    assert prep.auto_frontier_plan(blocks[0]) is not None
    expected = _kasteleyn(*shape)
    print(f"frontier grid {shape} -> {expected}", flush=True)
    assert perm_binary.binary_permanent(matrix) == expected
    assert perm_binary.binary_permanent_frontier(matrix) == expected


def test_frontier_switches_between_small_and_large_layers(native):
    """All-ones blocks of order 14 reach C(14, 7) = 3432 states (beyond the
    single-workgroup kernel's 1024) in the middle of each block and one
    state between blocks, so the DP moves between the two paths several
    times; the count is 14!^2 * 3!."""
    import perm_binary

    n = 14 + 3 + 14
    M = [[0] * n for _ in range(n)]
    for lo, size in ((0, 14), (14, 3), (17, 14)):
        for r in range(lo, lo + size):
            for c in range(lo, lo + size):
                M[r][c] = 1
    assert perm_binary.binary_permanent_frontier(M, row_order=range(n), max_states=10_000) == (
        math.factorial(14) ** 2 * 6)
    with pytest.raises(perm_binary.FrontierStateLimitExceeded) as caught:
        perm_binary.binary_permanent_frontier(M, row_order=range(n), max_states=3431)
    assert caught.value.processed_rows == 7 and caught.value.observed_states == 3432


def test_frontier_integral_row_order_and_support_mode(native):
    """tests/test_binary_frontier.py:125-150 on the device."""
    import numpy as np
    import perm_binary

    assert perm_binary.binary_permanent_frontier([[1, 0], [0, 1]], row_order=np.array([np.int64(1), np.int32(0)])) == 1
    assert perm_binary.binary_permanent_frontier([[-7, 0], [4, 9]], strict=False) == 1


def test_frontier_with_64_live_columns_and_beyond(native):
    """The upper-triangular all-ones matrix (permanent 1) keeps every column
    live from the first row: at order 64 that is all 64 slots (the
    multi-kernel path, whose keys may have every bit set); at order 70 the
    device state mask cannot hold it and the call raises."""
    import perm_binary
    from permhip import binary_prep as prep

    def upper(n):
        return [[int(c >= r) for c in range(n)] for r in range(n)]

    masks = prep.row_masks(upper(64))
    assert prep.frontier_layers(masks, range(64))[2] == 64
    assert perm_binary.binary_permanent_frontier(upper(64), row_order=range(64)) == 1
    with pytest.raises(ValueError, match="live"):
        perm_binary.binary_permanent_frontier(upper(70), row_order=range(70))


def test_frontier_entry_point_edges(native):
    """Order 0 gives the empty product for every modulus; no modulus, too
    many, or an even one are rejected; a row with no slots gives zero."""
    assert native.binary_frontier_residues([], [], 0, [2147483647, 2147483629], 10)[0] == [1, 1]
    with pytest.raises(RuntimeError, match="frontier"):
        native.binary_frontier_residues([1], [1], 1, [], 10)
    with pytest.raises(RuntimeError, match="odd"):
        native.binary_frontier_residues([1], [1], 1, [2147483646], 10)
    assert native.binary_frontier_residues([0b1, 0b0], [0b1, 0b0], 1, [2147483647], 10) == ([0], None)


def _band(n, half_width):
    return [[int(abs(r - c) <= half_width) for c in range(n)] for r in range(n)]


def test_dense_large_blocks_take_the_uncapped_dp(native):
    """A band of half-width 3 and order 70 has about 7 entries a row, so
    auto does not try the capped frontier; the block (order >= 64) goes to
    the uncapped DP, the part the reference's subset DP plays
    (perm_binary.py:230-265, 997-1003).  Checked against the oracle's
    Python frontier DP; backend='python' gives the same, explicit native
    refuses the block."""
    import perm_binary
    from permhip import binary_prep as prep

    band = _band(70, 3)
    masks = prep.row_masks(band)
    assert prep.auto_frontier_plan(masks) is None
    expect = oracle.frontier_count(masks, range(70), max_states=10 ** 7)
    assert perm_binary.binary_permanent(band) == expect
    assert perm_binary.binary_permanent(band, backend="python") == expect
    for p in (2147483647, (1 << 61) - 1):
        assert perm_binary.binary_permanent_mod_p(band, p) == expect % p
        assert perm_binary.binary_permanent_mod_p(band, p, backend="python") == expect % p
    with pytest.raises(ValueError, match="less than 64"):
        perm_binary.binary_permanent(band, backend="native")
    with pytest.raises(ValueError, match="less than 64"):
        perm_binary.binary_permanent_mod_p(band, 2147483647, backend="native")


def test_python_backend_name_matches_the_definition(native):
    import perm_binary

    rng = random.Random(31)
    for n in range(0, 13):
        B = [[int(rng.random() < 0.5) for _ in range(n)] for _ in range(n)]
        d = oracle.subset_permanent(B) if n else 1
        assert perm_binary.binary_permanent(B, backend="python") == d
        if n:
            assert perm_binary.binary_permanent_mod_p(B, 1000003, backend="python") == d % 1000003
