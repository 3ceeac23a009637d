"""Frontier DP host side and its oracle (no device).

The oracle's frontier_count (perm_binary.py:389-441) is pinned against the
permutation definition on random small matrices and against closed forms;
the library's row order (phip_frontier_order) against widths recomputed
from the order and against the optimum on grids; the live-column slot
relabelling against the oracle's layer sizes and counts.
"""

import math
import random

import numpy as np
import pytest

import perm_binary
import oracle
from permhip import binary_prep as prep


def _random_masks(rng, n, density):
    return tuple(sum(1 << c for c in range(n) if rng.random() < density) for _ in range(n))


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


# --- the oracle, pinned ----------------------------------------------------


def test_oracle_frontier_matches_definition_on_random_small_matrices():
    rng = random.Random(20261203)
    for n in range(1, 7):
        for _ in range(12):
            masks = _random_masks(rng, n, 0.55)
            matrix = prep.unpack_masks(masks)
            expected = oracle.definition_permanent(matrix)
            order = list(range(n))
            assert oracle.frontier_count(masks, order) == expected
            rng.shuffle(order)
            assert oracle.frontier_count(masks, order) == expected


def test_oracle_frontier_closed_forms():
    masks = prep.row_masks(_block_diagonal_ones(10, 5))
    assert oracle.frontier_count(masks, range(50), max_states=1000) == math.factorial(10) ** 5
    cycle = prep.row_masks(_cycle_adjacency(21))
    assert oracle.frontier_count(cycle, range(21)) == 2
    tri = prep.row_masks(_tridiagonal(30))
    assert oracle.frontier_count(tri, range(30)) == _fibonacci(31)
    hall = prep.row_masks([[1, 0, 0, 0], [1, 0, 0, 0], [0, 1, 1, 1], [0, 1, 1, 1]])
    assert oracle.frontier_count(hall, range(4)) == 0
    assert oracle.frontier_count((), ()) == 1


def test_oracle_frontier_state_cap_fields():
    # The reference's own case (tests/test_binary_frontier.py:79-90).
    ones = prep.row_masks([[1] * 8 for _ in range(8)])
    with pytest.raises(oracle.FrontierCapExceeded) as caught:
        oracle.frontier_count(ones, range(8), max_states=5)
    assert (caught.value.max_states, caught.value.observed_states, caught.value.processed_rows) == (5, 6, 1)
    sizes = []
    oracle.frontier_count(ones, range(8), layer_sizes=sizes)
    assert sizes == [math.comb(8, k) for k in range(1, 8)] + [1]


# --- greedy order ------------------------------------------------------------


def _widths(masks, order):
    """(maximum width, area) recomputed from an order: after each prefix the
    width is the number of columns touched but not yet complete."""
    n = len(masks)
    degree = [sum((int(m) >> c) & 1 for m in masks) for c in range(n)]
    seen = [0] * n
    widest = area = 0
    for r in order:
        for c in range(n):
            seen[c] += (int(masks[r]) >> c) & 1
        width = sum(1 for c in range(n) if 0 < seen[c] < degree[c])
        widest = max(widest, width)
        area += width
    return widest, area


@pytest.mark.parametrize("n,density", [(0, 0.5), (1, 1.0), (5, 0.5), (12, 0.3), (21, 0.15), (40, 0.08),
                                       (70, 0.05), (90, 0.03)])
def test_order_is_a_permutation_with_its_reported_widths(n, density):
    rng = random.Random(1000 + n)
    for _ in range(3):
        masks = _random_masks(rng, n, density)
        order, width, area = prep.frontier_order(masks)
        assert sorted(order) == list(range(n))
        assert (width, area) == _widths(masks, order)
        # Never worse than the rows' own order, which is one of the starts.
        assert (width, area) <= _widths(masks, range(n))


def _grid_masks(w, h):
    """Dimer matrix of a w x h grid: black cells x white cells."""
    cells = [(i, j) for i in range(w) for j in range(h)]
    white = {c: k for k, c in enumerate(c for c in cells if sum(c) % 2)}
    return tuple(sum(1 << white[(i + di, j + dj)] for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1))
                     if (i + di, j + dj) in white) for (i, j) in cells if sum((i, j)) % 2 == 0)


@pytest.mark.parametrize("w,h", [(4, 30), (6, 8), (8, 10), (10, 12), (12, 16)])
def test_grid_orders_reach_the_short_side(w, h):
    """A sweep along the long side keeps about one grid line of white cells
    live, min(w, h); no order can do better than that order of magnitude,
    and the reference's greedy reaches exactly min(w, h) on these grids."""
    masks = _grid_masks(w, h)
    order, width, area = prep.frontier_order(masks)
    assert (width, area) == _widths(masks, order)
    assert width <= min(w, h)


def test_order_is_deterministic_and_small_on_an_odd_cycle():
    masks = perm_binary._row_masks(_cycle_adjacency(21))
    first = perm_binary._frontier_greedy_order(masks)
    assert first == perm_binary._frontier_greedy_order(masks)
    assert sorted(first[0]) == list(range(21))
    assert first[1] <= 4


def test_tridiagonal_band_has_width_two():
    order, width, _ = prep.frontier_order(prep.row_masks(_tridiagonal(300)))
    assert width == 2 and sorted(order) == list(range(300))


def test_order_rejects_columns_outside_the_matrix():
    with pytest.raises(ValueError, match="outside the matrix"):
        prep.frontier_order((1 << 3, 1))


def test_auto_plan_eligibility():
    assert prep.auto_frontier_plan(prep.row_masks(_cycle_adjacency(21))) is not None
    assert prep.auto_frontier_plan(prep.row_masks(_cycle_adjacency(19))) is None  # order < 20
    dense = prep.row_masks([[1] * 24 for _ in range(24)])
    assert prep.auto_frontier_plan(dense) is None  # average degree > 4
    plan = prep.auto_frontier_plan(prep.row_masks(_tridiagonal(80)))
    assert plan is not None and sorted(plan[1]) == list(range(80))


def test_auto_plan_prefers_the_narrower_orientation():
    # Rows 0..29 each hit column 0 plus one private column: column 0 is
    # live for the whole walk row by row, but the transpose has one dense
    # row and narrow columns.
    rng = random.Random(7)
    for _ in range(5):
        masks = _random_masks(rng, 30, 0.1)
        plan = prep.auto_frontier_plan(masks)
        direct = prep.frontier_order(masks)
        transposed = prep.transpose_masks(masks)
        if transposed == tuple(masks):
            continue
        t = prep.frontier_order(transposed)
        best_width = min(direct[1], t[1])
        if best_width > 24 or 2 * best_width > 30:
            assert plan is None
        else:
            use_t = (t[1], t[2]) < (direct[1], direct[2])
            assert plan == ((transposed, t[0]) if use_t else (tuple(masks), direct[0]))


# --- slot relabelling ------------------------------------------------------------


def _slot_dp(row_slots, expiring, layer_sizes):
    """The device's recurrence over slot masks, on the CPU."""
    states = {0: 1}
    for rm, em in zip(row_slots, expiring):
        nxt = {}
        for used, count in states.items():
            avail = rm & ~used
            while avail:
                bit = avail & -avail
                avail ^= bit
                target = used | bit
                if target & em != em:
                    continue
                target &= ~em
                nxt[target] = nxt.get(target, 0) + count
        layer_sizes.append(len(nxt))
        states = nxt
        if not states:
            return 0
    return states.get(0, 0)


@pytest.mark.parametrize("n,density", [(8, 0.4), (20, 0.15), (40, 0.05), (90, 0.015)])
def test_slot_relabelling_keeps_layers_and_counts(n, density):
    rng = random.Random(n)
    checked = 0
    for trial in range(6):
        masks = _random_masks(rng, n, density)
        if trial % 2:
            # A permuted diagonal keeps every row and column nonempty.
            perm = list(range(n))
            rng.shuffle(perm)
            masks = tuple(m | (1 << perm[r]) for r, m in enumerate(masks))
        order, width, _ = prep.frontier_order(masks)
        layers = prep.frontier_layers(masks, order)
        if layers is None:
            assert any(all(not (int(m) >> c) & 1 for m in masks) for c in range(n))
            continue
        row_slots, expiring, slots = layers
        assert slots <= width + max(bin(int(m)).count("1") for m in masks)
        assert all(e & ~r == 0 for r, e in zip(row_slots, expiring))
        expected_sizes, got_sizes = [], []
        try:
            expected = oracle.frontier_count(masks, order, max_states=50_000, layer_sizes=expected_sizes)
        except oracle.FrontierCapExceeded:
            continue
        assert _slot_dp(row_slots, expiring, got_sizes) == expected
        assert got_sizes == expected_sizes
        checked += 1
    assert checked


def test_slot_relabelling_reuses_slots_on_a_long_band():
    masks = prep.row_masks(_tridiagonal(200))
    row_slots, expiring, slots = prep.frontier_layers(masks, range(200))
    assert slots <= 4
    sizes = []
    assert _slot_dp(row_slots, expiring, sizes) == _fibonacci(201)


# --- public API, host-only cases and validation ---------------------------------


def test_frontier_empty_singleton_and_infeasible_conventions():
    assert perm_binary.binary_permanent_frontier([]) == 1
    assert perm_binary.binary_permanent_frontier([[0]]) == 0
    hall = [[1, 0, 0, 0], [1, 0, 0, 0], [0, 1, 1, 1], [0, 1, 1, 1]]
    assert perm_binary.binary_permanent_frontier(hall, max_states=1) == 0


@pytest.mark.parametrize("max_states", [0, -1, 1.5, True, np.bool_(True), None])
def test_frontier_rejects_invalid_state_caps(max_states):
    error = ValueError if type(max_states) is int and max_states <= 0 else TypeError
    with pytest.raises(error, match="max_states"):
        perm_binary.binary_permanent_frontier([[1]], max_states=max_states)


@pytest.mark.parametrize("row_order", [[0], [0, 0], [0, 2], [True, False], [0.0, 1.0], 5, ["0", "1"]])
def test_frontier_rejects_invalid_row_orders(row_order):
    with pytest.raises((TypeError, ValueError), match="row_order"):
        perm_binary.binary_permanent_frontier([[1, 1], [1, 1]], row_order=row_order)


def test_frontier_snapshots_row_order_length():
    class Shrinking(list):
        def __len__(self):
            return 2

        def __iter__(self):
            return iter([0])

    with pytest.raises(ValueError, match="length changed"):
        perm_binary.binary_permanent_frontier([[1, 1], [1, 1]], row_order=Shrinking([0, 1]))


def test_frontier_preserves_strict_support_semantics():
    with pytest.raises(ValueError, match="0 or 1"):
        perm_binary.binary_permanent_frontier([[2]])


def test_matching_blocks_above_63_rows():
    masks = prep.row_masks(_block_diagonal_ones(10, 8))
    feasible, blocks = prep.matching_blocks(masks)
    assert feasible and [len(b) for b in blocks] == [10] * 8
    assert all(all(m == (1 << 10) - 1 for m in b) for b in blocks)


def _kasteleyn(m, n):
    value = 1.0
    for j in range(1, (m + 1) // 2 + 1):
        for k in range(1, (n + 1) // 2 + 1):
            value *= 4 * math.cos(math.pi * j / (m + 1)) ** 2 + 4 * math.cos(math.pi * k / (n + 1)) ** 2
    return round(value)


@pytest.mark.parametrize("shape", [(6, 8), (8, 10)])
def test_oracle_frontier_counts_domino_tilings(shape):
    """Kasteleyn's product formula for domino tilings pins the oracle on the
    auto plan of an open grid's biadjacency block."""
    from permhip import grids

    masks = prep.row_masks(grids.grid_biadjacency(shape))
    plan = prep.auto_frontier_plan(masks)
    assert plan is not None
    assert oracle.frontier_count(plan[0], plan[1]) == _kasteleyn(*shape)


@pytest.mark.parametrize("row_order", [[0, 1], [0, 1, 1], [0, 1, 3], [0, -1, 2], [0, True, 2],
                                       [0, np.bool_(True), 2], [0, 1.0, 2], 7])
def test_frontier_rejects_the_reference_row_orders(row_order):
    """The row orders of tests/test_binary_frontier.py:104-122."""
    with pytest.raises((TypeError, ValueError), match="row_order"):
        perm_binary.binary_permanent_frontier([[1, 0, 0], [0, 1, 0], [0, 0, 1]], row_order=row_order)


def test_frontier_hostile_matrix_semantics():
    """Rows and matrices whose iteration disagrees with their length
    (tests/test_binary_frontier.py:144-172)."""

    class ShortRow:
        def __len__(self):
            return 2

        def __iter__(self):
            return iter([1])

    with pytest.raises(ValueError, match="row width changed"):
        perm_binary.binary_permanent_frontier([ShortRow(), ShortRow()])

    class ShortMatrix:
        def __len__(self):
            return 2

        def __iter__(self):
            return iter([[1, 0]])

    with pytest.raises(ValueError, match="row count changed"):
        perm_binary.binary_permanent_frontier(ShortMatrix())
    with pytest.raises(ValueError, match="0 or 1"):
        perm_binary.binary_permanent_frontier([[-7, 0], [4, 9]])
