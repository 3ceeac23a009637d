"""Host preprocessing in the library (no device): the perfect matching and
the irreducible blocks against a brute-force decomposition, and the cached
prime sequences against a fresh scan.

The brute force is independent of the library: an entry (r, c) lies on a
perfect matching exactly when the matrix without row r and column c has
one (scipy's maximum_bipartite_matching decides it), and the irreducible
blocks are the connected components of the bipartite graph of those
entries (Dulmage-Mendelsohn fine decomposition)."""

import random

import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components, maximum_bipartite_matching

from permhip import _native, binary_prep, planner


def _random_masks(rng, n, density):
    return [sum(1 << j for j in range(n) if rng.random() < density) for _ in range(n)]


def _has_perfect_matching(A):
    n = A.shape[0]
    if n == 0:
        return True
    m = maximum_bipartite_matching(csr_matrix(A), perm_type="column")
    return bool((m >= 0).all())


def _brute_blocks(A):
    """None without a perfect matching, else the set of blocks, each a
    (frozenset of rows, frozenset of columns)."""
    n = A.shape[0]
    if not _has_perfect_matching(A):
        return None
    allowed = []
    for r in range(n):
        for c in range(n):
            if A[r, c]:
                sub = np.delete(np.delete(A, r, axis=0), c, axis=1)
                if _has_perfect_matching(sub):
                    allowed.append((r, c))
    # Nodes: rows 0..n-1, columns n..2n-1.
    data = np.ones(len(allowed))
    rows = [r for r, _ in allowed]
    cols = [n + c for _, c in allowed]
    graph = csr_matrix((data, (rows, cols)), shape=(2 * n, 2 * n))
    _, label = connected_components(graph, directed=False)
    out = {}
    for v in range(2 * n):
        out.setdefault(label[v], (set(), set()))[0 if v < n else 1].add(v % n)
    return {(frozenset(r), frozenset(c)) for r, c in out.values()}


def _library_blocks(masks):
    n = len(masks)
    ptr, idx = binary_prep._csr(masks)
    got = _native.binary_blocks_csr(ptr, idx, n)
    if got is None:
        return None, None
    match, row_block, col_block, count = got
    blocks = {}
    for r in range(n):
        blocks.setdefault(row_block[r], (set(), set()))[0].add(r)
    for c in range(n):
        blocks.setdefault(col_block[c], (set(), set()))[1].add(c)
    assert sorted(blocks) == list(range(count))
    return match, {(frozenset(r), frozenset(c)) for r, c in blocks.values()}


@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 8, 13, 21, 34])
def test_blocks_equal_brute_force(n):
    rng = random.Random(900 + n)
    for density in (0.05, 0.15, 0.3, 0.6, 1.0):
        for _ in range(6):
            masks = _random_masks(rng, n, density)
            if rng.random() < 0.6 and n:
                # Force a perfect matching through a random permutation.
                perm = list(range(n))
                rng.shuffle(perm)
                masks = [m | (1 << perm[i]) for i, m in enumerate(masks)]
            A = np.array([[(m >> j) & 1 for j in range(n)] for m in masks], dtype=np.int8).reshape(n, n)
            expect = _brute_blocks(A)
            match, got = _library_blocks(masks)
            assert got == expect, (n, density)
            feasible, mask_blocks = binary_prep.matching_blocks(masks)
            assert feasible == (expect is not None)
            if expect is None:
                assert binary_prep.perfect_matching(masks) is None
                continue
            # A valid perfect matching.
            assert sorted(match) == list(range(n)) and all(A[r, c] for r, c in enumerate(match))
            assert binary_prep.perfect_matching(masks) == tuple(match)
            # The mask entry point: blocks by smallest row, local columns in
            # ascending order, each block the submatrix on its rows/columns.
            ordered = sorted(expect, key=lambda b: min(b[0]))
            assert len(mask_blocks) == len(ordered)
            for (rows, cols), local in zip(ordered, mask_blocks):
                cs = sorted(cols)
                want = tuple(sum(1 << k for k, c in enumerate(cs) if A[r, c]) for r in sorted(rows))
                assert local == want


@pytest.mark.parametrize("n", [47, 63, 64, 90, 200])
def test_large_orders_and_both_entry_points(n):
    """Mask and CSR entry points agree up to order 63; above it the CSR
    path is checked against a planted block structure."""
    rng = random.Random(n)
    sizes = []
    left = n
    while left:
        s = min(left, rng.randint(1, 12))
        sizes.append(s)
        left -= s
    perm_r = list(range(n))
    perm_c = list(range(n))
    rng.shuffle(perm_r)
    rng.shuffle(perm_c)
    A = np.zeros((n, n), dtype=np.int8)
    start = 0
    for s in sizes:
        # A full s x s square is fully indecomposable; entries above the
        # diagonal blocks keep the structure block triangular.
        A[start:start + s, start:start + s] = 1
        A[start:start + s, start + s:] = np.array([[rng.random() < 0.1 for _ in range(n - start - s)]
                                                 for _ in range(s)], dtype=np.int8).reshape(s, n - start - s)
        start += s
    B = A[np.ix_(perm_r, perm_c)]
    masks = [int(sum(1 << j for j in range(n) if B[i, j])) for i in range(n)]
    feasible, blocks = binary_prep.matching_blocks(masks)
    assert feasible
    assert sorted(len(b) for b in blocks) == sorted(sizes)
    assert all(all(m == (1 << len(b)) - 1 for m in b) for b in blocks)
    _, got = _library_blocks(masks)
    inv_r = {r: i for i, r in enumerate(perm_r)}
    inv_c = {c: i for i, c in enumerate(perm_c)}
    expect, start = set(), 0
    for s in sizes:
        rows = frozenset(i for i in range(n) if start <= perm_r[i] < start + s)
        cols = frozenset(j for j in range(n) if start <= perm_c[j] < start + s)
        expect.add((rows, cols))
        start += s
    assert got == expect
    assert inv_r and inv_c
    # Removing one entry of a 1x1 block (if any) or any diagonal square
    # entry count keeps it consistent; a zero row has no matching.
    masks[0] = 0
    assert binary_prep.matching_blocks(masks) == (False, ())
    assert binary_prep.perfect_matching(masks) is None


def test_block_structure_invariants():
    rng = random.Random(77)
    n = 30
    # Block upper-triangular by construction: blocks are the diagonal squares.
    sizes = [4, 1, 7, 3, 9, 6]
    A = np.zeros((n, n), dtype=np.uint8)
    start = 0
    for s in sizes:
        A[start:start + s, start:start + s] = 1
        A[start:start + s, start + s:] = rng.random() < 0.5
        start += s
    masks = [int(sum(1 << j for j in range(n) if A[i, j])) for i in range(n)]
    ok, blocks = binary_prep.matching_blocks(masks)
    assert ok and sorted(len(b) for b in blocks) == sorted(sizes)
    assert all(all(m == (1 << len(b)) - 1 for m in b) for b in blocks)


def test_rejections():
    with pytest.raises(RuntimeError, match="beyond the order"):
        _native.binary_blocks([0b100, 0b1])
    assert _native.binary_blocks([0b01, 0b01]) == (False, ())
    ptr = np.array([0, 1, 2], dtype=np.int32)
    with pytest.raises(RuntimeError, match="outside the matrix"):
        _native.binary_blocks_csr(ptr, np.array([0, 2], dtype=np.int32), 2)
    with pytest.raises(RuntimeError, match="repeated column"):
        _native.binary_blocks_csr(np.array([0, 2, 3], dtype=np.int32), np.array([1, 1, 0], dtype=np.int32), 2)
    assert _native.binary_blocks_csr(ptr, np.array([0, 0], dtype=np.int32), 2) is None
    assert _native.binary_blocks_csr(np.zeros(1, dtype=np.int32), np.zeros(0, dtype=np.int32), 0) == ([], [], [], 0)


def test_cached_prime_sequences_match_scans():
    fresh, c = [], 1 << 30
    while len(fresh) < 40:
        if planner.is_prime(c):
            fresh.append(c)
        c += 1
    assert [planner._ascending_prime(i) for i in range(40)] == fresh
    down, c = [], (1 << 31) - 1
    while len(down) < 40:
        if binary_prep.is_prime_u64(c):
            down.append(c)
        c -= 2
    assert [binary_prep.descending_prime(i) for i in range(40)] == down


def test_normalize_ndarray_fast_path_matches_the_entry_loop():
    """Integer and boolean arrays are validated vectorised; the rows and the
    first offending entry's message equal those of the entry-by-entry loop
    (perm_binary.py:173-215) on the same data as nested lists."""
    import numpy as np

    from permhip import binary_prep as prep

    rng = np.random.default_rng(0)
    for _ in range(200):
        n = int(rng.integers(0, 12))
        A = rng.integers(-2, 3, (n, n))
        for strict in (True, False):
            for arr in (A, A.astype(np.int8), A != 0, np.abs(A).astype(np.uint8)):
                outcomes = []
                for form in (arr.tolist(), arr):
                    try:
                        outcomes.append(("ok", prep.normalize_binary_matrix(form, strict=strict)))
                    except (TypeError, ValueError) as exc:
                        outcomes.append((type(exc), str(exc)))
                assert outcomes[0] == outcomes[1]


def test_bound_from_masks_equals_bound_of_unpacked_rows():
    rng = random.Random(4242)
    for n in (0, 1, 2, 7, 8, 9, 31, 64, 65, 130):
        for density in (0.0, 0.1, 0.5, 1.0):
            masks = _random_masks(rng, n, density)
            rows = binary_prep.unpack_masks(masks)
            assert binary_prep.column_degrees(masks) == [sum(c) for c in zip(*rows)]
            assert binary_prep.normalized_bound_masks(masks) == binary_prep.normalized_bound(rows), (n, density)


def test_normalize_list_rows_keep_entry_loop_semantics():
    """Rows of plain 0/1 ints are taken as they are; any other row is
    validated entry by entry, so the first offending entry in row-major
    order names the error and non-strict mode maps nonzero to 1."""
    rng = random.Random(8)
    plants = [True, False, np.int64(1), np.int8(0), 2, -1, 1.5, "1", None]
    for _ in range(400):
        n = rng.randrange(1, 7)
        rows = [[rng.randrange(2) for _ in range(n)] for _ in range(n)]
        for _ in range(rng.randrange(3)):
            rows[rng.randrange(n)][rng.randrange(n)] = rng.choice(plants)
        if rng.random() < 0.3:
            rows = [tuple(r) for r in rows]
        for strict in (True, False):
            expected = None
            for r, row in enumerate(rows):
                for c, v in enumerate(row):
                    if isinstance(v, (float, str)) or v is None:
                        expected = (TypeError, f"entry ({r}, {c}) is {type(v).__name__}")
                    elif strict and int(v) not in (0, 1):
                        expected = (ValueError, f"entry ({r}, {c}) is {int(v)}")
                    if expected:
                        break
                if expected:
                    break
            if expected:
                with pytest.raises(expected[0]) as info:
                    binary_prep.normalize_binary_matrix(rows, strict=strict)
                assert str(info.value).endswith(expected[1])
            else:
                got = binary_prep.normalize_binary_matrix(rows, strict=strict)
                assert got == tuple(tuple(1 if int(v) else 0 for v in row) for row in rows)
                assert all(type(v) is int for row in got for v in row)


def test_row_masks_set_the_bits_of_nonzero_entries():
    rng = random.Random(31)
    for n in (0, 1, 63, 64, 65, 100, 257):
        for values in ((0, 1), (0, 1, 2, 255), (0, 1, -1, 300), (False, True)):
            rows = [tuple(rng.choice(values) for _ in range(n)) for _ in range(n)]
            expected = tuple(sum(1 << c for c in range(n) if rows[r][c]) for r in range(n))
            assert binary_prep.row_masks(rows) == expected, (n, values)
    ragged = [(1, 0) * 40, (1,) * 81]
    assert binary_prep.row_masks(ragged) == (sum(1 << c for c in range(0, 80, 2)), (1 << 81) - 1)
