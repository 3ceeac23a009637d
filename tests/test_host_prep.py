"""Host preprocessing in the library (phip_binary_blocks, no device): the
perfect matching and irreducible blocks must equal the Python restatement
of perm_binary.py:709-752 block for block, in the same order; and the cached
prime sequences must equal a fresh scan."""

import random

import numpy as np
import pytest

from permhip import _native, binary_prep, planner


def _random_masks(rng, n, density):
    return [sum(1 << j for j in range(n) if rng.random() < density) for _ in range(n)]


@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 8, 13, 21, 34, 47, 63])
def test_native_blocks_equal_python(n):
    rng = random.Random(900 + n)
    for density in (0.05, 0.15, 0.3, 0.6, 1.0):
        for _ in range(20):
            masks = _random_masks(rng, n, density)
            if rng.random() < 0.5 and n:
                # Force a perfect matching through a random permutation.
                perm = list(range(n))
                rng.shuffle(perm)
                masks = [m | (1 << perm[i]) for i, m in enumerate(masks)]
            assert binary_prep.matching_blocks(masks) == binary_prep.matching_blocks_reference(masks), (n, density)


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
