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
