"""Argument validation of the drop-in objects and modules, matching the
reference's pinned messages and exception types (CPU only: every case is
rejected before any device call)."""

import numpy as np
import pytest

INT64_MIN = -(1 << 63)
INT64_MAX = (1 << 63) - 1


@pytest.mark.parametrize(
    "matrix,modulus,threads,error,message",
    [
        (([1],), 101, 1, TypeError, "list of lists"),
        (1, 101, 1, TypeError, "list of lists"),
        ([(1,)], 101, 1, ValueError, "square"),
        ([[1, 2]], 101, 1, ValueError, "square"),
        ([[1], [2, 3]], 101, 1, ValueError, "square"),
        ([[0] * 64 for _ in range(64)], 101, 1, ValueError, "less than 64"),
        ([[1]], 0, 1, ValueError, "positive"),
        ([[1]], -3, 1, ValueError, "positive"),
        ([[1]], 1 << 63, 1, OverflowError, None),
        ([[1]], 101.0, 1, TypeError, None),
        ([[1.0]], 101, 1, TypeError, None),
        ([[1 << 63]], 101, 1, OverflowError, None),
        ([[INT64_MIN - 1]], 101, 1, OverflowError, None),
        ([[1]], 101, 1 << 31, OverflowError, None),
        ([[1]], 101, 1.0, TypeError, None),
    ],
)
def test_modular_adapter_validation(matrix, modulus, threads, error, message):
    """tests/test_generic_native.py:186-211 of the reference."""
    from permhip.backends import perm_c_ext

    for fn in (perm_c_ext.permanent_mod_p, perm_c_ext.ryser_permanent_mod_p):
        with pytest.raises(error, match=message):
            fn(matrix, modulus, threads)


@pytest.mark.parametrize(
    "matrix,prime,threads,error,message",
    [
        ([[1, 0]], 101, 0, ValueError, "square"),
        ([[1, 2], [0, 1]], 101, 0, ValueError, "0 or 1"),
        ([[1, 0], [0, 1]], 9, 0, ValueError, "prime"),
        ([[1, 0, 0], [0, 1, 0], [0, 0, 1]], 3, 0, ValueError, "greater than"),
        ([[1]], 2, 0, ValueError, "odd"),
        ([[1, 0], [0, 1]], 101, -1, ValueError, "num_threads"),
        ([[1]], True, 0, TypeError, "p must be an integer"),
        ([[1]], 101, True, TypeError, "num_threads must be an integer"),
    ],
)
def test_binary_adapter_validation(matrix, prime, threads, error, message):
    """tests/test_binary_native.py:95-122 of the reference."""
    from permhip.backends import perm_binary_ext

    with pytest.raises(error, match=message):
        perm_binary_ext.binary_permanent_mod_p(matrix, prime, threads)


def test_binary_adapter_order_limit():
    from permhip.backends import perm_binary_ext

    with pytest.raises(ValueError, match="at most 63"):
        perm_binary_ext.binary_permanent_mod_p([[0] * 64 for _ in range(64)], 101)


def test_module_level_validation():
    import perm
    import perm_binary
    import perm_mod3

    with pytest.raises(ValueError, match="square"):
        perm.matrix_permanent([[1, 2]])
    with pytest.raises(TypeError, match="integers"):
        perm.matrix_permanent([[1.5]])
    with pytest.raises(ValueError, match="method"):
        perm.matrix_permanent([[1]], method="fast")
    with pytest.raises(TypeError, match="method"):
        perm.matrix_permanent([[1]], method=3)
    with pytest.raises(ValueError, match="non-negative"):
        perm.matrix_permanent([[1]], num_threads=-1)
    assert perm.matrix_permanent([]) == 1
    assert perm.matrix_permanent([[-7]]) == -7
    with pytest.raises(ValueError, match="0 or 1"):
        perm_binary.binary_permanent([[2]])
    with pytest.raises(TypeError, match="parallel"):
        perm_binary.binary_permanent([[1]], parallel=1)
    with pytest.raises(ValueError, match="backend"):
        perm_binary.binary_permanent([[1]], backend="cuda")
    assert perm_binary.binary_permanent([]) == 1
    assert perm_binary.binary_permanent([[1, 0], [1, 0]]) == 0  # no perfect matching
    with pytest.raises(ValueError, match="odd prime"):
        perm_binary.binary_permanent_mod_p([[1]], 9)
    with pytest.raises(ValueError, match="square"):
        perm_mod3.permanent_mod3_fast(np.zeros((2, 3), dtype=np.int64))
    with pytest.raises(TypeError, match="integers"):
        perm_mod3.permanent_mod3_fast(np.zeros((2, 2)))
    assert perm_mod3.permanent_mod3_fast(np.zeros((0, 0), dtype=np.int64)) == 1
    assert perm_mod3.permanent_mod3_fast(np.array([[5]])) == 2


def test_binary_host_preprocessing_blocks_and_bounds():
    from permhip import binary_prep

    # Block triangular: blocks are found and bounded independently.
    rows = [[1, 1, 0, 0], [1, 1, 0, 0], [1, 0, 1, 1], [0, 1, 1, 1]]
    feasible, blocks = binary_prep.matching_blocks(binary_prep.row_masks(rows))
    assert feasible and sorted(len(b) for b in blocks) == [2, 2]
    assert binary_prep.crt_primes_above_bound(1_393_140_695_040_000_000) == (2147483647, 2147483629)
    assert not binary_prep.is_prime_u64(341_550_071_728_321)
    assert binary_prep.ceil_nth_root(10 ** 12, 3) == 10 ** 4
    assert binary_prep.ceil_nth_root(10 ** 12 + 1, 3) == 10 ** 4 + 1


INT64_MIN = -(1 << 63)


@pytest.mark.parametrize(
    "matrix,modulus,threads,error,message",
    [
        (([1],), 101, 1, TypeError, "list of lists"),
        (1, 101, 1, TypeError, "list of lists"),
        ([(1,)], 101, 1, ValueError, "square"),
        ([[1, 2]], 101, 1, ValueError, "square"),
        ([[1], [2, 3]], 101, 1, ValueError, "square"),
        ([[0] * 64 for _ in range(64)], 101, 1, ValueError, "less than 64"),
        ([[1]], 0, 1, ValueError, "positive"),
        ([[1]], -3, 1, ValueError, "positive"),
        ([[1]], 1 << 63, 1, OverflowError, None),
        ([[1]], 101.0, 1, TypeError, None),
        ([[1.0]], 101, 1, TypeError, None),
        ([[1 << 63]], 101, 1, OverflowError, None),
        ([[INT64_MIN - 1]], 101, 1, OverflowError, None),
        ([[1]], 101, 1 << 31, OverflowError, None),
        ([[1]], 101, 1.0, TypeError, None),
    ],
)
def test_perm_c_ext_object_rejects_like_the_reference(matrix, modulus, threads, error, message):
    """The perm_c_ext-shaped object validates as the reference's extension
    (tests/test_generic_native.py:186-211), before any device call."""
    from permhip.backends import perm_c_ext

    for implementation in (perm_c_ext.permanent_mod_p, perm_c_ext.ryser_permanent_mod_p):
        with pytest.raises(error, match=message):
            implementation(matrix, modulus, threads)


@pytest.mark.parametrize(
    "matrix,prime,threads,error,message",
    [
        ([[1, 0]], 101, 0, ValueError, "square"),
        ([[1, 2], [0, 1]], 101, 0, ValueError, "0 or 1"),
        ([[1, 0], [0, 1]], 9, 0, ValueError, "prime"),
        ([[1, 0, 0], [0, 1, 0], [0, 0, 1]], 3, 0, ValueError, "greater than"),
        ([[1]], 2, 0, ValueError, "odd"),
        ([[1, 0], [0, 1]], 101, -1, ValueError, "num_threads"),
        ([[1]], True, 0, TypeError, "p must be an integer"),
        ([[1]], 101, True, TypeError, "num_threads must be an integer"),
        ([[0] * 64 for _ in range(64)], 101, 0, ValueError, "at most 63"),
    ],
)
def test_perm_binary_ext_object_rejects_like_the_reference(matrix, prime, threads, error, message):
    """The _binary_ext-shaped object validates as the reference's extension
    (tests/test_binary_native.py:95-122), before any device call."""
    from permhip.backends import perm_binary_ext

    with pytest.raises(error, match=message):
        perm_binary_ext.binary_permanent_mod_p(matrix, prime, threads)


@pytest.mark.parametrize("order", [0, 2, 9, 25, 64])
def test_necklace_stats_reject_unsupported_orders(order):
    """perm_c_ext.schur_necklace_orbit_stats refuses orders that are not odd
    primes in [3, 64) before any device work (perm_c_omp.c:1066-1071)."""
    from permhip.backends import perm_c_ext

    with pytest.raises(ValueError, match="odd prime"):
        perm_c_ext.schur_necklace_orbit_stats(order)
    with pytest.raises(TypeError):
        perm_c_ext.schur_necklace_orbit_stats(5.0)
