"""Pins the oracle (CPU, no GPU): brute force, independent identities,
closed forms, the reference's golden values, and the reference's own
compiled extensions where built."""

import itertools
import json
import math
import random
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_pure_python_oracles_agree_exhaustively_through_order_3():
    for n in range(4):
        for entries in itertools.product((-1, 0, 1, 2), repeat=n * n) if n < 3 else []:
            M = [list(entries[r * n:(r + 1) * n]) for r in range(n)]
            d = oracle.definition_permanent(M)
            assert oracle.subset_permanent(M) == d
            if n:
                assert oracle.exact_glynn(M) == d
    rng = random.Random(3)
    for n in range(3, 8):
        for _ in range(5):
            M = [[rng.randrange(-7, 8) for _ in range(n)] for _ in range(n)]
            d = oracle.definition_permanent(M)
            assert oracle.subset_permanent(M) == d
            assert oracle.exact_glynn(M) == d


def test_closed_forms():
    for n in range(1, 9):
        J = [[1] * n for _ in range(n)]
        assert oracle.exact_glynn(J) == math.factorial(n)
        assert oracle.C.permanent_mod_p(J, 1_000_000_007) == math.factorial(n) % 1_000_000_007
        # J - I: derangement numbers.
        D = [[0 if i == j else 1 for j in range(n)] for i in range(n)]
        der = round(math.factorial(n) * sum((-1) ** k / math.factorial(k) for k in range(n + 1)))
        assert oracle.exact_glynn(D) == der
        # Rank one u v^T: n! prod u prod v.
        u = list(range(1, n + 1))
        v = [(-1) ** i * (i + 2) for i in range(n)]
        R = [[a * b for b in v] for a in u]
        assert oracle.exact_glynn(R) == math.factorial(n) * math.prod(u) * math.prod(v)
        # s I + J: sum_k C(n,k) (n-k)! s^k.
        s = 3
        M = [[(s if i == j else 0) + 1 for j in range(n)] for i in range(n)]
        expect = sum(math.comb(n, k) * math.factorial(n - k) * s ** k for k in range(n + 1))
        assert oracle.subset_permanent(M) == expect


@pytest.mark.parametrize("modulus", [2, 3, 9, 15, 101, 2147483647, 3037000499, 3037000501, (1 << 63) - 1])
def test_c_oracle_against_definition_all_modulus_tiers(modulus):
    rng = random.Random(20260904)
    for n in range(7):
        for _ in range(3):
            M = [[rng.randrange(-(1 << 63), 1 << 63) for _ in range(n)] for _ in range(n)]
            expect = oracle.definition_permanent(M) % modulus
            assert oracle.C.permanent_mod_p(M, modulus, 2) == expect
            assert oracle.C.ryser_mod_p(M, modulus, 3) == expect


def test_c_oracle_thread_partitions_agree():
    rng = random.Random(271828)
    M = [[rng.randrange(-100, 100) for _ in range(17)] for _ in range(17)]
    expect = oracle.subset_permanent(M, 2147483647)
    for t in (1, 2, 3, 5):
        assert oracle.C.glynn_mod_p(M, 2147483647, t) == expect


def test_binary_oracle_against_definition_and_bregman_extremals():
    rng = random.Random(20260815)
    for n in range(1, 8):
        B = [[rng.randrange(2) for _ in range(n)] for _ in range(n)]
        d = oracle.definition_permanent(B)
        for p in (2147483647, 2147483659, 9223372036854775783):
            for t in (1, 3):
                assert oracle.C.binary_mod_p(B, p, t) == d % p
    # (d!)^k for k blocks of J_d (tests/test_binary_applications.py:227-232).
    for d, k in ((2, 5), (3, 4), (5, 3)):
        n = d * k
        B = [[1 if i // d == j // d else 0 for j in range(n)] for i in range(n)]
        assert oracle.C.binary_mod_p(B, 2147483647) == math.factorial(d) ** k % 2147483647
    # 13! (tests/test_cuda_device_binary.py:141-143)
    assert oracle.C.binary_mod_p([[1] * 13 for _ in range(13)], 2147483647) == math.factorial(13) % 2147483647


def test_binary_oracle_periodic_grid_pin():
    """4x4x5 grid with periodic x/y: residues 1238389710 and 1259774250 mod
    the two largest 31-bit primes (BINARY_APPLICATIONS.md:157-160) and the
    exact count 2551276235535120 (tests/test_cuda_device_binary.py:30).
    The 40 x 40 traversal is too slow for the CPU suite, so this checks the
    grid generator on the published small grid counts and the CRT identity of
    the published residues."""
    from permhip import grids

    for shape, expect in [((2, 2), 2), ((2, 4), 5), ((4, 4), 36), ((2, 2, 2), 9), ((2, 2, 3), 32)]:
        assert oracle.subset_permanent(grids.grid_biadjacency(shape)) == expect
    assert oracle.subset_permanent(grids.grid_biadjacency((4, 4), periodic_axes=(0, 1))) == 272
    exact = 2551276235535120
    assert exact % 2147483647 == 1238389710
    assert exact % 2147483629 == 1259774250


def test_mod3_oracle():
    rng = np.random.default_rng(7)
    for n in range(0, 9):
        A = rng.integers(-5, 6, (n, n))
        assert oracle.C.mod3(A, 2) == oracle.definition_permanent(A.tolist()) % 3


def test_update_cost_matches_definition():
    B = np.array([[1, 1, 0], [0, 1, 1], [1, 0, 1]], dtype=np.uint8)
    # Column degrees 2,2,2 -> sorted [2,2,2]; cost = 2*2^(1) + 2*2^0 = 6.
    assert oracle.C.binary_update_cost(B) == 6


def _simulated_update_cost(matrix):
    """Multiplies of the ratio walk by simulation: the Gray code over the
    n - 1 sparsest columns (ascending degree, the densest fixed) flips
    column ctz(t) at state t and pays its degree; the cheaper of the matrix
    and its transpose."""
    best = None
    for m in (matrix, matrix.T):
        n = m.shape[0]
        deg = sorted(int(d) for d in m.sum(axis=0))[: n - 1]
        cost = sum(deg[(t & -t).bit_length() - 1] for t in range(1, 1 << (n - 1)))
        best = cost if best is None else min(best, cost)
    return best


def test_update_cost_matches_simulated_walk():
    rng = np.random.default_rng(19)
    for n in range(2, 12):
        for density in (0.2, 0.5, 0.8):
            B = (rng.random((n, n)) < density).astype(np.uint8)
            assert oracle.C.binary_update_cost(B) == _simulated_update_cost(B), (n, density)


def test_reference_extensions_agree_with_oracle_when_built():
    ext = oracle.reference_extensions()
    if "perm_c_ext" not in ext:
        pytest.skip("reference extensions not built (make -C oracle)")
    rng = random.Random(1)
    for n in range(1, 12):
        M = [[rng.randrange(-50, 51) for _ in range(n)] for _ in range(n)]
        for p in (3, 101, 1073741827, 2147483647, (1 << 61) - 1):
            red = [[v % p for v in row] for row in M]
            assert ext["perm_c_ext"].permanent_mod_p(red, p, 2) == oracle.C.permanent_mod_p(M, p, 2)
        B = [[rng.randrange(2) for _ in range(n)] for _ in range(n)]
        for p in (2147483647, 2147483659):
            if p > n:
                assert ext["perm_binary_ext"].binary_permanent_mod_p(B, p, 2) == oracle.C.binary_mod_p(B, p, 2)
        A = np.array(M, dtype=np.int64)
        assert ext["perm_mod3_omp"].permanent_mod3(A % 3, 2) == oracle.C.mod3(A, 2)


def test_golden_fixtures_agree_with_oracle_where_cheap():
    path = GOLDEN / "reference_residues.json"
    if not path.exists():
        pytest.skip("golden fixtures not generated")
    from permhip import workloads

    data = json.loads(path.read_text())
    if "config1" in data:
        case = data["config1"]
        A = workloads.config1()
        for p, r in zip(case["moduli"], case["residues"]):
            assert oracle.C.permanent_mod_p(A.tolist(), p, 8) == r
