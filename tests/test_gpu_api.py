"""The drop-in modules (perm, perm_binary, perm_mod3) on the device against
the oracle, closed forms and the reference's golden values."""

import json
import math
import random
from pathlib import Path

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "reference_residues.json"

OEIS_A003112 = {
    1: 1, 3: -3, 5: -5, 7: -105, 9: 81, 11: 6765, 13: 175747, 15: 30375,
    17: 25219857, 19: 142901109, 21: 4548104883, 23: -31152650265,
    25: -5198937484375, 27: 65230244418933, 29: -1300425712598285,
    31: 126691467546591,
}


@pytest.mark.parametrize("n", sorted(OEIS_A003112))
def test_schur_permanents_match_oeis(native, n):
    import perm

    assert perm.permanent(n) == OEIS_A003112[n]
    assert perm.schur_permanent(n) == OEIS_A003112[n]


def test_reference_schur_residue_order_33(native):
    """Known answer of the reference's device tests
    (tests/test_cuda_device_schur_glynn.py:141-146).  The prime is above 2^31
    and 33 is not prime, so the 32 x 32 cofactor minor runs on the wide Glynn
    kernel in 32-bit unsigned Montgomery arithmetic."""
    p, omega = 4_294_967_161, 2_662_262
    assert pow(omega, 33, p) == 1
    assert native.schur_permanent_modp(33, omega, p) == 1_428_739_342


def test_even_and_trivial_schur_orders(native):
    import perm

    assert [perm.permanent(n) for n in (0, 1, 2, 4, 10)] == [1, 1, 0, 0, 0]


def test_matrix_permanent_random_against_subset_dp(native):
    import perm

    rng = random.Random(20261203)
    for n in range(0, 13):
        for lo, hi in ((-10, 10), (0, 1), (-(10 ** 12), 10 ** 12)):
            A = [[rng.randrange(lo, hi + 1) for _ in range(n)] for _ in range(n)]
            assert perm.matrix_permanent(A) == oracle.subset_permanent(A), (n, lo)


def test_matrix_permanent_numpy_and_bigint_inputs(native):
    import perm

    A = np.array([[2 ** 62, -3], [5, 2 ** 62]], dtype=np.int64)
    assert perm.matrix_permanent(A) == 2 ** 124 - 15
    U = np.array([[2 ** 63, 1], [1, 2 ** 63]], dtype=np.uint64)
    assert perm.matrix_permanent(U) == 2 ** 126 + 1
    big = [[1 << 80, 3, 0], [2, 1 << 70, 5], [7, 0, -(1 << 90)]]
    assert perm.matrix_permanent(big) == oracle.definition_permanent(big)
    assert type(perm.matrix_permanent([[3]])) is int


def test_closed_forms(native):
    import perm

    for n in (5, 9, 14, 20):
        assert perm.matrix_permanent([[1] * n for _ in range(n)]) == math.factorial(n)
        u = list(range(1, n + 1))
        v = [(-1) ** i * (i % 5 + 1) for i in range(n)]
        R = [[a * b for b in v] for a in u]
        assert perm.matrix_permanent(R) == math.factorial(n) * math.prod(u) * math.prod(v)
        s = 7
        M = [[(s if i == j else 0) + 1 for j in range(n)] for i in range(n)]
        assert perm.matrix_permanent(M) == sum(
            math.comb(n, k) * math.factorial(n - k) * s ** k for k in range(n + 1)
        )


def test_low_rank_method_matches_generic(native):
    import perm

    n = 12
    u = [[(i * 7 + 3 * j) % 11 - 5 for j in range(2)] for i in range(n)]
    w = [[(i * 5 + j) % 7 - 3 for i in range(n)] for j in range(2)]
    A = [[sum(u[i][k] * w[k][j] for k in range(2)) for j in range(n)] for i in range(n)]
    assert perm.matrix_permanent(A, method="low_rank") == perm.matrix_permanent(A, method="generic")
    with pytest.raises(perm.LowRankLimitExceeded):
        rng = random.Random(1)
        perm.matrix_permanent([[rng.randrange(9) for _ in range(8)] for _ in range(8)], method="low_rank")


def test_golden_reference_residues(native):
    """Residues written by the reference's own compiled kernels."""
    if not GOLDEN.exists():
        pytest.skip("golden fixtures not generated")
    import perm
    from permhip import _native, workloads

    data = json.loads(GOLDEN.read_text())
    cases = {"config1": workloads.config1(), "headline": workloads.headline()}
    mats = workloads.batch4096()
    for key in data:
        if key.startswith("batch4096["):
            cases[key] = mats[int(key[len("batch4096["):-1])]
    for key, A in cases.items():
        if key not in data:
            continue
        case = data[key]
        assert _native.permanent_residues(A, case["moduli"]) == case["residues"], key
        assert perm.matrix_permanent(A.tolist()) == int(case["exact"]), key
    if "schur31" in data:
        case = data["schur31"]
        for p, r in zip(case["moduli"], case["residues"]):
            assert perm.schur_permanent_mod_p_fast(31, p) == r
    if "binary36" in data:
        import perm_binary
        import perm_mod3

        case = data["binary36"]
        B = workloads.binary36()
        assert _native.binary_permanent_residues(B, case["moduli"]) == case["residues"]
        assert perm_mod3.permanent_mod3_fast(B) == case["mod3"]
        assert perm_binary.binary_permanent(B.tolist()) % 3 == case["mod3"]


def test_binary_known_counts(native):
    import perm_binary
    from permhip import grids

    for shape, expect in [((2, 2), 2), ((2, 4), 5), ((4, 4), 36), ((4, 6), 281), ((2, 2, 2), 9),
                          ((2, 2, 3), 32), ((2, 2, 4), 121), ((2, 3, 4), 1845)]:
        assert perm_binary.binary_permanent(grids.grid_biadjacency(shape)) == expect, shape
    assert perm_binary.binary_permanent(grids.grid_biadjacency((4, 4), periodic_axes=(0, 1))) == 272
    assert perm_binary.binary_permanent(grids.grid_biadjacency((4, 4, 2), periodic_axes=(0, 1))) == 589185
    assert perm_binary.binary_permanent([[1] * 13 for _ in range(13)]) == math.factorial(13)
    for d, k in ((2, 5), (3, 4), (5, 3)):
        n = d * k
        B = [[1 if i // d == j // d else 0 for j in range(n)] for i in range(n)]
        assert perm_binary.binary_permanent(B) == math.factorial(d) ** k
        assert perm_binary.binary_permanent_bound(B) == math.factorial(d) ** k
    der = [1, 0]
    for n in range(2, 12):
        der.append((n - 1) * (der[-1] + der[-2]))
    for n in range(2, 12):
        D = [[0 if i == j else 1 for j in range(n)] for i in range(n)]
        assert perm_binary.binary_permanent(D) == der[n]


def test_periodic_4x4x5_grid_exact_count(native):
    """2,551,276,235,535,120 with residues 1,238,389,710 and 1,259,774,250
    modulo 2^31-1 and 2147483629 (BINARY_APPLICATIONS.md:155-162)."""
    import perm_binary
    from permhip import _native, grids

    G = grids.grid_biadjacency((4, 4, 5), periodic_axes=(0, 1))
    assert len(G) == 40
    assert perm_binary._crt_primes_above_bound(perm_binary.binary_permanent_bound(G)) == (2147483647, 2147483629)
    assert _native.binary_permanent_residues(np.array(G, dtype=np.uint8), [2147483647, 2147483629]) == [
        1238389710, 1259774250]
    assert perm_binary.binary_permanent(G) == 2551276235535120


def test_binary_random_and_mod_p(native):
    import perm_binary

    rng = random.Random(314159)
    for n in range(2, 12):
        for _ in range(3):
            B = [[rng.randrange(2) for _ in range(n)] for _ in range(n)]
            d = oracle.subset_permanent(B)
            assert perm_binary.binary_permanent(B) == d
            assert perm_binary.binary_permanent_mod_p(B, 101) == d % 101
            assert perm_binary.binary_permanent(B, strict=False) == d


def test_mod3_module(native):
    import perm_mod3

    rng = np.random.default_rng(123)
    for n in (0, 1, 2, 5, 10, 15, 20, 26):
        A = rng.integers(-10, 11, (n, n))
        assert perm_mod3.permanent_mod3_fast(A) == oracle.C.mod3(A, 8), n
    U = np.array([[2 ** 64 - 1, 1], [1, 1]], dtype=np.uint64)
    assert perm_mod3.permanent_mod3_fast(U) == (2 ** 64 - 1 + 1) % 3


def test_adapter_objects_as_reference_seams(native):
    from permhip.backends import perm_binary_ext, perm_c_ext, perm_mod3_omp

    assert perm_c_ext.permanent_mod_p([[2, -1], [3, 4]], 101, 0) == 5
    assert perm_c_ext.ryser_permanent_mod_p([[2, -1], [3, 4]], 101) == 5
    assert perm_c_ext.get_max_threads() >= 1
    assert perm_binary_ext.binary_permanent_mod_p([[1] * 8 for _ in range(8)], 2147483647, (1 << 31) - 1) == 40320
    assert perm_mod3_omp.permanent_mod3(np.array([[1, 1], [1, 1]])) == 2


@pytest.mark.parametrize("direct_cost", [0.0, 2.0 ** 30])
def test_served_results_equal_alone(native, direct_cost):
    """Every served answer is exact, with plans only (direct_cost 0) and with
    small requests on an idle server taking the single-matrix path."""
    from permhip.serve import PermanentServer

    rng = np.random.default_rng(55)
    reqs = []
    for i in range(40):
        if i % 2:
            n = int(rng.integers(3, 16))
            reqs.append(("integer", rng.integers(-50, 51, (n, n))))
        else:
            n = int(rng.integers(3, 18))
            reqs.append(("binary", (rng.random((n, n)) < 0.5).astype(np.int64)))
    server = PermanentServer(max_inflight=2, plan_cost=1e6, direct_cost=direct_cost)
    futures = [server.submit(m, k) for k, m in reqs]
    mixed = [f.result(timeout=60) for f in futures]
    # One at a time on the idle server: direct when small enough.
    alone = [server.submit(m, k).result(timeout=60) for k, m in reqs[:6]]
    server.close()
    assert alone == mixed[:6]
    for (kind, m), v in zip(reqs, mixed):
        assert v == oracle.subset_permanent(m.tolist()), kind


# --- the reference's GPU-side modules (perm_cuda, perm_binary_cuda) ----------

SCHUR_KNOWN = {3: -3, 5: -5, 7: -105, 9: 81, 11: 6765, 13: 175747, 15: 30375, 17: 25219857,
               19: 142901109, 21: 4548104883, 33: 868088125376401545}


@pytest.mark.parametrize("n", sorted(SCHUR_KNOWN))
def test_perm_cuda_schur_permanents(native, n):
    """Device CRT plan (descending 32-bit primes) end to end against
    OEIS A003112 and the reference's checked-in n = 33 value."""
    import perm_cuda

    assert perm_cuda.is_cuda_available()
    assert perm_cuda.permanent_cuda(n, verbose=False) == SCHUR_KNOWN[n]


def test_perm_cuda_residue_functions(native, oracle_c):
    import perm_cuda

    rng = np.random.default_rng(99)
    for n in (2, 3, 7, 12, 16):
        A = rng.integers(-(10 ** 12), 10 ** 12, (n, n))
        for p in (101, 1073741827, 4294967291, (1 << 61) - 1):
            assert perm_cuda.permanent_modp_cuda(A, p) == oracle_c.permanent_mod_p(A.tolist(), p, 8), (n, p)
        assert perm_cuda.permanent_mod3_cuda(A) == oracle_c.mod3(A, 8), n
    for n, p in ((5, 11), (9, 19), (15, 31)):
        from permhip import planner

        w = planner.nth_root_of_unity_mod_p(n, p)
        S = planner.build_schur_matrix_mod_p(n, w, p)
        assert perm_cuda.schur_permanent_modp_cuda(n, w, p) == oracle_c.permanent_mod_p(S, p, 8)


def test_perm_binary_cuda_counts(native):
    import perm_binary_cuda
    from permhip import grids

    assert perm_binary_cuda.is_binary_cuda_available()
    assert perm_binary_cuda.binary_permanent_cuda(grids.grid_biadjacency((4, 6))) == 281
    assert perm_binary_cuda.binary_permanent_cuda(grids.grid_biadjacency((2, 3, 4))) == 1845
    assert perm_binary_cuda.binary_permanent_cuda(
        grids.grid_biadjacency((4, 4, 2), periodic_axes=(0, 1))) == 589185
    assert perm_binary_cuda.binary_permanent_cuda([[1] * 13 for _ in range(13)]) == math.factorial(13)
    rng = random.Random(2718)
    for n in range(2, 14):
        B = [[rng.randrange(2) for _ in range(n)] for _ in range(n)]
        d = oracle.subset_permanent(B)
        assert perm_binary_cuda.binary_permanent_cuda(B) == d
        assert perm_binary_cuda.binary_permanent_mod_p_cuda(B, 2147483629) == d % 2147483629


def test_mod3_demo_functions(native, capsys):
    """perm_mod3's demo entry points (perm_mod3.py:386-457) run the mod-3
    kernel against the general residue kernel modulo 3 and report agreement."""
    import perm_mod3

    perm_mod3.test_random_matrices()
    out = capsys.readouterr().out
    assert out.count("match=True") == 4 and "match=False" not in out
    perm_mod3.benchmark_mod3(n=12, num_trials=3)
    out = capsys.readouterr().out
    assert "Correctness verified." in out and "MISMATCH" not in out


def test_full_size_invariants_of_the_headline_matrix(native):
    """At the headline size (32 x 32, 8 primes, 2^31 Gray states each):
    perm(A) = perm(A^T) = perm(P A Q) for permutation matrices P, Q; the
    permanent is linear in each row (row 5 split as u + v) and scales with
    a row (x3).  (Its residues are pinned to the reference's golden values
    in test_golden_reference_residues.)"""
    import perm
    from permhip import workloads

    A = workloads.headline()
    base = perm.matrix_permanent(A)
    assert perm.matrix_permanent(A.T.copy()) == base
    rng = np.random.default_rng(1)
    P, Q = rng.permutation(32), rng.permutation(32)
    assert perm.matrix_permanent(A[P][:, Q].copy()) == base
    u = rng.integers(-10, 11, 32)
    U, V = A.copy(), A.copy()
    U[5], V[5] = u, A[5] - u
    assert perm.matrix_permanent(U) + perm.matrix_permanent(V) == base
    S = A.copy()
    S[7] *= 3
    assert perm.matrix_permanent(S) == 3 * base


def test_full_size_invariants_of_binary36(native):
    """binary36 (36 x 36, 4 primes, 2^35 Gray states each): the count is
    invariant under transposition and row/column permutations, and its
    residue mod 3 equals the mod-3 kernel's."""
    import perm_binary
    import perm_mod3
    from permhip import workloads

    B = np.array(workloads.binary36(), dtype=np.int64)
    base = perm_binary.binary_permanent(B)
    assert perm_binary.binary_permanent(B.T.copy()) == base
    rng = np.random.default_rng(2)
    assert perm_binary.binary_permanent(B[rng.permutation(36)][:, rng.permutation(36)].copy()) == base
    assert perm_mod3.permanent_mod3_fast(B) == base % 3


def test_bench_roofline_mix_runs(native):
    """The headline roofline's synthetic mix kernel exists and runs."""
    import bench

    mix = bench.instruction_mix(bench.HEADLINE_N, 4, 8, True)
    rounds = [round(mix[t]) for t in native.MIX_TYPES]
    assert native.mix_microbench(rounds, iters=16) > 0
