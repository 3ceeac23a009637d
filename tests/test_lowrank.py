"""Low-rank branch of matrix_permanent on the CPU (host code; the device is
replaced by the oracle-backed fake where the generic path is reached).

Pins, each independent of the module's own arithmetic: the permutation
definition / subset DP on random rank-r products, the rank-one closed form
n! prod(u) prod(v), the two-row-type closed form m!(n-m)! [t^m]
prod_j (z_j + t v_j), and sI+J = sum_k C(n,k) (n-k)! s^k; plus the
reference's dispatch cases (tests/test_low_rank_dispatch.py:61-158)."""

import math
import time

import numpy as np
import pytest

import oracle
from test_host_fake_device import fake_device  # noqa: F401  (fixture)

PRIMES = [1073741827, 1073741831, 1073741833, (1 << 61) - 1]


def _two_type_closed_form(n, m, v, z):
    """per of the matrix whose first m rows are v and the rest z."""
    poly = [1]
    for vj, zj in zip(v, z):
        nxt = [0] * (len(poly) + 1)
        for i, c in enumerate(poly):
            nxt[i] += c * zj
            nxt[i + 1] += c * vj
        poly = nxt
    return math.factorial(m) * math.factorial(n - m) * poly[m]


def test_random_low_rank_products_match_definition():
    from permhip import lowrank

    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(300):
        n = int(rng.integers(1, 11))
        r = int(rng.integers(1, 5))
        A = (rng.integers(-6, 7, (n, r)) @ rng.integers(-6, 7, (r, n))).tolist()
        got = lowrank.low_rank_residues(A, PRIMES, 4, required=False)
        if got is None:
            continue
        exact = oracle.subset_permanent(A)
        assert got == [exact % p for p in PRIMES]
        checked += 1
    assert checked > 250


def test_closed_forms_at_large_orders():
    from permhip import lowrank

    n = 60
    u = [-2] + [i % 3 + 1 for i in range(1, n)]
    v = [i % 5 + 1 for i in range(n)]
    A = [[x * y for y in v] for x in u]
    expect = math.factorial(n) * math.prod(u) * math.prod(v)
    assert lowrank.low_rank_residues(A, PRIMES, 1, required=True) == [expect % p for p in PRIMES]

    n, m = 48, 19
    rng = np.random.default_rng(12)
    v = [int(x) for x in rng.integers(-9, 10, n)]
    z = [int(x) for x in rng.integers(-9, 10, n)]
    B = [list(v) for _ in range(m)] + [list(z) for _ in range(n - m)]
    expect = _two_type_closed_form(n, m, v, z)
    assert lowrank.low_rank_residues(B, PRIMES, 2, required=True) == [expect % p for p in PRIMES]


def test_auto_answers_rank_one_and_two_fast_without_the_device(fake_device, monkeypatch):  # noqa: F811
    import perm

    def no_device(*a, **k):
        raise AssertionError("auto must take the low-rank path")

    monkeypatch.setattr(fake_device, "permanent_residues", no_device)
    n = 60
    u = [-2] + [i % 3 + 1 for i in range(1, n)]
    v = [i % 5 + 1 for i in range(n)]
    A = [[x * y for y in v] for x in u]
    t0 = time.perf_counter()
    got = perm.matrix_permanent(A)
    assert time.perf_counter() - t0 < 1.0
    assert got == math.factorial(n) * math.prod(u) * math.prod(v)
    assert got == perm.matrix_permanent(A, method="low_rank")

    n, m = 48, 20
    rng = np.random.default_rng(13)
    v = [int(x) for x in rng.integers(-9, 10, n)]
    z = [int(x) for x in rng.integers(-9, 10, n)]
    B = [list(v) for _ in range(m)] + [list(z) for _ in range(n - m)]
    t0 = time.perf_counter()
    got = perm.matrix_permanent(B)
    assert time.perf_counter() - t0 < 1.0
    assert got == _two_type_closed_form(n, m, v, z) == perm.matrix_permanent(B, method="low_rank")

    # Orders of 64 and more are answered when the rank is small, refused
    # with the native message otherwise.
    big = [[(i % 4 + 1) * (j % 3 - 1) for j in range(70)] for i in range(70)]
    assert perm.matrix_permanent(big) == 0  # a column of zeros
    w = [(j % 3) + 1 for j in range(70)]
    big = [[(i % 4 + 1) * w[j] for j in range(70)] for i in range(70)]
    assert perm.matrix_permanent(big) == math.factorial(70) * math.prod(i % 4 + 1 for i in range(70)) * math.prod(w)
    with pytest.raises(ValueError, match="less than 64"):
        perm.matrix_permanent([[int(i == j) for j in range(64)] for i in range(64)])


def test_exceptional_first_prime_falls_back_to_the_device(fake_device, monkeypatch):  # noqa: F811
    import perm
    from permhip import lowrank, planner

    n = 20
    first = planner.generic_crt_primes(1)[0][0]
    A = [[1 + first * (i == j) for j in range(n)] for i in range(n)]
    expect = sum(math.comb(n, k) * math.factorial(n - k) * first ** k for k in range(n + 1))
    assert lowrank.factor_residues(A, [first], 4)[0][1].shape[0] == 1  # rank one mod the first prime
    calls = []
    real = fake_device.permanent_residues

    def counting(matrix, moduli):
        calls.append(list(moduli))
        return real(matrix, moduli)

    monkeypatch.setattr(fake_device, "permanent_residues", counting)
    monkeypatch.setattr(lowrank, "rank_limit", lambda order, moduli: 4)
    monkeypatch.setattr(lowrank, "permanents_from_factors",
                        lambda *a: pytest.fail("every prime is factored before any coefficient work"))
    assert perm.matrix_permanent(A) == expect
    assert calls and calls[0][0] == first


def test_explicit_limits_and_base_cases(fake_device, monkeypatch):  # noqa: F811
    import perm
    from permhip import lowrank

    prime = 1_073_741_827
    assert perm.matrix_permanent([[prime, 0], [0, -7]], method="low_rank") == -7 * prime
    with pytest.raises(perm.LowRankLimitExceeded, match="rank is at least 5"):
        perm.matrix_permanent([[int(i == j) for j in range(5)] for i in range(5)], method="low_rank")
    assert perm.matrix_permanent([[1, 2], [3, 4]], method="low_rank") == 10
    assert perm.matrix_permanent([], method="low_rank") == 1
    monkeypatch.setattr(lowrank, "LOW_RANK_MAX_STATES", 2)
    with pytest.raises(perm.LowRankLimitExceeded, match="states"):
        perm.matrix_permanent([[1, 0], [0, 1]], method="low_rank")


def test_dispatch_skips_probe_for_small_orders_and_generic(fake_device, monkeypatch):  # noqa: F811
    import perm
    from permhip import lowrank

    monkeypatch.setattr(lowrank, "low_rank_residues", lambda *a, **k: pytest.fail("must not probe"))
    assert perm.matrix_permanent([[1, 2], [3, 4]]) == 10
    small = [[1] * 12 for _ in range(12)]
    assert perm.matrix_permanent(small) == math.factorial(12)
    ones = [[1] * 24 for _ in range(24)]
    assert perm.matrix_permanent(ones, method="generic") == math.factorial(24)


def test_only_budget_exceptions_trigger_the_fallback(fake_device, monkeypatch):  # noqa: F811
    import perm
    from permhip import lowrank

    def broken(*a, **k):
        raise RuntimeError("factor failed")

    monkeypatch.setattr(lowrank, "_rref_factor", broken)
    assert lowrank.rank_limit(40, 8) > 0
    with pytest.raises(RuntimeError, match="factor failed"):
        perm.matrix_permanent([[1] * 40 for _ in range(40)])


def test_rank_limit_cost_model():
    from permhip import lowrank

    assert lowrank.rank_limit(19, 5) == 0
    # The device is far cheaper than a rank-4 coefficient grid at order 24,
    # never available at order 64 and above.
    assert lowrank.rank_limit(24, 6) < 4
    assert lowrank.rank_limit(63, 20) == 4
    assert lowrank.rank_limit(100, 60) == 3  # rank 4 needs C(103, 3) > 50,000 states
    for n in range(20, 64):
        k = 8
        lim = lowrank.rank_limit(n, k)
        for r in range(1, lim + 1):
            assert lowrank.low_rank_seconds(n, r, k) < lowrank.device_seconds(n, k)
