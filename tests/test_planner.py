"""Host planning and reconstruction against the reference's pinned values
(CPU only)."""

import math
import random

import pytest

import oracle
from permhip import planner

# OEIS A003112 (tests/test_reference_values.py:15-43 of the reference).
OEIS_A003112 = {
    1: 1, 3: -3, 5: -5, 7: -105, 9: 81, 11: 6765, 13: 175747, 15: 30375,
    17: 25219857, 19: 142901109, 21: 4548104883, 23: -31152650265,
    25: -5198937484375, 27: 65230244418933, 29: -1300425712598285,
    31: 126691467546591, 33: 868088125376401545, 35: -15139017417029296875,
}
PROJECT_EXTENSIONS = {
    37: -339696021338069660245,
    39: -29260070859488572505025,
    41: 1617562409111471411255401,
    43: -23811413912733485902431027,
    45: -470483360490107435783203125,
    47: -52876272639843868105982022705,
}
# tests/test_schur_cpu_glynn.py:32-43 of the reference.
THEOREM4_DIVISORS = {
    3: 3, 9: 81, 15: 30375, 21: 750141, 25: 15625, 27: 1594323,
    33: 235782657, 35: 1313046875, 39: 3502727631, 45: 3 ** 20 * 5 ** 9,
}


@pytest.mark.parametrize(
    "order,expected_count",
    [(31, 3), (37, 3), (39, 3), (41, 4), (43, 4), (45, 3), (47, 4)],
)
def test_schur_plan_coverage_and_minimal_pass_count(order, expected_count):
    target = 2 * planner.schur_permanent_bound(order) + 1
    primes, product, congruence = planner.plan_schur_exact_crt(order, target, planner.SCHUR_CRT_PRIME_FLOOR)
    theorem = congruence[1] if congruence else 1
    assert len(primes) == expected_count
    assert min(primes) >= 1 << 30
    assert all(p % order == 1 and planner.is_prime(p) for p in primes)
    assert all(math.gcd(p, theorem) == 1 for p in primes)
    assert theorem * product >= target
    assert theorem * math.prod(primes[:-1]) < target


@pytest.mark.parametrize("order,divisor", sorted(THEOREM4_DIVISORS.items()))
def test_theorem4_divisors(order, divisor):
    residue, modulus = planner.graham_lehmer_schur_congruence(order)
    assert (residue, modulus) == (0, divisor)
    value = {**OEIS_A003112, **PROJECT_EXTENSIONS}.get(order)
    if value is not None:
        assert value % divisor == 0


def test_graham_lehmer_prime_orders_match_known_values():
    for order, value in {**OEIS_A003112, **PROJECT_EXTENSIONS}.items():
        if order >= 5 and planner.is_prime(order):
            residue, modulus = planner.graham_lehmer_schur_congruence(order)
            assert modulus == order ** 3
            assert value % modulus == residue == math.factorial(order) % modulus
    assert planner.graham_lehmer_schur_congruence(2) is None
    assert planner.graham_lehmer_schur_congruence(8) is None


def test_schur_bound_is_exact_ceiling():
    for n in range(0, 60):
        b = planner.schur_permanent_bound(n)
        assert (b - 1) ** 2 < n ** n <= b ** 2 or n == 0


@pytest.mark.parametrize("n", [3, 5, 7, 9, 11, 13, 15])
def test_schur_reconstruction_through_oracle_residues(n):
    """Planner + minor identity + CRT reproduce OEIS values with the C oracle
    computing each residue on the full S_n."""
    bound = planner.schur_permanent_bound(n)
    primes, product, cong = planner.plan_schur_exact_crt(n, 2 * bound + 1, n + 1)
    residues = []
    for p in primes:
        w = planner.nth_root_of_unity_mod_p(n, p)
        residues.append(oracle.C.permanent_mod_p(planner.build_schur_matrix_mod_p(n, w, p), p, 2))
        # Cofactor identity on the minor.
        minor = planner.schur_minor_mod_p(n, w, p)
        assert n * oracle.C.permanent_mod_p(minor, p, 2) % p == residues[-1]
    moduli = list(primes)
    if cong:
        residues.append(cong[0])
        moduli.append(cong[1])
    value = planner.signed_lift(
        planner.chinese_remainder_theorem(residues, moduli), bound, product * (cong[1] if cong else 1)
    )
    assert value == OEIS_A003112[n]


def test_roots_of_unity_have_exact_order():
    for n, p in [(31, 1073741971), (5, 11), (7, 29), (13, 53)]:
        if not planner.is_prime(p) or (p - 1) % n:
            continue
        w = planner.nth_root_of_unity_mod_p(n, p)
        assert pow(w, n, p) == 1
        assert all(pow(w, k, p) != 1 for k in range(1, n))
    with pytest.raises(ValueError):
        planner.nth_root_of_unity_mod_p(7, 31)


def test_generic_prime_plan():
    for rows, count in [([[10] * 20 for _ in range(20)], None), ([[1]], 1)]:
        bound = planner.matrix_permanent_bound(rows)
        primes, product = planner.generic_crt_primes(bound)
        assert product >= 2 * bound + 1
        assert math.prod(primes[:-1]) < 2 * bound + 1
        assert primes[0] == 1073741827
        assert all(planner.is_prime(p) for p in primes)
    from permhip import workloads

    for A, k in ((workloads.config1(), 5), (workloads.headline(), 8)):
        primes, _ = planner.generic_crt_primes(planner.matrix_permanent_bound(A.tolist()))
        assert len(primes) == k


def test_crt_and_signed_lift():
    rng = random.Random(4)
    moduli = [1073741827, 1073741831, 1073741833]
    M = math.prod(moduli)
    for _ in range(100):
        x = rng.randrange(-(M // 2), M // 2)
        r = planner.chinese_remainder_theorem([x % m for m in moduli], moduli)
        assert planner.signed_lift(r, M // 2, M) == x
    with pytest.raises(ValueError, match="coprime"):
        planner.chinese_remainder_theorem([1, 2], [6, 9])
    assert planner.chinese_remainder_theorem([], []) == 0


def test_is_prime_against_trial_division():
    for v in range(0, 5000):
        trial = v > 1 and all(v % d for d in range(2, int(v ** 0.5) + 1))
        assert planner.is_prime(v) == trial
    assert not planner.is_prime(341_550_071_728_321)
    assert planner.is_prime((1 << 61) - 1)


def test_reference_device_golden_residues_match_known_permanents():
    """The reference's device known answers (tests/test_cuda_device_schur_glynn.py:
    141-169) are the pinned exact permanents reduced: they check each other,
    and the omegas are roots of unity of the stated order."""
    exact = {**OEIS_A003112, **PROJECT_EXTENSIONS}
    for n, omega, p, expected in [
        (33, 2_662_262, 4_294_967_161, 1_428_739_342),
        (41, 782_766_686, 4_294_966_153, 3_889_040_436),
        (47, 1_002_938_914, 4_294_967_161, 4_173_217_525),
    ]:
        assert planner.is_prime(p) and p % n == 1
        assert pow(omega, n, p) == 1 and omega != 1
        assert exact[n] % p == expected, n


def test_crt_known_moduli_solves_the_congruences():
    """The cached-constant CRT returns the unique solution in [0, prod):
    it satisfies every congruence and equals the incremental solution."""
    import math
    import random

    rng = random.Random(5)
    primes = [planner._ascending_prime(i) for i in range(12)]
    for k in range(0, 13):
        moduli = tuple(primes[:k])
        for _ in range(20):
            res = [rng.randrange(m) for m in moduli]
            got = planner.crt_known_moduli(res, moduli)
            assert 0 <= got < math.prod(moduli)
            assert all(got % m == r for r, m in zip(res, moduli))
            assert got == planner.chinese_remainder_theorem(res, moduli)
    assert planner.crt_known_moduli([3, 4], (5, 7)) == 18
    with pytest.raises(ValueError, match="pairwise coprime"):
        planner.crt_known_moduli([1, 2], (6, 9))
    with pytest.raises(ValueError, match="same length"):
        planner.crt_known_moduli([1], (5, 7))


def test_bulk_bounds_equal_single_bounds():
    """matrix_permanent_bounds (stacked per order) equals
    matrix_permanent_bound on every array: mixed orders, zero rows
    (max(1, 0) = 1), entries near int64's ends (the single version's exact
    fallback) and the empty matrix."""
    import numpy as np

    rng = np.random.default_rng(8)
    arrays = []
    for _ in range(300):
        n = int(rng.integers(0, 12))
        bound = int(rng.choice((1, 10, 1000, 1 << 40)))
        a = rng.integers(-bound, bound + 1, (n, n)).astype(np.int64)
        if n and rng.random() < 0.2:
            a[int(rng.integers(n))] = 0
        arrays.append(a)
    big = np.full((3, 3), -(1 << 63), dtype=np.int64)
    big[1, 1] = (1 << 63) - 1
    arrays.append(big)
    assert planner.matrix_permanent_bounds(arrays) == [planner.matrix_permanent_bound(a) for a in arrays]
    assert planner.matrix_permanent_bound(big) == (3 << 63) * ((3 << 63) - 1) * (3 << 63)
