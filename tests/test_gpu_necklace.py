"""Necklace-orbit evaluation of Schur cofactor minors on the device against
the Glynn path, the oracle and known perm(S_n) (bit-exact residues)."""

import pytest

pytestmark = pytest.mark.gpu

# OEIS A003112 and the reference's checked-in values (schur_permanents.txt).
KNOWN = {5: -5, 7: -105, 11: 6765, 13: 175747, 17: 25219857, 19: 142901109,
         23: -31152650265, 29: -1300425712598285, 31: 126691467546591,
         37: -339696021338069660245, 41: 1617562409111471411255401}


def _with_tuning(native, key, value, fn):
    old = native.get_tuning(key)
    native.set_tuning(key, value)
    try:
        return fn()
    finally:
        native.set_tuning(key, old)


@pytest.mark.parametrize("n", [5, 7, 11, 13, 17, 19, 23, 29, 31])
def test_necklace_matches_glynn_and_known_values(native, n):
    from permhip import planner

    for p in planner.find_primes_congruent_to_1_mod_n(n, 2, 1 << 30) + [
            planner.find_primes_congruent_to_1_mod_n(n, 1, 1000)[0]]:
        w = planner.nth_root_of_unity_mod_p(n, p)
        neck = _with_tuning(native, native.TUNE_NECKLACE_MIN_ORDER, 5,
                            lambda: native.schur_permanent_modp(n, w, p))
        glynn = _with_tuning(native, native.TUNE_NECKLACE_MIN_ORDER, 64,
                             lambda: native.schur_permanent_modp(n, w, p))
        assert neck == glynn == KNOWN[n] % p, (n, p)


def test_necklace_against_oracle_small_orders(native, oracle_c):
    from permhip import planner

    for n in (5, 7, 11, 13):
        for p in planner.find_primes_congruent_to_1_mod_n(n, 3, 200):
            w = planner.nth_root_of_unity_mod_p(n, p)
            S = planner.build_schur_matrix_mod_p(n, w, p)
            got = _with_tuning(native, native.TUNE_NECKLACE_MIN_ORDER, 5,
                               lambda: native.schur_permanent_modp(n, w, p))
            assert got == oracle_c.permanent_mod_p(S, p, 8), (n, p)


@pytest.mark.parametrize("depth", [2, 19, 20, 28])
def test_task_depth_does_not_change_residues(native, depth):
    """The task depth is clamped to [L/2, L - 14]; for n = 37 that is 18..22."""
    from permhip import planner

    n = 37
    p = planner.find_primes_congruent_to_1_mod_n(n, 1, 1 << 30)[0]
    w = planner.nth_root_of_unity_mod_p(n, p)
    got = _with_tuning(native, native.TUNE_NECKLACE_DEPTH, depth,
                       lambda: native.schur_permanent_modp(n, w, p))
    assert got == KNOWN[n] % p


@pytest.mark.parametrize("n", [37, 41])
def test_exact_schur_permanents_through_necklaces(native, n):
    import perm

    assert perm.permanent(n) == KNOWN[n]


@pytest.mark.parametrize("n", [5, 7, 11, 13, 17, 19, 23, 29, 31])
def test_necklace_with_32_bit_primes(native, n):
    """Primes in [2^31, 2^32) take the unsigned Montgomery kernel: the two
    smallest = 1 (mod n) above 2^31 and the largest below 2^32 (the
    reference's device planner, perm_cuda.py:289-302)."""
    import perm_cuda
    from permhip import planner

    primes = planner.find_primes_congruent_to_1_mod_n(n, 2, 1 << 31) + perm_cuda._find_primes_for_schur(n, 1)
    for p in primes:
        assert (1 << 31) < p < (1 << 32) and p % n == 1
        w = planner.nth_root_of_unity_mod_p(n, p)
        neck = _with_tuning(native, native.TUNE_NECKLACE_MIN_ORDER, 5,
                            lambda: native.schur_permanent_modp(n, w, p))
        glynn = _with_tuning(native, native.TUNE_NECKLACE_MIN_ORDER, 64,
                             lambda: native.schur_permanent_modp(n, w, p))
        assert neck == glynn == KNOWN[n] % p, (n, p)


@pytest.mark.parametrize("n,omega,p,expected", [
    (41, 782_766_686, 4_294_966_153, 3_889_040_436),
    (47, 1_002_938_914, 4_294_967_161, 4_173_217_525),
])
def test_reference_device_golden_residues(native, n, omega, p, expected):
    """Known answers of the reference's device tests
    (tests/test_cuda_device_schur_glynn.py:147-169): 32-bit primes at the
    two orders its necklace kernel serves."""
    assert pow(omega, n, p) == 1
    assert native.schur_permanent_modp(n, omega, p) == expected


def _rotation_complement_orbits(length):
    """Orbits of {0,1}^length under rotation and complement, by Burnside's
    lemma over the dihedral-free group Z_length x Z_2 (rotation by k fixes
    2^gcd(k, length) words; rotation by k composed with complement fixes
    2^gcd(k, length) words when length / gcd(k, length) is even, else none)."""
    import math

    fixed = 0
    for k in range(length):
        g = math.gcd(k, length)
        fixed += 2 ** g
        if (length // g) % 2 == 0:
            fixed += 2 ** g
    return fixed // (2 * length)


@pytest.mark.parametrize("n", [3, 5, 7, 11, 13, 17, 19, 23, 29, 31])
def test_necklace_orbit_stats_cover_every_sign_word(native, n):
    """perm_c_ext.schur_necklace_orbit_stats (perm_c_omp.c:1062-1080;
    the reference's test tests/test_schur_cpu_glynn.py:101-130): one
    representative per orbit, and the orbit weights cover all 2^(n-1)
    sign words exactly once."""
    from permhip.backends import perm_c_ext

    reps, words = perm_c_ext.schur_necklace_orbit_stats(n)
    assert reps == _rotation_complement_orbits(n - 1)
    assert words == 1 << (n - 1)


def test_necklace_orbit_stats_at_order_41(native):
    reps, words = native.schur_necklace_orbit_stats(41)
    assert reps == _rotation_complement_orbits(40)
    assert words == 1 << 40
