"""Seeded random sweep across every residue path: orders 5-22, entries from
0/1 up to +-2^40, and moduli from every tier (the binary planner's and the
integer planner's 31-bit primes, odd 31- and 32-bit moduli, 2^61 - 1, odd
moduli up to 2^63, even moduli), several per call, against the oracle's C
restatement."""

import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _modulus(rng):
    from permhip import binary_prep, planner

    kind = rng.randrange(7)
    if kind == 0:
        return binary_prep.descending_prime(rng.randrange(16))
    if kind == 1:
        return planner._ascending_prime(rng.randrange(16))
    if kind == 2:
        return rng.randrange(3, 1 << 31) | 1
    if kind == 3:
        return rng.randrange(1 << 31, 1 << 32) | 1
    if kind == 4:
        return (1 << 61) - 1
    if kind == 5:
        return rng.randrange(1 << 40, 1 << 63) | 1
    return rng.randrange(2, 1 << 20) * 2


@pytest.mark.parametrize("seed", range(4))
def test_random_sweep_against_oracle(native, oracle_c, seed):
    rng = random.Random(777 + seed)
    for case in range(75):
        n = rng.randrange(5, 21)
        r = rng.choice([1, 2, 10, 100, 10 ** 6, 2 ** 40])
        if rng.random() < 0.3:
            density = rng.choice([0.3, 0.5, 0.7])
            A = np.array([[rng.random() < density for _ in range(n)] for _ in range(n)], dtype=np.int64)
        else:
            A = np.array([[rng.randrange(-r, r + 1) for _ in range(n)] for _ in range(n)], dtype=np.int64)
        moduli = [_modulus(rng) for _ in range(rng.choice([1, 2, 3, 4, 5, 6, 8, 9]))]
        got = native.permanent_residues(A, moduli)
        for m, g in zip(moduli, got):
            assert g == oracle_c.permanent_mod_p(A.tolist(), m, 16), (seed, case, n, r, m)
