"""Host logic of perm_cuda / perm_binary_cuda (the reference's GPU-side
Python modules) with library objects replaced by fakes, and the device CRT
plan against the reference's pinned values (tests/test_cuda_wrappers.py,
tests/test_binary_cuda_wrappers.py, tests/test_cuda_crt_planning.py).
No device is needed: a fake stands in for the library, as in the reference's
own tests; residues the fakes return come from the oracle."""

import ctypes
import math

import numpy as np
import pytest

import oracle
import perm_binary
import perm_binary_cuda
import perm_cuda

# tests/test_cuda_crt_planning.py:12-31 of the reference.
EXPECTED_GPU_PRIMES = {
    41: [4294966153, 4294965251, 4294964923],
    43: [4294966163, 4294965733, 4294965131, 4294964959],
    45: [4294966591, 4294966231, 4294965691],
    47: [4294967161, 4294964341, 4294962367, 4294961897],
}
THEOREM4_DIVISORS = {
    3: 3, 9: 81, 15: 30375, 21: 750141, 25: 15625, 27: 1594323,
    33: 235782657, 35: 1313046875, 39: 3502727631, 45: 3 ** 20 * 5 ** 9,
}
# Known perm(S_n): OEIS A003112 and the reference's schur_permanents.txt.
SCHUR = {
    3: -3, 5: -5, 7: -105, 9: 81, 11: 6765, 13: 175747, 15: 30375, 17: 25219857,
    31: 126691467546591, 33: 868088125376401545, 35: -15139017417029296875,
    41: 1617562409111471411255401, 45: -470483360490107435783203125,
    47: -52876272639843868105982022705,
}


class FakeMod3Library:
    def __init__(self, result=0):
        self.result = result
        self.values = None

    @staticmethod
    def cuda_runtime_available():
        return 1

    def permanent_mod3_matrix(self, pointer, n):
        self.values = np.ctypeslib.as_array(pointer, shape=(n * n,)).copy()
        return self.result


class FakeModPLibrary:
    def __init__(self, matrix_result=0, schur_result=0):
        self.matrix_result = matrix_result
        self.schur_result = schur_result
        self.values = None
        self.schur_args = None

    @staticmethod
    def cuda_runtime_available():
        return 1

    def permanent_modp_cuda(self, pointer, n, p):
        typed = ctypes.cast(pointer, ctypes.POINTER(ctypes.c_longlong))
        self.values = np.ctypeslib.as_array(typed, shape=(n * n,)).copy()
        return self.matrix_result

    def schur_permanent_modp_cuda(self, n, omega, p):
        self.schur_args = (n, omega, p)
        return self.schur_result


class FakeBinaryLibrary:
    """Answers with the oracle's subset DP."""

    def __init__(self, *, available=True, result=None):
        self.available = available
        self.result = result
        self.calls = []

    def cuda_runtime_available(self):
        return int(self.available)

    def binary_permanent_modp_cuda(self, pointer, order, modulus):
        typed = ctypes.cast(pointer, ctypes.POINTER(ctypes.c_uint8))
        matrix = np.ctypeslib.as_array(typed, shape=(order * order,)).reshape(order, order).astype(int).tolist()
        self.calls.append((matrix, order, modulus))
        if self.result is not None:
            return self.result
        return oracle.subset_permanent(matrix, modulus)


# --- perm_cuda: argument handling --------------------------------------------

def test_mod3_reduces_before_int32_conversion(monkeypatch):
    fake = FakeMod3Library(result=2)
    monkeypatch.setattr(perm_cuda, "_mod3_cuda", fake)
    huge = 1 << 100
    assert perm_cuda.permanent_mod3_cuda(np.array([[huge + 2, -huge], [3, -4]], dtype=object)) == 2
    np.testing.assert_array_equal(fake.values, [0, 2, 0, 2])


def test_modp_reduces_before_int64_conversion(monkeypatch):
    p = 4_294_967_291
    values = [1 << 100, -(1 << 100), (1 << 63) - 1, -(1 << 63)]
    fake = FakeModPLibrary(matrix_result=17)
    monkeypatch.setattr(perm_cuda, "_modp_cuda", fake)
    assert perm_cuda.permanent_modp_cuda(np.array(values, dtype=object).reshape(2, 2), p) == 17
    np.testing.assert_array_equal(fake.values, [v % p for v in values])


def test_schur_arguments_are_validated_and_normalized(monkeypatch):
    fake = FakeModPLibrary(schur_result=4)
    monkeypatch.setattr(perm_cuda, "_modp_cuda", fake)
    assert perm_cuda.schur_permanent_modp_cuda(3, -2, 7) == 4
    assert fake.schur_args == (3, 5, 7)


@pytest.mark.parametrize(
    "call,name,library",
    [
        (lambda: perm_cuda.permanent_mod3_cuda(np.eye(2, dtype=int)), "_mod3_cuda", FakeMod3Library(result=-1)),
        (lambda: perm_cuda.permanent_modp_cuda(np.eye(2, dtype=int), 5), "_modp_cuda",
         FakeModPLibrary(matrix_result=-1)),
        (lambda: perm_cuda.schur_permanent_modp_cuda(3, 2, 7), "_modp_cuda", FakeModPLibrary(schur_result=-1)),
    ],
)
def test_failure_sentinels_raise(monkeypatch, call, name, library):
    monkeypatch.setattr(perm_cuda, name, library)
    with pytest.raises(RuntimeError, match="failed"):
        call()


def test_trivial_sizes_need_no_library(monkeypatch):
    monkeypatch.setattr(perm_cuda, "_mod3_cuda", None)
    monkeypatch.setattr(perm_cuda, "_modp_cuda", None)
    assert perm_cuda.permanent_mod3_cuda([[10]]) == 1
    assert perm_cuda.permanent_modp_cuda([[-2]], 5) == 3
    assert perm_cuda.permanent_modp_cuda([], 1) == 0
    assert perm_cuda.schur_permanent_modp_cuda(1, 999, 5) == 1


def test_missing_library_raises_instead_of_computing_on_the_host(monkeypatch):
    monkeypatch.setattr(perm_cuda, "_modp_cuda", None)
    monkeypatch.setattr(perm_cuda, "_mod3_cuda", None)
    with pytest.raises(RuntimeError, match="not loaded"):
        perm_cuda.permanent_modp_cuda(np.eye(3, dtype=int), 7)
    with pytest.raises(RuntimeError, match="not loaded"):
        perm_cuda.permanent_mod3_cuda(np.eye(3, dtype=int))
    with pytest.raises(RuntimeError, match="not loaded"):
        perm_cuda.schur_permanent_modp_cuda(5, 4, 11)


@pytest.mark.parametrize(
    "matrix",
    [[1, 2], [[1, 2]], [[1, 2], [3]], np.zeros((2, 2, 1), dtype=int), np.zeros((64, 64), dtype=int)],
)
def test_matrix_shape_and_dimension_validation(matrix):
    with pytest.raises(ValueError):
        perm_cuda.permanent_modp_cuda(matrix, 5)


def test_integer_and_modulus_validation():
    with pytest.raises(TypeError):
        perm_cuda.permanent_mod3_cuda([[1.5]])
    with pytest.raises(ValueError):
        perm_cuda.permanent_modp_cuda([[1]], 0)
    with pytest.raises(ValueError):
        perm_cuda.permanent_modp_cuda([[1]], 1 << 63)
    with pytest.raises(ValueError):
        perm_cuda.schur_permanent_modp_cuda(-1, 1, 3)


def test_schur_bound_is_the_exact_ceiling():
    radicand = 27 ** 27
    bound = perm_cuda.schur_permanent_bound(27)
    assert bound * bound >= radicand > (bound - 1) * (bound - 1)
    assert bound - int(math.ceil(27 ** (27 / 2))) == 1_563


# --- perm_cuda: device CRT plan ---------------------------------------------

@pytest.mark.parametrize("n", sorted(EXPECTED_GPU_PRIMES))
def test_large_prime_plan_matches_reference(n):
    bound = perm_cuda.schur_permanent_bound(n)
    primes = perm_cuda._plan_schur_crt_primes(n, bound)
    theorem = perm_cuda._graham_lehmer_schur_congruence(n)[1]
    assert primes == EXPECTED_GPU_PRIMES[n]
    assert all(1 < q < 2 ** 32 and (q - 1) % n == 0 and perm_cuda._is_prime(q) for q in primes)
    assert all(math.gcd(q, theorem) == 1 for q in primes)
    assert theorem * math.prod(primes) >= 2 * bound + 1 > theorem * math.prod(primes[:-1])


@pytest.mark.parametrize("n", sorted(EXPECTED_GPU_PRIMES))
def test_roots_have_exact_order(n):
    for q in EXPECTED_GPU_PRIMES[n]:
        w = perm_cuda._compute_omega(n, q)
        assert pow(w, n, q) == 1
        assert all(pow(w, n // f, q) != 1 for f in perm_cuda._distinct_prime_factors(n))


def test_plan_stops_on_the_exact_coverage_condition():
    first, second = perm_cuda._find_primes_for_schur(41, 2)
    combined = 41 ** 3 * first
    assert perm_cuda._plan_schur_crt_primes(41, (combined - 1) // 2) == [first]
    assert perm_cuda._plan_schur_crt_primes(41, (combined + 1) // 2) == [first, second]


@pytest.mark.parametrize("n,divisor", sorted(THEOREM4_DIVISORS.items()))
def test_theorem4_divisors(n, divisor):
    assert perm_cuda._graham_lehmer_schur_congruence(n) == (0, divisor)
    if n in SCHUR:
        assert SCHUR[n] % divisor == 0


@pytest.mark.parametrize("n", [5, 7, 11, 13, 17, 31, 41, 47])
def test_prime_order_congruence_against_known_values(n):
    residue, modulus = perm_cuda._graham_lehmer_schur_congruence(n)
    assert modulus == n ** 3 and residue == math.factorial(n) % modulus
    assert SCHUR[n] % modulus == residue


@pytest.mark.parametrize("n", [0, 1, 2, 4, 6, 8, 10, 46])
def test_no_congruence_for_trivial_orders(n):
    assert perm_cuda._graham_lehmer_schur_congruence(n) is None


def test_primality_and_root_inputs_are_checked():
    assert not perm_cuda._is_prime(2 ** 32 - 1)
    with pytest.raises(ValueError, match="prime"):
        perm_cuda._compute_omega(3, 10)
    with pytest.raises(ValueError, match="mod"):
        perm_cuda._compute_omega(5, 13)


@pytest.mark.parametrize("n", [41, 45, 47])
def test_reconstruction_keeps_theorem_modulus_off_the_device(monkeypatch, n):
    used = []
    monkeypatch.setattr(perm_cuda, "is_cuda_available", lambda: True)

    def residue(order, omega, q):
        assert order == n and pow(omega, n, q) == 1
        used.append(q)
        return SCHUR[n] % q

    monkeypatch.setattr(perm_cuda, "schur_permanent_modp_cuda", residue)
    assert perm_cuda.permanent_cuda(n, verbose=False) == SCHUR[n]
    assert used == EXPECTED_GPU_PRIMES[n]


def test_theorem_alone_covers_small_prime_order(monkeypatch):
    monkeypatch.setattr(perm_cuda, "is_cuda_available", lambda: pytest.fail("no device check needed"))
    monkeypatch.setattr(perm_cuda, "schur_permanent_modp_cuda", lambda *a: pytest.fail("no device residue"))
    assert perm_cuda.permanent_cuda(5, verbose=False) == -5


def test_permanent_cuda_refuses_without_a_device(monkeypatch):
    monkeypatch.setattr(perm_cuda, "is_cuda_available", lambda: False)
    with pytest.raises(RuntimeError):
        perm_cuda.permanent_cuda(9, verbose=False)
    assert [perm_cuda.permanent_cuda(n, verbose=False) for n in (0, 1, 2, 8)] == [1, 1, 0, 0]


# --- perm_binary_cuda ---------------------------------------------------------

def test_binary_availability_requires_probe(monkeypatch):
    monkeypatch.setattr(perm_binary_cuda, "_binary_cuda", None)
    assert not perm_binary_cuda.is_binary_cuda_available()
    fake = FakeBinaryLibrary(available=False)
    monkeypatch.setattr(perm_binary_cuda, "_binary_cuda", fake)
    assert not perm_binary_cuda.is_binary_cuda_available()
    fake.available = True
    assert perm_binary_cuda.is_binary_cuda_available()


def test_binary_modular_wrapper_normalizes(monkeypatch):
    fake = FakeBinaryLibrary()
    monkeypatch.setattr(perm_binary_cuda, "_binary_cuda", fake)
    m = np.array([[True, True, False], [False, True, True], [True, False, True]], dtype=np.bool_)
    assert perm_binary_cuda.binary_permanent_mod_p_cuda(m, 101) == oracle.definition_permanent(m.astype(int).tolist())
    assert fake.calls == [(m.astype(int).tolist(), 3, 101)]


def test_binary_exact_wrapper_uses_two_descending_primes(monkeypatch):
    fake = FakeBinaryLibrary()
    monkeypatch.setattr(perm_binary_cuda, "_binary_cuda", fake)
    J13 = [[1] * 13 for _ in range(13)]
    assert perm_binary_cuda.binary_permanent_cuda(J13) == math.factorial(13)
    assert [c[2] for c in fake.calls] == list(perm_binary._crt_primes_above_bound(math.factorial(13)))
    assert len(fake.calls) == 2


def test_binary_blocks_factor_before_dispatch(monkeypatch):
    fake = FakeBinaryLibrary()
    monkeypatch.setattr(perm_binary_cuda, "_binary_cuda", fake)
    # Two 3x3 blocks on the diagonal, one fixed 1x1 block.
    m = np.zeros((7, 7), dtype=int)
    m[:3, :3] = 1
    m[3:6, 3:6] = [[1, 1, 0], [0, 1, 1], [1, 0, 1]]
    m[6, 6] = 1
    m[0, 6] = 1
    assert perm_binary_cuda.binary_permanent_cuda(m) == 6 * 2
    assert sorted(c[1] for c in fake.calls) == [3, 3]  # one prime covers each small block


def test_binary_infeasible_and_trivial_need_no_device(monkeypatch):
    monkeypatch.setattr(perm_binary_cuda, "_binary_cuda", None)
    assert perm_binary_cuda.binary_permanent_cuda([]) == 1
    assert perm_binary_cuda.binary_permanent_cuda([[0]]) == 0
    assert perm_binary_cuda.binary_permanent_cuda([[1, 1], [0, 0]]) == 0
    assert perm_binary_cuda.binary_permanent_mod_p_cuda([[1, 1], [0, 0]], 7) == 0
    with pytest.raises(RuntimeError, match="unavailable"):
        perm_binary_cuda.binary_permanent_cuda([[1, 1], [1, 1]])


@pytest.mark.parametrize("result,match", [(-1, "failed"), (101, "invalid residue")])
def test_binary_sentinels(monkeypatch, result, match):
    monkeypatch.setattr(perm_binary_cuda, "_binary_cuda", FakeBinaryLibrary(result=result))
    with pytest.raises(RuntimeError, match=match):
        perm_binary_cuda.binary_permanent_mod_p_cuda([[1, 1], [1, 1]], 101)


def test_binary_prime_validation():
    with pytest.raises(ValueError):
        perm_binary_cuda.binary_permanent_mod_p_cuda([[1, 1], [1, 1]], 2147483659)  # prime >= 2^31
    with pytest.raises(ValueError):
        perm_binary_cuda.binary_permanent_mod_p_cuda([[1, 1], [1, 1]], 9)
    with pytest.raises(ValueError):
        perm_binary_cuda.binary_permanent_mod_p_cuda([[1, 1, 1]] * 3, 3)
    with pytest.raises(ValueError):
        perm_binary_cuda.binary_permanent_mod_p_cuda([[2, 1], [1, 1]], 101)
