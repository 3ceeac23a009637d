"""The allocation-free entry point (phip_permanent_residues_noalloc) against
the allocating one (phip_permanent_residues), through the C ABI: two calls
in a row, the second smaller, then the refusals."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_I64P = ctypes.POINTER(ctypes.c_int64)


def _call(lib, name, A, moduli, stride=0):
    a = np.ascontiguousarray(A, dtype=np.int64)
    n = int(a.shape[-1])
    mods = np.ascontiguousarray(moduli, dtype=np.int64)
    out = np.full(len(mods), -7, dtype=np.int64)
    status = getattr(lib, name)(a.ctypes.data_as(_I64P), stride, n, mods.ctypes.data_as(_I64P), len(mods),
                                out.ctypes.data_as(_I64P))
    return status, [int(v) for v in out]


def test_noalloc_matches_the_allocating_entry_point(native):
    from permhip import planner, workloads

    lib = native.load()
    cases = []
    A = workloads.headline()
    primes, _ = planner.generic_crt_primes(planner.matrix_permanent_bound(A.tolist()))
    cases.append((A, primes))  # n = 32, 8 primes: one paired set of 8
    B = np.random.default_rng(11).integers(-100, 101, (20, 20))
    cases.append((B, [1073741827, 2147483647, 1000003]))  # smaller: n = 20, 3 moduli
    for M, mods in cases:
        s1, want = _call(lib, "phip_permanent_residues", M, mods)
        s2, got = _call(lib, "phip_permanent_residues_noalloc", M, mods)
        assert s1 == 0 and s2 == 0, native.last_error()
        assert got == want
    # Per-modulus matrices (prime_stride = n * n), fewer moduli again.
    C = np.random.default_rng(12).integers(-9, 10, (2, 12, 12))
    s1, want = _call(lib, "phip_permanent_residues", C, [101, 1009], stride=144)
    s2, got = _call(lib, "phip_permanent_residues_noalloc", C, [101, 1009], stride=144)
    assert s1 == s2 == 0 and got == want
    # 0/1 table (zero-skipping kernels) and the oracle.
    import oracle

    D = (np.random.default_rng(13).random((24, 24)) < 0.5).astype(np.int64)
    s2, got = _call(lib, "phip_permanent_residues_noalloc", D, [2147483647, 2147483629])
    assert s2 == 0
    assert got == [oracle.C.permanent_mod_p(D.tolist(), p, 4) for p in (2147483647, 2147483629)]
    # Five primes: the block walk with exact accumulation (limb columns in
    # the item slots), against the allocating entry point.
    five = [2147483647, 2147483629, 2147483587, 2147483579, 2147483563]
    s1, want = _call(lib, "phip_permanent_residues", D, five)
    s2, got = _call(lib, "phip_permanent_residues_noalloc", D, five)
    assert s1 == s2 == 0 and got == want
    assert got[:2] == [oracle.C.permanent_mod_p(D.tolist(), p, 4) for p in five[:2]]


def test_noalloc_refusals(native):
    lib = native.load()
    A = np.ones((10, 10), dtype=np.int64)
    for mods, n_override, what in (([1 << 31], None, "odd moduli"), ([100], None, "odd moduli"),
                                   ([101] * 65, None, "1..64 moduli"), ([101], 4, "orders 5..46")):
        M = np.ones((n_override, n_override), dtype=np.int64) if n_override else A
        status, _ = _call(lib, "phip_permanent_residues_noalloc", M, mods)
        assert status == -1
        assert what in native.last_error()
    # A refusal does not disturb the next call.
    s, got = _call(lib, "phip_permanent_residues_noalloc", A, [1000003])
    assert s == 0 and got == [3628800 % 1000003]
