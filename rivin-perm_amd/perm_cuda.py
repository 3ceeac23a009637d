"""Device residues behind the interface of the reference's ``perm_cuda``.

Same public functions, argument checks, base cases and sentinel rule as
perm_cuda.py in igorrivin/rivin-perm, computed on the MI355X by
``libpermhip.so``:

    permanent_mod3_cuda(matrix)            perm_cuda.py:132-151
    schur_permanent_modp_cuda(n, omega, p) perm_cuda.py:154-179
    permanent_modp_cuda(matrix, p)         perm_cuda.py:182-202
    permanent_cuda(n, verbose=True)        perm_cuda.py:466-538

``permanent_cuda`` uses the reference's device CRT plan (perm_cuda.py:205-463):
descending primes q == 1 (mod n) below 2^32, an n-th root of unity of exact
order per prime, the Graham--Lehmer theorem modulus added only at
reconstruction, and a centred (signed) CRT.

The module globals ``_modp_cuda`` and ``_mod3_cuda`` hold objects with the
reference library's symbol names (``cuda_runtime_available``,
``permanent_modp_cuda``, ``schur_permanent_modp_cuda``,
``permanent_mod3_matrix``) bound to the ``phip_*`` entry points, so code that
swaps in its own library object, as the reference's tests do, keeps working.
"""

from __future__ import annotations

import ctypes
import math
import operator
import time
from typing import List, Optional, Tuple

import numpy as np

from permhip import _native
from permhip import planner as _planner

_I64P = ctypes.POINTER(ctypes.c_longlong)
_I32P = ctypes.POINTER(ctypes.c_int)


class _ModpLibrary:
    """The reference's modp symbols on libpermhip."""

    def __init__(self, lib):
        self._lib = lib

    def cuda_runtime_available(self) -> int:
        return int(_native._touch_device(self._lib))

    def permanent_modp_cuda(self, pointer, n, p) -> int:
        _native._check_fork()
        return int(self._lib.phip_permanent_modp(ctypes.cast(pointer, _I64P), int(n), int(p)))

    def schur_permanent_modp_cuda(self, n, omega, p) -> int:
        _native._check_fork()
        return int(self._lib.phip_schur_permanent_modp(int(n), int(omega), int(p)))


class _Mod3Library:
    """The reference's mod-3 symbols on libpermhip."""

    def __init__(self, lib):
        self._lib = lib

    def cuda_runtime_available(self) -> int:
        return int(_native._touch_device(self._lib))

    def permanent_mod3_matrix(self, pointer, n) -> int:
        _native._check_fork()
        return int(self._lib.phip_permanent_mod3_matrix(ctypes.cast(pointer, _I32P), int(n)))


_mod3_cuda = None
_modp_cuda = None
_load_error: Optional[Exception] = None


def _load_cuda_libs() -> None:
    """Bind the library objects (perm_cuda.py:22-53); a missing library
    leaves them None and the device functions raise."""
    global _mod3_cuda, _modp_cuda, _load_error
    try:
        lib = _native.load()
    except (OSError, RuntimeError) as exc:
        _load_error = exc
        return
    _modp_cuda = _ModpLibrary(lib)
    _mod3_cuda = _Mod3Library(lib)


_load_cuda_libs()


def is_cuda_available() -> bool:
    """True when the library is loaded and a device answers its probe
    (perm_cuda.py:62-72)."""
    probe = getattr(_modp_cuda, "cuda_runtime_available", None)
    if probe is None:
        return False
    try:
        return bool(probe())
    except (OSError, TypeError, ValueError, _native.ForkedProcessError):
        return False


def _require_current_cuda_backend(library, name: str) -> None:
    if getattr(library, "cuda_runtime_available", None) is None:
        raise RuntimeError(f"{name} device library is outdated; rebuild libpermhip.so")


def _as_index(value, name: str) -> int:
    """Exact integer via __index__; floats are refused (perm_cuda.py:82-87)."""
    try:
        return operator.index(value)
    except TypeError as exc:
        raise TypeError(f"{name} must be an integer") from exc


def _validate_modulus(p) -> int:
    modulus = _as_index(p, "p")
    if modulus <= 0 or modulus > (1 << 63) - 1:
        raise ValueError("p must be a positive signed 64-bit integer")
    return modulus


def _validate_dimension(n) -> int:
    order = _as_index(n, "n")
    if order < 0 or order >= 64:
        raise ValueError("n must satisfy 0 <= n < 64")
    return order


def _normalized_square_matrix(matrix, modulus: int, dtype) -> Tuple[int, np.ndarray]:
    """Square integer matrix reduced mod `modulus` in Python integers before
    the fixed-width copy (perm_cuda.py:104-122)."""
    try:
        cells = np.asarray(matrix, dtype=object)
    except (TypeError, ValueError) as exc:
        raise ValueError("matrix must be a rectangular square array") from exc
    if cells.ndim == 1 and cells.size == 0:
        cells = cells.reshape(0, 0)
    if cells.ndim != 2 or cells.shape[0] != cells.shape[1]:
        raise ValueError("matrix must be a square two-dimensional array")
    n = _validate_dimension(cells.shape[0])
    flat = [_as_index(v, f"matrix entry {i}") % modulus for i, v in enumerate(cells.flat)]
    out = np.array(flat, dtype=dtype).reshape(n, n) if n else np.zeros((0, 0), dtype=dtype)
    return n, np.ascontiguousarray(out)


def _raise_on_cuda_failure(result, operation: str) -> int:
    value = int(result)
    if value < 0:
        detail = _native.last_error() if _native.available() else ""
        raise RuntimeError(f"{operation} failed" + (f": {detail}" if detail else ""))
    return value


def permanent_mod3_cuda(matrix) -> int:
    """perm(matrix) mod 3 on the device (perm_cuda.py:132-151)."""
    n, reduced = _normalized_square_matrix(matrix, 3, np.int32)
    if n == 0:
        return 1
    if n == 1:
        return int(reduced[0, 0])
    if _mod3_cuda is None:
        raise RuntimeError(f"device mod-3 library not loaded: {_load_error}")
    _require_current_cuda_backend(_mod3_cuda, "mod3")
    result = _mod3_cuda.permanent_mod3_matrix(reduced.ctypes.data_as(_I32P), n)
    return _raise_on_cuda_failure(result, "device permanent modulo 3")


def schur_permanent_modp_cuda(n, omega, p) -> int:
    """perm(S_n) mod p for S_n[j][k] = omega^(jk) (perm_cuda.py:154-179)."""
    n = _validate_dimension(n)
    p = _validate_modulus(p)
    omega = _as_index(omega, "omega") % p
    if n <= 1:
        return 1 % p
    if _modp_cuda is None:
        raise RuntimeError(f"device modp library not loaded: {_load_error}")
    _require_current_cuda_backend(_modp_cuda, "modp")
    result = _modp_cuda.schur_permanent_modp_cuda(n, omega, p)
    return _raise_on_cuda_failure(result, "device Schur permanent")


def permanent_modp_cuda(matrix, p) -> int:
    """perm(matrix) mod p on the device (perm_cuda.py:182-202)."""
    p = _validate_modulus(p)
    n, reduced = _normalized_square_matrix(matrix, p, np.int64)
    if n == 0:
        return 1 % p
    if n == 1:
        return int(reduced[0, 0])
    if _modp_cuda is None:
        raise RuntimeError(f"device modp library not loaded: {_load_error}")
    _require_current_cuda_backend(_modp_cuda, "modp")
    result = _modp_cuda.permanent_modp_cuda(reduced.ctypes.data_as(_I64P), n, p)
    return _raise_on_cuda_failure(result, "device permanent modulo p")


# ---------------------------------------------------------------------------
# Device CRT plan for Schur orders (perm_cuda.py:205-463)
# ---------------------------------------------------------------------------

# Every device prime stays below 2^32, so two reduced residues multiply below
# 2^64 (perm_cuda.py:205-208).
_CRT_PRIME_LIMIT = 1 << 32


def _is_prime(value) -> bool:
    """Deterministic for every 64-bit value (perm_cuda.py:211-243)."""
    return _planner.is_prime(_as_index(value, "value"))


def _distinct_prime_factors(value) -> List[int]:
    value = _as_index(value, "value")
    if value < 1:
        raise ValueError("value must be positive")
    return _planner._distinct_prime_factors(value)


def _primitive_root(p) -> int:
    """Smallest generator of (Z/pZ)* (perm_cuda.py:270-286)."""
    p = _as_index(p, "p")
    if not _is_prime(p):
        raise ValueError("p must be prime")
    return _planner.primitive_root_mod_p(p)


def _schur_prime_candidates(n: int):
    """Odd q == 1 (mod n) below 2^32 in descending order (perm_cuda.py:289-302)."""
    top = _CRT_PRIME_LIMIT - 1
    q = top - (top - 1) % n  # largest q <= top with q == 1 (mod n)
    step = n
    if n % 2:
        step = 2 * n  # q == 1 (mod n) alternates parity; keep the odd ones
        if q % 2 == 0:
            q -= n
    while q >= 2:
        yield q
        q -= step


def _find_primes_for_schur(n, count) -> List[int]:
    """The `count` largest primes q == 1 (mod n) below 2^32 (perm_cuda.py:305-322)."""
    n = _validate_dimension(n)
    count = _as_index(count, "count")
    if n == 0:
        raise ValueError("n must be positive")
    if count < 0:
        raise ValueError("count must be non-negative")
    found: List[int] = []
    if count == 0:
        return found
    for q in _schur_prime_candidates(n):
        if _is_prime(q):
            found.append(q)
            if len(found) == count:
                return found
    raise ValueError(f"Could not find {count} suitable CRT primes for n={n}")


def _graham_lehmer_schur_congruence(n) -> Optional[Tuple[int, int]]:
    """(residue, modulus) from Graham--Lehmer for odd n >= 3
    (perm_cuda.py:325-369)."""
    return _planner.graham_lehmer_schur_congruence(_validate_dimension(n))


def _plan_schur_crt_primes(n, bound) -> List[int]:
    """Device primes that, with the theorem modulus, cover 2*bound + 1
    exactly (perm_cuda.py:372-404)."""
    n = _validate_dimension(n)
    bound = _as_index(bound, "bound")
    if n == 0:
        raise ValueError("n must be positive")
    if bound < 0:
        raise ValueError("bound must be non-negative")
    congruence = _graham_lehmer_schur_congruence(n)
    covered = congruence[1] if congruence is not None else 1
    target = 2 * bound + 1
    chosen: List[int] = []
    if covered >= target:
        return chosen
    for q in _schur_prime_candidates(n):
        if _is_prime(q):
            chosen.append(q)
            covered *= q
            if covered >= target:
                return chosen
    raise ValueError(f"Could not cover the CRT bound for n={n}")


def _compute_omega(n, p) -> int:
    """An element of exact multiplicative order n modulo the prime p
    (perm_cuda.py:407-426)."""
    n = _as_index(n, "n")
    p = _as_index(p, "p")
    if n < 1:
        raise ValueError("n must be positive")
    if (p - 1) % n:
        raise ValueError("p must satisfy p ≡ 1 (mod n)")
    omega = pow(_primitive_root(p), (p - 1) // n, p)
    if pow(omega, n, p) != 1 or any(pow(omega, n // f, p) == 1 for f in _distinct_prime_factors(n)):
        raise RuntimeError("failed to construct an n-th root of exact order")
    return omega


def _crt_reconstruct(residues, moduli, bound) -> int:
    """Centred CRT: the representative in (-M/2, M/2] (perm_cuda.py:429-448)."""
    value = _planner.chinese_remainder_theorem(list(residues), list(moduli))
    modulus = math.prod(moduli)
    return value - modulus if value > modulus // 2 else value


def schur_permanent_bound(n) -> int:
    """ceil(sqrt(n^n)) exactly (perm_cuda.py:451-463)."""
    return _planner.schur_permanent_bound(_validate_dimension(n))


def permanent_cuda(n, verbose: bool = True) -> int:
    """Exact perm(S_n) from device residues (perm_cuda.py:466-538)."""
    n = _validate_dimension(n)
    if n <= 1:
        return 1
    if n % 2 == 0:
        return 0
    bound = schur_permanent_bound(n)
    device_primes = _plan_schur_crt_primes(n, bound)
    congruence = _graham_lehmer_schur_congruence(n)
    if device_primes and not is_cuda_available():
        raise RuntimeError("device library or device not available")
    if verbose:
        print(f"Computing perm(S_{n}) using {len(device_primes)} primes on the GPU")
        print(f"GPU primes: {device_primes}")
        if congruence is not None:
            print(f"Graham-Lehmer CRT modulus: {congruence[1]}")
        print(f"Bound: {bound:.2e}")
    residues: List[int] = []
    t_all = time.perf_counter()
    for i, q in enumerate(device_primes):
        omega = _compute_omega(n, q)
        t0 = time.perf_counter()
        r = schur_permanent_modp_cuda(n, omega, q)
        residues.append(r)
        if verbose:
            print(f"  Prime {i + 1}/{len(device_primes)}: p={q}, omega={omega}, "
                  f"perm≡{r} (mod {q}), time={time.perf_counter() - t0:.1f}s")
    moduli = list(device_primes)
    if congruence is not None:
        residues.append(congruence[0])
        moduli.append(congruence[1])
    result = _crt_reconstruct(residues, moduli, bound)
    if verbose:
        elapsed = time.perf_counter() - t_all
        print(f"perm(S_{n}) = {result}")
        print(f"Total time: {elapsed:.1f}s ({elapsed / 60:.1f} min)")
    return result


__all__ = [
    "is_cuda_available",
    "permanent_cuda",
    "permanent_mod3_cuda",
    "permanent_modp_cuda",
    "schur_permanent_bound",
    "schur_permanent_modp_cuda",
]
