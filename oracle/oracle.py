"""ORACLE — test infrastructure only (checker for tests/, smoke() and the
bench's cpu_baseline leg; never imported by the product).

Three layers:

* ``definition_permanent`` / ``subset_permanent`` / ``exact_glynn``: small
  pure-Python oracles independent of the kernels' algorithms
  (tests/test_generic_native.py:26-48 of the reference use the same two
  definitions).
* ``frontier_count``: the live-column frontier DP over column masks in
  Python integers (perm_binary.py:389-441), with its state cap.  Pinned in
  tests/test_frontier_host.py against ``definition_permanent`` on random
  small matrices, the closed forms 10!^5 (block diagonal) and 2 (odd
  cycle), the Hall-obstruction zero and the reference's state-cap fields.
* ``C``: ctypes binding of ``_build/liboracle.so``, the C + OpenMP
  restatement of the reference's residue kernels (perm_oracle.c).
* ``reference_extensions()``: the reference's own OpenMP extensions compiled
  from its sources into ``_ref/`` by ``oracle/Makefile`` (when built).
"""

from __future__ import annotations

import ctypes
import importlib.util
import itertools
import math
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
REF_DIR = HERE / "_ref"

_I64P = ctypes.POINTER(ctypes.c_int64)
_U64P = ctypes.POINTER(ctypes.c_uint64)
_U8P = ctypes.POINTER(ctypes.c_uint8)


def build(reference: bool = True) -> None:
    """Compile the oracle (and the reference extensions when mounted)."""
    target = "all" if reference else "oracle"
    subprocess.run(["make", "-s", "-C", str(HERE), target], check=True)


class _C:
    def __init__(self):
        self._lib = None

    @property
    def lib(self):
        if self._lib is None:
            if not LIB_PATH.exists():
                build(reference=False)
            lib = ctypes.CDLL(str(LIB_PATH))
            for name in ("oracle_glynn_mod_p", "oracle_ryser_mod_p", "oracle_permanent_mod_p"):
                fn = getattr(lib, name)
                fn.restype = ctypes.c_int64
                fn.argtypes = [_I64P, ctypes.c_int, ctypes.c_int64, ctypes.c_int]
            lib.oracle_binary_mod_p.restype = ctypes.c_int64
            lib.oracle_binary_mod_p.argtypes = [_U8P, ctypes.c_int, ctypes.c_int64, ctypes.c_int]
            lib.oracle_binary_update_cost.restype = None
            lib.oracle_binary_update_cost.argtypes = [_U64P, ctypes.c_int, _U64P, _U64P]
            lib.oracle_mod3.restype = ctypes.c_int
            lib.oracle_mod3.argtypes = [_I64P, ctypes.c_int, ctypes.c_int]
            lib.oracle_max_threads.restype = ctypes.c_int
            lib.oracle_affine_mod_p.restype = None
            lib.oracle_affine_mod_p.argtypes = [_I64P, _I64P, ctypes.c_int, ctypes.c_int64, ctypes.c_int, _I64P]
            self._lib = lib
        return self._lib

    @staticmethod
    def _reduced(matrix, p: int) -> np.ndarray:
        rows = [[int(x) % p for x in row] for row in matrix]
        return np.ascontiguousarray(np.array(rows, dtype=np.int64).reshape(len(rows), len(rows)))

    def permanent_mod_p(self, matrix, p: int, threads: int = 0) -> int:
        a = self._reduced(matrix, p)
        return int(self.lib.oracle_permanent_mod_p(a.ctypes.data_as(_I64P), a.shape[0], p, threads))

    def glynn_mod_p(self, matrix, p: int, threads: int = 0) -> int:
        a = self._reduced(matrix, p)
        return int(self.lib.oracle_glynn_mod_p(a.ctypes.data_as(_I64P), a.shape[0], p, threads))

    def ryser_mod_p(self, matrix, p: int, threads: int = 0) -> int:
        a = self._reduced(matrix, p)
        return int(self.lib.oracle_ryser_mod_p(a.ctypes.data_as(_I64P), a.shape[0], p, threads))

    def binary_mod_p(self, matrix, p: int, threads: int = 0) -> int:
        a = np.ascontiguousarray(np.asarray(matrix, dtype=np.uint8))
        n = a.shape[0] if a.ndim == 2 else 0
        return int(self.lib.oracle_binary_mod_p(a.ctypes.data_as(_U8P), n, p, threads))

    def binary_update_cost(self, matrix) -> int:
        """Cost of the cheaper orientation, as the reference chooses it."""
        a = np.asarray(matrix, dtype=np.uint8)
        n = a.shape[0]
        costs = []
        for m in (a, a.T):
            cols = np.array(
                [sum(1 << i for i in range(n) if m[i, j]) for j in range(n)], dtype=np.uint64
            )
            hi, lo = ctypes.c_uint64(), ctypes.c_uint64()
            self.lib.oracle_binary_update_cost(
                cols.ctypes.data_as(_U64P), n, ctypes.byref(hi), ctypes.byref(lo)
            )
            costs.append((hi.value << 64) | lo.value)
        return min(costs)

    def mod3(self, matrix, threads: int = 0) -> int:
        a = np.ascontiguousarray(np.asarray(matrix, dtype=np.int64))
        return int(self.lib.oracle_mod3(a.ctypes.data_as(_I64P), a.shape[0], threads))

    def affine_mod_p(self, constant, linear, p: int, threads: int = 0) -> tuple:
        """Coefficients of per(C + x L) mod odd p, ascending degree."""
        n = len(constant)
        if n == 0:
            return (1 % p,)
        c = self._reduced(constant, p)
        l = self._reduced(linear, p)
        out = np.zeros(n + 1, dtype=np.int64)
        self.lib.oracle_affine_mod_p(c.ctypes.data_as(_I64P), l.ctypes.data_as(_I64P), n, p, threads,
                                     out.ctypes.data_as(_I64P))
        return tuple(int(v) for v in out)

    def max_threads(self) -> int:
        return int(self.lib.oracle_max_threads())


C = _C()


# ---------------------------------------------------------------------------
# Pure-Python oracles
# ---------------------------------------------------------------------------

def definition_permanent(matrix) -> int:
    """Sum over permutations of products (exact integers)."""
    n = len(matrix)
    return sum(
        math.prod(int(matrix[r][c]) for r, c in enumerate(cols))
        for cols in itertools.permutations(range(n))
    )


def subset_permanent(matrix, modulus=None) -> int:
    """Row-by-row matching DP over column subsets (independent of both the
    Ryser and Glynn signed-sum identities)."""
    n = len(matrix)
    counts = [0] * (1 << n)
    counts[0] = 1
    for used in range(1, 1 << n):
        row = matrix[bin(used).count("1") - 1]
        total = 0
        rest = used
        while rest:
            bit = rest & -rest
            total += counts[used ^ bit] * int(row[bit.bit_length() - 1])
            rest ^= bit
        counts[used] = total % modulus if modulus else total
    return counts[-1] if modulus is None else counts[-1] % modulus


def exact_glynn(matrix) -> int:
    """Glynn's formula over the integers (no modulus), Gray-code order."""
    n = len(matrix)
    if n == 0:
        return 1
    a = [[int(x) for x in row] for row in matrix]
    rs = [sum(row) for row in a]
    total = 0
    sign = 1
    for s in range(1 << (n - 1)):
        if s:
            j = (s & -s).bit_length() - 1
            g = s ^ (s >> 1)
            d = -2 if (g >> j) & 1 else 2
            for i in range(n):
                rs[i] += d * a[i][j]
            sign = -sign
        total += sign * math.prod(rs)
    assert total % (1 << (n - 1)) == 0
    return total >> (n - 1)


class FrontierCapExceeded(Exception):
    """A layer of frontier_count would hold more than max_states states."""

    def __init__(self, max_states: int, observed: int, rows: int):
        super().__init__(max_states, observed, rows)
        self.max_states, self.observed_states, self.processed_rows = max_states, observed, rows


def frontier_count(row_masks, row_order, max_states: int = 1 << 62, layer_sizes=None) -> int:
    """Perfect matchings of a 0/1 matrix by the live-column frontier DP
    (perm_binary.py:389-441): a state is the set of used columns that some
    later row still has; each row adds one available column, and the
    columns whose last row it is must all be used and then leave the state.
    Raises FrontierCapExceeded(max_states, max_states + 1, row position + 1)
    when a new state would make a layer larger than max_states.  When
    layer_sizes is a list, the size of every layer is appended to it."""
    n = len(row_masks)
    if n == 0:
        return 1
    last = [-1] * n
    for pos, r in enumerate(row_order):
        m = int(row_masks[r])
        for c in range(n):
            if (m >> c) & 1:
                last[c] = pos
    if min(last) < 0:
        return 0
    expiring = [0] * n
    for c, pos in enumerate(last):
        expiring[pos] |= 1 << c
    states = {0: 1}
    for pos, r in enumerate(row_order):
        nxt = {}
        for used, count in states.items():
            for c in range(n):
                bit = 1 << c
                if not (int(row_masks[r]) & bit) or (used & bit):
                    continue
                target = used | bit
                if target & expiring[pos] != expiring[pos]:
                    continue
                target &= ~expiring[pos]
                if target not in nxt and len(nxt) >= max_states:
                    raise FrontierCapExceeded(max_states, max_states + 1, pos + 1)
                nxt[target] = nxt.get(target, 0) + count
        if layer_sizes is not None:
            layer_sizes.append(len(nxt))
        states = nxt
        if not states:
            return 0
    return states.get(0, 0)


# ---------------------------------------------------------------------------
# Reference extensions (compiled from the reference's C sources)
# ---------------------------------------------------------------------------

def reference_extensions():
    """{name: module} for the reference's OpenMP extensions in _ref/, or {}."""
    found = {}
    for name in ("perm_c_ext", "perm_binary_ext", "perm_mod3_omp", "perm_polynomial_ext"):
        path = REF_DIR / f"{name}.so"
        if not path.exists():
            continue
        spec = importlib.util.spec_from_file_location(name, path)
        module = importlib.util.module_from_spec(spec)
        try:
            spec.loader.exec_module(module)
        except ImportError:
            continue
        found[name] = module
        sys.modules.setdefault(f"oracle_ref_{name}", module)
    return found
