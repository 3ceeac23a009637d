"""Exact permanents of many matrices in one persistent device launch.

``PermanentBatch`` plans every matrix on the host exactly as the single
matrix API does (row-sum bound, ascending CRT primes from 2^30, signed lift;
perm.py:939-953, 1115-1128, 1207-1210), uploads all (matrix, prime) jobs
once, and replays the captured persistent launch for each run.  Matrices the
persistent kernel does not serve (order outside [5, 46] or entries beyond
int64) go through the single-matrix device path.
"""

from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

from . import _native, planner

BATCH_MIN_ORDER = 5
BATCH_MAX_ORDER = 46


def _int64_rows(matrix):
    """(rows, int64 array or None when an entry does not fit int64).  Rows
    are lists of Python ints, or the int64 array itself for integer array
    input (indexable the same way)."""
    if (isinstance(matrix, np.ndarray) and matrix.ndim == 2 and matrix.shape[0] == matrix.shape[1]
            and matrix.dtype.kind in "iu" and matrix.dtype.itemsize <= 8
            and not (matrix.dtype == np.uint64 and matrix.size and int(matrix.max()) > (1 << 63) - 1)):
        arr = np.ascontiguousarray(matrix, dtype=np.int64)
        return arr, arr
    rows = [[int(x) for x in row] for row in matrix]
    n = len(rows)
    if any(len(r) != n for r in rows):
        raise ValueError("Matrix must be square")
    flat = [v for r in rows for v in r]
    if any(v < -(1 << 63) or v > (1 << 63) - 1 for v in flat):
        return rows, None
    return rows, np.array(flat, dtype=np.int64).reshape(n, n)


class PermanentBatch:
    """Plan once, run many times; results are exact Python integers."""

    def __init__(
        self,
        matrices: Sequence,
        *,
        use_graph: bool = True,
        blocks_per_cu: int | None = None,
        autotune: bool | dict = False,
    ):
        """``autotune`` True (or a dict of ``autotune.tune_plan`` keyword
        arguments) times candidate chunk sizes and workgroups per CU on this
        plan and keeps the fastest; the choice is in ``self.tuning``."""
        self.rows: List[List[List[int]]] = []
        self.bounds: List[int] = []
        self.primes: List[Tuple[int, ...]] = []
        self.products: List[int] = []
        device_mats, device_primes, self.device_index = [], [], []
        self.other_index = []
        converted = [_int64_rows(m) for m in matrices]
        # Bounds of the int64 matrices in one pass per order.
        fitting = [i for i, (_, arr) in enumerate(converted) if arr is not None]
        bulk = dict(zip(fitting, planner.matrix_permanent_bounds([converted[i][1] for i in fitting])))
        for i, (rows, arr) in enumerate(converted):
            n = len(rows)
            if not n:
                bound = 1
            elif i in bulk:
                bound = bulk[i]
            else:
                bound = planner.matrix_permanent_bound(rows)
            primes, product = planner.generic_crt_primes(bound) if n > 1 else ([], 1)
            primes = tuple(primes)
            self.rows.append(rows)
            self.bounds.append(bound)
            self.primes.append(primes)
            self.products.append(product)
            if arr is not None and BATCH_MIN_ORDER <= n <= BATCH_MAX_ORDER:
                self.device_index.append(i)
                device_mats.append(arr)
                device_primes.append(primes)
            else:
                self.other_index.append(i)
        self.plan = _native.Batch(device_mats, device_primes) if device_mats else None
        if self.plan is not None:
            self.plan.set_option(_native.BATCH_USE_GRAPH, 1 if use_graph else 0)
            if blocks_per_cu is not None:
                self.plan.set_option(_native.BATCH_BLOCKS_PER_CU, blocks_per_cu)
        self.tuning = None
        if autotune and self.plan is not None:
            from . import autotune as _autotune

            kwargs = dict(autotune) if isinstance(autotune, dict) else {}
            kwargs.setdefault("cache", _autotune.default_cache_path())
            self.tuning = _autotune.tune_plan(self.plan, **kwargs)

    @property
    def mulmods(self) -> float:
        """Algorithmic modular multiplications, n 2^(n-1) per residue."""
        return sum(len(p) * len(r) * 2.0 ** (len(r) - 1) for r, p in zip(self.rows, self.primes) if len(r) > 1)

    def residues(self) -> List[List[int]]:
        """Per-matrix residue lists (device plan plus single-matrix path)."""
        out: List[List[int]] = [[] for _ in self.rows]
        if self.plan is not None:
            flat = self.plan.run()
            pos = 0
            for i in self.device_index:
                k = len(self.primes[i])
                out[i] = flat[pos:pos + k]
                pos += k
        for i in self.other_index:
            rows = self.rows[i]
            if len(rows) <= 1:
                continue
            _, arr = _int64_rows(rows)
            if arr is None:
                arr = np.array([[[v % p for v in r] for r in rows] for p in self.primes[i]], dtype=np.int64)
            out[i] = _native.permanent_residues(arr, self.primes[i])
        return out

    def run(self) -> List[int]:
        """Exact permanents in input order."""
        results = []
        for rows, bound, primes, product, res in zip(
            self.rows, self.bounds, self.primes, self.products, self.residues()
        ):
            n = len(rows)
            if n == 0:
                results.append(1)
            elif n == 1:
                results.append(int(rows[0][0]))
            else:
                value = planner.crt_known_moduli(res, primes)
                results.append(planner.signed_lift(value, bound, product))
        return results

    def stats(self) -> dict:
        return self.plan.stats() if self.plan is not None else {}

    def close(self) -> None:
        if self.plan is not None:
            self.plan.close()
            self.plan = None


def batch_matrix_permanents(matrices: Sequence) -> List[int]:
    """Exact permanents of many integer matrices (one persistent launch)."""
    batch = PermanentBatch(matrices)
    try:
        return batch.run()
    finally:
        batch.close()
