"""Launch-shape autotuner of the persistent Glynn engine.

Two knobs decide how a persistent launch covers the Gray code; neither
changes the arithmetic, only the schedule:

* work items (log2): the planner sizes every job's chunk -- the 2^cbits
  consecutive Gray states one lane walks -- so that the plan holds about
  2^items_log2 one-wavefront items.  Longer chunks spend less on each lane's
  start (row sums at gray(base), one lane reduction); more items balance the
  tail better over the compute units.
* workgroups per CU: 256-lane workgroups (4 wavefronts) per compute unit in
  the persistent grid, i.e. the waves per CU that fetch items.  More waves
  hide latency as long as their registers and LDS let them stay resident;
  each class's grid is capped at what its kernel keeps resident.

``tune_plan`` times candidate settings on the caller's own plan by replaying
it, checks that every candidate returns the same residues, and leaves the plan
on the fastest.  ``tune_single`` does the same for the single-matrix entry
points (``perm.matrix_permanent``), whose launches read the process-wide
tuning.  Results may be cached in a JSON file keyed by device, library build
and workload shape.

The search is coordinate descent from the current setting: workgroups per
CU, then work items, then workgroups again if the work items moved.  A
candidate replaces the incumbent only when it is faster by more than
``tolerance``, so timing noise does not move the setting around.

The reference has no counterpart: its launches take a fixed chunk and grid
(perm_cuda_kernel.cu, permanent_binary_cuda.cu launch code).
"""

from __future__ import annotations

import json
import os
import statistics
import time
from dataclasses import asdict, dataclass, field
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native

ITEMS_LOG2_CANDIDATES = (12, 13, 14, 15, 16, 17, 18)
BLOCKS_PER_CU_CANDIDATES = (1, 2, 3, 4)
LANES_LOG2_CANDIDATES = (17, 18, 19, 20, 21, 22)
WAVES_PER_BLOCK = 4


@dataclass
class Trial:
    items_log2: int
    blocks_per_cu: int
    seconds: float
    chunk_bits: Tuple[int, int] = (-1, -1)


@dataclass
class TuneResult:
    """Chosen setting, its time and the default's, and every trial."""

    items_log2: int
    blocks_per_cu: int
    seconds: float
    default_seconds: float
    chunk_bits: Tuple[int, int] = (-1, -1)
    trials: List[Trial] = field(default_factory=list)
    tuning_seconds: float = 0.0
    cached: bool = False

    @property
    def waves_per_cu(self) -> int:
        return self.blocks_per_cu * WAVES_PER_BLOCK

    @property
    def speedup(self) -> float:
        return self.default_seconds / self.seconds if self.seconds > 0 else 1.0

    def as_dict(self) -> dict:
        d = asdict(self)
        d["waves_per_cu"] = self.waves_per_cu
        d["speedup_vs_default"] = self.speedup
        d["trials"] = len(self.trials)
        return d


# --------------------------------------------------------------------------
# Cache.


def _library_identity() -> str:
    path = str(_native.LIBRARY_PATH)
    try:
        st = os.stat(path)
        return f"{os.path.basename(path)}:{st.st_size}:{int(st.st_mtime)}"
    except OSError:
        return "unknown"


def _device_identity() -> str:
    try:
        return _native.device_info()
    except Exception:
        return "unknown"


def plan_signature(plan) -> str:
    """Workload shape of a plan: states, jobs, residues, first class."""
    st = plan.stats()
    parts = [st.get(k, -1) for k in ("states", "jobs", "residues", "group", "set")]
    return "plan:" + "|".join(f"{float(v):.6g}" for v in parts)


def _cache_key(signature: str) -> str:
    return f"{_device_identity()}|{_library_identity()}|{signature}"


def _load_cache(path: str) -> Dict[str, dict]:
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _store_cache(path: str, key: str, result: TuneResult) -> None:
    data = _load_cache(path)
    data[key] = {
        "items_log2": result.items_log2,
        "blocks_per_cu": result.blocks_per_cu,
        "seconds": result.seconds,
        "default_seconds": result.default_seconds,
    }
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def default_cache_path() -> Optional[str]:
    """PERMHIP_AUTOTUNE_CACHE names a JSON cache file; unset means none."""
    return os.environ.get("PERMHIP_AUTOTUNE_CACHE") or None


# --------------------------------------------------------------------------
# Search.


def _descend(
    measure: Callable[[int, int], float],
    start: Tuple[int, int],
    items: Sequence[int],
    blocks: Sequence[int],
    tolerance: float,
    out_of_time: Callable[[], bool],
) -> Tuple[Tuple[int, int], float, Dict[Tuple[int, int], float]]:
    """Coordinate descent over (items_log2, blocks_per_cu)."""
    seen: Dict[Tuple[int, int], float] = {}

    def timed(setting):
        if setting not in seen:
            seen[setting] = measure(*setting)
        return seen[setting]

    best = start
    best_t = timed(start)

    def sweep(axis: int, values: Sequence[int]):
        nonlocal best, best_t
        for v in values:
            if out_of_time():
                return
            cand = (v, best[1]) if axis == 0 else (best[0], v)
            if cand == best:
                continue
            t = timed(cand)
            if t < best_t * (1.0 - tolerance):
                best, best_t = cand, t

    sweep(1, blocks)
    items_before = best[0]
    sweep(0, items)
    if best[0] != items_before:
        sweep(1, blocks)
    return best, best_t, seen


def tune_plan(
    plan,
    *,
    items: Sequence[int] = ITEMS_LOG2_CANDIDATES,
    blocks: Sequence[int] = BLOCKS_PER_CU_CANDIDATES,
    reps: int = 3,
    max_seconds: float = 30.0,
    tolerance: float = 0.01,
    cache: Optional[str] = None,
    clock: Callable[[], float] = time.perf_counter,
) -> TuneResult:
    """Tune a ``_native.Batch`` in place; returns the chosen setting.

    Each candidate is timed as the median of ``reps`` replays of the whole
    plan (after one replay that re-captures its graph).  Raises RuntimeError
    if two settings give different residues."""
    if reps < 1:
        raise ValueError("reps must be at least 1")
    t_begin = clock()
    st0 = plan.stats()
    start = (int(st0.get("items_log2", 15)), int(st0.get("blocks_per_cu", 2)))
    key = _cache_key(plan_signature(plan)) if cache else None
    if cache:
        hit = _load_cache(cache).get(key)
        if hit:
            plan.set_option(_native.BATCH_ITEMS_LOG2, int(hit["items_log2"]))
            plan.set_option(_native.BATCH_BLOCKS_PER_CU, int(hit["blocks_per_cu"]))
            st = plan.stats()
            return TuneResult(
                items_log2=int(hit["items_log2"]),
                blocks_per_cu=int(hit["blocks_per_cu"]),
                seconds=float(hit["seconds"]),
                default_seconds=float(hit["default_seconds"]),
                chunk_bits=(int(st.get("chunk_bits_min", -1)), int(st.get("chunk_bits_max", -1))),
                tuning_seconds=clock() - t_begin,
                cached=True,
            )

    reference: List[Optional[list]] = [None]
    trials: List[Trial] = []

    def measure(items_log2: int, blocks_per_cu: int) -> float:
        plan.set_option(_native.BATCH_ITEMS_LOG2, items_log2)
        plan.set_option(_native.BATCH_BLOCKS_PER_CU, blocks_per_cu)
        got = plan.run()
        if reference[0] is None:
            reference[0] = got
        elif got != reference[0]:
            raise RuntimeError(
                f"autotune: residues differ at items_log2={items_log2}, blocks_per_cu={blocks_per_cu}"
            )
        times = []
        for _ in range(reps):
            t0 = clock()
            plan.run(fetch=False)
            times.append(clock() - t0)
        t = statistics.median(times)
        st = plan.stats()
        trials.append(
            Trial(items_log2, blocks_per_cu, t,
                  (int(st.get("chunk_bits_min", -1)), int(st.get("chunk_bits_max", -1))))
        )
        return t

    # More workgroups per CU than the kernels keep resident only queue
    # behind the others; those counts are not tried.
    resident = int(st0.get("resident_blocks_per_cu", -1))
    if resident > 0:
        fitting = tuple(b for b in blocks if b <= resident)
        blocks = fitting or (min(blocks),)
        if start[1] > resident:
            start = (start[0], max(blocks))
    best, best_t, seen = _descend(
        measure, start, items, blocks, tolerance, lambda: clock() - t_begin > max_seconds
    )
    plan.set_option(_native.BATCH_ITEMS_LOG2, best[0])
    plan.set_option(_native.BATCH_BLOCKS_PER_CU, best[1])
    st = plan.stats()
    result = TuneResult(
        items_log2=best[0],
        blocks_per_cu=best[1],
        seconds=best_t,
        default_seconds=seen[start],
        chunk_bits=(int(st.get("chunk_bits_min", -1)), int(st.get("chunk_bits_max", -1))),
        trials=trials,
        tuning_seconds=clock() - t_begin,
    )
    if cache:
        _store_cache(cache, key, result)
    return result


def tune_single(
    n: int = 32,
    *,
    lanes: Sequence[int] = LANES_LOG2_CANDIDATES,
    blocks: Sequence[int] = BLOCKS_PER_CU_CANDIDATES,
    reps: int = 3,
    max_seconds: float = 30.0,
    tolerance: float = 0.01,
    seed: int = 0,
    cache: Optional[str] = None,
    clock: Callable[[], float] = time.perf_counter,
) -> TuneResult:
    """Tune the single-matrix entry points on a seeded order-n matrix with
    entries in [-10, 10] and its CRT primes; sets the process-wide tuning
    (lanes log2 = work items log2 + 6, workgroups per CU) to the best found.
    ``items_log2`` in the result is lanes log2 - 6."""
    from . import planner

    t_begin = clock()
    key = _cache_key(f"single:{n}") if cache else None
    if cache:
        hit = _load_cache(cache).get(key)
        if hit:
            _native.set_tuning(_native.TUNE_LANES_LOG2, int(hit["items_log2"]) + 6)
            _native.set_tuning(_native.TUNE_BATCH_BLOCKS, int(hit["blocks_per_cu"]))
            return TuneResult(
                items_log2=int(hit["items_log2"]),
                blocks_per_cu=int(hit["blocks_per_cu"]),
                seconds=float(hit["seconds"]),
                default_seconds=float(hit["default_seconds"]),
                tuning_seconds=clock() - t_begin,
                cached=True,
            )
    rng = np.random.default_rng(seed)
    A = rng.integers(-10, 11, size=(n, n)).astype(np.int64)
    primes, _ = planner.generic_crt_primes(planner.matrix_permanent_bound(A.tolist()))
    start = (_native.get_tuning(_native.TUNE_LANES_LOG2) - 6, _native.get_tuning(_native.TUNE_BATCH_BLOCKS))
    reference: List[Optional[list]] = [None]
    trials: List[Trial] = []

    def measure(items_log2: int, blocks_per_cu: int) -> float:
        _native.set_tuning(_native.TUNE_LANES_LOG2, items_log2 + 6)
        _native.set_tuning(_native.TUNE_BATCH_BLOCKS, blocks_per_cu)
        got = _native.permanent_residues(A, primes)
        if reference[0] is None:
            reference[0] = got
        elif got != reference[0]:
            raise RuntimeError(
                f"autotune: residues differ at lanes_log2={items_log2 + 6}, blocks_per_cu={blocks_per_cu}"
            )
        times = []
        for _ in range(reps):
            t0 = clock()
            _native.permanent_residues(A, primes)
            times.append(clock() - t0)
        t = statistics.median(times)
        trials.append(Trial(items_log2, blocks_per_cu, t))
        return t

    try:
        best, best_t, seen = _descend(
            measure, start, [v - 6 for v in lanes], blocks, tolerance, lambda: clock() - t_begin > max_seconds
        )
    except BaseException:
        _native.set_tuning(_native.TUNE_LANES_LOG2, start[0] + 6)
        _native.set_tuning(_native.TUNE_BATCH_BLOCKS, start[1])
        raise
    _native.set_tuning(_native.TUNE_LANES_LOG2, best[0] + 6)
    _native.set_tuning(_native.TUNE_BATCH_BLOCKS, best[1])
    result = TuneResult(
        items_log2=best[0],
        blocks_per_cu=best[1],
        seconds=best_t,
        default_seconds=seen[start],
        trials=trials,
        tuning_seconds=clock() - t_begin,
    )
    if cache:
        _store_cache(cache, key, result)
    return result
