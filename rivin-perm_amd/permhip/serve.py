"""Continuous-batching service for exact permanents.

Requests (integer matrices or 0/1 matrices) arrive at any time.  A scheduler
thread plans each request on the host (bounds, CRT primes, matching blocks
for 0/1 inputs), then admits waiting requests into device plans as in-flight
slots free up: the oldest waiting request always goes first (no starvation),
and the rest of a plan is filled with the cheapest waiting requests until the
plan's estimated cost (sum over jobs of n 2^(n-1) per prime) or its device
memory budget is reached.  Up to ``max_inflight`` plans run concurrently on
their own streams, so requests that arrive while others are on the GPU start
as soon as the device drains the earlier plans' work queues.  All residues of
a request are produced inside one plan and stay on the device until that
plan finishes; the host then reconstructs each request's integer by CRT.
On an idle server (nothing waiting or in flight) a request below
``direct_cost`` runs at once through the single-matrix device path instead,
which skips creating a plan.  The scheduler sleeps until a request arrives
when idle, and polls in-flight plans every 0.2 ms otherwise.

Failures stay with the requests they belong to: a request the device path
cannot take (an integer matrix of order 64 or more) is refused in
``submit``; an error while a request runs outside the plans fails that
request's future; an error while a plan is created, launched or polled
fails the futures of every request in that plan.  The scheduler thread
keeps serving the other requests in every case.

Latency (completion minus arrival) and throughput are recorded per request.
"""

from __future__ import annotations

import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _native, binary_prep, planner
from .batch import BATCH_MAX_ORDER, BATCH_MIN_ORDER


@dataclass
class _Part:
    matrix: np.ndarray  # int64 (n, n)
    primes: List[int]
    kind: str  # 'signed' (integer, signed lift) or 'count' (0/1 block, nonnegative)
    bound: int
    residues: Optional[List[int]] = None

    @property
    def cost(self) -> float:
        n = self.matrix.shape[0]
        return len(self.primes) * n * 2.0 ** (n - 1)

    @property
    def device_bytes(self) -> int:
        n = self.matrix.shape[0]
        return len(self.primes) * (4 * n * n + 48) + 64 * 16


@dataclass
class _Request:
    rid: int
    kind: str
    arrival: float
    future: Future
    parts: List[_Part] = field(default_factory=list)
    constant: int = 1  # product of parts computed outside the device plans
    done_parts: int = 0
    finished: float = 0.0
    failed: bool = False

    @property
    def cost(self) -> float:
        return sum(p.cost for p in self.parts)


# Largest order the device residue kernels take (the limit of
# perm_c_omp.c:885-888 and perm_modp_cuda.cu:744-747).
MAX_ORDER = 63
# Bytes of partial sums one work item holds (kItemSlots 8-byte slots).
_ITEM_PARTIAL_BYTES = 12 * 8


def _plan_integer(matrix) -> tuple:
    rows = [[int(x) for x in r] for r in matrix]
    n = len(rows)
    if any(len(r) != n for r in rows):
        raise ValueError("Matrix must be square")
    if n > MAX_ORDER:
        raise ValueError("matrix dimension must be less than 64")
    if n == 0:
        return [], 1
    if n == 1:
        return [], rows[0][0]
    arr = np.array(rows, dtype=np.int64)
    bound = planner.matrix_permanent_bound(arr)
    primes, _ = planner.generic_crt_primes(bound)
    return [_Part(arr, primes, "signed", bound)], 1


def _plan_binary(matrix) -> tuple:
    rows = binary_prep.normalize_binary_matrix(matrix, strict=True)
    if not rows:
        return [], 1
    if len(rows) == 1:
        return [], rows[0][0]
    feasible, blocks = binary_prep.matching_blocks(binary_prep.row_masks(rows))
    if not feasible:
        return [], 0
    parts = []
    for block in blocks:
        if len(block) == 1:
            continue
        if len(block) >= 64:
            raise ValueError("native binary-permanent blocks must have order less than 64")
        b = binary_prep.unpack_masks(block)
        bound = binary_prep.normalized_bound(b)
        primes = list(binary_prep.crt_primes_above_bound(bound))
        parts.append(_Part(np.array(b, dtype=np.int64), primes, "count", bound))
    return parts, 1


def _finish_part(part: _Part) -> int:
    if part.kind == "signed":
        product = 1
        for p in part.primes:
            product *= p
        value = planner.crt_known_moduli(part.residues, tuple(part.primes))
        return planner.signed_lift(value, part.bound, product)
    value = binary_prep.crt_nonnegative(part.residues, part.primes)
    if value > part.bound:
        raise RuntimeError("binary residues reconstructed outside the bound")
    return value


class PermanentServer:
    """Background scheduler; ``submit`` returns a Future of the exact integer."""

    def __init__(self, *, max_inflight: int = 3, plan_cost: float = 4e11,
                 plan_memory: int = 1 << 30, blocks_per_cu: int = 3, direct_cost: float = 2.0 ** 30):
        _native.require_device()
        self.max_inflight = max_inflight
        self.direct_cost = direct_cost
        self.plan_cost = plan_cost
        self.plan_memory = plan_memory
        self.blocks_per_cu = blocks_per_cu
        # Device bytes of one plan that do not depend on its requests: the
        # work items' partial sums.  The planner spreads a plan over about
        # 2^items_log2 items, plus up to 2^(items_log2 - 2) more per extra
        # (group, set, variant) class; 4 * 2^items_log2 items covers the
        # classes a served mix uses.
        items_log2 = _native.get_tuning(_native.TUNE_BATCH_ITEMS_LOG2)
        self.plan_overhead_bytes = 4 * (1 << items_log2) * _ITEM_PARTIAL_BYTES
        self.errors = 0  # requests failed by device or scheduler errors
        self._incoming: "queue.Queue[_Request]" = queue.Queue()
        self._waiting: List[_Request] = []
        self._inflight = []  # (Batch, [(request, part)])
        self._next_id = 0
        self._stop = threading.Event()
        self._wake = threading.Event()  # set by submit and close
        self._lock = threading.Lock()
        self.completed: List[_Request] = []
        self.plans_launched = 0
        self._thread = threading.Thread(target=self._loop, name="permhip-serve", daemon=True)
        self._thread.start()

    # -- client side ---------------------------------------------------
    def submit(self, matrix, kind: str = "integer") -> Future:
        if kind not in ("integer", "binary"):
            raise ValueError("kind must be 'integer' or 'binary'")
        fut: Future = Future()
        with self._lock:
            rid = self._next_id
            self._next_id += 1
        req = _Request(rid, kind, time.perf_counter(), fut)
        try:
            parts, constant = _plan_integer(matrix) if kind == "integer" else _plan_binary(matrix)
        except Exception as exc:  # validation errors go to the caller's future
            fut.set_exception(exc)
            return fut
        req.parts, req.constant = parts, constant
        self._incoming.put(req)
        self._wake.set()
        return fut

    def close(self, wait: bool = True) -> None:
        self._stop.set()
        self._wake.set()
        if wait:
            self._thread.join()

    # -- scheduler -----------------------------------------------------
    def _fail(self, req: _Request, exc: BaseException) -> None:
        """Fail one request's future (once); the scheduler carries on."""
        if req.failed or req.future.done():
            return
        req.failed = True
        req.finished = time.perf_counter()
        self.errors += 1
        req.future.set_exception(exc)
        self.completed.append(req)

    def _complete(self, req: _Request) -> None:
        if req.failed:
            return
        try:
            value = req.constant
            for part in req.parts:
                value *= _finish_part(part)
            req.finished = time.perf_counter()
            req.future.set_result(value)
        except Exception as exc:
            req.finished = time.perf_counter()
            req.future.set_exception(exc)
        self.completed.append(req)

    def _run_outside_plans(self, req: _Request) -> None:
        """Parts the persistent kernel does not serve (order < 5 or > 46) go
        through the single-matrix device launch."""
        for part in req.parts:
            n = part.matrix.shape[0]
            if not BATCH_MIN_ORDER <= n <= BATCH_MAX_ORDER:
                part.residues = _native.permanent_residues(part.matrix, part.primes)
                req.done_parts += 1

    def _take_incoming(self, req: _Request) -> None:
        """Parts outside the plans' orders, and a small request on an idle
        device, run at once through the single-matrix path; everything else
        waits for a plan.  An error fails this request only."""
        try:
            self._run_outside_plans(req)
            if (req.done_parts < len(req.parts) and not self._inflight and not self._waiting
                    and req.cost <= self.direct_cost):
                # Idle device and a small request: no plan to build.
                for part in req.parts:
                    if part.residues is None:
                        part.residues = _native.permanent_residues(part.matrix, part.primes)
                        req.done_parts += 1
        except Exception as exc:
            self._fail(req, exc)
            return
        if req.done_parts == len(req.parts):
            self._complete(req)
        else:
            self._waiting.append(req)

    def _launch_plan(self, chosen: List[_Request]) -> None:
        """One device plan for the chosen requests' remaining parts; an
        error while it is created or launched fails those requests."""
        mats, primes, owners = [], [], []
        for req in chosen:
            for part in req.parts:
                if part.residues is None:
                    mats.append(part.matrix)
                    primes.append(part.primes)
                    owners.append((req, part))
        plan = None
        try:
            plan = _native.Batch(mats, primes)
            plan.set_option(_native.BATCH_USE_GRAPH, 0)
            plan.set_option(_native.BATCH_BLOCKS_PER_CU, self.blocks_per_cu)
            plan.launch()
        except Exception as exc:
            if plan is not None:
                try:
                    plan.close()
                except Exception:
                    pass
            for req in chosen:
                self._fail(req, exc)
            return
        self.plans_launched += 1
        self._inflight.append((plan, owners))

    def _admit(self) -> None:
        while True:
            try:
                req = self._incoming.get_nowait()
            except queue.Empty:
                break
            self._take_incoming(req)
        while self._waiting and len(self._inflight) < self.max_inflight:
            chosen = [self._waiting.pop(0)]  # oldest first
            cost = chosen[0].cost
            mem = self.plan_overhead_bytes + sum(p.device_bytes for p in chosen[0].parts)
            for req in sorted(self._waiting, key=lambda r: r.cost):
                if cost + req.cost > self.plan_cost:
                    break
                rmem = sum(p.device_bytes for p in req.parts)
                if mem + rmem > self.plan_memory:
                    continue
                chosen.append(req)
                cost += req.cost
                mem += rmem
            for req in chosen[1:]:
                self._waiting.remove(req)
            self._launch_plan(chosen)

    def _poll(self) -> bool:
        progressed = False
        still = []
        for plan, owners in self._inflight:
            try:
                flat = plan.poll(block=False)
            except Exception as exc:
                # The plan's results are lost: fail its requests, drop it.
                progressed = True
                for req, _ in owners:
                    self._fail(req, exc)
                try:
                    plan.close()
                except Exception:
                    pass
                continue
            if flat is None:
                still.append((plan, owners))
                continue
            progressed = True
            pos = 0
            for req, part in owners:
                k = len(part.primes)
                part.residues = flat[pos:pos + k]
                pos += k
                req.done_parts += 1
                if req.done_parts == len(req.parts):
                    self._complete(req)
            plan.close()
        self._inflight = still
        return progressed

    def _fail_all(self, exc: BaseException) -> None:
        """An error outside any one request or plan: fail every request the
        scheduler holds, so no future is left unresolved."""
        for plan, owners in self._inflight:
            for req, _ in owners:
                self._fail(req, exc)
            try:
                plan.close()
            except Exception:
                pass
        self._inflight = []
        for req in self._waiting:
            self._fail(req, exc)
        self._waiting = []

    def _loop(self) -> None:
        while True:
            # Cleared before the queue is drained: a request that arrives
            # after the drain sets the event again, so no wake-up is lost.
            self._wake.clear()
            try:
                self._admit()
                progressed = self._poll()
            except Exception as exc:
                self._fail_all(exc)
                progressed = True
            idle = not self._inflight and not self._waiting and self._incoming.empty()
            if self._stop.is_set() and idle:
                return
            if idle:
                self._wake.wait()
            elif not progressed:
                self._wake.wait(0.0002)


def latency_floor(count: int = 40, seed: int = 6) -> dict:
    """Median end-to-end latency of small requests submitted one at a time
    to an idle server (integer 12x12 entries U[-50,50], 0/1 20x20 density
    0.5), next to the median host planning time of the same requests
    (bound, primes, and for 0/1 matrices the matching and block split)."""
    import statistics

    rng = np.random.default_rng(seed)
    server = PermanentServer(max_inflight=1)
    out = {}
    try:
        for kind, n in (("integer", 12), ("binary", 20)):
            lat, plan = [], []
            for _ in range(count):
                if kind == "integer":
                    M = rng.integers(-50, 51, (n, n)).astype(np.int64)
                    planner_fn = _plan_integer
                else:
                    M = (rng.random((n, n)) < 0.5).astype(np.int64)
                    planner_fn = _plan_binary
                t0 = time.perf_counter()
                planner_fn(M)
                plan.append(time.perf_counter() - t0)
                t0 = time.perf_counter()
                server.submit(M, kind).result()
                lat.append(time.perf_counter() - t0)
            out[f"{kind}_n{n}"] = {"p50_ms": 1e3 * statistics.median(lat),
                                   "host_planning_p50_ms": 1e3 * statistics.median(plan)}
    finally:
        server.close()
    return out


def benchmark(seconds: float = 2.0, rate: float = 200.0, seed: int = 5, verify: int = 24, progress=None) -> dict:
    """Replay the seeded Poisson stream against a live server; report
    latency percentiles, throughput and backlog, and check a sample of
    answers against the single-request path.  ``progress(text)``, when
    given, is called every 1000 submitted and every 1000 answered
    requests."""
    from . import workloads

    stream = workloads.serve_stream(rate=rate, seconds=seconds, seed=seed)
    server = PermanentServer()
    futures = []
    t0 = time.perf_counter()
    for arrival, kind, matrix in stream:
        delay = t0 + arrival - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        futures.append((kind, matrix, server.submit(matrix, kind)))
        if progress and len(futures) % 1000 == 0:
            progress(f"serve: {len(futures)} requests submitted at {time.perf_counter() - t0:.1f} s")
    submit_end = time.perf_counter()
    results = []
    for _, _, f in futures:
        results.append(f.result())
        if progress and len(results) % 1000 == 0:
            progress(f"serve: {len(results)} answered at {time.perf_counter() - t0:.1f} s")
    t_end = time.perf_counter()
    server.close()
    lat = sorted(r.finished - r.arrival for r in server.completed)

    def pct(q):
        return lat[min(len(lat) - 1, int(q * len(lat)))] if lat else None

    # Same answer alone as in the mix.
    rng = np.random.default_rng(seed + 1)
    idx = rng.choice(len(futures), size=min(verify, len(futures)), replace=False)
    solo = PermanentServer(max_inflight=1)
    mismatches = 0
    for i in idx:
        kind, matrix, _ = futures[i]
        if solo.submit(matrix, kind).result() != results[i]:
            mismatches += 1
    solo.close()
    if mismatches:
        raise RuntimeError(f"{mismatches} served answers differ from the same request alone")
    mulmods = 0.0
    for kind, matrix, _ in futures:
        n = matrix.shape[0]
        mulmods += n * 2.0 ** (n - 1)
    return {
        "workload": (
            f"Poisson {rate:g} req/s for {seconds:g} s (default_rng({seed})), 50/50 integer n~U[10,34] "
            "entries U[-50,50] / 0-1 n~U[16,38] density U[0.3,0.7]"
        ),
        "requests": len(futures),
        "p50_latency_s": pct(0.50),
        "p99_latency_s": pct(0.99),
        "throughput_req_s": len(futures) / (t_end - t0),
        "arrival_window_s": submit_end - t0,
        "drain_after_last_arrival_s": t_end - submit_end,
        "plans_launched": server.plans_launched,
        "verified_alone_vs_mix": int(len(idx)),
        "offered_gray_mulmods_per_prime": mulmods,
        "small_request_latency": latency_floor(),
    }
