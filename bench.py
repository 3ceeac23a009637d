#!/usr/bin/env python3
"""Benchmark of the exact modular permanent path on one MI355X.

Workload (one step): the exact permanent of the 32 x 32 integer matrix
A = default_rng(20261203).integers(-10, 11), i.e. 8 Glynn residues modulo the
ascending 31-bit primes from 2^30 (2^31 Gray states each) plus host CRT.  The
matrix and its tables are resident in HBM (a device plan created before the
timed region); a step replays the captured launch and reconstructs the
integer.

Printed as one JSON line (rank 0 only; this build runs on one GPU):
  value        reference-equivalent Gray-code mulmods per second: n 2^(n-1)
               per residue, the reference's n-factor row product per state
               (perm_c_omp.c:40-47).  The kernel executes no modular
               reduction in its walk (exact 16-bit row sums, exact group
               products, each state's product exact in radix-2^31 limbs,
               one walk for all eight primes); `executed` reports what it
               does run per state and per second
  roofline     the dominant kernel's executed VALU lane operations per
               second (rocprof SQ_INSTS_VALU per launch x 64 lanes, over
               the kernel time measured live with HIP events) against the
               nominal INT32 VALU peak, 256 CUs x 64 lanes x 2.4 GHz
  cpu_baseline the reference's own OpenMP kernel (perm_c_omp.c, compiled
               from its sources into oracle/_ref) on every host core: one
               full residue of the headline matrix, scaled by the prime count
  plain / batch / serve   the per-matrix plain launch, the 4096-matrix
               uneven batch in one persistent launch, and the served stream
"""

from __future__ import annotations

import argparse
import re
import json
import math
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "rivin-perm_amd"))

METRIC = "Gray-code mulmods/s & exact perms/s at n=32 vs plain launch; batch p99 latency"
HEADLINE_SEED = 20261203
HEADLINE_N = 32
BATCH_SEED = 4


def headline_matrix() -> np.ndarray:
    return np.random.default_rng(HEADLINE_SEED).integers(-10, 11, (HEADLINE_N, HEADLINE_N)).astype(np.int64)


def batch_matrices(count: int = 4096):
    rng = np.random.default_rng(BATCH_SEED)
    mats = []
    for _ in range(count):
        n = int(rng.integers(12, 29))
        mats.append(rng.integers(-100, 101, (n, n)).astype(np.int64))
    return mats


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# VALU instructions per Gray state of the exact-accumulation kernel at
# n = 32 in groups of 4 (eight groups), counted in the compiled loop of
# batch_glynn_kernel<4, 8, false, false, true> (an eight-state block):
# flips, group products, magnitudes and sign, the limb products and splits,
# the column accumulations.
EXACT_MIX_N32_G4 = {
    "v_pk_add_u16/v_pk_sub_i16": 18,
    "v_mul_i32_i24(_sdwa)": 24,
    "v_sub_u32+v_max_i32": 16,
    "v_bitop3_b32 (sign, three-input XOR)": 4,
    "v_mad_u64_u32": 28,
    "v_and_b32": 17.1,
    "v_lshrrev_b64": 10.1,
    "v_alignbit_b32": 7,
    "v_mad_i64_i32": 8,
    "other": 5.3,
}  # 137.5 in all; rocprof counts 133.1 per state (SQ_INSTS_VALU)

# The signed-limb form of the same walk (variant 6: signed group products
# multiplied in signed limbs, no magnitudes or sign per state), as counted
# in the compiled loop.
EXACT_SIGNED_MIX_N32_G4 = {
    "v_pk_add_u16/v_pk_sub_i16": 14,
    "v_mul_i32_i24(_sdwa)": 24,
    "v_mad_i64_i32 (20 signed limb products, 8 accumulations)": 28,
    "v_mad_u64_u32 (unsigned limb columns)": 8,
    "v_and_b32": 18,
    "v_ashrrev_i64": 7,
    "v_lshrrev_b64": 4,
    "v_alignbit_b32": 7,
    "other": 1,
}  # 111 in all


def instruction_mix(n: int, group: int, k: int = 1, paired: bool = False) -> dict:
    """VALU instructions per Gray state of one walk of gray_walk<n, group>
    serving a set of k primes (packed rows; the unrolled 16-state step
    amortised, as counted in the compiled loop).

    Shared by the set: row-sum flips (one packed op per two rows; only the
    chunk's middle state flips in a lane-dependent direction, which is
    negligible per state), the exact group products, and for k > 1 the exact
    64-bit products of adjacent groups.  Per prime: the Montgomery tree (two
    v_mad_i64_i32 and one v_mul_lo_u32 per product; one of each per reduced
    pair product) and the in-order signed accumulation (one v_mad_i64_i32).
    Paired sets replace, per two states, the tree's two last products and two
    accumulations by two exact products, one reduction and one accumulation,
    and negate one group per even state."""
    if group == 0:
        raise ValueError("the roofline mix covers packed row sums")
    groups = -(-n // group)
    words = -(-n // 2)
    mix = {
        "v_pk_add_u16": words,
        "v_mul_i32_i24_sdwa": n - groups,
        "v_add_u32": 2,  # loop control and register moves
    }
    if k == 1:
        mix["v_mad_i64_i32"] = 2 * (groups - 1) + 1
        mix["v_mul_lo_u32"] = groups - 1
    else:
        pairs = groups // 2
        leaves = pairs + groups % 2
        if paired:
            mix["v_mad_i64_i32"] = pairs + k * (pairs + 2 * (leaves - 2) + 2)
            mix["v_mul_lo_u32"] = k * (pairs + leaves - 2 + 0.5)
            mix["v_add_u32"] += 0.5
        else:
            mix["v_mad_i64_i32"] = pairs + k * (pairs + 2 * (leaves - 1) + 1)
            mix["v_mul_lo_u32"] = k * (pairs + leaves - 1)
    return mix


# Nominal INT32 VALU peak of one MI355X: 256 CUs, 4 SIMDs of 16 lanes each
# completing one 32-bit operation per lane per cycle, 2.4 GHz peak clock.
NOMINAL_INT32_LANE_OPS = 256 * 64 * 2.4e9


def roofline(kernel_ms: float, states_per_launch: float, compute_units: int, rates: dict, mix: dict) -> dict:
    """Executed VALU lane operations per second of the dominant kernel
    against the nominal INT32 peak.  The instruction count per launch is
    rocprof's SQ_INSTS_VALU from the committed counter profile of this
    kernel at this shape (profiles/glynn_headline_n32_pmc.txt); the time is
    this run's HIP-event kernel time.  The fastest 32-bit instruction
    measured here is reported beside the nominal peak: adds and shifts
    issue faster than one per lane per cycle on gfx950, so a stream rich in
    them can pass the nominal figure."""
    insts = _profile_value("SQ_INSTS_VALU")
    nominal = compute_units * 64 * 2.4e9
    fastest = max(rates.get("v_add_u32", 0.0), rates.get("v_ashrrev_i32", 0.0))
    out = {
        "bound": "valu",
        "achieved": None,
        "peak": nominal / 1e12,
        "unit": "Tlane-op/s (INT32 VALU)",
        "frac": None,
        "traffic": traffic_from_profiles(),
        "kernel_ms": kernel_ms,
        "peak_measured_fastest_int32": fastest / 1e12,
        "valu_insts_per_state_model": sum(mix.values()),
        "instructions_per_state_model": mix,
        "source": "profiles/glynn_headline_n32_pmc.txt (SQ_INSTS_VALU per launch)",
    }
    if insts is not None:
        achieved = insts * 64 / (kernel_ms * 1e-3)
        out.update({
            "achieved": achieved / 1e12,
            "frac": achieved / nominal,
            "frac_of_measured_fastest": achieved / fastest if fastest else None,
            "valu_wave_insts_per_launch": insts,
            "valu_insts_per_state_measured": insts * 64 / states_per_launch,
        })
    return out


def _profile_value(key: str):
    """A figure of the committed rocprof counter summary of the headline
    kernel, or None."""
    path = ROOT / "profiles" / "glynn_headline_n32_pmc.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith(key + ":"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def traffic_from_profiles():
    """HBM bytes per launch of the headline kernel (FETCH_SIZE + WRITE_SIZE,
    copy-calibrated), else None."""
    return _profile_value("hbm_bytes_per_launch")


def binary36_leg() -> dict:
    """The binary configuration: the exact count of binary36 end to end
    (matching, bound, its descending 31-bit primes in one zero-skipping
    launch, CRT) and the mod-3 kernel on the same matrix, whose residue
    must equal the count mod 3."""
    import perm_binary
    import perm_mod3
    from permhip import workloads

    B = np.array(workloads.binary36(), dtype=np.int64)

    def warm_median(fn, reps=3):
        first = fn()  # the first call also loads the kernels
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            if fn() != first:
                raise RuntimeError("binary36 results differ between calls")
            times.append(time.perf_counter() - t0)
        return first, statistics.median(times)

    count, dt = warm_median(lambda: perm_binary.binary_permanent(B))
    m3, dt3 = warm_median(lambda: perm_mod3.permanent_mod3_fast(B))
    if m3 != count % 3:
        raise RuntimeError("binary36 count and mod-3 residue disagree")
    return {
        "workload": "binary36: 36x36 Bernoulli(0.5), default_rng(3), resampled until no empty row or column",
        "s_exact": dt,
        "value": str(count),
        "s_mod3": dt3,
        "mod3": m3,
        "timing": "median of 3 calls after a first one, end to end (host planning, launch, CRT)",
    }


def binary36_cpu_baseline(threads: int) -> dict:
    """The reference's perm_binary_ext.binary_permanent_mod_p and
    perm_mod3_omp.permanent_mod3 (compiled from their sources into
    oracle/_ref; the oracle's restatements when those are not built) on
    leading blocks of binary36 -- 30x30 for one binary prime, 32x32 mod 3 --
    with the device's time on the same blocks beside them."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle  # checker/baseline only
    import perm_mod3
    from permhip import _native, workloads

    exts = oracle.reference_extensions()
    B = np.array(workloads.binary36(), dtype=np.int64)
    p = 2_147_483_647
    b30 = B[:30, :30].copy()
    b32 = B[:32, :32].copy()
    out = {}
    if "perm_binary_ext" in exts:
        kind, fn = "reference", (lambda: exts["perm_binary_ext"].binary_permanent_mod_p(b30.tolist(), p, threads))
    else:
        kind, fn = "port", (lambda: oracle.C.binary_mod_p(b30.astype(np.uint8), p, threads))
    t0 = time.perf_counter()
    cpu_r = fn()
    cpu_s = time.perf_counter() - t0
    _native.binary_permanent_residues(b30.astype(np.uint8), [p])
    t0 = time.perf_counter()
    gpu_r = _native.binary_permanent_residues(b30.astype(np.uint8), [p])[0]
    gpu_s = time.perf_counter() - t0
    if gpu_r != cpu_r:
        raise RuntimeError("binary residue differs from the CPU baseline's")
    out["binary_mod_p"] = {"kind": kind, "cores": threads, "sample": f"leading 30x30 block, p = {p}",
                           "cpu_s": cpu_s, "gpu_s": gpu_s, "speedup": cpu_s / gpu_s}
    if "perm_mod3_omp" in exts:
        kind, fn = "reference", (lambda: exts["perm_mod3_omp"].permanent_mod3(b32 % 3, threads))
    else:
        kind, fn = "port", (lambda: oracle.C.mod3(b32, threads))
    t0 = time.perf_counter()
    cpu_m = fn()
    cpu_s = time.perf_counter() - t0
    perm_mod3.permanent_mod3_fast(b32)
    t0 = time.perf_counter()
    gpu_m = perm_mod3.permanent_mod3_fast(b32)
    gpu_s = time.perf_counter() - t0
    if gpu_m != cpu_m:
        raise RuntimeError("mod-3 residue differs from the CPU baseline's")
    out["mod3"] = {"kind": kind, "cores": threads, "sample": "leading 32x32 block",
                   "cpu_s": cpu_s, "gpu_s": gpu_s, "speedup": cpu_s / gpu_s}
    return out


def host_cpu() -> dict:
    """Host cores this process may use and the CPU model (/proc/cpuinfo)."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cores": cores, "model": model, "os_cpu_count": os.cpu_count()}


def cpu_baseline(matrix: np.ndarray, primes, threads: int) -> dict:
    """The reference's OpenMP Glynn (perm_c_ext.permanent_mod_p, compiled
    from perm_c_omp.c) for one full residue of the headline matrix: 32 x 32,
    the first CRT prime, 2^31 Gray states, on every host core.  The exact
    permanent needs the same walk for each of the primes, so its CPU time is
    this residue's time times the prime count."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle  # checker/baseline only

    exts = oracle.reference_extensions()
    p = primes[0]
    reduced = [[int(v) % p for v in row] for row in matrix]
    if "perm_c_ext" in exts:
        ext = exts["perm_c_ext"]
        kind = "reference"
        fn = lambda: ext.permanent_mod_p(reduced, p, threads)  # noqa: E731
    else:
        kind = "port"
        fn = lambda: oracle.C.permanent_mod_p(reduced, p, threads)  # noqa: E731
    t0 = time.perf_counter()
    residue = fn()
    dt = time.perf_counter() - t0
    n = len(reduced)
    mulmods = n * 2.0 ** (n - 1)
    cpu = host_cpu()
    return {
        "value": mulmods / dt,
        "unit": "mulmod/s",
        "cores": threads,
        "kind": kind,
        "sample": (
            f"perm_c_ext.permanent_mod_p on the {n}x{n} headline matrix modulo its first CRT prime ({p}): "
            f"one full residue, 2^{n - 1} Glynn states, {dt:.2f} s on {threads} threads"
        ),
        "residue": int(residue),
        "residue_s": dt,
        "exact_n32_s": dt * len(primes),
        "cpu_model": cpu["model"],
        "host_cores": cpu["cores"],
    }


POLY_N = 26
POLY_SEED = 9


def polynomial_board() -> list:
    rng = np.random.default_rng(POLY_SEED)
    return rng.integers(0, 2, (POLY_N, POLY_N)).tolist()


def polynomial_leg() -> dict:
    """Exact rook polynomial of a seeded 26x26 0/1 board through the
    perm_polynomial mirror, and the coefficient-valued Glynn kernel's rate
    (n - 1)(n + 2) modular products per Gray state per prime: each of the
    n - 1 affine factors after the first costs one fma2 (two products) per
    lower coefficient plus two plain products."""
    import perm_polynomial
    from permhip import _native

    A = polynomial_board()
    t0 = time.perf_counter()
    rook = perm_polynomial.rook_polynomial(A)
    dt = time.perf_counter() - t0
    n = POLY_N
    ones = np.ones((n, n), dtype=np.int64)
    board = np.array(A, dtype=np.int64)
    products = (n - 1) * (n + 2) * 2.0 ** (n - 1) * 4
    rates = {}
    # 30-bit primes (the plan's; signed lazy Montgomery) and 31-bit primes
    # (unsigned Montgomery with one lazy reduction per coefficient).
    for name, primes in (("30-bit", perm_polynomial._descending_primes(1 << 100)[:4]),
                         ("31-bit", [2_147_483_647, 2_147_483_629, 2_147_483_587, 2_147_483_579])):
        _native.profile_enable(True)
        _native.affine_permanent_residues(ones, board, primes)
        ms, _ = _native.profile_read()
        _native.profile_enable(False)
        rates[name] = {"kernel_ms_4_primes": ms, "products_per_s": products / (ms * 1e-3)}
    return {
        "workload": f"rook polynomial of a {n}x{n} 0/1 board, default_rng({POLY_SEED}).integers(0, 2)",
        "s_exact": dt,
        "r_n_permanent": str(rook[n]),
        "kernel_ms_4_primes": rates["30-bit"]["kernel_ms_4_primes"],
        "products_per_s": rates["30-bit"]["products_per_s"],
        "by_prime_width": rates,
    }


def polynomial_cpu_baseline(threads: int) -> dict:
    """The reference's own perm_polynomial_ext.affine_permanent_mod (built
    from its source into oracle/_ref) on per(J + x A) of the same board,
    one prime, or the oracle's restatement when that is not built."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle  # checker/baseline only

    A = polynomial_board()
    n = POLY_N
    ones = [[1] * n for _ in range(n)]
    p = 2_147_483_647
    ext = oracle.reference_extensions().get("perm_polynomial_ext")
    if ext is not None:
        kind, fn = "reference", (lambda: ext.affine_permanent_mod(ones, A, p, threads))
    else:
        kind, fn = "port", (lambda: oracle.C.affine_mod_p(ones, A, p, threads))
    t0 = time.perf_counter()
    fn()
    dt = time.perf_counter() - t0
    products = (n - 1) * (n + 2) * 2.0 ** (n - 1)
    return {"value": products / dt, "unit": "modular products/s", "cores": threads, "kind": kind,
            "sample": f"affine_permanent_mod(J, A, {p}) on the {n}x{n} board, one prime, {dt:.2f} s"}


def published_workloads() -> dict:
    """End-to-end times (host planning, device residues, CRT) of workloads
    the reference publishes timings for, next to its numbers
    (BINARY_APPLICATIONS.md:155-172, README.md:594-596)."""
    import perm
    import perm_binary
    from permhip import grids

    out = {}
    G = grids.grid_biadjacency((4, 4, 5), periodic_axes=(0, 1))
    t0 = time.perf_counter()
    count = perm_binary.binary_permanent(G)
    dt = time.perf_counter() - t0
    if count != 2551276235535120:
        raise RuntimeError("40x40 periodic grid count is wrong")
    out["binary_40x40_periodic_grid_exact"] = {
        "s": dt, "value": str(count),
        "reference_s": {"GH200 CUDA end-to-end": 22.73, "GH200 CPU 64 threads": 242.38},
    }
    # perm(S_37) and perm(S_41) as checked in by the reference (schur_permanents.txt).
    known = {37: -339696021338069660245, 41: 1617562409111471411255401}
    for n, ref in ((37, {"64-core CPU C/OpenMP": 18.43}), (41, {"GH200 CUDA 40x40-minor necklace": 47.39})):
        t0 = time.perf_counter()
        value = perm.permanent(n)
        dt = time.perf_counter() - t0
        if value != known[n]:
            raise RuntimeError(f"perm(S_{n}) is wrong")
        out[f"schur_{n}_exact"] = {"s": dt, "value": str(value), "reference_s": ref}
    # The reference's device entry point at order 41 plans descending 32-bit
    # primes = 1 (mod 41) (perm_cuda.py:466-538); the mirror runs the same
    # plan on the unsigned-Montgomery necklace kernel.
    import perm_cuda

    t0 = time.perf_counter()
    value = perm_cuda.permanent_cuda(41, verbose=False)
    dt = time.perf_counter() - t0
    if value != known[41]:
        raise RuntimeError("perm_cuda.permanent_cuda(41) is wrong")
    out["perm_cuda_permanent_cuda_41"] = {
        "s": dt, "value": str(value),
        "primes": perm_cuda._plan_schur_crt_primes(41, perm_cuda.schur_permanent_bound(41)),
        "reference_s": {"GH200 CUDA 40x40-minor necklace": 47.39},
    }
    return out


def wide_moduli() -> dict:
    """Kernel throughput of the wide-modulus Glynn kernels, one modulus of
    each arithmetic, on a 30x30 matrix (2^29 Gray states, 30 mulmods each);
    every residue is checked against the exact permanent reduced."""
    import perm
    from permhip import _native

    n = 30
    A = np.random.default_rng(7).integers(-(10 ** 6), 10 ** 6 + 1, (n, n)).astype(np.int64)
    exact = perm.matrix_permanent(A)
    cases = {
        "mont32 q=2^32-5": (1 << 32) - 5,
        "m61 q=2^61-1": (1 << 61) - 1,
        "pm61 q=2^61-31": (1 << 61) - 31,
        "mont64 q=2^61+1": (1 << 61) + 1,
        "mont64 q=2^63-25": (1 << 63) - 25,
    }
    mulmods = n * 2.0 ** (n - 1)
    out = {}
    for name, q in cases.items():
        best = None
        for _ in range(3):
            _native.profile_enable(True)
            ms0, _ = _native.profile_read()
            got = _native.permanent_residues(A, [q])[0]
            ms1, _ = _native.profile_read()
            _native.profile_enable(False)
            if got != exact % q:
                raise RuntimeError(f"wide residue mod {q} is wrong")
            best = ms1 - ms0 if best is None else min(best, ms1 - ms0)
        out[name] = {"kernel_ms": best, "mulmod_per_s": mulmods / (best * 1e-3)}
    return out


def _launch_shape(plan) -> dict:
    """Chunk size and workgroups per CU a plan runs with, and how the
    autotuner chose them."""
    st = plan.stats()
    out = {
        "items_log2": int(st["items_log2"]),
        "chunk_bits": [int(st["chunk_bits_min"]), int(st["chunk_bits_max"])],
        "blocks_per_cu": int(st["blocks_per_cu"]),
        "resident_blocks_per_cu": int(st.get("resident_blocks_per_cu", -1)),
        "waves_per_cu": 4 * (min(int(st["blocks_per_cu"]), int(st["resident_blocks_per_cu"]))
                             if int(st.get("resident_blocks_per_cu", -1)) > 0 else int(st["blocks_per_cu"])),
        "chosen_by": "defaults",
    }
    if plan.tuning is not None:
        t = plan.tuning.as_dict()
        out["chosen_by"] = "autotune"
        out["autotune"] = {k: t[k] for k in ("seconds", "default_seconds", "speedup_vs_default", "trials",
                                              "tuning_seconds")}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 4096-matrix batch leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--no-published", action="store_true",
                    help="skip the workloads the reference publishes timings for")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="mean Poisson arrival rate (req/s) of the served leg")
    ap.add_argument("--kernel-only", action="store_true", help="headline launches only (for profilers)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="run one plain device copy of BYTES (HBM counter calibration) and exit")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default chunk size and workgroups per CU (no tuning trials)")
    ap.add_argument("--items-log2", type=int, default=None,
                    help="headline plan: fixed work-item target (log2), no tuning")
    ap.add_argument("--blocks-per-cu", type=int, default=None,
                    help="headline plan: fixed workgroups per CU, no tuning")
    ap.add_argument("--rook-residues", type=int, default=0, metavar="N",
                    help="coefficients of per(J + x A) for the seeded N x N 0/1 board modulo 4 primes below "
                         "2^30, once, and exit (one affine kernel launch, for profilers)")
    ap.add_argument("--wide-residue", type=int, default=0, metavar="BITS",
                    help="perm(A) of the seeded 30x30 wide-moduli matrix modulo 2^61 - 1 (BITS=61) or "
                         "2^32 - 5 (BITS=32), once, and exit (one wide Glynn launch, for profilers)")
    ap.add_argument("--binary36", type=int, default=0, metavar="REPS",
                    help="only the binary36 workload's 4 residues, REPS times (for profilers); "
                         "--skip-zeros-permille sets the zero-skipping threshold")
    ap.add_argument("--skip-zeros-permille", type=int, default=None)
    ap.add_argument("--serve-breakdown", type=int, default=0, metavar="SAMPLES",
                    help="time SAMPLES requests alone per (kind, order) of the served stream, scale by the stream's "
                         "counts, print the device-time table and exit")
    ap.add_argument("--skip-blocks", type=int, default=None, choices=(0, 1),
                    help="0/1 tables on the zero-skipping kernels: 1 walks in blocks of 32 states (the default), "
                         "0 tests every state")
    ap.add_argument("--skip-exact-min-primes", type=int, default=None, metavar="K",
                    help="0/1 block tables of at most eight row groups with K primes or more take the block walk "
                         "with exact accumulation (default 5; 0 keeps every table on the per-prime walk)")
    ap.add_argument("--frontier", type=int, default=0, metavar="REPS",
                    help="only the frontier DP on the open 12x16 grid's biadjacency (order 96), REPS times")
    ap.add_argument("--batch-only", type=int, default=0, metavar="REPS",
                    help="only the 4096-matrix batch at the default launch shape: one untimed run, then "
                         "REPS timed runs (for profilers)")
    ap.add_argument("--schur-residue", type=int, default=0, metavar="N",
                    help="compute perm(S_N) modulo the first prime = 1 (mod N) above 2^30 and exit "
                         "(one necklace or Glynn launch, for profilers)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py runs on one GPU only: there is no multi-GPU path "
            "(the residues would shard by prime, or each residue by Gray range, with a host gather)",
            file=sys.stderr,
        )
        return 2
    rank = int(os.environ.get("RANK", "0"))

    from permhip import _native, planner
    from permhip.batch import PermanentBatch

    if args.skip_blocks is not None:
        _native.set_tuning(_native.TUNE_SKIP_BLOCKS, args.skip_blocks)
    if args.skip_exact_min_primes is not None:
        _native.set_tuning(_native.TUNE_SKIP_EXACT_MIN_PRIMES, args.skip_exact_min_primes)

    if args.copy_calibration:
        _native.copy_calibration(args.copy_calibration, 1)
        log(f"copied {args.copy_calibration} bytes")
        return 0
    if args.rook_residues:
        import perm_polynomial

        n = args.rook_residues
        board = np.random.default_rng(POLY_SEED).integers(0, 2, (n, n)).astype(np.int64)
        primes = perm_polynomial._descending_primes(1 << 100)[:4]
        t0 = time.perf_counter()
        rows = _native.affine_permanent_residues(np.ones((n, n), dtype=np.int64), board, primes)
        log(f"per(J + xA), n={n}, primes {primes}: {len(rows)} coefficient vectors in {time.perf_counter() - t0:.3f} s")
        return 0
    if args.wide_residue:
        q = (1 << 61) - 1 if args.wide_residue == 61 else (1 << 32) - 5
        A = np.random.default_rng(7).integers(-(10 ** 6), 10 ** 6 + 1, (30, 30)).astype(np.int64)
        t0 = time.perf_counter()
        r = _native.permanent_residues(A, [q])[0]
        log(f"perm(A) mod {q} = {r} in {time.perf_counter() - t0:.3f} s")
        return 0
    if args.binary36:
        from permhip import binary_prep, workloads

        if args.skip_zeros_permille is not None:
            _native.set_tuning(_native.TUNE_SKIP_ZEROS_PERMILLE, args.skip_zeros_permille)
        B = np.array(workloads.binary36(), dtype=np.int64)
        primes = [binary_prep.descending_prime(i) for i in range(4)]
        for rep in range(args.binary36):
            t0 = time.perf_counter()
            r = _native.permanent_residues(B, primes)
            log(f"binary36 residues {r} in {(time.perf_counter() - t0) * 1e3:.1f} ms (rep {rep})")
        return 0
    if args.serve_breakdown:
        import collections

        import perm
        import perm_binary
        from permhip import workloads

        by_order = collections.defaultdict(list)
        for _, kind, A in workloads.serve_stream(args.serve_rate, args.serve_seconds):
            by_order[(kind, A.shape[0])].append(A)
        totals = collections.Counter()
        print("kind | order | requests | ms per request (mean of the first SAMPLES, alone) | estimated s")
        for (kind, n), mats in sorted(by_order.items()):
            fn = perm.matrix_permanent if kind == "integer" else perm_binary.binary_permanent
            fn(mats[0])  # warm the plan shapes
            sample = mats[: args.serve_breakdown]
            t0 = time.perf_counter()
            for A in sample:
                fn(A)
            each = (time.perf_counter() - t0) / len(sample)
            totals[kind] += each * len(mats)
            print(f"{kind} | {n} | {len(mats)} | {each * 1e3:.2f} | {each * len(mats):.2f}", flush=True)
        print(f"total | | | | integer {totals['integer']:.1f} s, binary {totals['binary']:.1f} s, "
              f"all {sum(totals.values()):.1f} s")
        return 0
    if args.batch_only:
        from permhip.batch import PermanentBatch

        batch = PermanentBatch(batch_matrices())
        try:
            warm = batch.run()
            for rep in range(args.batch_only):
                t0 = time.perf_counter()
                got = batch.run()
                dt = time.perf_counter() - t0
                if got != warm:
                    raise RuntimeError("batch results changed between runs")
                st = batch.stats()
                log(f"batch of {len(warm)} matrices in {dt * 1e3:.1f} ms (kernels {st['last_ms']:.1f} ms, "
                    f"{st['mulmods'] / dt:.3e} mulmod/s, rep {rep})")
        finally:
            batch.close()
        return 0
    if args.frontier:
        import perm_binary
        from permhip import grids

        G = grids.grid_biadjacency((12, 16))
        for rep in range(args.frontier):
            t0 = time.perf_counter()
            v = perm_binary.binary_permanent_frontier(G)
            log(f"frontier 12x16 grid: {v} in {(time.perf_counter() - t0) * 1e3:.1f} ms (rep {rep})")
        return 0
    if args.schur_residue:
        n = args.schur_residue
        p = planner.find_primes_congruent_to_1_mod_n(n, 1, 1 << 30)[0]
        t0 = time.perf_counter()
        r = _native.schur_permanent_modp(n, planner.nth_root_of_unity_mod_p(n, p), p)
        log(f"perm(S_{n}) mod {p} = {r} in {time.perf_counter() - t0:.2f} s")
        return 0

    info = _native.device_info()
    log(f"device: {info}")
    A = headline_matrix()
    bound = planner.matrix_permanent_bound(A.tolist())
    primes, product = planner.generic_crt_primes(bound)
    mulmods_per_perm = HEADLINE_N * 2.0 ** (HEADLINE_N - 1) * len(primes)

    # Profiler runs (--kernel-only) launch only the measured kernel: no tuning
    # trials, the setting comes from the flags.
    fixed = args.items_log2 is not None or args.blocks_per_cu is not None or args.kernel_only
    plan = PermanentBatch([A], autotune=False if (args.no_autotune or fixed) else {"reps": 3, "max_seconds": 20.0})
    if args.items_log2 is not None:
        plan.plan.set_option(_native.BATCH_ITEMS_LOG2, args.items_log2)
    if args.blocks_per_cu is not None:
        plan.plan.set_option(_native.BATCH_BLOCKS_PER_CU, args.blocks_per_cu)
    if plan.tuning is not None:
        log(f"headline autotune: {plan.tuning.as_dict()}")
    log(f"headline plan: n={HEADLINE_N} primes={len(primes)} stats={plan.stats()}")
    exact = None
    for _ in range(args.warmup):
        exact = plan.run()[0]
    if args.kernel_only:
        for _ in range(args.steps):
            plan.run()
        log("kernel-only run done")
        return 0

    # Timed region: K exact permanents from resident tables.
    _native.load()
    step_times = []
    t_all0 = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        value = plan.run()[0]
        step_times.append(time.perf_counter() - t0)
        if exact is not None and value != exact:
            raise RuntimeError("headline permanent changed between steps")
        exact = value
    total = time.perf_counter() - t_all0
    ms_per_step = 1e3 * total / args.steps
    mulmods_per_s = mulmods_per_perm * args.steps / total
    log(f"headline: {ms_per_step:.2f} ms/step, {mulmods_per_s:.3e} mulmod/s, perm={exact}")

    # Kernel-only time of the persistent launch: HIP events on the plan's
    # stream around the class launches (no graph), one untimed run, then the
    # mean of ten, as rocprof's kernel-trace average is a mean.
    plan.plan.set_option(_native.BATCH_USE_GRAPH, 0)
    plan.run()
    kernel_ms = []
    for _ in range(10):
        plan.run()
        kernel_ms.append(plan.stats()["last_ms"])
    plan.plan.set_option(_native.BATCH_USE_GRAPH, 1)
    k_ms = statistics.fmean(kernel_ms)
    kernel_rate = mulmods_per_perm / (k_ms * 1e-3)

    rates = _native.valu_microbench()
    log("VALU lane-ops/s: " + ", ".join(f"{k}={v:.3e}" for k, v in rates.items()))
    st = plan.stats()
    group, kset, variant = int(st["group"]), int(st["set"]), int(st.get("variant", 0))
    cus = int(re.search(r"CUs=(\d+)", info).group(1)) if re.search(r"CUs=(\d+)", info) else 256
    exact_variant = variant in (3, 4, 6)
    if variant == 6:
        mix = dict(EXACT_SIGNED_MIX_N32_G4)
    elif exact_variant:
        mix = dict(EXACT_MIX_N32_G4)
    else:
        mix = instruction_mix(HEADLINE_N, group, kset, variant == 1)
    states_per_launch = float(st["states"]) / len(primes)  # one walk serves every prime of the table
    roof = roofline(k_ms, states_per_launch, cus, rates, mix)
    roof["kernel"] = (
        f"batch_glynn_kernel<{group}, {kset}, {str(variant == 1).lower()}, {str(variant == 2).lower()}, "
        f"{str(exact_variant).lower()}> (n={HEADLINE_N}, {len(primes)} primes in "
        f"{int(st['jobs'])} walk(s) of the Gray code)"
    )
    roof["instruction_rates"] = rates
    # What the kernel executes against what `value` counts.
    executed = {
        "reference_equivalent_mulmods_per_state": HEADLINE_N * len(primes),
        "gray_states_per_s": states_per_launch / (k_ms * 1e-3),
        "kernel_reference_equivalent_mulmods_per_s": kernel_rate,
    }
    if exact_variant:
        # Exact accumulation: no modular arithmetic in the walk; each state's
        # product of 8 exact group products is formed in radix-2^31 limbs
        # (28 multiply-adds) and added to 8 int64 columns, for all primes.
        limb_mads = 28  # 4 pair products, two 2 x 2 and one 4 x 4 limb products, either form
        executed.update({
            "variant": "exact accumulation, signed limbs" if variant == 6 else "exact accumulation",
            "montgomery_reductions_per_state": 0,
            "limb_multiply_adds_per_state": limb_mads,
            "limb_multiply_adds_per_s": limb_mads * states_per_launch / (k_ms * 1e-3),
        })
    else:
        # One Montgomery reduction is one v_mul_lo_u32 of the mix.
        executed.update({
            "variant": "paired prime set" if variant == 1 else "prime set",
            "montgomery_reductions_per_state": mix["v_mul_lo_u32"],
            "montgomery_reductions_per_s": mix["v_mul_lo_u32"] * states_per_launch / (k_ms * 1e-3),
        })

    # The drop-in call: perm.matrix_permanent(A) end to end (validation,
    # bound, primes, a one-matrix plan in the device context's buffers, the
    # launch, CRT), as a caller of the reference's API sees it.
    import perm

    drop_in = []
    for _ in range(3):
        t0 = time.perf_counter()
        v = perm.matrix_permanent(A.tolist())
        drop_in.append(time.perf_counter() - t0)
        if v != exact:
            raise RuntimeError("perm.matrix_permanent disagrees with the resident plan")

    # Plain launch: one synchronous C-ABI call per (matrix, prime) -- host
    # tables, copy, launch, reduce, copy back -- as the reference wrappers
    # call their device code (perm_cuda.py:196-198, perm_binary_cuda.py:133-139).
    plain = []
    for _ in range(max(3, args.steps // 4)):
        t0 = time.perf_counter()
        res = [_native.permanent_residues(A, [p])[0] for p in primes]
        plain.append(time.perf_counter() - t0)
    plain_exact = planner.signed_lift(planner.chinese_remainder_theorem(res, primes), bound, product)
    if plain_exact != exact:
        raise RuntimeError("plain launch and resident plan disagree")
    plain_s = statistics.median(plain)
    result = {
        "metric": METRIC,
        "value": mulmods_per_s,
        "unit": "mulmod/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded integer matrices)",
        "config": {
            "workload": (
                "exact permanent of a 32x32 integer matrix, entries U[-10,10] "
                f"(numpy default_rng({HEADLINE_SEED})), 8 CRT primes from 2^30, resident device plan"
            ),
            "n": HEADLINE_N,
            "primes": len(primes),
            "parallelism": "single GPU",
        },
        "value_counts": "reference-equivalent mulmods: n 2^(n-1) per residue (perm_c_omp.c:40-47)",
        "perms_per_s": args.steps / total,
        "exact_permanent": str(exact),
        "executed": executed,
        "drop_in_call_s": statistics.median(drop_in),
        "plain_launch": {"s_per_perm": plain_s, "perms_per_s": 1.0 / plain_s},
        "roofline": roof,
        "launch_shape": _launch_shape(plan),
        "device": info,
    }

    if not args.no_batch:
        mats = batch_matrices()
        t0 = time.perf_counter()
        batch = PermanentBatch(mats)
        t_create = time.perf_counter() - t0
        if not args.no_autotune:
            from permhip import autotune

            batch.tuning = autotune.tune_plan(batch.plan, reps=1, max_seconds=25.0)
        warm = batch.run()
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            got = batch.run()
            runs.append(time.perf_counter() - t0)
            if got != warm:
                raise RuntimeError("batch results changed between runs")
        b_s = statistics.median(runs)
        st = batch.stats()
        # Plain launches of the same matrices: one call per (matrix, prime),
        # and one call per matrix carrying all of its primes.
        plans = []
        for M in mats:
            b = planner.matrix_permanent_bound(M.tolist())
            ps, prod = planner.generic_crt_primes(b)
            plans.append((M, b, ps, prod))
        t0 = time.perf_counter()
        plain_res = []
        for M, b, ps, prod in plans:
            r = [_native.permanent_residues(M, [p])[0] for p in ps]
            plain_res.append(planner.signed_lift(planner.chinese_remainder_theorem(r, ps), b, prod))
        t_plain = time.perf_counter() - t0
        if plain_res != warm:
            raise RuntimeError("batch and plain per-(matrix, prime) results disagree")
        t0 = time.perf_counter()
        for M, b, ps, prod in plans:
            _native.permanent_residues(M, ps)
        t_plain_matrix = time.perf_counter() - t0
        result["batch"] = {
            "workload": "4096 integer matrices, n ~ U[12,28], entries U[-100,100], default_rng(4)",
            "matrices_per_s": len(mats) / b_s,
            "s_per_batch": b_s,
            "plan_create_s": t_create,
            "one_shot_matrices_per_s": len(mats) / (t_create + b_s),
            "one_shot_speedup_vs_plain": t_plain / (t_create + b_s),
            "one_shot_speedup_vs_plain_all_primes_per_call": t_plain_matrix / (t_create + b_s),
            "mulmod_per_s": st.get("mulmods", 0) / b_s,
            "jobs": st.get("jobs"),
            "items": st.get("items"),
            "plain_matrices_per_s": len(mats) / t_plain,
            "speedup_vs_plain": t_plain / b_s,
            "plain_all_primes_per_call_matrices_per_s": len(mats) / t_plain_matrix,
            "speedup_vs_plain_all_primes_per_call": t_plain_matrix / b_s,
            "bit_exact_vs_plain": True,
            "launch_shape": _launch_shape(batch),
        }
        log(f"batch: {result['batch']}")
        batch.close()

    if not args.no_serve:
        try:
            from permhip import serve
        except ImportError:
            serve = None
        if serve is not None:
            result["serve"] = serve.benchmark(seconds=args.serve_seconds, rate=args.serve_rate, progress=log)
            log(f"serve: {result['serve']}")

    if not args.no_published:
        result["binary36"] = binary36_leg()
        log(f"binary36: {result['binary36']}")
        result["polynomial"] = polynomial_leg()
        log(f"polynomial: {result['polynomial']}")
        result["wide_moduli"] = wide_moduli()
        log(f"wide moduli: {result['wide_moduli']}")
        result["published_workloads"] = published_workloads()
        log(f"published workloads: {result['published_workloads']}")

    if not args.no_cpu:
        threads = host_cpu()["cores"]
        result["cpu_baseline"] = cpu_baseline(A, primes, threads)
        if result["cpu_baseline"]["residue"] != exact % primes[0]:
            raise RuntimeError("the CPU baseline's residue differs from the device's")
        log(f"cpu baseline: {result['cpu_baseline']}")
        if "polynomial" in result:
            result["polynomial"]["cpu_baseline"] = polynomial_cpu_baseline(threads)
            log(f"polynomial cpu baseline: {result['polynomial']['cpu_baseline']}")
        if "binary36" in result:
            result["binary36"]["cpu_baseline"] = binary36_cpu_baseline(threads)
            log(f"binary36 cpu baseline: {result['binary36']['cpu_baseline']}")

    if rank == 0:
        print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
