#!/usr/bin/env python3
"""Writes tests/golden/reference_residues.json from the reference's own
OpenMP extensions (perm_c_omp.c, perm_binary_omp.c, perm_mod3_omp.c), compiled
from the reference sources into oracle/_ref by oracle/Makefile.

Only the reference's compiled C kernels are called; bounds, prime plans and
CRT come from this build's planner.  The output is data: seeded inputs are
identified by their generator (permhip.workloads) plus a SHA-256 of the
matrix bytes, with the moduli and the residues the reference returned.

Usage: python tests/golden/make_golden.py [--threads 8] [--only NAME ...]
NAME is one of config1, schur31, batch_sample, batch_classes, serve_slice,
binary36, headline.
"""

from __future__ import annotations

import argparse
import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "rivin-perm_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

import oracle  # noqa: E402
from permhip import planner, workloads  # noqa: E402

OUT = Path(__file__).resolve().parent / "reference_residues.json"


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int64).tobytes()).hexdigest()


def descending_primes(bound: int):
    primes, product, c = [], 1, (1 << 31) - 1
    while product <= bound:
        if planner.is_prime(c):
            primes.append(c)
            product *= c
        c -= 2
    return primes


SERVE_SLICE = 200


def binary_bound(rows) -> int:
    """min(n!, product of row sums, product of column sums): an upper bound
    on the permanent of a 0/1 matrix."""
    import math

    n = len(rows)
    return min(math.factorial(n), math.prod(sum(r) for r in rows), math.prod(sum(c) for c in zip(*rows)))


def crt(residues, primes) -> int:
    x, m = 0, 1
    for r, p in zip(residues, primes):
        x += m * ((r - x) * pow(m, -1, p) % p)
        m *= p
    return x


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--only", nargs="*", default=None)
    args = ap.parse_args()
    ext = oracle.reference_extensions()
    if not {"perm_c_ext", "perm_binary_ext", "perm_mod3_omp"} <= set(ext):
        print("build the reference extensions first: make -C oracle", file=sys.stderr)
        return 1
    data = json.loads(OUT.read_text()) if OUT.exists() else {}
    T = args.threads

    def want(name):
        return args.only is None or name in args.only

    def save():
        OUT.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")

    def integer_case(name, A, generator):
        rows = A.tolist()
        bound = planner.matrix_permanent_bound(rows)
        primes, product = planner.generic_crt_primes(bound)
        res = []
        t0 = time.time()
        for p in primes:
            red = [[int(v) % p for v in row] for row in rows]
            res.append(int(ext["perm_c_ext"].permanent_mod_p(red, p, T)))
            print(f"  {name} p={p} r={res[-1]} ({time.time() - t0:.0f}s)", flush=True)
        exact = planner.signed_lift(planner.chinese_remainder_theorem(res, primes), bound, product)
        data[name] = {
            "generator": generator,
            "n": len(rows),
            "sha256_int64": digest(A),
            "moduli": primes,
            "residues": res,
            "exact": str(exact),
            "source": "perm_c_ext.permanent_mod_p (reference, compiled from perm_c_omp.c)",
        }
        save()

    if want("config1"):
        integer_case("config1", workloads.config1(), "permhip.workloads.config1()")
    if want("schur31"):
        n = 31
        bound = planner.schur_permanent_bound(n)
        primes, product, cong = planner.plan_schur_exact_crt(n, 2 * bound + 1, 1 << 30)
        res = []
        for p in primes:
            w = planner.nth_root_of_unity_mod_p(n, p)
            res.append(int(ext["perm_c_ext"].schur_permanent_mod_p(n, w, p, T)))
        data["schur31"] = {
            "generator": "S_31, omega = g^((p-1)/31) for the smallest generator g",
            "n": n,
            "moduli": primes,
            "residues": res,
            "theorem": list(cong),
            "source": "perm_c_ext.schur_permanent_mod_p (reference)",
        }
        save()
    if want("batch_sample"):
        mats = workloads.batch4096()
        picks = [i for i, m in enumerate(mats) if m.shape[0] >= 23][:6]
        for i in picks:
            integer_case(f"batch4096[{i}]", mats[i], f"permhip.workloads.batch4096()[{i}]")
    if want("batch_classes"):
        # Two matrices of every order 12..28 of the 4096-matrix batch: their
        # prime counts (4..10) and row bounds cover every (group, prime set,
        # variant) class the full batch plan uses.
        mats = workloads.batch4096()
        for n in range(12, 29):
            for i in [i for i, m in enumerate(mats) if m.shape[0] == n][:2]:
                name = f"batch4096[{i}]"
                if name not in data:
                    integer_case(name, mats[i], f"permhip.workloads.batch4096()[{i}]")
        # Then two matrices of every kernel class of the full plan that the
        # picks above leave uncovered (the plan's classes come from the
        # library's host planner, no device needed).
        from permhip import _native

        primes = [planner.generic_crt_primes(planner.matrix_permanent_bound(m.tolist()))[0] for m in mats]
        classes = _native.batch_plan_classes(mats, primes)
        of_matrix, pos = [], 0
        for ps in primes:
            of_matrix.append(set(classes[pos:pos + len(ps)]))
            pos += len(ps)
        covered = set()
        for key in data:
            if key.startswith("batch4096["):
                covered |= of_matrix[int(key[len("batch4096["):-1])]
        for cls in sorted(set(classes) - covered):
            for i in [i for i in range(len(mats)) if cls in of_matrix[i]][:2]:
                integer_case(f"batch4096[{i}]", mats[i], f"permhip.workloads.batch4096()[{i}]")
    if want("serve_slice"):
        # The first SERVE_SLICE requests of the served stream: integer
        # requests up to order 28 and 0/1 requests up to order 30 (the rest
        # take the reference's CPU path too long to pin here).
        stream = workloads.serve_stream()[:SERVE_SLICE]
        out = {}
        t0 = time.time()
        for i, (arrival, kind, A) in enumerate(stream):
            n = A.shape[0]
            if kind == "integer" and n <= 28:
                rows = A.tolist()
                bound = planner.matrix_permanent_bound(rows)
                primes, product = planner.generic_crt_primes(bound)
                res = [int(ext["perm_c_ext"].permanent_mod_p([[int(v) % p for v in r] for r in rows], p, T))
                       for p in primes]
                exact = planner.signed_lift(planner.chinese_remainder_theorem(res, primes), bound, product)
            elif kind == "binary" and n <= 30:
                rows = A.astype(np.int64).tolist()
                primes = descending_primes(binary_bound(rows))
                res = [int(ext["perm_binary_ext"].binary_permanent_mod_p(rows, p, T)) for p in primes]
                exact = crt(res, primes)
            else:
                continue
            out[str(i)] = {"kind": kind, "n": n, "sha256_int64": digest(A), "moduli": primes,
                           "residues": res, "exact": str(exact)}
            print(f"  serve[{i}] {kind} n={n} ({time.time() - t0:.0f}s)", flush=True)
        data["serve_slice"] = {
            "generator": f"permhip.workloads.serve_stream()[:{SERVE_SLICE}]",
            "requests": out,
            "source": "perm_c_ext.permanent_mod_p / perm_binary_ext.binary_permanent_mod_p (reference)",
        }
        save()
    if want("binary36"):
        B = workloads.binary36()
        rows = B.tolist()
        primes = descending_primes(2 ** 106)
        res = []
        t0 = time.time()
        for p in primes:
            res.append(int(ext["perm_binary_ext"].binary_permanent_mod_p(rows, p, T)))
            print(f"  binary36 p={p} r={res[-1]} ({time.time() - t0:.0f}s)", flush=True)
        mod3 = int(ext["perm_mod3_omp"].permanent_mod3(B.astype(np.int64) % 3, T))
        data["binary36"] = {
            "generator": "permhip.workloads.binary36()",
            "n": 36,
            "sha256_int64": digest(B),
            "moduli": primes,
            "residues": res,
            "mod3": mod3,
            "source": "perm_binary_ext.binary_permanent_mod_p, perm_mod3_omp.permanent_mod3 (reference)",
        }
        save()
    if want("headline"):
        integer_case("headline", workloads.headline(), "permhip.workloads.headline()")
    return 0


if __name__ == "__main__":
    sys.exit(main())
