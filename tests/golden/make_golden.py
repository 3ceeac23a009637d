#!/usr/bin/env python3
"""Writes tests/golden/reference_residues.json from the reference's own
OpenMP extensions (perm_c_omp.c, perm_binary_omp.c, perm_mod3_omp.c), compiled
from the reference sources into oracle/_ref by oracle/Makefile.

Only the reference's compiled C kernels are called; bounds, prime plans and
CRT come from this build's planner.  The output is data: seeded inputs are
identified by their generator (permhip.workloads) plus a SHA-256 of the
matrix bytes, with the moduli and the residues the reference returned.

Usage: python tests/golden/make_golden.py [--threads 8] [--only NAME ...]
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--only", nargs="*", default=None)
    args = ap.parse_args()
    ext = oracle.reference_extensions()
    if len(ext) != 3:
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
