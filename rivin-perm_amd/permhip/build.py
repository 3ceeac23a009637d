"""Build recipe for libpermhip.so (HIP kernels for gfx950 plus the C ABI).

Compiles every translation unit under ``csrc/`` with hipcc into
``_build/`` and links ``permhip/libpermhip.so`` in-tree, so the library
travels with the source tree.  The persistent Glynn kernel is compiled once
per (row-sum group, prime-set size) class so the classes build in parallel.

Usage: ``python -m permhip.build`` from ``rivin-perm_amd/`` or
``permhip.build.build()``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
ROOT = PACKAGE_DIR.parent
CSRC = ROOT / "csrc"
INCLUDE = ROOT.parent / "include"
BUILD_DIR = ROOT / "_build"
LIBRARY = PACKAGE_DIR / "libpermhip.so"

ARCH = "gfx950"
# (row-sum group, prime-set size, variant) classes of the persistent Glynn
# kernel.  Variant 0 is the plain kernel; sets of 4 and 8 primes below
# 1.25 * 2^30 also have the paired-state variant (1), 0/1 tables in groups
# of 4 or 6 the zero-skipping variant (2) for sets of 2, 4 and 8, and every
# group the exact-accumulation variant (3, eight limb columns per item), and
# groups of 2 rows its compact form (4: the product tree sized for 22-bit
# group values, ten limb columns), and 0/1 block tables in groups of 6
# the block walk with exact accumulation (5, eight limb columns; groups of 6
# only: a block table falls back to groups of 4 only above order 35, where
# it has more than eight groups), and groups of 3 or 4 rows the signed-limb
# exact form (6, eight limb columns).
BATCH_VARIANT_SUFFIX = {0: "", 1: "_paired", 2: "_skip", 3: "_exact", 4: "_compact", 5: "_skipexact", 6: "_signed"}
BATCH_CLASSES = (
    ((0, 1, 0),)
    + tuple((g, k, 0) for g in (2, 3, 4, 6) for k in (1, 2, 4, 8))
    + tuple((g, k, 1) for g in (2, 3, 4, 6) for k in (4, 8))
    + tuple((g, k, 2) for g in (4, 6) for k in (2, 4, 8))
    + tuple((g, 8, 3) for g in (2, 3, 4, 6))
    + tuple((g, k, 3) for g in (2, 3, 4) for k in (10, 12))
    + ((2, 10, 4),)
    + ((6, 8, 5),)
    + tuple((g, 8, 6) for g in (3, 4))
)
PLAIN_SOURCES = (
    "device_context.hip",
    "residues.hip",
    "mod3.hip",
    "capi.hip",
    "batch_plan.hip",
    "microbench.hip",
    "necklace.hip",
    "wide.hip",
    "affine.hip",
    "host_prep.cpp",
    "frontier.hip",
    "noalloc.hip",
)


def _hipcc() -> str:
    explicit = os.environ.get("HIPCC")
    if explicit:
        return explicit
    for candidate in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if candidate and Path(candidate).exists():
            return candidate
    raise RuntimeError("hipcc not found; set HIPCC")


def _flags() -> list:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-fPIC",
        "-std=c++17",
        # Packed FP32 instructions are kept out of every code object.
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", str(CSRC),
        "-I", str(INCLUDE),
    ]


def _compile(args):
    source, obj, defines = args
    cmd = [_hipcc(), *_flags(), *defines, "-c", str(source), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(
            f"hipcc failed for {source.name}:\n{' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}"
        )
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    """Compile all sources and link ``libpermhip.so``; returns its path."""
    BUILD_DIR.mkdir(parents=True, exist_ok=True)
    units = []
    # Largest prime sets first: they take longest to compile.
    for group, k, variant in sorted(BATCH_CLASSES, key=lambda gkv: -gkv[1]):
        units.append(
            (
                CSRC / "batch.hip",
                BUILD_DIR / f"batch_g{group}_k{k}{BATCH_VARIANT_SUFFIX[variant]}.o",
                [f"-DPHIP_BATCH_GROUP={group}", f"-DPHIP_BATCH_SET={k}", f"-DPHIP_BATCH_PAIRED={int(variant == 1)}",
                 f"-DPHIP_BATCH_SKIP_ZEROS={int(variant in (2, 5))}", f"-DPHIP_BATCH_EXACT={int(variant in (3, 4, 5, 6))}",
                 f"-DPHIP_BATCH_COMPACT={int(variant == 4)}", f"-DPHIP_BATCH_SIGNED={int(variant == 6)}"],
            )
        )
    for name in PLAIN_SOURCES:
        units.append((CSRC / name, BUILD_DIR / (Path(name).stem + ".o"), []))
    workers = jobs or min(len(units), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        objects = list(pool.map(_compile, units))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
