"""The wide-modulus residue arithmetic (csrc/wide_arith.h) on the CPU.

The header's products are written in 32x32 -> 64 steps so that its host
build runs the device arithmetic; this test compiles it with g++ and checks
every policy against Python integers on edge values and random operands:
Mont32 (odd q < 2^32, R = 2^32), M61 (q = 2^61 - 1), PM61 (q = 2^61 - c,
odd c < 2^28) and Mont64 (odd q < 2^63, R = 2^64)."""

import random
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "rivin-perm_amd" / "csrc"

HARNESS = r"""
#include <cstdio>
#include <cstdint>
#include "wide_arith.h"
using namespace phip;
int main() {
    char kind[8];
    unsigned long long q, a, b;
    while (std::scanf("%7s %llu %llu %llu", kind, &q, &a, &b) == 4) {
        unsigned long long m = 0, s = 0, ng = 0;
        if (kind[0] == 'c') {
            std::printf("%d\n", wide_class(q));
            continue;
        }
        if (kind[0] == '3') {
            Mont32 ar = make_mont32(q);
            m = ar.mul((uint32_t)a, (uint32_t)b); s = ar.add((uint32_t)a, (uint32_t)b); ng = ar.neg((uint32_t)a);
        } else if (kind[0] == 'm') {
            M61 ar = make_m61(q);
            m = ar.mul(a, b); s = ar.add(a, b); ng = ar.neg(a);
        } else if (kind[0] == 'p') {
            PM61 ar = make_pm61(q);
            m = ar.mul(a, b); s = ar.add(a, b); ng = ar.neg(a);
        } else {
            Mont64 ar = make_mont64(q);
            m = ar.mul(a, b); s = ar.add(a, b); ng = ar.neg(a);
        }
        std::printf("%llu %llu %llu\n", m, s, ng);
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("wide")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", str(CSRC), str(src), "-o", str(exe)], check=True)

    def run(lines):
        out = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True,
                             check=True).stdout.split("\n")
        return [list(map(int, row.split())) for row in out if row.strip()]

    return run


def _operands(q, rng, count=40):
    edge = [0, 1, 2, q - 1, q - 2, q // 2, (q + 1) // 2]
    return edge + [rng.randrange(q) for _ in range(count)]


def _check(harness, kind, q, r_bits, rng):
    ops = _operands(q, rng)
    pairs = [(a, b) for a in ops[:12] for b in ops[:12]] + [(rng.randrange(q), rng.randrange(q)) for _ in range(300)]
    rows = harness([f"{kind} {q} {a} {b}" for a, b in pairs])
    rinv = pow(1 << r_bits, -1, q) if r_bits else 1
    for (a, b), (m, s, ng) in zip(pairs, rows):
        assert m == a * b * rinv % q, (kind, q, a, b)
        assert s == (a + b) % q, (kind, q, a, b)
        assert ng == (-a) % q, (kind, q, a)


def test_mont32(harness):
    rng = random.Random(1)
    for q in (2147483659, 4294967291, 4294966153, 4294967161, 3221225473, 2 ** 32 - 1, 2 ** 31 + 1):
        _check(harness, "3", q, 32, rng)


def test_m61(harness):
    _check(harness, "m", (1 << 61) - 1, 0, random.Random(4))


def test_pm61(harness):
    rng = random.Random(2)
    for c in (1, 3, 31, 1 << 10 | 1, (1 << 28) - 1, 12345679):
        _check(harness, "p", (1 << 61) - c, 0, rng)


def test_mont64(harness):
    rng = random.Random(3)
    for q in ((1 << 63) - 25, (1 << 62) + 1, (1 << 61) + 1, 10 ** 18 + 9, (1 << 40) + 15):
        _check(harness, "6", q, 64, rng)


def test_classes(harness):
    cases = {
        3: 0, 2 ** 31 + 1: 0, 2 ** 32 - 1: 0,
        2 ** 61 - 1: 3, 2 ** 61 - 3: 1, 2 ** 61 - (2 ** 28 - 1): 1,
        2 ** 61 - (2 ** 28 + 1): 2, 2 ** 61 + 1: 2, 2 ** 61 - 2: 2, 2 ** 63 - 25: 2, 2 ** 32 + 1: 2,
    }
    rows = harness([f"c {q} 0 0" for q in cases])
    assert [r[0] for r in rows] == list(cases.values())
