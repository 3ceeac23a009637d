"""The radix-2^31 limb products of the exact-accumulation Glynn sink
(csrc/exact_limbs.h) on the CPU: the header is compiled with g++ and every
product is checked against Python integers, on edge magnitudes (0, 1,
2^31 - 1 everywhere) and random ones, including the limb ranges the sink
relies on (every limb below 2^31)."""

import random
import shutil
import subprocess

import pytest

from test_wide_arith import CSRC

HARNESS = r"""
#include <cstdio>
#include <cstdint>
#include "exact_limbs.h"
using namespace phip;
int main() {
    unsigned m[8];
    while (std::scanf("%u %u %u %u %u %u %u %u", &m[0], &m[1], &m[2], &m[3], &m[4], &m[5], &m[6], &m[7]) == 8) {
        uint32_t in[8], out[8], x[4], p0[2], p1[2];
        for (int i = 0; i < 8; ++i) in[i] = m[i];
        limb_product8(in, out);
        limb_mul_1x1(in[0], in[1], p0);
        limb_mul_1x1(in[2], in[3], p1);
        limb_mul_2x2(p0, p1, x);
        for (int i = 0; i < 8; ++i) std::printf("%u ", out[i]);
        for (int i = 0; i < 4; ++i) std::printf("%u ", x[i]);
        std::printf("\n");
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("limbs")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", str(CSRC), str(src), "-o", str(exe)], check=True)
    return exe


def test_limb_products_match_python(harness):
    rng = random.Random(31)
    top = (1 << 31) - 1
    cases = [[0] * 8, [1] * 8, [top] * 8, [top, 1] * 4, [1 << 30] * 8]
    for _ in range(3000):
        kind = rng.random()
        if kind < 0.3:
            cases.append([rng.choice((0, 1, 2, top, top - 1, 1 << 30, (1 << 30) - 1)) for _ in range(8)])
        else:
            cases.append([rng.randrange(1 << rng.randrange(1, 32)) for _ in range(8)])
    text = "\n".join(" ".join(map(str, c)) for c in cases) + "\n"
    out = subprocess.run([str(harness)], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
    for c, line in zip(cases, out):
        vals = [int(v) for v in line.split()]
        z, x = vals[:8], vals[8:]
        assert all(0 <= v < (1 << 31) for v in vals)
        prod8 = 1
        for v in c:
            prod8 *= v
        prod4 = c[0] * c[1] * c[2] * c[3]
        assert sum(v << (31 * i) for i, v in enumerate(z)) == prod8, c
        assert sum(v << (31 * i) for i, v in enumerate(x)) == prod4, c
