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


def _truncated_harness_source():
    cases = []
    for limbs in (8, 10, 12):
        for count in range(1, 17):
            cases.append(
                f"        if (L == {limbs} && c == {count}) {{\n"
                f"            using P = LimbProduct<0, {count}, {limbs}>;\n"
                f"            uint32_t out[P::kLimbs];\n"
                f"            P::run(in, out);\n"
                f"            for (int i = 0; i < P::kLimbs; ++i) std::printf(\"%u \", out[i]);\n"
                f"        }}\n"
            )
    return (
        "#include <cstdio>\n#include <cstdint>\n#include \"exact_limbs.h\"\nusing namespace phip;\n"
        "int main() {\n    int L, c;\n"
        "    while (std::scanf(\"%d %d\", &L, &c) == 2) {\n"
        "        uint32_t in[16] = {};\n"
        "        for (int i = 0; i < c; ++i) { unsigned v; std::scanf(\"%u\", &v); in[i] = v; }\n"
        + "".join(cases)
        + "        std::printf(\"\\n\");\n    }\n    return 0;\n}\n"
    )


@pytest.fixture(scope="module")
def truncated_harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("limbs_trunc")
    src = d / "t.cpp"
    src.write_text(_truncated_harness_source())
    exe = d / "t"
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", str(CSRC), str(src), "-o", str(exe)], check=True)
    return exe


def test_truncated_products_match_python(truncated_harness):
    """LimbProduct<0, c, L>: the product of c values below 2^31 modulo
    2^(31 L), in min(c, L) limbs, for every count 1..16 and L in {8, 10,
    12} (the exact sink's column counts).  Values at 2^31 - 1 make every
    column as full as it gets, which is where a wrong split point would
    overflow a 64-bit column word."""
    rng = random.Random(1231)
    top = (1 << 31) - 1
    cases = []
    for limbs in (8, 10, 12):
        for count in range(1, 17):
            cases.append((limbs, [top] * count))
            cases.append((limbs, [1] * count))
            cases.append((limbs, [0] + [top] * (count - 1)))
            for _ in range(60):
                if rng.random() < 0.4:
                    vals = [rng.choice((1, top, top - 1, 1 << 30, (1 << 30) - 1, 3)) for _ in range(count)]
                else:
                    vals = [rng.randrange(1 << rng.randrange(1, 32)) for _ in range(count)]
                cases.append((limbs, vals))
    text = "".join(f"{L} {len(v)} {' '.join(map(str, v))}\n" for L, v in cases)
    out = subprocess.run([str(truncated_harness)], input=text, capture_output=True, text=True,
                         check=True).stdout.split("\n")
    for (limbs, vals), line in zip(cases, out):
        got = [int(x) for x in line.split()]
        assert len(got) == min(len(vals), limbs)
        assert all(0 <= x < (1 << 31) for x in got)
        prod = 1
        for v in vals:
            prod *= v
        want = prod % (1 << (31 * len(got)))
        assert sum(x << (31 * i) for i, x in enumerate(got)) == want, (limbs, vals)
