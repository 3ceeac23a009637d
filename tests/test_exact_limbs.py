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


def _compact_limbs(count, value_bits=22):
    return 1 if count <= 1 else (count * value_bits + 30) // 31


def _compact_fits(vals, lo, hi, limbs_cap, value_bits=22):
    """The compact tree's condition, written from its definition: every node
    (split at the largest power of two below its count) whose limb count is
    below the cap holds a product of at most 31 * limbs bits."""
    count = hi - lo
    if count <= 1:
        return True
    limbs = _compact_limbs(count, value_bits)
    if limbs < limbs_cap:
        prod = 1
        for v in vals[lo:hi]:
            prod *= v
        if prod.bit_length() > 31 * limbs:
            return False
    split = 1
    while 2 * split < count:
        split *= 2
    return (_compact_fits(vals, lo, lo + split, limbs_cap, value_bits)
            and _compact_fits(vals, lo + split, hi, limbs_cap, value_bits))


def _compact_harness_source():
    cases = []
    for count in range(1, 17):
        cases.append(
            f"        if (c == {count}) {{\n"
            f"            using P = LimbProduct<0, {count}, 10, 22>;\n"
            f"            uint32_t out[P::kLimbs];\n"
            f"            P::run(in, out);\n"
            f"            for (int i = 0; i < P::kLimbs; ++i) std::printf(\"%u \", out[i]);\n"
            f"        }}\n"
        )
    return (
        "#include <cstdio>\n#include <cstdint>\n#include \"exact_limbs.h\"\nusing namespace phip;\n"
        "int main() {\n    int c;\n"
        "    while (std::scanf(\"%d\", &c) == 1) {\n"
        "        uint32_t in[16] = {};\n"
        "        for (int i = 0; i < c; ++i) { unsigned v; std::scanf(\"%u\", &v); in[i] = v; }\n"
        + "".join(cases)
        + "        std::printf(\"\\n\");\n    }\n    return 0;\n}\n"
    )


@pytest.fixture(scope="module")
def compact_harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("limbs_compact")
    src = d / "c.cpp"
    src.write_text(_compact_harness_source())
    exe = d / "c"
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", str(CSRC), str(src), "-o", str(exe)], check=True)
    return exe


def test_compact_products_match_python(compact_harness):
    """LimbProduct<0, c, 10, 22> (the compact exact classes' tree): nodes of
    c >= 2 values keep ceil(22 c / 31) limbs.  Whenever every node's product
    fits its limbs (the planner's condition, restated above) the result is
    the product modulo 2^(31 min(10, ceil(22 c / 31))), in limbs below 2^31;
    values of up to 22 bits always fit, and wider values that still fit put
    the nodes at their limits.  A value set that breaks a node's limit gives
    a different number, so the condition is not vacuous."""
    rng = random.Random(2222)
    fitting, broken = [], []
    for count in range(1, 17):
        fitting.append([(1 << 22) - 1] * count)
        fitting.append([1] * count)
        for _ in range(400):
            if rng.random() < 0.5:
                vals = [rng.randrange(1 << 22) for _ in range(count)]
            else:
                vals = [rng.randrange(1 << rng.randrange(1, 32)) | 1 for _ in range(count)]
            (fitting if _compact_fits(vals, 0, count, 10) else broken).append(vals)
    assert len(broken) > 100
    cases = fitting + broken
    text = "".join(f"{len(v)} {' '.join(map(str, v))}\n" for v in cases)
    out = subprocess.run([str(compact_harness)], input=text, capture_output=True, text=True,
                         check=True).stdout.split("\n")
    mismatches = 0
    for i, (vals, line) in enumerate(zip(cases, out)):
        got = [int(x) for x in line.split()]
        limbs = min(10, _compact_limbs(len(vals)))
        assert len(got) == limbs
        assert all(0 <= x < (1 << 31) for x in got)
        prod = 1
        for v in vals:
            prod *= v
        value = sum(x << (31 * k) for k, x in enumerate(got))
        if i < len(fitting):
            assert value == prod % (1 << (31 * limbs)), vals
        elif value != prod % (1 << (31 * limbs)):
            mismatches += 1
    assert mismatches > 0


SIGNED_HARNESS = r"""
#include <cstdio>
#include <cstdint>
#include "exact_limbs.h"
using namespace phip;
int main() {
    long long v[8];
    while (std::scanf("%lld %lld %lld %lld %lld %lld %lld %lld", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6],
                      &v[7]) == 8) {
        int32_t g[8], out[8], p0[2], p1[2], x[4];
        for (int i = 0; i < 8; ++i) g[i] = (int32_t)v[i];
        slimb_product8(g, out);
        slimb_mul_1x1(g[0], g[1], p0);
        slimb_mul_1x1(g[2], g[3], p1);
        slimb_mul_2x2(p0, p1, x);
        for (int i = 0; i < 8; ++i) std::printf("%d ", out[i]);
        for (int i = 0; i < 4; ++i) std::printf("%d ", x[i]);
        std::printf("\n");
    }
    return 0;
}
"""


def _signed_value(limbs):
    return sum(v << (31 * i) for i, v in enumerate(limbs))


def test_signed_limb_products_match_python(tmp_path):
    """slimb_product8 (signed limbs, the signed-limb exact classes) against
    Python integers: values below 2^31 in magnitude whose four-value
    products stay below 2^122 (the planner's condition), at the edges
    (every value +-(2^30.5), four-value products just below 2^122) and at
    random; limbs 0..6 in [0, 2^31), the top limb in [-2^31, 2^31)."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "s.cpp"
    src.write_text(SIGNED_HARNESS)
    exe = tmp_path / "s"
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", str(CSRC), str(src), "-o", str(exe)], check=True)
    rng = random.Random(122)
    edge = int(2 ** 30.5) - 1  # four of them: just below 2^122
    cases = [[0] * 8, [1] * 8, [-1] * 8, [edge] * 8, [-edge] * 8, [edge, -edge] * 4, [(1 << 31) - 1] + [1] * 7,
             [-(1 << 31) + 1] * 2 + [1] * 6]
    for _ in range(3000):
        bits = rng.randrange(1, 31)
        row = [rng.randrange(-(1 << bits) + 1, 1 << bits) for _ in range(8)]
        if rng.random() < 0.3:
            row = [rng.choice((-1, 1)) * rng.choice((edge, edge - 1, 1 << 30, 1, 0)) for _ in range(8)]
        cases.append(row)
    for c in cases:
        for q in (c[:4], c[4:]):
            assert abs(q[0] * q[1] * q[2] * q[3]) < 1 << 122
    text = "\n".join(" ".join(map(str, c)) for c in cases) + "\n"
    out = subprocess.run([str(exe)], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
    for c, line in zip(cases, out):
        vals = [int(v) for v in line.split()]
        z, x = vals[:8], vals[8:]
        assert all(0 <= v < (1 << 31) for v in z[:7] + x[:3]), c
        assert -(1 << 31) <= z[7] < (1 << 31) and -(1 << 31) <= x[3] < (1 << 31), c
        prod8 = 1
        for v in c:
            prod8 *= v
        assert _signed_value(z) == prod8, c
        assert _signed_value(x) == c[0] * c[1] * c[2] * c[3], c


def test_planner_takes_signed_limbs_where_quads_fit():
    """Host planner (no device): the headline table (groups of 4, eight
    primes) takes the signed-limb class; tuning 0, or four-group bound
    products of 2^123 (row bounds 208), keep the magnitude class."""
    import numpy as np

    from permhip import _native, planner, workloads

    A = workloads.headline()
    primes, _ = planner.generic_crt_primes(planner.matrix_permanent_bound(A.tolist()))
    assert set(_native.batch_plan_classes([A], [primes])) == {(4, 8, 6)}
    old = _native.get_tuning(_native.TUNE_EXACT_SIGNED)
    assert old == 1
    try:
        _native.set_tuning(_native.TUNE_EXACT_SIGNED, 0)
        assert set(_native.batch_plan_classes([A], [primes])) == {(4, 8, 3)}
    finally:
        _native.set_tuning(_native.TUNE_EXACT_SIGNED, old)
    rng = np.random.default_rng(208)
    mags = np.array([6] * 16 + [7] * 16)
    B = np.array([rng.permutation(mags) * rng.choice((-1, 1), 32) for _ in range(32)], dtype=np.int64)
    assert set(_native.batch_plan_classes([B], [primes])) == {(4, 8, 3)}
    C = np.full((32, 32), 6, dtype=np.int64)
    assert set(_native.batch_plan_classes([C], [primes])) == {(4, 8, 6)}
