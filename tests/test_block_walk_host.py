"""Host side of the block zero-skipping walk: the library's block tables
(phip_block_tables, the tables a plan uploads) driven through a lane-level
model of the device walk (64 lanes, the LDS ring of survivor entries, the
owner flags whose running count hands each lane its entry, the k-th
set bit of a survivor mask) give the permanent of random 0/1 matrices, as
the oracle's subset DP and permutation sum compute it.  CPU only: the GPU
tests run the kernel itself against the oracle."""

import itertools
import math

import numpy as np
import pytest

import oracle

RING = 128
MASK32 = 0xFFFFFFFF


def _byte(word, b):
    return (int(word) >> (8 * b)) & 0xFF


def _signed_byte(v):
    return v - 256 if v >= 128 else v


class _Tables:
    def __init__(self, A, mod3=False):
        from permhip import _native

        words, self.rows, self.even, self.group = _native.block_tables(A, mod3)
        self.n = n = A.shape[0]
        self.W = W = (n + 3) // 4
        w = [int(x) for x in words]
        self.state0 = w[:W]
        self.deltas = [w[W * (1 + j):W * (2 + j)] for j in range(n - 1)]
        inner = w[n * W:n * W + 12 * W]
        self.inner = [[(inner[k * 12 + (y & 3)] + inner[k * 12 + 4 + (y >> 2)]) & MASK32 for y in range(32)]
                      for k in range(W)]
        self.masks = w[n * W + 12 * W:]
        assert len(self.masks) == (self.rows * 64 if mod3 else (self.rows + 1) * 32)


def _walk_item(T, cbits, first_lane, lanes_total):
    """One wavefront's item: the sum of its 64 lanes' accumulated terms."""
    n, W = T.n, T.W
    lanes = range(64)
    active = [first_lane + l < lanes_total for l in lanes]
    base = [((first_lane + l) if active[l] else 0) << cbits for l in lanes]
    outer, par = [], []
    for l in lanes:
        gray = base[l] ^ (base[l] >> 1)
        o = list(T.state0)
        for j in range(cbits - 1, n - 1):
            if (gray >> j) & 1:
                o = [(o[k] - T.deltas[j][k]) & MASK32 for k in range(W)]
        outer.append(o)
        par.append(bin(gray).count("1") & 1)
    ring = [[0] * RING for _ in range(W + 2)]
    state = {"ehead": 0, "etail": 0, "thead": 0, "ttail": 0}
    acc = [0] * 64

    def consume(cnt):
        etail, ttail, ehead = state["etail"], state["ttail"], state["ehead"]
        flags = [0] * 64
        for rel_e in range(1, min(ehead - etail, 65)):
            rel = (ring[W + 1][(etail + rel_e) % RING] - ttail) & MASK32
            if rel < 64:
                flags[rel] = 1
        running = list(itertools.accumulate(flags))
        last = None
        for l in lanes:
            slot = (etail + running[l]) % RING
            k = (ttail + l - ring[W + 1][slot]) & MASK32
            mask = ring[W][slot]
            m = mask
            for _ in range(min(k, 32)):
                m &= m - 1
            y = (m & -m).bit_length() - 1 if m else 31
            words = [ring[kk][slot] for kk in range(W)]
            sign = words[0] >> 31
            words[0] &= 0x7FFFFFFF
            sign ^= bin(y).count("1") & 1
            if l == 63:
                last = running[l] + (1 if k + 1 == bin(mask).count("1") else 0)
            if l < cnt:
                signed = [((words[kk] + T.inner[kk][y]) & MASK32) ^ 0x80808080 for kk in range(W)]
                term = math.prod(_signed_byte(_byte(signed[i // 4], i % 4)) for i in range(n))
                acc[l] += -term if sign else term
        if cnt == 64:
            state["etail"] = etail + last
        state["ttail"] = (ttail + cnt) & MASK32

    def block():
        surv = []
        for l in lanes:
            m = 0
            for i in range(T.rows):
                b = _byte(outer[l][i // 4], i % 4)
                assert b % 4 == 0, "row bytes must be aligned table offsets"
                m |= T.masks[32 * i + b // 4]
            surv.append(~m & MASK32)
        has = [active[l] and surv[l] != 0 for l in lanes]
        if not any(has):
            return
        counts = [bin(surv[l]).count("1") if has[l] else 0 for l in lanes]
        incl = list(itertools.accumulate(counts))
        rank = 0
        for l in lanes:
            if has[l]:
                slot = (state["ehead"] + rank) % RING
                rank += 1
                for kk in range(W):
                    ring[kk][slot] = (outer[l][kk] >> 1) | ((par[l] << 31) if kk == 0 else 0)
                ring[W][slot] = surv[l]
                ring[W + 1][slot] = (state["thead"] + incl[l] - counts[l]) & MASK32
        state["ehead"] += rank
        assert state["ehead"] - state["etail"] < RING
        state["thead"] = (state["thead"] + incl[63]) & MASK32
        while (state["thead"] - state["ttail"]) & MASK32 >= 64:
            consume(64)

    ub = [b >> 5 for b in base]
    block()
    for u in range(1, 1 << (cbits - 5)):
        jj = (u & -u).bit_length() - 1
        col = T.deltas[5 + jj]
        for l in lanes:
            subtract = not (((ub[l] + u) >> (jj + 1)) & 1)
            outer[l] = [((outer[l][k] - col[k]) if subtract else (outer[l][k] + col[k])) & MASK32 for k in range(W)]
            par[l] ^= 1
        block()
    pending = (state["thead"] - state["ttail"]) & MASK32
    if pending:
        consume(pending)
    return sum(acc)


def _model_permanent(A, cbits):
    T = _Tables(A)
    lanes_total = 1 << (T.n - 1 - cbits)
    total = sum(_walk_item(T, cbits, first, lanes_total) for first in range(0, lanes_total, 64))
    scale = 1 << (T.n - 1)  # Glynn's 2^(n-1)
    assert total % scale == 0
    return total // scale


def _random_binary(rng, n, density):
    return (rng.random((n, n)) < density).astype(np.int64)


@pytest.mark.parametrize("n,cbits", [(8, 6), (9, 7), (10, 6), (11, 8), (12, 6)])
def test_block_walk_model_matches_oracle(n, cbits):
    rng = np.random.default_rng(900 + n)
    for density in (0.35, 0.6):
        A = _random_binary(rng, n, density)
        want = oracle.subset_permanent(A.tolist())
        assert _model_permanent(A, cbits) == want, (n, cbits, density)


def test_block_walk_model_many_entries():
    """Order 14 with chunks of 2^9 states: 16 outer states per lane, so the
    ring wraps several times and batches span entries of many lanes."""
    rng = np.random.default_rng(77)
    A = _random_binary(rng, 14, 0.65)
    assert _model_permanent(A, 9) == oracle.subset_permanent(A.tolist())


def test_block_tables_layout():
    """Rows of even degree lead, the tested rows cover them in whole pairs
    of row words, every row byte is a multiple of 4, and each tested row's
    masks partition the 32 inner states over its possible zero sums."""
    from permhip import _native

    rng = np.random.default_rng(5)
    for n in (8, 13, 21, 30, 37, 46):
        A = _random_binary(rng, n, 0.5)
        T = _Tables(A)
        degrees = sorted(int(d) for d in A.sum(axis=1))
        assert T.even == sum(1 for d in degrees if d % 2 == 0)
        words = (n + 3) // 4
        assert T.rows == 4 * min(words, 2 * ((T.even + 7) // 8))
        assert T.group in (4, 6)
        for i in range(4 * words):
            assert _byte(T.state0[i // 4], i % 4) % 4 == 0
        for i in range(T.rows):
            union = 0
            for c in range(6):
                assert union & T.masks[32 * (i + 1) + c] == 0
                union |= T.masks[32 * (i + 1) + c]
            assert union == (MASK32 if i < T.even else 0), (n, i)
        # Only the words a row's zero sums select are set.
        used = {32 * (i + 1) + c for i in range(T.even) for c in range(6)}
        assert all(T.masks[k] == 0 for k in range(len(T.masks)) if k not in used)
    with pytest.raises(RuntimeError):
        _native.block_tables(np.full((9, 9), 2, dtype=np.int64))
    with pytest.raises(RuntimeError):
        _native.block_tables(np.ones((7, 7), dtype=np.int64))


def test_block_tables_definition_small():
    """A matrix small enough for the permutation sum: the model agrees."""
    rng = np.random.default_rng(3)
    A = _random_binary(rng, 8, 0.7)
    assert _model_permanent(A, 7) == oracle.definition_permanent(A.tolist())


def _mod3_permanent(A):
    """perm(A) mod 3 through the library's modulo-3 block tables and a lane
    model of that walk: masks of 64 words per row, the nonzero states (no
    tested row a multiple of 3) taken in place."""
    B = np.where(A % 3 == 2, -1, A % 3).astype(np.int64)
    T = _Tables(B, mod3=True)
    n, W = T.n, T.W
    cbits = max(6, n - 1 - 6)
    lanes_total = 1 << (n - 1 - cbits)
    total = 0
    for lane in range(lanes_total):
        base = lane << cbits
        gray = base ^ (base >> 1)
        outer = list(T.state0)
        for j in range(cbits - 1, n - 1):
            if (gray >> j) & 1:
                outer = [(outer[k] - T.deltas[j][k]) & MASK32 for k in range(W)]
        par = bin(gray).count("1") & 1
        ub = base >> 5
        for u in range(1 << (cbits - 5)):
            if u:
                jj = (u & -u).bit_length() - 1
                col = T.deltas[5 + jj]
                subtract = not (((ub + u) >> (jj + 1)) & 1)
                outer = [((outer[k] - col[k]) if subtract else (outer[k] + col[k])) & MASK32 for k in range(W)]
                par ^= 1
            m = 0
            for i in range(T.rows):
                b = _byte(outer[i // 4], i % 4)
                assert b % 4 == 0
                m |= T.masks[64 * i + b // 4]
            surv = ~m & MASK32
            while surv:
                y = (surv & -surv).bit_length() - 1
                surv &= surv - 1
                signed = [(((outer[k] >> 1) + T.inner[k][y]) & MASK32) ^ 0x80808080 for k in range(W)]
                vals = [_signed_byte(_byte(signed[i // 4], i % 4)) for i in range(n)]
                assert all(v % 3 for v in vals), "a taken state has a row sum that is a multiple of 3"
                term = math.prod(vals)
                total += -term if (par ^ (bin(y).count("1") & 1)) else term
    # Glynn's 2^(n-1) is (-1)^(n-1) modulo 3.
    return (total * (-1) ** (n - 1)) % 3


@pytest.mark.parametrize("n", [8, 9, 11, 12])
def test_mod3_block_tables_model_matches_oracle(oracle_c, n):
    """The modulo-3 tables (entries -1, 0, 1; every row tested) give perm
    mod 3 of random integer matrices, as the oracle's bipedal Ryser does."""
    rng = np.random.default_rng(950 + n)
    for _ in range(3):
        A = rng.integers(-4, 5, (n, n)).astype(np.int64)
        want = oracle_c.mod3(np.where(A % 3 < 0, A % 3 + 3, A % 3).astype(np.int64))
        assert _mod3_permanent(A) == want, (n, A.tolist())


def test_planner_walks_the_better_orientation():
    """A 0/1 matrix whose rows all have odd degree (no row can sum to zero)
    but whose columns all have even degree: the planner walks its
    transpose on the zero-skipping kernels (host planner, no device)."""
    from permhip import _native

    rng = np.random.default_rng(31)
    n = 14
    A = (rng.random((n, n)) < 0.5).astype(np.int64)
    for i in range(n - 1):
        if A[i].sum() % 2 == 0:
            A[i, n - 1] ^= 1
    for j in range(n - 1):
        if A[:, j].sum() % 2 == 1:
            A[n - 1, j] ^= 1
    if A[n - 1].sum() % 2 == 0:
        A[n - 1, n - 1] ^= 1
    assert all(d % 2 == 1 for d in A.sum(axis=1)) and all(d % 2 == 0 for d in A.sum(axis=0))
    primes = [2147483647, 2147483629]
    assert {cls[2] for cls in _native.batch_plan_classes([A], [primes])} == {2}
    assert {cls[2] for cls in _native.batch_plan_classes([A.T.copy()], [primes])} == {2}


def test_planner_takes_the_exact_block_walk():
    """0/1 block tables of at most eight row groups with five primes or
    more take the block walk with exact accumulation (variant 5, eight limb
    columns); fewer primes, more groups (groups of 4 rows above order 32)
    or the threshold at 0 keep the per-prime walk (host planner, no
    device)."""
    from permhip import _native

    primes = [2147483647, 2147483629, 2147483587, 2147483579, 2147483563]
    rng = np.random.default_rng(36)
    A = (rng.random((36, 36)) < 0.5).astype(np.int64)
    classes = lambda M, k: {cls for cls in _native.batch_plan_classes([M], [primes[:k]])}
    assert classes(A, 2) == {(6, 2, 2)}
    for k in (3, 4):
        assert classes(A, k) == {(6, 4, 2)}
    assert classes(A, 5) == {(6, 8, 5)}
    # All-ones J_40: degree 40 rows group by 4 (40^6 >= 2^31), ten groups.
    assert classes(np.ones((40, 40), dtype=np.int64), 5) == {(4, 8, 2)}
    old = _native.get_tuning(_native.TUNE_SKIP_EXACT_MIN_PRIMES)
    assert old == 5
    try:
        _native.set_tuning(_native.TUNE_SKIP_EXACT_MIN_PRIMES, 0)
        assert classes(A, 5) == {(6, 8, 2)}
        _native.set_tuning(_native.TUNE_SKIP_EXACT_MIN_PRIMES, 2)
        assert classes(A, 2) == {(6, 8, 5)}
    finally:
        _native.set_tuning(_native.TUNE_SKIP_EXACT_MIN_PRIMES, old)
