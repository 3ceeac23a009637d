"""Persistent batch: bit-exact against the single-matrix launch and the
oracle, across graph replay, launch shapes and uneven batch compositions."""

import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mixed_batch(seed, count):
    rng = np.random.default_rng(seed)
    mats, moduli = [], []
    primes31 = [1073741827, 1073741831, 2147483647, 2147483629, 1000003, 65537]
    for i in range(count):
        n = int(rng.integers(5, 23))
        kind = i % 4
        if kind == 0:
            M = rng.integers(-10, 11, (n, n))       # groups of 3
        elif kind == 1:
            M = rng.integers(0, 2, (n, n))          # groups of 4
        elif kind == 2:
            M = rng.integers(-100, 101, (n, n))     # groups of 2
        else:
            M = rng.integers(-(10 ** 9), 10 ** 9, (n, n))  # wide rows
        mats.append(M.astype(np.int64))
        k = int(rng.integers(1, 4))
        moduli.append([int(q) for q in rng.choice(primes31, size=k, replace=False)])
    return mats, moduli


def test_batch_matches_plain_and_oracle(native, oracle_c):
    mats, moduli = _mixed_batch(11, 48)
    batch = native.Batch(mats, moduli)
    flat = batch.run()
    pos = 0
    for i, (M, ms) in enumerate(zip(mats, moduli)):
        got = flat[pos:pos + len(ms)]
        pos += len(ms)
        assert got == native.permanent_residues(M, ms), i
        if M.shape[0] <= 14:
            assert got == [oracle_c.permanent_mod_p(M.tolist(), q, 4) for q in ms], i
    # Replays, graph on/off and launch shapes give identical residues.
    assert batch.run() == flat
    batch.set_option(native.BATCH_USE_GRAPH, 0)
    assert batch.run() == flat
    for blocks in (1, 3, 8):
        batch.set_option(native.BATCH_BLOCKS_PER_CU, blocks)
        assert batch.run() == flat, blocks
    batch.close()


@pytest.mark.parametrize("items_log2", [6, 12, 18])
def test_item_size_does_not_change_residues(native, items_log2):
    mats, moduli = _mixed_batch(5, 16)
    old = native.get_tuning(native.TUNE_BATCH_ITEMS_LOG2)
    try:
        native.set_tuning(native.TUNE_BATCH_ITEMS_LOG2, items_log2)
        batch = native.Batch(mats, moduli)
        got = batch.run()
        batch.close()
    finally:
        native.set_tuning(native.TUNE_BATCH_ITEMS_LOG2, old)
    expect = []
    for M, ms in zip(mats, moduli):
        expect += native.permanent_residues(M, ms)
    assert got == expect


def test_single_matrix_launch_shapes_agree(native):
    """Partition determinism of the plain launch over chunk sizes."""
    rng = random.Random(9)
    A = np.array([[rng.randrange(-10, 11) for _ in range(24)] for _ in range(24)], dtype=np.int64)
    primes = [1073741827, 1073741831, 2147483647]
    base = native.permanent_residues(A, primes)
    old = native.get_tuning(native.TUNE_LANES_LOG2), native.get_tuning(native.TUNE_MIN_CBITS)
    try:
        for lanes, cbits in [(8, 4), (12, 6), (16, 10), (20, 15)]:
            native.set_tuning(native.TUNE_LANES_LOG2, lanes)
            native.set_tuning(native.TUNE_MIN_CBITS, cbits)
            assert native.permanent_residues(A, primes) == base, (lanes, cbits)
    finally:
        native.set_tuning(native.TUNE_LANES_LOG2, old[0])
        native.set_tuning(native.TUNE_MIN_CBITS, old[1])


def test_batch_rejects_unsupported_jobs(native):
    with pytest.raises(RuntimeError, match="orders"):
        native.Batch([np.eye(4, dtype=np.int64)], [[101]])
    with pytest.raises(RuntimeError, match="odd"):
        native.Batch([np.eye(6, dtype=np.int64)], [[1 << 40]])
    # A later valid plan still works.
    b = native.Batch([np.eye(6, dtype=np.int64)], [[101]])
    assert b.run() == [1]
    b.close()


def test_permanent_batch_exact_integers(native):
    from permhip.batch import PermanentBatch
    import oracle

    rng = np.random.default_rng(21)
    mats = [rng.integers(-100, 101, (n, n)).astype(np.int64) for n in (2, 3, 5, 9, 12, 13)]
    mats.append(np.array([[1 << 70, 2], [3, 4]], dtype=object))  # beyond int64
    got = PermanentBatch(mats).run()
    for M, v in zip(mats, got):
        assert v == oracle.subset_permanent([[int(x) for x in r] for r in M])


def test_item_resizing_of_a_live_plan(native):
    """Re-sizing the chunks of an uploaded plan keeps every residue."""
    mats, moduli = _mixed_batch(17, 24)
    batch = native.Batch(mats, moduli)
    flat = batch.run()
    seen = set()
    for items_log2 in (0, 9, 14, 20, 30, 15):
        batch.set_option(native.BATCH_ITEMS_LOG2, items_log2)
        st = batch.stats()
        assert st["items_log2"] == items_log2
        assert 4 <= st["chunk_bits_min"] <= st["chunk_bits_max"] <= 45
        seen.add((st["chunk_bits_min"], st["chunk_bits_max"], st["items"]))
        assert batch.run() == flat, items_log2
    assert len(seen) > 1
    with pytest.raises(RuntimeError, match="work items"):
        batch.set_option(native.BATCH_ITEMS_LOG2, 31)
    assert batch.run() == flat
    batch.close()


def test_autotuned_batch_matches_default(native):
    from permhip import autotune
    from permhip.batch import PermanentBatch

    rng = np.random.default_rng(31)
    mats = [rng.integers(-50, 51, (n, n)).astype(np.int64) for n in rng.integers(14, 23, size=40)]
    plain = PermanentBatch(mats)
    expect = plain.run()
    plain.close()
    tuned = PermanentBatch(mats, autotune={"reps": 1, "max_seconds": 20.0})
    r = tuned.tuning
    assert r is not None and len(r.trials) >= 2
    st = tuned.stats()
    assert (st["items_log2"], st["blocks_per_cu"]) == (r.items_log2, r.blocks_per_cu)
    assert st["waves"] % (autotune.WAVES_PER_BLOCK * st["blocks_per_cu"]) == 0
    assert tuned.run() == expect
    tuned.close()


def test_autotuned_single_matrix_path(native):
    import perm
    from permhip import autotune

    old = native.get_tuning(native.TUNE_LANES_LOG2), native.get_tuning(native.TUNE_BATCH_BLOCKS)
    rng = np.random.default_rng(41)
    A = rng.integers(-10, 11, (22, 22)).astype(np.int64)
    before = perm.matrix_permanent(A)
    try:
        r = autotune.tune_single(22, reps=1, max_seconds=10.0)
        assert native.get_tuning(native.TUNE_LANES_LOG2) == r.items_log2 + 6
        assert native.get_tuning(native.TUNE_BATCH_BLOCKS) == r.blocks_per_cu
        assert perm.matrix_permanent(A) == before
    finally:
        native.set_tuning(native.TUNE_LANES_LOG2, old[0])
        native.set_tuning(native.TUNE_BATCH_BLOCKS, old[1])


def test_resident_workgroups_are_reported(native):
    """The plan reports the most workgroups per CU any of its class kernels
    keeps resident (each class's grid is capped at its own count), and the
    autotuner does not try more than that."""
    import numpy as np

    from permhip import autotune

    A = np.random.default_rng(20261203).integers(-10, 11, (32, 32)).astype(np.int64)
    primes = [1073741827, 1073741831, 1073741833, 1073741839, 1073741843, 1073741857, 1073741891, 1073741909]
    plan = native.Batch([A], [primes])
    try:
        resident = plan.stats()["resident_blocks_per_cu"]
        assert 1 <= resident <= 8
        result = autotune.tune_plan(plan, items=(14, 15), blocks=(1, 2, 3, 4), reps=1, max_seconds=10)
        assert result.blocks_per_cu <= resident
        assert all(t.blocks_per_cu <= resident for t in result.trials)
    finally:
        plan.close()
