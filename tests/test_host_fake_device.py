"""Host logic of the drop-in modules on the CPU, with the device calls
replaced by an oracle-backed fake (the reference tests its dispatch seams
the same way: tests/test_core.py:190-206, test_cuda_wrappers.py:10-64).

These tests cover planning, CRT, validation, block factoring, batching and
the serving scheduler; they say nothing about the kernels, which the gpu
tests check against the same oracle."""

import math
import random

import numpy as np
import pytest

import oracle


class FakeBatch:
    def __init__(self, matrices, moduli_lists):
        self.mats = [np.asarray(m, dtype=np.int64) for m in matrices]
        self.moduli = [list(ms) for ms in moduli_lists]
        self.total = sum(len(ms) for ms in self.moduli)
        self.options = {}
        self._pending = None
        for m in self.mats:
            if not 5 <= m.shape[0] <= 46:
                raise RuntimeError("phip_batch_create failed: batch orders must lie in [5, 46]")

    def _compute(self):
        out = []
        for m, ms in zip(self.mats, self.moduli):
            out += [oracle.C.permanent_mod_p(m.tolist(), q, 2) for q in ms]
        return out

    def set_option(self, option, value):
        self.options[option] = value

    def run(self, fetch=True):
        return self._compute() if fetch else None

    def launch(self):
        self._pending = self._compute()

    def poll(self, block=False):
        out, self._pending = self._pending, None
        return out

    def stats(self):
        return {"last_ms": 0.0, "jobs": float(self.total)}

    def close(self):
        pass


@pytest.fixture
def fake_device(monkeypatch):
    from permhip import _native, planner

    def residues(matrix, moduli):
        a = np.asarray(matrix, dtype=np.int64)
        moduli = list(moduli)
        if a.ndim == 2:
            return [oracle.C.permanent_mod_p(a.tolist(), m, 2) for m in moduli]
        return [oracle.C.permanent_mod_p(a[i].tolist(), m, 2) for i, m in enumerate(moduli)]

    def binary(matrix, primes):
        return [oracle.C.binary_mod_p(np.asarray(matrix, dtype=np.uint8), p, 2) for p in primes]

    def schur(n, omega, p):
        return oracle.C.permanent_mod_p(planner.build_schur_matrix_mod_p(n, omega, p), p, 2)

    monkeypatch.setattr(_native, "require_device", lambda: None)
    monkeypatch.setattr(_native, "device_info", lambda: "fake gfx950 CUs=4")
    monkeypatch.setattr(_native, "permanent_residues", residues)
    monkeypatch.setattr(_native, "binary_permanent_residues", binary)
    monkeypatch.setattr(_native, "schur_permanent_modp", schur)
    monkeypatch.setattr(_native, "permanent_mod3", lambda a: oracle.C.mod3(np.asarray(a), 2))
    monkeypatch.setattr(_native, "Batch", FakeBatch)
    return _native


OEIS_SMALL = {1: 1, 3: -3, 5: -5, 7: -105, 9: 81, 11: 6765, 13: 175747, 15: 30375, 17: 25219857}


def test_schur_planning_and_crt(fake_device):
    import perm

    for n, v in OEIS_SMALL.items():
        assert perm.permanent(n) == v, n
        assert perm.schur_permanent(n) == v, n
    assert [perm.permanent(n) for n in (0, 2, 4)] == [1, 0, 0]
    for n, p in ((7, 29), (13, 53)):
        w = perm.nth_root_of_unity_mod_p(n, p)
        S = perm.build_schur_matrix_mod_p(n, w, p)
        assert perm.schur_permanent_mod_p(n, p) == oracle.C.permanent_mod_p(S, p)
    with pytest.raises(ValueError, match="mod"):
        perm.schur_permanent_mod_p(7, 31)


def test_matrix_permanent_host_paths(fake_device):
    import perm

    rng = random.Random(5)
    for n in range(0, 11):
        A = [[rng.randrange(-30, 31) for _ in range(n)] for _ in range(n)]
        assert perm.matrix_permanent(A) == oracle.subset_permanent(A)
        assert perm.matrix_permanent(A, method="generic", parallel=False) == oracle.subset_permanent(A)
    big = [[1 << 80, 3, 0], [2, 1 << 70, 5], [7, 0, -(1 << 90)]]
    assert perm.matrix_permanent(big) == oracle.definition_permanent(big)
    assert perm.matrix_permanent(np.array([[2 ** 63, 1], [1, 2 ** 63]], dtype=np.uint64)) == 2 ** 126 + 1
    assert perm.ryser_permanent_mod_p([[1, 2], [3, 4]], 7) == 10 % 7
    assert perm.ryser_permanent_mod_p([[1, 2], [3, 4]], (1 << 70) + 1) == 10


def test_low_rank_agrees(fake_device):
    import perm

    n = 10
    A = [[(i + 1) * (j - 3) + (i % 3) * (j % 2) for j in range(n)] for i in range(n)]
    assert perm.matrix_permanent(A, method="low_rank") == oracle.subset_permanent(A)
    with pytest.raises(perm.LowRankLimitExceeded):
        rng = random.Random(2)
        perm.matrix_permanent([[rng.randrange(50) for _ in range(7)] for _ in range(7)], method="low_rank")


def test_binary_host_paths(fake_device):
    import perm_binary
    from permhip import grids

    for shape, expect in [((2, 2), 2), ((2, 4), 5), ((4, 4), 36), ((2, 2, 3), 32), ((2, 3, 4), 1845)]:
        assert perm_binary.binary_permanent(grids.grid_biadjacency(shape)) == expect
    rng = random.Random(77)
    for n in range(0, 12):
        B = [[1 if rng.random() < 0.35 else 0 for _ in range(n)] for _ in range(n)]
        d = oracle.subset_permanent(B) if n else 1
        assert perm_binary.binary_permanent(B) == d
        if n:
            assert perm_binary.binary_permanent_mod_p(B, 101) == d % 101
    # Block-diagonal input factors into blocks.
    J3 = [[1, 1, 1]] * 3
    B = [r + [0] * 3 for r in J3] + [[0] * 3 + r for r in J3]
    assert perm_binary.binary_permanent(B) == 36


def test_mod3_host_paths(fake_device):
    import perm_mod3

    rng = np.random.default_rng(3)
    for n in range(0, 9):
        A = rng.integers(-9, 10, (n, n))
        expect = oracle.definition_permanent(A.tolist()) % 3 if n else 1
        assert perm_mod3.permanent_mod3_fast(A) == expect
        assert perm_mod3.permanent_mod3_ryser(A) == expect


def test_permanent_batch_host(fake_device):
    from permhip.batch import PermanentBatch

    rng = np.random.default_rng(8)
    mats = [rng.integers(-100, 101, (n, n)) for n in (1, 2, 4, 5, 7, 9, 11)]
    got = PermanentBatch(mats).run()
    assert got == [oracle.subset_permanent(m.tolist()) for m in mats]

    tuned = PermanentBatch(mats, autotune={"reps": 1, "items": (14, 15), "blocks": (1, 2)})
    assert tuned.tuning is not None
    assert tuned.plan.options[fake_device.BATCH_ITEMS_LOG2] == tuned.tuning.items_log2
    assert tuned.plan.options[fake_device.BATCH_BLOCKS_PER_CU] == tuned.tuning.blocks_per_cu
    assert tuned.run() == got


def test_serve_scheduler(fake_device):
    from permhip.serve import PermanentServer

    rng = np.random.default_rng(9)
    reqs = []
    for i in range(30):
        n = int(rng.integers(1, 11))
        if i % 2:
            reqs.append(("integer", rng.integers(-50, 51, (n, n))))
        else:
            reqs.append(("binary", (rng.random((n, n)) < 0.6).astype(np.int64)))
    server = PermanentServer(max_inflight=2, plan_cost=1e4)
    futs = [server.submit(m, k) for k, m in reqs]
    bad = server.submit(np.array([[2]]), "binary")
    got = [f.result(timeout=30) for f in futs]
    with pytest.raises(ValueError, match="0 or 1"):
        bad.result(timeout=30)
    server.close()
    assert got == [oracle.subset_permanent(m.tolist()) for _, m in reqs]
    assert len(server.completed) == len(reqs)
    assert math.isfinite(sum(r.finished - r.arrival for r in server.completed))


def test_serve_scheduler_survives_failures(fake_device, monkeypatch):
    """A request or plan that fails resolves its own futures with the error;
    requests before and after it still complete and the scheduler thread
    stays alive."""
    from permhip import serve
    from permhip.serve import PermanentServer

    rng = np.random.default_rng(10)
    server = PermanentServer(max_inflight=2, plan_cost=1e4, direct_cost=0.0)
    first = rng.integers(-9, 10, (7, 7))
    ok1 = server.submit(first, "integer")
    assert ok1.result(timeout=30) == oracle.subset_permanent(first.tolist())

    # Orders the device path does not take are refused at submit.
    huge = server.submit(np.ones((64, 64), dtype=np.int64), "integer")
    with pytest.raises(ValueError, match="less than 64"):
        huge.result(timeout=30)

    # A plan whose creation fails fails only its own requests.
    class Boom(FakeBatch):
        def __init__(self, *a, **k):
            raise RuntimeError("phip_batch_create failed: hipMalloc failed")

    monkeypatch.setattr(fake_device, "Batch", Boom)
    doomed = server.submit(rng.integers(-9, 10, (8, 8)), "integer")
    with pytest.raises(RuntimeError, match="hipMalloc"):
        doomed.result(timeout=30)

    # A plan whose poll fails, likewise.
    class BadPoll(FakeBatch):
        def poll(self, block=False):
            raise RuntimeError("phip_batch_query failed: device lost")

    monkeypatch.setattr(fake_device, "Batch", BadPoll)
    lost = server.submit(rng.integers(-9, 10, (9, 9)), "integer")
    with pytest.raises(RuntimeError, match="device lost"):
        lost.result(timeout=30)

    # A request whose outside-plan part fails (order < 5 runs outside the
    # plans through the single-matrix path).
    monkeypatch.setattr(fake_device, "Batch", FakeBatch)
    real = fake_device.permanent_residues

    def flaky(matrix, moduli):
        if np.asarray(matrix).shape[-1] == 3:
            raise RuntimeError("phip_permanent_residues failed: injected")
        return real(matrix, moduli)

    monkeypatch.setattr(fake_device, "permanent_residues", flaky)
    small = server.submit(rng.integers(-9, 10, (3, 3)), "integer")
    with pytest.raises(RuntimeError, match="injected"):
        small.result(timeout=30)

    # Later requests complete normally.
    later = [rng.integers(-9, 10, (n, n)) for n in (6, 10, 4)]
    futs = [server.submit(m, "integer") for m in later]
    assert [f.result(timeout=30) for f in futs] == [oracle.subset_permanent(m.tolist()) for m in later]
    assert server._thread.is_alive()
    assert server.errors == 3
    server.close()
    assert serve.MAX_ORDER == 63
