"""The 4096-matrix batch in one plan and a slice of the served stream, at
their real sizes, against residues written by the reference's own compiled
kernels (tests/golden/make_golden.py: perm_c_ext.permanent_mod_p and
perm_binary_ext.binary_permanent_mod_p, built from the reference sources).

Requests beyond what the reference's CPU path pins here (integer orders
above 28, 0/1 orders above 30) are checked by the permanent's invariance
under transposition and row and column permutations, served in a second,
differently mixed stream."""

import hashlib
import json
import time
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "reference_residues.json"


def _digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int64).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def golden():
    return json.loads(GOLDEN.read_text())


def test_batch4096_one_plan_matches_reference_residues(native, golden):
    """The whole 4096-matrix plan (every matrix and prime, one persistent
    launch per class) reproduces the reference's residues and exact values
    for the golden matrices, and those matrices cover every kernel class the
    plan uses."""
    from permhip import workloads
    from permhip.batch import PermanentBatch

    mats = workloads.batch4096()
    batch = PermanentBatch(mats)
    try:
        stats = batch.stats()
        # At least one job per matrix (a table's primes split into sets, or
        # one exact job for all of them).
        assert stats["jobs"] >= len(mats) and stats["residues"] == sum(len(p) for p in batch.primes)
        residues = batch.residues()
        exact = batch.run()
    finally:
        batch.close()
    classes = native.batch_plan_classes([mats[i] for i in batch.device_index],
                                        [batch.primes[i] for i in batch.device_index])
    slot_of = {}
    pos = 0
    for i in batch.device_index:
        slot_of[i] = range(pos, pos + len(batch.primes[i]))
        pos += len(batch.primes[i])
    assert pos == len(classes)
    keys = [k for k in golden if k.startswith("batch4096[")]
    assert len(keys) >= 30
    covered = set()
    for key in keys:
        i = int(key[len("batch4096["):-1])
        g = golden[key]
        assert g["n"] == mats[i].shape[0]
        if "sha256_int64" in g:
            assert g["sha256_int64"] == _digest(mats[i])
        assert list(batch.primes[i]) == g["moduli"], key
        assert residues[i] == g["residues"], key
        assert exact[i] == int(g["exact"]), key
        covered.update(classes[s] for s in slot_of[i])
    assert set(classes) == covered, sorted(set(classes) - covered)


def _serve(server, stream, pace: bool):
    futures = []
    t0 = time.perf_counter()
    for arrival, kind, A in stream:
        if pace:
            delay = t0 + arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
        futures.append(server.submit(A, kind))
    return [f.result(timeout=600) for f in futures]


def test_served_slice_matches_reference_and_invariance(native, golden):
    """The first 200 requests of the served stream (Poisson 200 req/s,
    integer n in [10, 34], 0/1 n in [16, 38]) through the continuous-batching
    server at their arrival times."""
    from permhip import workloads
    from permhip.serve import PermanentServer

    slice_ = golden["serve_slice"]
    count = int(slice_["generator"].rsplit(":", 1)[1].rstrip("]"))
    stream = workloads.serve_stream()[:count]
    server = PermanentServer()
    try:
        got = _serve(server, stream, pace=True)
    finally:
        server.close()
    assert server.errors == 0 and server.plans_launched > 0
    requests = slice_["requests"]
    pinned = 0
    for i, (_, kind, A) in enumerate(stream):
        g = requests.get(str(i))
        if g is None:
            continue
        assert g["kind"] == kind and g["n"] == A.shape[0]
        assert g["sha256_int64"] == _digest(A)
        assert got[i] == int(g["exact"]), (i, kind, A.shape[0])
        pinned += 1
    assert pinned >= 120
    assert {k for _, k, _ in stream} == {"integer", "binary"}

    # Every request again, transposed and with rows and columns permuted,
    # in reverse arrival order and all at once.
    rng = np.random.default_rng(55)
    variant = []
    for _, kind, A in reversed(stream):
        n = A.shape[0]
        B = A.T[rng.permutation(n)][:, rng.permutation(n)]
        variant.append((0.0, kind, np.ascontiguousarray(B)))
    server = PermanentServer()
    try:
        again = _serve(server, variant, pace=False)
    finally:
        server.close()
    assert list(reversed(again)) == got
    large = [i for i, (_, k, A) in enumerate(stream) if str(i) not in requests]
    assert large and max(stream[i][2].shape[0] for i in large) >= 33
