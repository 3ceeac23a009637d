"""The golden fixtures of the 4096-matrix batch cover every kernel class
(row-sum group, prime-set size, variant) the batch's plan uses, so the GPU
test that runs the whole plan checks every class against residues written
by the reference.  Host planning only (no device)."""

import json
from pathlib import Path

from permhip import _native, planner, workloads

GOLDEN = Path(__file__).resolve().parent / "golden" / "reference_residues.json"


def test_batch_golden_covers_every_plan_class():
    golden = json.loads(GOLDEN.read_text())
    mats = workloads.batch4096()
    primes = [planner.generic_crt_primes(planner.matrix_permanent_bound(m.tolist()))[0] for m in mats]
    classes = _native.batch_plan_classes(mats, primes)
    assert len(classes) == sum(len(p) for p in primes)
    covered, pos, starts = set(), 0, []
    for p in primes:
        starts.append(pos)
        pos += len(p)
    for key in golden:
        if key.startswith("batch4096["):
            i = int(key[len("batch4096["):-1])
            assert golden[key]["moduli"] == primes[i]
            covered |= set(classes[starts[i]:starts[i] + len(primes[i])])
    assert set(classes) == covered, sorted(set(classes) - covered)
    # The classes a 4096-matrix plan of entries U[-100, 100] uses: packed row
    # sums in groups of 2 or 3; modular sets of 1, 2, 4 and 8 primes, exact
    # walks with 8 limb columns (at most eight groups, orders 16..24) and
    # 10 (the orders 24..28, whose row-bound products stay below 2^310),
    # the tables of 2-row groups among them on the compact classes (their
    # partial bound products fit 22-bit group values).
    assert {c[0] for c in classes} <= {2, 3}
    assert {c[1] for c in classes if c[2] not in (3, 4)} <= {1, 2, 4, 8}
    assert {c[1] for c in classes if c[2] == 3} == {8, 10}
    assert {(c[0], c[1]) for c in classes if c[2] == 4} == {(2, 10)}
