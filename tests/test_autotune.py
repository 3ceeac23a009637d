"""The launch-shape autotuner's search, checks and cache on the CPU.

A plan stand-in charges each replay a cost that depends on its options and
advances a virtual clock by it, so the search is deterministic; the GPU test
in test_gpu_batch.py runs the tuner on real plans."""

import json

import pytest

from permhip import _native, autotune


class ClockPlan:
    def __init__(self, cost, residues=(11, 22, 33), bad=None, items=15, blocks=2):
        self.cost = cost
        self.residues = list(residues)
        self.bad = bad or (lambda i, b: False)
        self.opts = {_native.BATCH_ITEMS_LOG2: items, _native.BATCH_BLOCKS_PER_CU: blocks}
        self.now = 0.0
        self.runs = 0

    def setting(self):
        return self.opts[_native.BATCH_ITEMS_LOG2], self.opts[_native.BATCH_BLOCKS_PER_CU]

    def set_option(self, option, value):
        self.opts[option] = int(value)

    def run(self, fetch=True):
        self.runs += 1
        self.now += self.cost(*self.setting())
        if not fetch:
            return None
        out = list(self.residues)
        if self.bad(*self.setting()):
            out[0] += 1
        return out

    def stats(self):
        i, b = self.setting()
        return {"items_log2": float(i), "blocks_per_cu": float(b), "chunk_bits_min": float(25 - i),
                "chunk_bits_max": float(25 - i), "states": 2.0 ** 31, "jobs": 1.0, "residues": 3.0,
                "group": 4.0, "set": 8.0}


def bowl(i, b):
    return 1.0 + 0.05 * (i - 13) ** 2 + 0.1 * (b - 3) ** 2


def test_descent_finds_the_minimum_and_leaves_the_plan_on_it():
    plan = ClockPlan(bowl)
    r = autotune.tune_plan(plan, reps=2, clock=lambda: plan.now)
    assert (r.items_log2, r.blocks_per_cu) == (13, 3)
    assert plan.setting() == (13, 3)
    assert r.waves_per_cu == 12
    assert r.chunk_bits == (12, 12)
    assert r.seconds == pytest.approx(bowl(13, 3))
    assert r.default_seconds == pytest.approx(bowl(15, 2))
    assert r.speedup > 1.0
    # Every trial timed once; start, 3 other block counts, 6 other item
    # targets, then the block counts again at items 13 except the incumbent.
    settings = [(t.items_log2, t.blocks_per_cu) for t in r.trials]
    assert len(settings) == len(set(settings))
    assert (15, 2) in settings and (13, 3) in settings
    d = r.as_dict()
    assert d["waves_per_cu"] == 12 and d["trials"] == len(r.trials)


def test_noise_below_tolerance_keeps_the_start():
    plan = ClockPlan(lambda i, b: 1.0 - 0.001 * (i == 12))
    r = autotune.tune_plan(plan, reps=1, tolerance=0.01, clock=lambda: plan.now)
    assert (r.items_log2, r.blocks_per_cu) == (15, 2)
    assert plan.setting() == (15, 2)


def test_differing_residues_raise():
    plan = ClockPlan(bowl, bad=lambda i, b: b == 4)
    with pytest.raises(RuntimeError, match="residues differ"):
        autotune.tune_plan(plan, reps=1, clock=lambda: plan.now)


def test_time_budget_stops_the_search():
    plan = ClockPlan(lambda i, b: 10.0)
    r = autotune.tune_plan(plan, reps=1, max_seconds=25.0, clock=lambda: plan.now)
    # Each trial costs 20 virtual seconds (one checked run plus one timed).
    assert len(r.trials) == 2
    assert plan.setting() == (r.items_log2, r.blocks_per_cu)


def test_reps_must_be_positive():
    with pytest.raises(ValueError):
        autotune.tune_plan(ClockPlan(bowl), reps=0)


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.json")
    plan = ClockPlan(bowl)
    first = autotune.tune_plan(plan, reps=1, cache=path, clock=lambda: plan.now)
    assert not first.cached
    data = json.loads(open(path).read())
    assert len(data) == 1
    entry = next(iter(data.values()))
    assert (entry["items_log2"], entry["blocks_per_cu"]) == (13, 3)

    again = ClockPlan(bowl)
    second = autotune.tune_plan(again, reps=1, cache=path, clock=lambda: again.now)
    assert second.cached and again.runs == 0
    assert again.setting() == (13, 3)
    assert (second.items_log2, second.blocks_per_cu) == (13, 3)

    # Another workload shape misses the cache.
    other = ClockPlan(bowl, residues=(1, 2))
    other.stats = lambda: dict(ClockPlan.stats(other), residues=2.0)
    third = autotune.tune_plan(other, reps=1, cache=path, clock=lambda: other.now)
    assert not third.cached and other.runs > 0
    assert len(json.loads(open(path).read())) == 2


def test_default_cache_path(monkeypatch):
    monkeypatch.delenv("PERMHIP_AUTOTUNE_CACHE", raising=False)
    assert autotune.default_cache_path() is None
    monkeypatch.setenv("PERMHIP_AUTOTUNE_CACHE", "/tmp/x.json")
    assert autotune.default_cache_path() == "/tmp/x.json"


def test_tune_single_sets_and_restores_tuning(monkeypatch):
    state = {_native.TUNE_LANES_LOG2: 20, _native.TUNE_BATCH_BLOCKS: 2}
    clock = [0.0]

    def residues(matrix, moduli):
        i, b = state[_native.TUNE_LANES_LOG2] - 6, state[_native.TUNE_BATCH_BLOCKS]
        clock[0] += bowl(i, b)
        return [len(moduli)] * len(moduli)

    monkeypatch.setattr(_native, "set_tuning", lambda k, v: state.__setitem__(k, int(v)))
    monkeypatch.setattr(_native, "get_tuning", lambda k: state[k])
    monkeypatch.setattr(_native, "permanent_residues", residues)
    r = autotune.tune_single(12, reps=1, lanes=(17, 18, 19, 20, 21), clock=lambda: clock[0])
    assert (r.items_log2, r.blocks_per_cu) == (13, 3)
    assert state == {_native.TUNE_LANES_LOG2: 19, _native.TUNE_BATCH_BLOCKS: 3}

    state.update({_native.TUNE_LANES_LOG2: 20, _native.TUNE_BATCH_BLOCKS: 2})

    def broken(matrix, moduli):
        clock[0] += 1.0
        if state[_native.TUNE_BATCH_BLOCKS] == 1:
            return [0] * len(moduli)
        return [1] * len(moduli)

    monkeypatch.setattr(_native, "permanent_residues", broken)
    with pytest.raises(RuntimeError, match="residues differ"):
        autotune.tune_single(12, reps=1, clock=lambda: clock[0])
    assert state == {_native.TUNE_LANES_LOG2: 20, _native.TUNE_BATCH_BLOCKS: 2}


def test_headline_instruction_mix_is_compiled_into_the_microbench():
    """bench.py's roofline times the headline kernel's instruction mix on a
    synthetic kernel compiled for fixed mixes; the mix the bench derives for
    the headline (paired and per-state sets of 8, and single primes) must be
    one of them, or the bench cannot report its roofline."""
    import re
    from pathlib import Path

    import bench

    root = Path(__file__).resolve().parent.parent
    src = (root / "rivin-perm_amd" / "csrc" / "microbench.hip").read_text()
    compiled = {tuple(int(v) for v in m.groups()) for m in re.finditer(
        r"mix_kernel<(\d+), (\d+), (\d+), (\d+), (\d+)>", src)}
    from permhip import _native

    for k, paired in ((8, True), (8, False), (1, False)):
        mix = bench.instruction_mix(bench.HEADLINE_N, 4, k, paired)
        rounds = tuple(round(mix[t]) for t in _native.MIX_TYPES)
        assert rounds in compiled, (k, paired, rounds)
