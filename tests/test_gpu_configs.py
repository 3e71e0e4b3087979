"""GPU at the two large workloads, against the oracle's digests.

* time-to-steady workload: 2048x2048 fBm seed 7, default parameters, 262144
  drops per cycle (the drop order is on by default at this size);
* batched ensemble: the 64 terrains of ensemble_spec(64, 1), sides 128..1024,
  per-terrain evapRate/depositionRate, 1 drop per 16 cells, as the benchmark
  runs it (autotuned launch shape, hipGraph replay).

tests/golden/config_hashes.json holds SHA-256 digests of the oracle's fields
after cycles 0 and 1 and its step statistics (tests/golden/make_config_hashes.py
made them; the oracle needs about a minute per 2048x2048 cycle).  Each cycle
must give the same bits.  Longer runs are checked through size-independent
properties: the integer mass ledger, bounded terrain, step accounting.
"""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "config_hashes.json")
FIELDS = ("height", "discharge", "momentumx", "momentumy", "erf_discharge")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def say(*a):
    print(*a, flush=True)


def test_steady_workload_cycles_match_oracle(shy, golden):
    g = golden["steady2048"]
    n, drops, seed = g["nx"], g["drops"], g["seed"]
    w = shy.World(n, n, seed=seed)
    w.generate_fbm(seed, 8)
    assert digest(w.field("height")) == g["initial_height"]
    for c, want in enumerate(g["cycles"]):
        say("2048^2 cycle", c)
        w.reset_stats()
        w.erode(drops)
        for f in FIELDS:
            assert digest(w.field(f)) == want[f], (c, f)
        assert w.stats() == want["stats"], c
    # 28 more cycles from the graph: ledger and bounds
    s0 = w.field("height").astype(np.float64).sum()
    w.reset_stats()
    say("2048^2 cycles 2..29")
    w.erode_cycles(28, drops, graph=True)
    h = w.field("height")
    st = w.stats()
    s1 = h.astype(np.float64).sum()
    ledger = (st["gain40"] - st["export40"]) * 2.0 ** -40
    say("ledger", s1 - s0, ledger, st)
    assert abs((s1 - s0) - ledger) <= 1e-6 * s0
    assert abs((s1 - s0) - st["walk_dh"] * 2.0 ** -20) <= 1e-6 * s0
    assert np.isfinite(h).all() and -0.5 < h.min() and h.max() < 1.5
    assert st["walked"] + st["skipped"] == 28 * drops
    assert st["exits"] + st["ends_age"] + st["ends_vol"] == st["walked"]
    d = w.field("discharge")
    assert np.isfinite(d).all() and d.min() >= 0.0
    w.close()


def make_ensemble(shy, golden):
    specs = []
    for g in golden["ensemble64"]:
        specs.append(dict(nx=g["nx"], ny=g["ny"], drops=g["drops"], seed=g["seed"],
                          params=shy.Params(**g["params"])))
    e = shy.Ensemble(specs)
    for t, s in enumerate(specs):
        e.generate_fbm(t, s["seed"], 8)
    return e, specs


def test_ensemble_workload_matches_oracle(shy, golden):
    """All 64 terrains, bit for bit, after an autotune (which must leave the
    state where it found it) and two graph-replayed cycles."""
    gs = golden["ensemble64"]
    spec = shy.ensemble_spec(64, 1)
    assert [(s["nx"], s["drops"], s["seed"]) for s in spec] == \
        [(g["nx"], g["drops"], g["seed"]) for g in gs]
    e, specs = make_ensemble(shy, golden)
    for t, g in enumerate(gs):
        assert digest(e.field(t, "height")) == g["initial_height"], t
    chosen, ms = e.autotune(1)
    say("autotune", chosen, ms)
    for c in range(2):
        say("ensemble cycle", c)
        e.reset_stats()
        e.erode(1, graph=True)
        steps = e.terrain_steps()
        for t, g in enumerate(gs):
            want = g["cycles"][c]
            for f in FIELDS:
                assert digest(e.field(t, f)) == want[f], (c, t, f)
            assert steps[t] == want["stats"]["steps"], (c, t)
        tot = {k: sum(g["cycles"][c]["stats"][k] for g in gs) for k in gs[0]["cycles"][c]["stats"]}
        assert e.stats() == tot, c
    e.close()


def test_ensemble_workload_ledger(shy, golden):
    """20 cycles of the 64-terrain ensemble: per terrain, the height change
    equals the walk's integer commits (the cascade moves mass only between
    cells of one terrain), and the whole ensemble's change equals the
    evaporation gain less the sediment carried off the maps."""
    e, specs = make_ensemble(shy, golden)
    s0 = np.array([e.field(t, "height").astype(np.float64).sum() for t in range(len(specs))])
    e.reset_stats()
    say("ensemble 20 cycles")
    e.erode(20, graph=True)
    st = e.stats()
    s1 = np.array([e.field(t, "height").astype(np.float64).sum() for t in range(len(specs))])
    ledger = (st["gain40"] - st["export40"]) * 2.0 ** -40
    say("ledger", (s1 - s0).sum(), ledger)
    assert abs((s1 - s0).sum() - ledger) <= 1e-6 * s0.sum()
    assert abs((s1 - s0).sum() - st["walk_dh"] * 2.0 ** -20) <= 1e-6 * s0.sum()
    assert st["walked"] + st["skipped"] == 20 * sum(s["drops"] for s in specs)
    assert st["exits"] + st["ends_age"] + st["ends_vol"] == st["walked"]
    for t in range(len(specs)):
        h = e.field(t, "height")
        assert np.isfinite(h).all() and -0.5 < h.min() and h.max() < 1.5, t
    assert e.terrain_steps().sum() == st["steps"]
    e.close()
