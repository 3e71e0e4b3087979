"""GPU: batched ensembles of uneven terrains in one launch per kernel.

Each terrain of the ensemble must evolve exactly as the oracle's
wave-synchronous cycle does for that terrain alone (its own size, parameters,
seed and terrain number in the spawn hash), under every launch shape the
autotuner may choose.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def make_specs(shy, sizes, seed=21):
    rng = np.random.default_rng(seed)
    specs = []
    for i, (nx, ny) in enumerate(sizes):
        p = shy.Params(evapRate=float(rng.uniform(0.0005, 0.002)),
                       depositionRate=float(rng.uniform(0.05, 0.2)))
        specs.append(dict(nx=nx, ny=ny, drops=nx * ny // 16, seed=1000 + i, params=p))
    return specs


def oracle_terrain(orc, spec, t, cycles):
    nx, ny = spec["nx"], spec["ny"]
    h = orc.fbm(spec["seed"], nx, ny)
    z = np.zeros_like(h)
    d, mx, my, root, erfd = z.copy(), z.copy(), z.copy(), z.copy(), z.copy()
    sp = spec["params"]
    P = orc.Params.default(**{k: getattr(sp, k) for k, _ in sp._fields_})
    steps = 0
    for c in range(cycles):
        st = orc.wave_cycle(h, d, mx, my, root, erfd, P, spec["seed"], c, spec["drops"],
                            terrain=t)
        steps += st.steps
    return h, d, mx, my, erfd, steps


SIZES = [(96, 96), (64, 160), (128, 80), (40, 40), (200, 128)]


def test_ensemble_matches_oracle_per_terrain(shy, orc):
    specs = make_specs(shy, SIZES)
    e = shy.Ensemble(specs)
    for t, s in enumerate(specs):
        e.generate_fbm(t, s["seed"], 8)
    cycles = 4
    print("ensemble", len(specs), "terrains", flush=True)
    e.erode(cycles, graph=True)
    steps = e.terrain_steps()
    for t, s in enumerate(specs):
        h, d, mx, my, erfd, st = oracle_terrain(orc, s, t, cycles)
        for name, ref in (("height", h), ("discharge", d), ("momentumx", mx),
                          ("momentumy", my), ("erf_discharge", erfd)):
            assert np.array_equal(bits(e.field(t, name)), bits(ref)), (t, name)
        assert steps[t] == st
    assert e.stats()["steps"] == steps.sum()
    e.close()


@pytest.mark.parametrize("tuning", [dict(walk_block=64, walk_grid=0, tile=1),
                                    dict(walk_block=512, walk_grid=37, tile=2),
                                    dict(walk_block=128, walk_grid=5, tile=0)])
def test_ensemble_launch_shapes_bit_identical(shy, tuning):
    specs = make_specs(shy, SIZES, seed=5)
    out = []
    for tune in (None, tuning):
        e = shy.Ensemble(specs)
        for t, s in enumerate(specs):
            e.generate_fbm(t, s["seed"], 8)
        if tune:
            e.set_tuning(**tune)
        print("shape", tune, flush=True)
        e.erode(3, graph=False)
        out.append([e.field(t, "height") for t in range(len(specs))] +
                   [e.field(t, "discharge") for t in range(len(specs))])
        e.close()
    for a, b in zip(*out):
        assert np.array_equal(bits(a), bits(b))


def test_autotune_keeps_trajectory(shy):
    specs = make_specs(shy, SIZES, seed=9)
    res = []
    for tune in (False, True):
        e = shy.Ensemble(specs)
        for t, s in enumerate(specs):
            e.generate_fbm(t, s["seed"], 8)
        e.erode(2, graph=False)
        if tune:
            chosen, ms = e.autotune(1)
            print("autotune chose", chosen, ms, flush=True)
            assert ms > 0 and chosen["walk_block"] in (64, 128, 256, 512)
        e.erode(2, graph=True)
        res.append([e.field(t, "height") for t in range(len(specs))])
        e.close()
    for a, b in zip(*res):
        assert np.array_equal(bits(a), bits(b))


def test_single_terrain_context_rejects_ensemble_only_misuse(shy):
    specs = make_specs(shy, SIZES[:2])
    e = shy.Ensemble(specs)
    with pytest.raises(shy.ShyError) as err:
        shy._check(e._ctx, shy.lib().shy_erode(e._ctx, 10))
    assert err.value.code == shy.SHY_ERR_STATE
    e.close()
