"""The step-log partition's code paths, bit-exact against the oracle.

The cycle groups its step log by 4096-cell tile.  Which scatter runs depends
on the most tiles of one terrain: up to 1024 tiles (maps up to 2^22 cells;
every other GPU test), up to 4096 tiles (4096x4096 here) and the unstaged
scatter beyond that (4096x8192 here).  Also: a cycle without drops (empty
partition, the tile reduce still writes zero commits and tracks), a
parameter change between cycles (the per-terrain volume-track table is
rebuilt), and ensembles whose terrains share tiles, fit inside one tile or
have no drops, with the walk counting the tiles (blocks of 64 and 256) and
without (blocks of 512 span terrains; a persistent walk grid).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def run_pair(shy, orc, n, seed, plan, params=None, walk_block=None, ny=None):
    """plan: list of (drops, oracle params or None to keep) per cycle; an
    n x ny terrain (ny defaults to n)."""
    P = params if params is not None else orc.Params.default()
    ny = n if ny is None else ny
    w = shy.World(n, ny, seed=seed)
    if walk_block:
        w.set_walk_block(walk_block)
    w.generate_fbm(seed, 8)
    h = orc.fbm(seed, n, ny)
    assert np.array_equal(bits(w.field("height")), bits(h))
    z = np.zeros_like(h)
    d, mx, my, root, erfd = z.copy(), z.copy(), z.copy(), z.copy(), z.copy()
    tot = {}
    for c, (drops, newp) in enumerate(plan):
        if newp is not None:
            P = newp
            w.params = shy.Params(**{k: getattr(P, k) for k, _ in P._fields_})
        print(f"partition {n}x{n} cycle {c} drops {drops}", flush=True)
        w.erode(drops)
        s = orc.wave_cycle(h, d, mx, my, root, erfd, P, seed, c, drops).as_dict()
        for k, v in s.items():
            tot[k] = tot.get(k, 0) + v
        for name, ref in (("height", h), ("discharge", d), ("momentumx", mx),
                          ("momentumy", my), ("erf_discharge", erfd)):
            assert np.array_equal(bits(w.field(name)), bits(ref)), (n, c, name)
    assert w.stats() == tot
    w.close()


def test_partition_window_4096_tiles(shy, orc):
    # 4096x4096 = 4096 tiles in one terrain: the 4096-tile staged scatter
    run_pair(shy, orc, 4096, 3, [(8192, None), (8192, None)])


def test_partition_unstaged_scatter(shy, orc):
    # 4096x8192 = 8192 tiles: beyond LDS staging, the plain LDS-cursor scatter
    run_pair(shy, orc, 4096, 4, [(8192, None), (8192, None)], ny=8192)


def test_partition_empty_and_parameter_change(shy, orc):
    P1 = orc.Params.default()
    P2 = orc.Params.default()
    P2.evapRate = 0.0023
    P2.maxAge = 300.0
    P3 = orc.Params.default()
    P3.maxAge = 700.0  # more log rows than the buffers held
    run_pair(shy, orc, 256, 9, [(4096, P1), (0, None), (4096, None), (4096, P2),
                                (4096, None), (4096, P3), (4096, None)])


def ensemble_vs_oracle(shy, orc, specs, cycles, tuning):
    e = shy.Ensemble(specs)
    for t, sp in enumerate(specs):
        e.generate_fbm(t, sp["seed"], 8)
    if tuning:
        e.set_tuning(**tuning)
    print("partition ensemble", [(s["nx"], s["ny"], s["drops"]) for s in specs], tuning,
          flush=True)
    e.erode(cycles, graph=False)
    steps = e.terrain_steps()
    for t, sp in enumerate(specs):
        nx, ny = sp["nx"], sp["ny"]
        h = orc.fbm(sp["seed"], nx, ny)
        z = np.zeros_like(h)
        d, mx, my, root, erfd = z.copy(), z.copy(), z.copy(), z.copy(), z.copy()
        P = orc.Params.default()
        st = 0
        for c in range(cycles):
            st += orc.wave_cycle(h, d, mx, my, root, erfd, P, sp["seed"], c, sp["drops"],
                                 terrain=t).steps
        for name, ref in (("height", h), ("discharge", d), ("momentumx", mx),
                          ("momentumy", my), ("erf_discharge", erfd)):
            assert np.array_equal(bits(e.field(t, name)), bits(ref)), (t, name, tuning)
        assert steps[t] == st, (t, tuning)
    e.close()


# zero-drop terrains first, in the middle and last; three terrains inside one
# tile; a terrain of 4.5 tiles; drop counts that are not multiples of 64
MIXED = [(24, 24, 0), (32, 32, 200), (16, 40, 37), (20, 20, 90), (96, 96, 0),
         (80, 115, 700), (48, 48, 130), (30, 30, 0)]


@pytest.mark.parametrize("tuning", [None, dict(walk_block=64, walk_grid=0, tile=0),
                                    dict(walk_block=512, walk_grid=0, tile=0),
                                    dict(walk_block=128, walk_grid=7, tile=0)])
def test_partition_ensemble_shared_and_empty_terrains(shy, orc, tuning):
    specs = [dict(nx=nx, ny=ny, drops=dr, seed=300 + i)
             for i, (nx, ny, dr) in enumerate(MIXED)]
    ensemble_vs_oracle(shy, orc, specs, 3, tuning)


def test_drop_order_bit_identical(shy):
    # the spatial drop order only changes which thread walks which drop
    out = []
    for on in (True, False):
        w = shy.World(512, 512, seed=5)
        w.generate_fbm(5, 8)
        w.set_drop_order(on)
        print("drop order", on, flush=True)
        w.erode_cycles(3, 8000, graph=False)
        w.erode_cycles(2, 8000, graph=True)
        out.append([w.field(k) for k in ("height", "discharge", "momentumx", "momentumy",
                                         "erf_discharge")] + [w.stats()])
        w.close()
    for a, b in zip(out[0][:-1], out[1][:-1]):
        assert np.array_equal(bits(a), bits(b))
    assert out[0][-1] == out[1][-1]
    specs = [dict(nx=nx, ny=ny, drops=dr, seed=500 + i) for i, (nx, ny, dr) in enumerate(MIXED)]
    res = []
    for on in (True, False):
        e = shy.Ensemble(specs)
        for t, sp in enumerate(specs):
            e.generate_fbm(t, sp["seed"], 8)
        e.set_tuning(walk_block=64, walk_grid=0, tile=0)
        e.set_drop_order(on)
        e.erode(3, graph=True)
        res.append([e.field(t, "height") for t in range(len(specs))] + [e.terrain_steps()])
        e.close()
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_partition_more_blocks_than_cus(shy, orc):
    # walk blocks of 64: 20480 drops = 320 partition blocks, more than the CUs,
    # so the scatter runs rounds of 4096 (two blocks per CU) instead of 8192
    run_pair(shy, orc, 512, 6, [(20480, None), (20480, None)], walk_block=64)


def test_partition_block_steps_read_from_memory(shy, orc):
    # walk blocks of 1024 threads and 266240 drops: 260 partition blocks of
    # 1024 drops, more than the CUs, so the scatter runs rounds of 4096, whose
    # LDS holds the step counts of at most 1008 drops: it reads them from
    # memory (evapRate 0.1 keeps the oracle's walks short)
    P = orc.Params.default(evapRate=0.1)
    run_pair(shy, orc, 512, 8, [(266240, P), (266240, None)], params=P, walk_block=1024)
