"""The step-log partition's code paths, bit-exact against the oracle.

The cycle groups its step log by 2048-cell tile.  Which scatter runs depends
on the largest tile window of a partition block (the tiles of the terrains of
its drops): up to 1024 tiles (maps up to 2^21 cells; every other GPU test),
up to 4096 tiles (2048x2048 here) and the unstaged scatter beyond that
(4096x4096 here).  Also: a cycle without drops (empty partition, the tile
reduce still writes zero commits and tracks) and a parameter change between
cycles (the per-terrain volume-track table is rebuilt).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def run_pair(shy, orc, n, seed, plan, params=None):
    """plan: list of (drops, oracle params or None to keep) per cycle."""
    P = params if params is not None else orc.Params.default()
    w = shy.World(n, n, seed=seed)
    w.generate_fbm(seed, 8)
    h = orc.fbm(seed, n, n)
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
    # 2048x2048 = 2048 tiles in one block's window: the 4096-tile staged scatter
    run_pair(shy, orc, 2048, 3, [(16384, None), (16384, None)])


def test_partition_unstaged_scatter(shy, orc):
    # 4096x4096 = 8192 tiles: beyond LDS staging, the plain LDS-cursor scatter
    run_pair(shy, orc, 4096, 4, [(8192, None), (8192, None)])


def test_partition_empty_and_parameter_change(shy, orc):
    P1 = orc.Params.default()
    P2 = orc.Params.default()
    P2.evapRate = 0.0023
    P2.maxAge = 300.0
    P3 = orc.Params.default()
    P3.maxAge = 700.0  # more log rows than the buffers held
    run_pair(shy, orc, 256, 9, [(4096, P1), (0, None), (4096, None), (4096, P2),
                                (4096, None), (4096, P3), (4096, None)])
