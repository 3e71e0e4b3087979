"""GPU: the C ABI's error paths, state handling and the host mirror API."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_rejects_invalid_arguments(shy):
    with pytest.raises(shy.ShyError) as e:
        shy.World(0, 8)
    assert e.value.code == shy.SHY_ERR_ARG
    with pytest.raises(shy.ShyError):
        shy.World(8, 8, params=shy.Params(evapRate=1.0))
    with pytest.raises(shy.ShyError):
        shy.World(8, 8, params=shy.Params(maxAge=1e9))
    w = shy.World(16, 16)
    bad = np.zeros((16, 16), np.float32)
    bad[3, 4] = np.nan
    with pytest.raises(shy.ShyError) as e:
        w.upload("height", bad)
    assert e.value.code == shy.SHY_ERR_NONFINITE
    bad[3, 4] = np.inf
    with pytest.raises(shy.ShyError):
        w.upload("discharge", bad)
    with pytest.raises(shy.ShyError):
        w.set_walk_block(100)
    with pytest.raises(shy.ShyError):
        w.cascade((-3.0, 2.0))
    with pytest.raises(shy.ShyError):
        w.erode(-1)
    with pytest.raises(ValueError):
        w.upload("height", np.zeros((8, 8), np.float32))
    # the context still works after rejected calls
    w.generate_fbm(1)
    w.erode(64)
    assert w.stats()["walked"] + w.stats()["skipped"] == 64
    w.close()


def test_cells_roundtrip_and_tracks(shy):
    rng = np.random.default_rng(1)
    n = 40
    cells = rng.uniform(0, 1, (n, n, 8)).astype(np.float32)
    cells[..., 4:7] = 0.0
    w = shy.World(n, n)
    w.upload_cells(cells)
    out = w.cells()
    assert np.array_equal(bits(out[..., [0, 1, 2, 3, 7]]), bits(cells[..., [0, 1, 2, 3, 7]]))
    w.erode(200)
    out = w.cells()
    # tracks of the last cycle stay visible until the next one (world.h:56-61)
    assert out[..., 4].sum() > 0
    assert np.all(out[..., 4] >= 0)
    w.close()


def test_seed_and_cycle_resume_bit_exact(shy):
    """Checkpoint/resume: cells + (seed, cycle) restore the exact trajectory."""
    n, drops = 96, 700
    a = shy.World(n, n, seed=5)
    a.generate_fbm(5)
    a.erode_cycles(3, drops, graph=False)
    snap = a.cells()
    cyc = a.cycle
    a.erode_cycles(4, drops, graph=False)
    b = shy.World(n, n, seed=5)
    b.upload_cells(snap)
    b.set_seed(5, cyc)
    b.erode_cycles(4, drops, graph=True)
    for f in ("height", "discharge", "momentumx", "momentumy"):
        assert np.array_equal(bits(a.field(f)), bits(b.field(f))), f
    assert a.cycle == b.cycle == 7


def test_determinism_two_runs(shy):
    n, drops = 256, 4096
    out = []
    for _ in range(2):
        w = shy.World(n, n, seed=11)
        w.generate_fbm(11)
        w.erode_cycles(5, drops)
        out.append((w.field("height"), w.field("discharge"), w.stats()))
        w.close()
    assert np.array_equal(bits(out[0][0]), bits(out[1][0]))
    assert np.array_equal(bits(out[0][1]), bits(out[1][1]))
    assert out[0][2] == out[1][2]


def test_map_view_accessors(shy, orc):
    n = 32
    w = shy.World(n, n)
    w.generate_fbm(2)
    d = np.random.default_rng(0).uniform(0, 5, (n, n)).astype(np.float32)
    w.upload("discharge", d)
    h = w.field("height")
    assert w.map.oob((-1, 0)) and w.map.oob((0, n)) and not w.map.oob((n - 1, 0))
    assert w.map.height((3, 4)) == h[3, 4] and w.map.height((-1, 2)) == 0.0
    assert w.map.discharge((5, 6)) == orc.erf(np.float32(0.4) * d[5, 6])
    assert w.map.getCell((40, 2)) is None
    w.close()


def fnv_hash(*fields):
    v = np.stack([f.ravel() for f in fields], 1).astype(np.float32).view(np.uint8).ravel()
    x = 1469598103934665603
    for b in v.tolist():
        x ^= b
        x = (x * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return f"{x:016x}"


def node_major(a, tileres):
    """Cells in the order map.nodes then node.s visit them: node (i, j) at
    index i*mapsize + j (cellpool.h:421-426), x-major inside a node."""
    nx, ny = a.shape
    parts = [a[i:i + tileres, j:j + tileres].ravel()
             for i in range(0, nx, tileres) for j in range(0, ny, tileres)]
    return np.concatenate(parts)


def test_cpp_facade_demo_matches_python_path(shy):
    """include/shy_world.hpp (World::erode, Drop::descend, the node-level map
    chain map.get(p)->get(p), map.nodes / node.s iteration, getCell with
    implicit vec2 -> ivec2) through the C ABI gives the same bits as the
    Python host for the same loop."""
    import json
    import os
    import subprocess
    demo = os.path.join(os.path.dirname(shy.LIB_PATH), "world_demo")
    n, cycles, drops, seed = 64, 6, 512, 3
    # the demo reads getCell(p + vec2(dx, dy)) at p = (x + 0.5, y + 0.25): the
    # vec2 -> ivec2 conversion truncates toward zero as glm's does, so
    # x + 0.5 - 1 = -0.5 at x = 0 names cell 0 (a floor would leave the map)
    inb = sum(1 for x in range(n) for y in range(n) for dx in (-1, 0, 1) for dy in (-1, 0, 1)
              if (dx or dy) and 0 <= int(x + 0.5 + dx) < n and 0 <= int(y + 0.25 + dy) < n)
    for terrain, tileres in (("reference", 512), ("fbm", 32), ("reference", 24)):
        print("demo", terrain, tileres, flush=True)
        r = subprocess.run([demo, str(n), str(cycles), str(drops), str(seed), terrain,
                            str(tileres)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        out = json.loads(r.stdout)
        w = shy.World(n, n, seed=seed)
        if terrain == "fbm":
            w.generate_fbm(seed, 8)
        else:
            w.generate_reference(seed, tileres)
        for _ in range(cycles):
            w.erode(drops)
        h, d = w.field("height"), w.field("discharge")
        assert out["nodes"] == (-(-n // tileres)) ** 2
        assert out["hash"] == fnv_hash(h, d)
        assert out["hash_nodes"] == fnv_hash(node_major(h, tileres), node_major(d, tileres))
        assert out["hash_momentum"] == fnv_hash(w.field("momentumx"), w.field("momentumy"))
        assert out["neighbours"] == inb
        assert out["particle_steps"] == w.stats()["steps"]
        drop = w.drop((n // 2 + 0.5, n // 3 + 0.5))
        steps = 0
        while drop.descend():
            steps += 1
        assert out["drop_steps"] == steps
        w.close()


def test_drop_log_matches_the_cycle_ledger(shy):
    """shy_terrain_drop_log: one entry per drop of the last cycle, whose steps
    and end kinds add up to that cycle's ledger; empty before a cycle and
    after a parameter change."""
    n, drops = 128, 2000
    w = shy.World(n, n, seed=5)
    w.generate_fbm(5, 8)
    steps, ends = w.drop_log()
    assert len(steps) == 0
    w.erode_cycles(2, drops, graph=False)
    w.reset_stats()
    w.erode(drops)
    st = w.stats()
    steps, ends = w.drop_log()
    assert len(steps) == drops and len(ends) == drops
    assert int(steps.sum()) == st["steps"]
    for code, key in ((0, "ends_age"), (1, "ends_vol"), (2, "exits"), (3, "skipped")):
        assert int((ends == code).sum()) == st[key], key
    assert steps.max() <= 501 and (steps[ends == 3] == 0).all()
    w.params = shy.Params(evapRate=0.002)
    assert len(w.drop_log()[0]) == 0
    w.close()
