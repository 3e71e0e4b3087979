"""GPU: render-facing outputs and map dumps (map I/O of the reference's frame
loop, SimpleHydrology.cpp:341-354, and updatenode, cellpool.h:286-304)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def eroded_world(shy, n=96, seed=5, cycles=4, drops=2048):
    w = shy.World(n, n, seed=seed)
    w.generate_fbm(seed, 8)
    for _ in range(cycles):
        w.erode(drops)
    return w


def test_textures_match_the_frame_loop(shy, orc):
    w = eroded_world(shy)
    water = (0.1, 0.3, 0.7)
    print("textures", flush=True)
    dis, mom = w.textures(water)
    assert np.array_equal(bits(dis[..., 3]), bits(w.field("erf_discharge")))
    assert np.all(dis[..., :3] == np.float32(water))
    mx, my = w.field("momentumx"), w.field("momentumy")
    half, one = np.float32(0.5), np.float32(1.0)
    assert np.array_equal(bits(mom[..., 0]), bits(half * (one + orc.erf(mx))))
    assert np.array_equal(bits(mom[..., 1]), bits(half * (one + orc.erf(my))))
    assert np.all(mom[..., 2] == 0.5) and np.all(mom[..., 3] == 1.0)
    assert mom[..., :2].min() >= 0.0 and mom[..., :2].max() <= 1.0
    w.close()


def test_vertices_match_updatenode(shy, orc):
    w = eroded_world(shy, n=64, cycles=2)
    print("vertices", flush=True)
    v = w.vertices()
    h = w.field("height")
    n = h.shape[0]
    s = np.float32(80.0)
    x, y = np.meshgrid(np.arange(n, dtype=np.float32), np.arange(n, dtype=np.float32),
                       indexing="ij")
    P = np.stack([x, s * h, y], -1)
    hx = np.zeros_like(h)
    hx[:-1] = h[1:]                      # off the map reads 0
    hy = np.zeros_like(h)
    hy[:, :-1] = h[:, 1:]
    T = np.stack([np.ones_like(h), s * hx - s * h, np.zeros_like(h)], -1)
    B = np.stack([np.zeros_like(h), s * hy - s * h, np.ones_like(h)], -1)
    assert np.array_equal(bits(v[:, :, 0]), bits(P))
    # quad::_normal (cellpool.h:181-204) of every cell, from the oracle
    normals = np.stack([orc.normal(h, i, j) for i in range(n) for j in range(n)])
    assert np.array_equal(bits(v[:, :, 1]), bits(normals.reshape(n, n, 3)))
    assert np.array_equal(bits(w.normals()), bits(normals.reshape(n, n, 3)))
    assert np.array_equal(bits(v[:, :, 2]), bits(T))
    assert np.array_equal(bits(v[:, :, 3]), bits(B))
    w.close()


def test_save_load_continues_bit_for_bit(shy, tmp_path):
    n, drops = 80, 1500
    w = eroded_world(shy, n=n, cycles=3, drops=drops)
    p = shy.Params(evapRate=0.0015)
    w.params = p
    path = str(tmp_path / "w.shymap")
    print("save", flush=True)
    w.save(path)
    hd, cells = shy.read_map(path)
    assert (hd.nx, hd.ny, hd.cycle, hd.seed) == (n, n, 3, 5)
    assert hd.params.evapRate == np.float32(0.0015)
    assert np.array_equal(bits(cells), bits(w.cells()))
    w2 = shy.World(n, n, seed=123)
    w2.load(path)
    assert w2.cycle == 3 and w2.params.evapRate == np.float32(0.0015)
    for _ in range(3):
        w.erode(drops)
        w2.erode(drops)
    assert np.array_equal(bits(w.cells()), bits(w2.cells()))
    # a dump written by the Python writer loads the same way
    path2 = str(tmp_path / "py.shymap")
    shy.write_map(path2, cells, hd.params, seed=hd.seed, cycle=hd.cycle)
    w3 = shy.World(n, n)
    w3.load(path2)
    assert np.array_equal(bits(w3.cells()), bits(cells))
    # wrong size and foreign files are refused with a status, nothing loaded
    w4 = shy.World(n + 1, n)
    with pytest.raises(shy.ShyError) as e:
        w4.load(path)
    assert e.value.code == shy.SHY_ERR_ARG
    bad = tmp_path / "bad.shymap"
    bad.write_bytes(b"x" * 100)
    with pytest.raises(shy.ShyError) as e:
        w4.load(str(bad))
    assert e.value.code == shy.SHY_ERR_IO
    for ww in (w, w2, w3, w4):
        ww.close()


def test_multitile_pool_order_and_tiled_reference_world(shy, orc):
    """A 4x4 node grid (cellpool.h mapsize 4, tileres 64): the reference's
    world init over nodes is the seamless world at tileres 64, and the
    node-major pool order round-trips."""
    n, tr = 256, 64
    w = shy.World(n, n, seed=3)
    print("tiled reference world", flush=True)
    w.generate_reference(3, tr)
    assert np.array_equal(bits(w.field("height")), bits(orc.fnl_world(n, n, 3, tr)))
    for _ in range(3):
        w.erode(4096)
    pool = w.cells_tiled(tr)
    cells = w.cells()
    m = n // tr
    for i in range(m):
        for j in range(m):
            assert np.array_equal(pool[i * m + j], cells[i * tr:(i + 1) * tr, j * tr:(j + 1) * tr])
    w2 = shy.World(n, n)
    w2.upload_cells_tiled(pool, tr)
    assert np.array_equal(bits(w2.cells()), bits(cells))
    with pytest.raises(shy.ShyError) as e:
        w.cells_tiled(100)
    assert e.value.code == shy.SHY_ERR_ARG
    w.close()
    w2.close()
