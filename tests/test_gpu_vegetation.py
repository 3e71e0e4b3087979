"""GPU: vegetation (vegetation.h:44-186) against the oracle's restatement of
the same generation semantics, coupled with erosion through the root
density (water.h:86-87)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_vegetation_and_erosion_coupled_bit_exact(shy, orc):
    """Alternate World::erode and Vegetation::grow, as the reference's frame
    loop does (SimpleHydrology.cpp:319-320): plants, root density, heights and
    flow equal the oracle bit for bit."""
    n, drops, seed, steps = 96, 1024, 11, 40
    # a gentle terrain: plants need normal.y >= 0.8 (a height step below
    # about 0.0094 per cell at mapscale 80), which a 96^2 map spanning [0, 1]
    # has almost nowhere
    h = (np.float32(0.3) + np.float32(0.2) * orc.fbm(seed, n, n)).astype(np.float32)
    z = np.zeros((n, n), np.float32)
    d, mx, my, root, erfd = z.copy(), z.copy(), z.copy(), z.copy(), z.copy()
    P = orc.Params.default()
    plants = np.zeros((0, 3), np.float32)
    w = shy.World(n, n, seed=seed)
    w.upload("height", h)
    w.vegetation_enable(1 << 16)
    for c in range(steps):
        print("cycle", c, flush=True)
        w.erode(drops)
        orc.wave_cycle(h, d, mx, my, root, erfd, P, seed, c, drops)
        w.vegetation_grow()
        plants, dropped = orc.veg_grow(h, d, root, plants, seed, c)
        g = w.plants()
        assert np.array_equal(bits(g), bits(plants)), (c, g.shape, plants.shape)
        assert np.array_equal(bits(w.field("rootdensity")), bits(root)), c
    assert np.array_equal(bits(w.field("height")), bits(h))
    assert np.array_equal(bits(w.field("discharge")), bits(d))
    assert plants.shape[0] > 0 and root.max() > 0
    w.close()


def test_vegetation_capacity_upload_and_state(shy, orc):
    n = 32
    w = shy.World(n, n, seed=2)
    h = np.full((n, n), 0.5, np.float32)
    w.upload("height", h)
    with pytest.raises(shy.ShyError) as e:
        w.vegetation_grow()
    assert e.value.code == shy.SHY_ERR_STATE
    w.vegetation_enable(capacity=3)
    pl = np.array([[1, 2, 0.5], [30, 31, 1.0]], np.float32)
    w.set_plants(pl)
    assert np.array_equal(w.plants(), pl)
    for _ in range(60):           # a flat dry map: spawns every call, capped at 3
        w.vegetation_grow()
    assert w.plants().shape[0] <= 3
    with pytest.raises(shy.ShyError):
        w.set_plants(np.array([[40, 1, 0.0]], np.float32))   # off the map
    w.close()
