"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars:
* wave cycles, fBm inputs and erf fields: bit-exact (integer commits, same
  float op order on both sides);
* single-drop trajectories on a frozen map: bit-exact against the oracle's
  frozen trace with the project erf, and within 1e-5 relative of the
  reference ``Drop::descend`` restatement that uses glibc ``erff``.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P_FIELDS = ("height", "discharge", "momentumx", "momentumy")


def say(*a):
    print(*a, flush=True)


def random_fields(rng, nx, ny):
    d = rng.uniform(0, 12, (nx, ny)).astype(np.float32)
    mx = rng.normal(0, 2, (nx, ny)).astype(np.float32)
    my = rng.normal(0, 2, (nx, ny)).astype(np.float32)
    return d, mx, my


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_fbm_bit_exact(shy, orc):
    for n, seed in ((64, 3), (256, 1), (384, 9)):
        say("fbm", n, seed)
        w = shy.World(n, n, seed=seed)
        w.generate_fbm(seed, octaves=8)
        g = w.field("height")
        o = orc.fbm(seed, n, n)
        assert np.array_equal(bits(g), bits(o)), (n, seed, np.abs(g - o).max())
        assert g.min() == 0.0 and g.max() == 1.0
        w.close()


def test_reference_terrain_bit_exact(shy, orc):
    """quad::map::init's OpenSimplex2 terrain on the device against the oracle
    (itself pinned against the reference's FastNoiseLite build)."""
    for nx, ny, seed, tr in ((512, 512, 1, 512), (300, 200, 4242, 512), (64, 96, -3, 64)):
        say("reference terrain", nx, ny, seed, tr)
        w = shy.World(nx, ny, seed=1)
        w.generate_reference(seed, tr)
        g = w.field("height")
        o = orc.fnl_world(nx, ny, seed, tr)
        assert np.array_equal(bits(g), bits(o)), (nx, ny, seed, np.abs(g - o).max())
        w.close()


def test_erf_field_bit_exact(shy, orc):
    rng = np.random.default_rng(5)
    w = shy.World(128, 96)
    d = np.concatenate([rng.uniform(0, 12, 128 * 48), rng.exponential(2.0, 128 * 48)])
    d = d.astype(np.float32).reshape(128, 96)
    d[0, :8] = [0.0, 1e-30, 1e-7, 0.5 / 0.4, 2.0 / 0.4, 4.0 / 0.4, 9.99, 1e6]
    say("erf field")
    w.upload("discharge", d)
    g = w.field("erf_discharge")
    o = orc.erf(np.float32(0.4) * d)
    assert np.array_equal(bits(g), bits(o))
    w.close()


@pytest.mark.parametrize("n", [48, 128])
def test_trace_frozen(shy, orc, n):
    rng = np.random.default_rng(n)
    h = orc.fbm(11, n, n)
    d, mx, my = random_fields(rng, n, n)
    root = np.zeros((n, n), np.float32)
    w = shy.World(n, n)
    for f, a in zip(P_FIELDS, (h, d, mx, my)):
        w.upload(f, a)
    erfd = w.field("erf_discharge")
    P = orc.Params.default()
    starts = [(rng.integers(0, n), rng.integers(0, n)) for _ in range(24)] + [(0, 0), (n - 1, n - 1)]
    for (x, y) in starts:
        say("trace", n, x, y)
        g = w.trace(float(x), float(y), 700)
        o, _ = orc.trace_frozen(h, d, mx, my, root, erfd, P, float(x), float(y), 700,
                                use_ref_erf=False)
        assert len(g) == len(o)
        for f in ("age", "status"):
            assert np.array_equal(g[f], o[f])
        for f in ("pos", "speed", "volume", "sediment", "erosion", "h2"):
            assert np.array_equal(bits(g[f]), bits(o[f])), (x, y, f)
        r, _ = orc.trace_frozen(h, d, mx, my, root, erfd, P, float(x), float(y), 700,
                                use_ref_erf=True)
        assert len(r) == len(g)
        for f in ("pos", "speed", "volume"):
            assert np.array_equal(bits(g[f]), bits(r[f]))
        np.testing.assert_allclose(g["sediment"], r["sediment"], rtol=1e-5, atol=1e-12)
    w.close()


def run_oracle_cycles(orc, h, d, mx, my, root, erfd, seed, c0, cycles, drops, P):
    st = []
    for c in range(c0, c0 + cycles):
        st.append(orc.wave_cycle(h, d, mx, my, root, erfd, P, seed, c, drops).as_dict())
    return st


def assert_world_equals(w, h, d, mx, my, erfd, tag):
    for name, ref in (("height", h), ("discharge", d), ("momentumx", mx), ("momentumy", my),
                      ("erf_discharge", erfd)):
        g = w.field(name)
        if not np.array_equal(bits(g), bits(ref)):
            bad = np.argwhere(bits(g) != bits(ref))
            raise AssertionError(f"{tag}: field {name} differs at {len(bad)} cells, first "
                                 f"{bad[:3].tolist()} gpu={g[tuple(bad[0])]} orc={ref[tuple(bad[0])]}")


@pytest.mark.parametrize("n,drops,cycles", [(64, 256, 6), (256, 1024, 4), (200, 3000, 3)])
def test_wave_cycles_bit_exact(shy, orc, n, drops, cycles):
    seed = 1
    h = orc.fbm(seed, n, n)
    z = np.zeros((n, n), np.float32)
    d, mx, my, root, erfd = z.copy(), z.copy(), z.copy(), z.copy(), z.copy()
    P = orc.Params.default()
    w = shy.World(n, n, seed=seed)
    w.upload("height", h)
    for c in range(cycles):
        say("wave", n, drops, "cycle", c)
        w.erode(drops)
        ost = orc.wave_cycle(h, d, mx, my, root, erfd, P, seed, c, drops).as_dict()
        assert_world_equals(w, h, d, mx, my, erfd, f"n={n} cycle={c}")
    gst = w.stats()
    assert gst["steps"] > 0
    w.close()


@pytest.mark.parametrize("kw", [
    # volumes run down to 1e-37 (below 2^-60 from step 187 on)
    dict(minVol=1e-37, evapRate=0.2, depositionRate=0.0),
    dict(gravity=1e-25),               # gravity terms below 2^-60
    dict(evapRate=-0.1),               # volumes grow past 2^60
])
def test_wave_cycles_extreme_parameters(shy, orc, kw):
    """The walk divides the gravity term by the volume through the volume's
    reciprocal; where the operands leave the range in which that is exact,
    it divides plainly.  Parameters that drive them there still give the
    oracle's bits (the oracle divides plainly everywhere).  The tiny-volume
    case runs without deposition: the sediment a drop carries scales as
    1/volume (water.h:135), so at volumes of 1e-37 it would exceed what a
    commit can hold (2^30 units of 2^-20) by orders of magnitude."""
    n, drops, seed = 64, 256, 3
    h = orc.fbm(seed, n, n)
    z = np.zeros((n, n), np.float32)
    d, mx, my, root, erfd = z.copy(), z.copy(), z.copy(), z.copy(), z.copy()
    P = orc.Params.default(**kw)
    w = shy.World(n, n, seed=seed)
    w.params = shy.Params(**kw)
    w.upload("height", h)
    for c in range(2):
        say("extreme parameters", kw, "cycle", c)
        w.erode(drops)
        orc.wave_cycle(h, d, mx, my, root, erfd, P, seed, c, drops)
        assert_world_equals(w, h, d, mx, my, erfd, f"{kw} cycle={c}")
    w.close()


def test_tracks_bit_exact(shy, orc):
    """The cycle's discharge/momentum tracks (water.h:115-117), read back
    before the next cycle, equal the oracle's integer track sums; on an
    inclined plane with one drop they hold the closed form: each visited
    cell the pre-evaporation volume and volume x the downhill step."""
    import math
    cases = [("fbm", 128, 2048, 5, None), ("plane", 64, 1, 41, (2 / 1024, -3 / 1024))]
    for name, n, drops, seed, ab in cases:
        say("tracks", name, n, drops)
        if ab is None:
            h = orc.fbm(seed, n, n)
            P = orc.Params.default()
        else:
            x, y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
            h = (0.5 + ab[0] * x + ab[1] * y).astype(np.float32)
            P = orc.Params.default(depositionRate=0.0)
        z = np.zeros((n, n), np.float32)
        d, mx, my, root, erfd = z.copy(), z.copy(), z.copy(), z.copy(), z.copy()
        w = shy.World(n, n, seed=seed)
        w.params = shy.Params(depositionRate=P.depositionRate)
        w.upload("height", h)
        for c in range(2):
            w.erode(drops)
            _, trk = orc.wave_cycle(h, d, mx, my, root, erfd, P, seed, c, drops,
                                    want_tracks=True)
            for k, f in enumerate(("discharge_track", "momentumx_track", "momentumy_track")):
                want = (trk[k].astype(np.float64) * 2.0 ** -24).astype(np.float32)
                assert np.array_equal(bits(w.field(f)), bits(want)), (name, c, f)
        if ab is not None:
            u = -np.array(ab) / math.hypot(*ab)
            t = w.field("discharge_track")
            m = t > 0
            tx, ty = w.field("momentumx_track")[m], w.field("momentumy_track")[m]
            assert m.sum() > 10
            assert np.all(tx * u[0] + ty * u[1] > 0)
            assert np.allclose(np.hypot(tx, ty), math.sqrt(2) * t[m], rtol=1e-6)
        w.close()


def test_wave_stats_equal(shy, orc):
    n, drops, seed = 128, 2048, 7
    h = orc.fbm(seed, n, n)
    z = np.zeros((n, n), np.float32)
    d, mx, my, root, erfd = z.copy(), z.copy(), z.copy(), z.copy(), z.copy()
    P = orc.Params.default()
    w = shy.World(n, n, seed=seed)
    w.upload("height", h)
    tot = {}
    for c in range(3):
        say("stats cycle", c)
        w.erode(drops)
        s = orc.wave_cycle(h, d, mx, my, root, erfd, P, seed, c, drops).as_dict()
        for k, v in s.items():
            tot[k] = tot.get(k, 0) + v
    assert w.stats() == tot
    w.close()


def test_graph_replay_equals_eager(shy):
    n, drops = 192, 4096
    ws = []
    for graph in (False, True):
        w = shy.World(n, n, seed=3)
        w.generate_fbm(3)
        say("graph", graph)
        w.erode_cycles(7, drops, graph=graph)
        ws.append(w)
    for f in ("height", "discharge", "momentumx", "momentumy", "erf_discharge",
              "discharge_track"):
        assert np.array_equal(bits(ws[0].field(f)), bits(ws[1].field(f))), f
    assert ws[0].stats() == ws[1].stats()
    assert ws[0].cycle == ws[1].cycle == 7


@pytest.mark.parametrize("block", [64, 512, 1024])
def test_walk_block_invariance(shy, block):
    """Commits are order-free: the launch shape cannot change a bit."""
    n, drops = 160, 5000
    out = []
    for bs in (256, block):
        w = shy.World(n, n, seed=4)
        w.generate_fbm(4)
        w.set_walk_block(bs)
        say("block", bs)
        for _ in range(3):
            w.erode(drops)
        out.append((w.field("height"), w.field("discharge"), w.stats()))
    assert np.array_equal(bits(out[0][0]), bits(out[1][0]))
    assert np.array_equal(bits(out[0][1]), bits(out[1][1]))
    assert out[0][2] == out[1][2]


def test_cascade_live_matches_reference(shy, orc):
    rng = np.random.default_rng(2)
    n = 24
    h = rng.uniform(0, 1, (n, n)).astype(np.float32)
    h[5, 5] = 3.0
    h[0, 0] = 2.0
    P = orc.Params.default()
    w = shy.World(n, n)
    w.upload("height", h)
    cells = orc.to_aos(h)
    for (x, y) in [(5.2, 5.7), (0.0, 0.0), (n - 1, 3), (-0.5, 4.0), (12.9, n - 0.1)]:
        say("cascade", x, y)
        w.cascade((x, y))
        orc.seq_cascade(cells, P, x, y)
        assert np.array_equal(bits(w.field("height")), bits(cells[..., 0]))
    w.close()


def test_descend_live_matches_reference(shy, orc):
    rng = np.random.default_rng(8)
    n = 64
    h = orc.fbm(5, n, n)
    d, mx, my = random_fields(rng, n, n)
    P = orc.Params.default()
    w = shy.World(n, n)
    for f, a in zip(P_FIELDS, (h, d, mx, my)):
        w.upload(f, a)
    cells = orc.to_aos(h, d, mx, my)
    orc.set_seq_erf(True)  # the device evaluates erf(0.4 d) with the project erf
    try:
        for start in [(10, 20), (40, 3)]:
            say("descend", start)
            drop = w.drop(start)
            ref = dict(age=0, pos=start, speed=(0.0, 0.0), volume=1.0, sediment=0.0)
            for k in range(800):
                a = drop.descend()
                b = orc.seq_descend(cells, P, ref)
                assert a == b, k
                assert np.array_equal(bits(np.float32(drop.pos)), bits(ref["pos"])), k
                assert np.array_equal(bits(np.float32(drop.speed)), bits(ref["speed"])), k
                assert np.float32(drop.volume) == np.float32(ref["volume"])
                assert np.float32(drop.sediment) == np.float32(ref["sediment"])
                if not a:
                    break
            assert np.array_equal(bits(w.field("height")), bits(cells[..., 0]))
    finally:
        orc.set_seq_erf(False)
    w.close()
