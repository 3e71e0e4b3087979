"""GPU at the benchmark's full size: 1024x1024 fBm, 65536 drops per cycle.

Bit-exact against the oracle for the first cycles (the oracle needs a few
seconds per cycle at this size), then size-independent properties over a
longer run: the mass ledger, bounded terrain, and the step accounting.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, DROPS, SEED = 1024, 65536, 1


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_fullsize_two_cycles_bit_exact(shy, orc):
    w = shy.World(N, N, seed=SEED)
    w.generate_fbm(SEED, 8)
    h = orc.fbm(SEED, N, N)
    assert np.array_equal(bits(w.field("height")), bits(h))
    z = np.zeros_like(h)
    d, mx, my, root, erfd = z.copy(), z.copy(), z.copy(), z.copy(), z.copy()
    P = orc.Params.default()
    tot = {}
    for c in range(2):
        print("fullsize cycle", c, flush=True)
        w.erode(DROPS)
        s = orc.wave_cycle(h, d, mx, my, root, erfd, P, SEED, c, DROPS).as_dict()
        for k, v in s.items():
            tot[k] = tot.get(k, 0) + v
        for name, ref in (("height", h), ("discharge", d), ("momentumx", mx),
                          ("momentumy", my), ("erf_discharge", erfd)):
            assert np.array_equal(bits(w.field(name)), bits(ref)), (c, name)
    assert w.stats() == tot
    w.close()


def test_fullsize_ledger_and_stability(shy):
    w = shy.World(N, N, seed=SEED)
    w.generate_fbm(SEED, 8)
    s0 = w.field("height").astype(np.float64).sum()
    w.erode_cycles(60, DROPS, graph=True)
    h = w.field("height")
    st = w.stats()
    s1 = h.astype(np.float64).sum()
    ledger = (st["gain40"] - st["export40"]) * 2.0 ** -40
    print("ledger", s1 - s0, ledger, st, flush=True)
    assert abs((s1 - s0) - ledger) <= 1e-6 * s0
    assert abs((s1 - s0) - st["walk_dh"] * 2.0 ** -20) <= 1e-6 * s0
    assert np.isfinite(h).all() and -0.5 < h.min() and h.max() < 1.5
    assert st["walked"] + st["skipped"] == 60 * DROPS
    assert st["exits"] + st["ends_age"] + st["ends_vol"] == st["walked"]
    d = w.field("discharge")
    assert np.isfinite(d).all() and d.min() >= 0.0
    w.close()
