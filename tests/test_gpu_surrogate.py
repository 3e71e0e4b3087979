"""GPU: what the trained surrogate gains, and the warm-start gate.

The network (simplehydrology_amd/weights, trained by train_surrogate on 80
simulator-made 2048x2048 flow maps, 2 more held out) maps the state after 10 cold cycles to
the long-run mean discharge/momentum.  These tests run on a held-out terrain
of the training statistics (fBm base frequency 1/2048) at 1024x1024.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ts(shy):
    from simplehydrology_amd import train_surrogate
    return train_surrogate


@pytest.fixture(scope="module")
def trained(ts):
    from simplehydrology_amd import surrogate
    assert os.path.exists(ts.WEIGHTS), "trained weights missing"
    return ts.load_weights(surrogate.UNet(seed=0), ts.WEIGHTS)


N, SEED, C0 = 1024, 31337, 10


def cold_run(ts, n=N, seed=SEED, c0=C0, t0=100, t1=200):
    """state after c0 cycles, and the mean flow over [t0, t1)"""
    w = ts.flow_world(n, seed, base_freq=1.0 / 2048)
    drops = n * n // 16
    rec = ts.Recorder(w)
    w.erode_cycles(c0, drops, graph=True)
    state = rec.grab().clone()
    cells = w.cells()
    w.erode_cycles(t0 - c0, drops, graph=True)
    acc = torch.zeros(n, n, 3, device="cuda:0")
    for _ in range(t1 - t0):
        w.erode_cycles(1, drops, graph=False)
        acc += rec.grab()[..., 1:4]
    w.close()
    return state, cells, acc / (t1 - t0)


def test_trained_network_beats_the_state_it_starts_from(ts, trained):
    """Discharge (as erf(0.4 d), node::discharge) predicted from the state
    after 10 cycles is closer to the long-run mean than that state itself,
    and an untrained network is far worse than both."""
    from simplehydrology_amd import surrogate as sg
    print("cold run", N, SEED, flush=True)
    state, _, target = cold_run(ts)
    T = ts.erf_discharge(target[..., 0])
    d, _, _ = sg.predict_tiled(trained, *(state[..., k] for k in range(4)))
    e_net = ts.l1_rel(ts.erf_discharge(d), T)
    e_in = ts.l1_rel(ts.erf_discharge(state[..., 1]), T)
    d0, _, _ = sg.predict_tiled(sg.UNet(seed=3), *(state[..., k] for k in range(4)))
    e_raw = ts.l1_rel(ts.erf_discharge(d0), T)
    print(f"L1 network {e_net:.4f}  input state {e_in:.4f}  untrained {e_raw:.4f}", flush=True)
    assert e_net < 0.95 * e_in
    assert e_raw > 1.5 * e_in


def test_gate_accepts_trained_and_rejects_untrained_with_restore(ts, trained):
    """At a gate of 0.35 (the plain method's own spread against a smooth
    field) the trained network passes and an untrained one fails; at the
    stated 2% gate the trained one is rejected too.  A rejection restores the
    plain run's state bit for bit."""
    from simplehydrology_amd import surrogate as sg
    drops = N * N // 16
    assert ts.GATE_L1 == 0.02
    for net, gate, expect in ((trained, 0.35, True), (sg.UNet(seed=3), 0.35, False),
                              (trained, ts.GATE_L1, False)):
        print("warm start", "trained" if net is trained else "untrained", gate, flush=True)
        w = ts.flow_world(N, SEED, base_freq=1.0 / 2048)
        w.erode_cycles(C0, drops, graph=True)
        before, cyc = w.cells(), w.cycle
        r = ts.warm_start(net, w, drops, SEED, K=20, W=10, gate_l1=gate)
        print({k: v for k, v in r.items() if k != "window_mean"}, flush=True)
        assert r["accepted"] == expect
        if expect:
            assert w.cycle == cyc + 20
        else:
            # rejected: the plain run's state before the import is back
            assert w.cycle == cyc
            assert np.array_equal(w.cells().view(np.uint32), before.view(np.uint32))
        w.close()


def test_terrain_training_data_through_the_ensemble(ts):
    """The C4-scale data path at a small size: terrains generated in one
    ensemble hold, per terrain, what a single World of the same seed and
    terrain number computes (terrain 0 checked against a World), and a short
    training run on them lowers the loss."""
    from simplehydrology_amd import surrogate as sg
    import simplehydrology_amd as shy
    n, m, c0, t0, t1, seed0 = 128, 6, 5, 15, 25, 4242
    print("terrain data", n, m, flush=True)
    feats, targets, info = ts.generate_terrains(m, n, c0, t0, t1, seed0=seed0, per_batch=4,
                                                log=lambda *_: None)
    assert feats.shape == (m, n, n, 8) and targets.shape == (m, n, n, 3)
    assert info["particle_steps"] > 0
    # terrain 0 of the first ensemble equals a World of its seed (terrain id 0)
    p = shy.Params(depositionRate=0.0)
    w = shy.World(n, n, seed=seed0, params=p)
    w.generate_fbm(seed0, 8)
    rec = ts.Recorder(w)
    w.erode_cycles(c0, n * n // 16, graph=False)
    st = rec.grab()
    f = sg.window_batch(sg.features(st[..., 0], st[..., 1], st[..., 2], st[..., 3]), [(0, 0)], n)
    assert torch.equal(f[0].view(torch.int16), feats[0].view(torch.int16))
    w.erode_cycles(t0 - c0, n * n // 16, graph=False)
    acc = torch.zeros(n, n, 3, device="cuda:0")
    for _ in range(t1 - t0):
        w.erode_cycles(1, n * n // 16, graph=False)
        acc += rec.grab()[..., 1:4]
    acc /= (t1 - t0)
    want = sg.encode_flow(acc[..., 0], acc[..., 1], acc[..., 2]).half()
    assert torch.equal(want.view(torch.int16), targets[0].view(torch.int16))
    w.close()
    net = sg.UNet(seed=0)
    hist, _ = ts.train(net, feats, targets, steps=40, batch=4, lr=2e-3, log=lambda *_: None,
                       log_every=10, whole_maps=True)
    print("losses", [h["loss"] for h in hist], flush=True)
    assert hist[-1]["loss"] < hist[0]["loss"]
