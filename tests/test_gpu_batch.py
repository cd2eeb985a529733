"""Batched uneven windows on the GPU: every window of a batch equals its own
single-shot run and the oracle on the window's sub-model, in all three modes
(sequential plain runs, batched launches, one captured hipGraph); the graph is
replayed several times with eager work between replays."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _setup(op, seed=31):
    from a3dfd import make_desc
    from a3dfd.models import small_model
    NX, NY, NZ = 64, 52, 30
    m = small_model(op, NX, NY, NZ, seed=seed)
    wins = [(0, 0, 0, 30, 28, 26, 12, 10, 2), (20, 10, 0, 44, 40, 30, 20, 21, 2), (40, 30, 2, 24, 22, 20, 9, 8, 3),
            (5, 25, 0, 36, 27, 24, 18, 13, 1), (33, 0, 4, 31, 26, 22, 15, 12, 2)]
    desc = make_desc(op, NX, NY, NZ, 10.0, 10.0, 10.0, 1, 1, 1, 70, 1e-3, 25.0, nb=8)
    return desc, wins, m


@pytest.mark.parametrize("op", ["iso", "vti", "tti"])
def test_batch_modes_equal_single_runs_and_oracle(oracle, op):
    from a3dfd import Batch, Window
    desc, wins, m = _setup(op)
    print(f"\n[gpu-batch] op={op} {len(wins)} windows", flush=True)
    with Batch(desc, [Window(*w) for w in wins]) as b:
        b.set_model(**m)
        ncls = b.tune(steps=3)
        seq, _ = b.run(mode=0)
        bat, _ = b.run(mode=1)
        g1, _ = b.run(mode=2)
        g2, _ = b.run(mode=2)          # replay of the same graph
        bat2, _ = b.run(mode=1)        # eager work between replays
        g3, _ = b.run(mode=2)
        info = b.info()
        b.set_zseg(3, 2)               # other z segments: same records
        bz, _ = b.run(mode=1)
        gz, _ = b.run(mode=2)
        extra = [bz, gz]
        if op in ("iso", "vti"):
            for wzm in (1, 0):         # w updated by the stress sweep, then by the velocity sweep
                b.set_w_sweep(wzm)
                extra += [b.run(mode=1)[0], b.run(mode=2)[0]]
    print(f"[gpu-batch] {op}: {ncls} size classes, tile rows {info['tile_rows']}, z segments {info['zseg']}")
    for w, win in enumerate(wins):
        ox, oy, oz, wx, wy, wz, sx, sy, sz = win
        for other in [bat[w], g1[w], g2[w], bat2[w], g3[w]] + [x[w] for x in extra]:
            assert np.array_equal(other, seq[w])
        sub = {k: np.ascontiguousarray(a[oy:oy + wy, ox:ox + wx, oz:oz + wz]) for k, a in m.items()}
        ref, _, _ = oracle.run(oracle.make_shot(op, wx, wy, wz, 10.0, 10.0, 10.0, sx, sy, sz, 70, 1e-3, 25.0, nb=8),
                               32, **sub)
        assert np.abs(ref).max() > 0
        assert rel_l2(seq[w], ref) <= 1e-5
        assert np.array_equal(seq[w], ref)


def test_chunked_graphs_equal_single_graph():
    from a3dfd import Batch, Window
    desc, wins, m = _setup("tti", seed=33)
    with Batch(desc, [Window(*w) for w in wins]) as b:
        b.set_model(**m)
        one, _ = b.run(mode=2)
        b.set_graph_chunk(16)            # 70 steps -> 5 graphs
        chunked, _ = b.run(mode=2)
        again, _ = b.run(mode=2)
    for w in range(len(wins)):
        assert np.array_equal(chunked[w], one[w]) and np.array_equal(again[w], one[w])


@pytest.mark.parametrize("scheme", [0, 3])
def test_batch_with_learned_strip_equals_single_runs(oracle, scheme):
    """The learned strip on every window of a VTI batch (nb = 10, a network with an open
    damping head): batched launches and captured graphs give bitwise the sequential
    per-window runs, and a window equals the oracle's strip restatement on its sub-model."""
    from a3dfd import STRIP_NB, Batch, Window, make_desc
    from a3dfd.models import small_model
    NX, NY, NZ, nt = 64, 52, 30, 90
    m = small_model("vti", NX, NY, NZ, seed=41)
    wins = [(0, 0, 0, 30, 28, 26, 12, 10, 2), (20, 10, 0, 44, 40, 30, 20, 21, 2), (33, 0, 4, 31, 26, 22, 15, 12, 2)]
    desc = make_desc("vti", NX, NY, NZ, 10.0, 10.0, 10.0, 1, 1, 1, nt, 1e-3, 25.0, nb=STRIP_NB, scheme=scheme)
    from a3dfd import strip
    rng = np.random.default_rng(5)
    net = strip.init_params(5, out_bias=0.0)
    net[-130:] += rng.normal(0, 0.2, 130).astype(np.float32)
    with Batch(desc, [Window(*w) for w in wins]) as b:
        b.set_model(**m)
        b.set_strip_net(net, amp=10.0)
        b.tune(steps=3)
        seq, _ = b.run(mode=0)
        bat, _ = b.run(mode=1)
        g1, _ = b.run(mode=2)
        g2, _ = b.run(mode=2)
        b.set_w_sweep(1)               # w updated by the stress sweep
        bw, _ = b.run(mode=1)
        gw, _ = b.run(mode=2)
    for w in range(len(wins)):
        assert np.abs(seq[w]).max() > 0
        for other in (bat[w], g1[w], g2[w], bw[w], gw[w]):
            assert np.array_equal(other, seq[w])
    # the first window against the oracle's strip restatement on its sub-model (the
    # matrix-core network against the oracle's binary16 form: within 1e-5)
    ox, oy, oz, wx, wy, wz, sx, sy, sz = wins[0]
    sub = {k: np.ascontiguousarray(a[oy:oy + wy, ox:ox + wx, oz:oz + wz]) for k, a in m.items()}
    oracle.set_strip_net(net)
    try:
        ref, _, _ = oracle.run(oracle.make_shot("vti", wx, wy, wz, 10.0, 10.0, 10.0, sx, sy, sz, nt, 1e-3, 25.0,
                                                nb=STRIP_NB), 32, **sub)
    finally:
        oracle.set_strip_net(None)
    e = rel_l2(seq[0], ref)
    print(f"\n[gpu-batch] strip scheme {scheme}: window 0 rel L2 vs oracle {e:.2e}")
    assert e <= 1e-5


def test_cfg4_size_windows(oracle):
    """Four windows of cfg4's size range (160-320 x 160-320 x 200-300 cells, npml 40) cut
    from the 600x600x300 TTI model, scheme 3 as the batched bench runs: sequential plain
    runs, batched launches and captured graphs bitwise equal after the batch tuner picked
    tile rows and z segments at these sizes, and the second window against the oracle on
    its sub-model (40 steps)."""
    from a3dfd import Batch, Window, make_desc, models
    c = models.cfg4_model()
    wins = models.uneven_windows(4, c["nx"], c["ny"], c["nz"], seed=44)
    nt = 40
    m = {k: c[k] for k in ("vel", "eps", "delta", "theta", "phi")}
    desc = make_desc("tti", c["nx"], c["ny"], c["nz"], c["dx"], c["dy"], c["dz"], 1, 1, 1, nt, c["dt"], c["fm"],
                     scheme=3)
    print(f"\n[gpu-batch] cfg4-size windows {wins}", flush=True)
    with Batch(desc, [Window(*w) for w in wins]) as b:
        b.set_model(**m)
        ncls = b.tune(steps=6)
        info = b.info()
        seq, _ = b.run(mode=0)
        bat, _ = b.run(mode=1)
        g1, _ = b.run(mode=2)
        g2, _ = b.run(mode=2)
    print(f"[gpu-batch] {ncls} size classes, tile rows {info['tile_rows']}, z segments {info['zseg']}", flush=True)
    for w in range(len(wins)):
        assert np.abs(seq[w]).max() > 0
        for other in (bat[w], g1[w], g2[w]):
            assert np.array_equal(other, seq[w]), w
    ox, oy, oz, wx, wy, wz, sx, sy, sz = wins[1]
    sub = {k: np.ascontiguousarray(a[oy:oy + wy, ox:ox + wx, oz:oz + wz]) for k, a in m.items()}
    ref, _, _ = oracle.run(oracle.make_shot("tti", wx, wy, wz, c["dx"], c["dy"], c["dz"], sx, sy, sz, nt, c["dt"],
                                            c["fm"]), 32, **sub)
    e = rel_l2(seq[1], ref)
    print(f"[gpu-batch] window 1 ({wx}x{wy}x{wz}) vs oracle: {e:.2e}", flush=True)
    assert np.abs(ref).max() > 0 and e <= 1e-5
