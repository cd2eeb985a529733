"""Batched windows in the persistent kernel (a3d_batch_run mode 3): a resident grid runs
the time loop on the device, taking each sweep's tiles from a ticket counter and meeting
at a grid barrier after each sweep. Every window's record must equal its plain
single-shot run (mode 0) bit for bit, whatever the tile rows, the steps per launch (so
launch boundaries fall inside the loop or not), the z segments, the sweep that makes w
and the stress scheme; replays with other GPU work in between must too."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _setup(op, scheme, seed=41):
    from a3dfd import make_desc
    from a3dfd.models import small_model
    NX, NY, NZ = 64, 52, 30
    m = small_model(op, NX, NY, NZ, seed=seed)
    wins = [(0, 0, 0, 30, 28, 26, 12, 10, 2), (20, 10, 0, 44, 40, 30, 20, 21, 2), (40, 30, 2, 24, 22, 20, 9, 8, 3),
            (5, 25, 0, 36, 27, 24, 18, 13, 1), (33, 0, 4, 31, 26, 22, 15, 12, 2)]
    desc = make_desc(op, NX, NY, NZ, 10.0, 10.0, 10.0, 1, 1, 1, 70, 1e-3, 25.0, nb=8, scheme=scheme)
    return desc, wins, m


@pytest.mark.parametrize("op,scheme", [("iso", 0), ("iso", 3), ("vti", 0), ("vti", 4), ("tti", 1), ("tti", 4)])
def test_persistent_equals_sequential_runs(oracle, op, scheme):
    from a3dfd import Batch, Window
    desc, wins, m = _setup(op, scheme)
    cases = [(4, 100, 1, None, None), (4, 7, 4, None, None), (8, 16, 3, None, None), (4, 9, 16, (3, 2), None)]
    if op != "tti":
        cases += [(4, 11, 2, None, 1), (8, 100, 5, None, 0)]
    outs = []
    with Batch(desc, [Window(*w) for w in wins]) as b:
        b.set_model(**m)
        seq, _ = b.run(mode=0)
        for ty, steps, group, zseg, wz in cases:
            print(f"\n[gpu-persist] op={op} scheme={scheme} tile rows {ty} steps/launch {steps} tiles/ticket {group} "
                  f"zseg {zseg} w sweep {wz}", flush=True)
            b.set_persist(ty, steps, group)
            if zseg:
                b.set_zseg(*zseg)
            if wz is not None:
                b.set_w_sweep(wz)
            outs.append(b.run(mode=3)[0])
            outs.append(b.run(mode=1)[0])    # other GPU work between persistent runs
            outs.append(b.run(mode=3)[0])
    for w, win in enumerate(wins):
        assert np.abs(seq[w]).max() > 0
        for o in outs:
            assert np.array_equal(o[w], seq[w]), w
        ox, oy, oz, wx, wy, wz_, sx, sy, sz = win
        sub = {k: np.ascontiguousarray(a[oy:oy + wy, ox:ox + wx, oz:oz + wz_]) for k, a in m.items()}
        ref, _, _ = oracle.run(oracle.make_shot(op, wx, wy, wz_, 10.0, 10.0, 10.0, sx, sy, sz, 70, 1e-3, 25.0, nb=8),
                               32, **sub)
        assert rel_l2(seq[w], ref) <= 1e-5


def test_persistent_refuses_fp16_storage():
    from a3dfd import A3DError, Batch, Window
    desc, wins, m = _setup("vti", 3)
    desc.storage = 1
    with Batch(desc, [Window(*w) for w in wins]) as b:
        b.set_model(**m)
        with pytest.raises(A3DError):
            b.run(mode=3)
        with pytest.raises(A3DError):
            b.set_persist(6, 10, 4)
