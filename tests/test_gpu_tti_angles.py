"""GPU: TTI sweeps that evaluate the rotation terms from the tilt and azimuth
(a3d_set_tti_angles) instead of reading the four terms the reference's kernels compute per
point in double (Toa_gpu_3dtti_fd_1orderfunciton.cu:101-103,184-186). The FP32 sine and
cosine are pinned against the exact values; the sweeps' records and snapshots must stay
within the 1e-5 relative L2 parity bar of the FP32 oracle at non-zero tilt and azimuth
(every stress scheme, FP16 storage within its bounds), be the same bits for every launch
shape, and the batched modes must equal the windows' plain runs bit for bit."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

D = 10.0


def test_rot_sincos_accuracy():
    """Max. abs. error <= 8e-8 for |t| <= 7 (FP32 rounding alone is 3e-8), and the
    quadrant boundaries (multiples of pi/2) in both directions."""
    from a3dfd._lib import probe_rot_sincos
    edges = (np.float32(np.pi / 2) * np.arange(-4, 5, dtype=np.float32)[:, None]
             + np.float32(1e-6) * np.arange(-3, 4, dtype=np.float32)[None, :])
    t = np.concatenate([np.linspace(-7.0, 7.0, 1 << 20, dtype=np.float32), edges.ravel()]).astype(np.float32)
    s, c = probe_rot_sincos(t)
    es = np.abs(s.astype(np.float64) - np.sin(t.astype(np.float64))).max()
    ec = np.abs(c.astype(np.float64) - np.cos(t.astype(np.float64))).max()
    print(f"\n[gpu-tti-angles] sin max abs err {es:.3e}, cos {ec:.3e}")
    assert es <= 8e-8 and ec <= 8e-8
    assert np.all(np.abs(s) <= 1.0) and np.all(np.abs(c) <= 1.0)


def _run(args, nb, m, angles, launch=None, **kw):
    from a3dfd import Modeller, make_desc
    with Modeller(make_desc(*args, nb=nb, **kw)) as mdl:
        mdl.set_model(**m)
        mdl.set_tti_angles(angles)
        if launch:
            mdl.set_launch(*launch)
        out, snaps, _ = mdl.run(nsnap_max=2)
    return out, snaps


@pytest.mark.parametrize("scheme", [0, 1, 3, 4])
def test_tti_angles_within_parity_bar(oracle, scheme):
    """A multi-tile grid whose x-, y- and z-faces of the PML share tiles with interior cells."""
    from a3dfd.models import small_model
    nx, ny, nz, nb = 70, 41, 30, 12
    args = ("tti", nx, ny, nz, D, D, D, 36, 21, 4, 160, 1e-3, 25.0)
    m = small_model("tti", nx, ny, nz, seed=19)
    print(f"\n[gpu-tti-angles] scheme {scheme}", flush=True)
    g, gs = _run(args, nb, m, 1, scheme=scheme, snap_interval=159)
    t, ts = _run(args, nb, m, 0, scheme=scheme, snap_interval=159)
    r32, s32, _ = oracle.run(oracle.make_shot(*args, nb=nb, snap_interval=159), 32, nsnap_max=2, **m)
    e, es, et = rel_l2(g, r32), rel_l2(gs, s32), rel_l2(t, r32)
    print(f"[gpu-tti-angles] scheme {scheme}: angles vs FP32 oracle {e:.2e} (snapshot {es:.2e}); terms read {et:.2e}")
    assert np.abs(g).max() > 0
    assert e <= 1e-5 and es <= 1e-5
    assert not np.array_equal(g, t)   # the terms are really evaluated in the sweeps


def test_tti_angles_fp16_storage(oracle):
    """binary16 wavefield storage with the terms from the angles: within the FP16 bounds of
    the terms-read run against the FP64 oracle."""
    from a3dfd.models import small_model
    nx, ny, nz, nb = 24, 20, 22, 10
    args = ("tti", nx, ny, nz, D, D, D, 12, 10, 3, 150, 1e-3, 25.0)
    m = small_model("tti", nx, ny, nz, seed=13)
    g, gs = _run(args, nb, m, 1, storage=1, scheme=4, snap_interval=149)
    r64, s64, _ = oracle.run(oracle.make_shot(*args, nb=nb, snap_interval=149), 64, nsnap_max=2, **m)
    e, es = rel_l2(g, r64), rel_l2(gs, s64)
    print(f"\n[gpu-tti-angles] FP16 storage, angles: {e:.2e} (snapshot {es:.2e}) vs FP64")
    assert np.isfinite(g).all()
    assert e <= 6.2e-3 and es <= 7.6e-2


def test_tti_angles_launch_invariance():
    """Tile rows, y tiles and z segments: the same bits."""
    from a3dfd.models import small_model
    nx, ny, nz, nb = 45, 33, 31, 10
    args = ("tti", nx, ny, nz, D, D, D, 22, 16, 5, 130, 1e-3, 25.0)
    m = small_model("tti", nx, ny, nz, seed=23)
    outs = [_run(args, nb, m, 1, launch=L, scheme=4, snap_interval=65)
            for L in ((8, 1, 8, 1), (2, 5, 2, 5), (4, 3, 4, 3))]
    a, sa = outs[0]
    assert np.abs(a).max() > 0
    for b, sb in outs[1:]:
        assert np.array_equal(a, b) and np.array_equal(sa, sb)


def test_tti_angles_batched_windows(oracle):
    """Batched launches, the captured graph and the persistent kernel equal the windows'
    plain runs bit for bit; each window within the parity bar of the oracle on its sub-model."""
    from a3dfd import Batch, Window, make_desc
    from a3dfd.models import small_model
    NX, NY, NZ = 64, 52, 30
    m = small_model("tti", NX, NY, NZ, seed=37)
    wins = [(0, 0, 0, 30, 28, 26, 12, 10, 2), (20, 10, 0, 44, 40, 30, 20, 21, 2), (40, 30, 2, 24, 22, 20, 9, 8, 3)]
    desc = make_desc("tti", NX, NY, NZ, D, D, D, 1, 1, 1, 70, 1e-3, 25.0, nb=8, scheme=4)
    print("\n[gpu-tti-angles] batched", flush=True)
    with Batch(desc, [Window(*w) for w in wins]) as b:
        b.set_model(**m)
        b.set_tti_angles(1)
        seq, _ = b.run(mode=0)
        bat, _ = b.run(mode=1)
        gr, _ = b.run(mode=2)
        b.set_persist(4, 16, 1)
        per, _ = b.run(mode=3)
        b.set_tti_angles(0)
        terms, _ = b.run(mode=0)
    for w, win in enumerate(wins):
        assert np.array_equal(bat[w], seq[w]) and np.array_equal(gr[w], seq[w]) and np.array_equal(per[w], seq[w])
        assert not np.array_equal(terms[w], seq[w])
        ox, oy, oz, wx, wy, wz, sx, sy, sz = win
        sub = {k: np.ascontiguousarray(a[oy:oy + wy, ox:ox + wx, oz:oz + wz]) for k, a in m.items()}
        ref, _, _ = oracle.run(oracle.make_shot("tti", wx, wy, wz, D, D, D, sx, sy, sz, 70, 1e-3, 25.0, nb=8), 32, **sub)
        assert rel_l2(seq[w], ref) <= 1e-5


def test_tti_angles_refused_elsewhere():
    from a3dfd import A3DError, Modeller, make_desc
    for op, kw in (("vti", {}), ("tti", dict(storage=2))):
        desc = make_desc(op, 30, 26, 24, D, D, D, 12, 10, 3, 20, 1e-3, 25.0, nb=10, **kw)
        with Modeller(desc) as mdl:
            with pytest.raises(A3DError):
                mdl.set_tti_angles(1)
            mdl.set_tti_angles(0)
