"""GPU: the z velocity w updated by the stress sweep (a3d_set_w_sweep(ctx, 1); ISO and
VTI) instead of the velocity sweep. The stress sweep makes w from its own column's old
P (ISO) or Q (VTI) with the velocity sweep's arithmetic, so the records and snapshots
must be bitwise equal to the reference two-sweep order, and bitwise equal to the FP32
oracle where the two-sweep run is."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

D = 10.0


def _run(op, nx, ny, nz, nt, src, wz, launch, seed=41, nb=10, nsnap=3, strip=None, **kw):
    from a3dfd import Modeller, make_desc
    from a3dfd.models import small_model
    m = small_model(op, nx, ny, nz, seed=seed)
    desc = make_desc(op, nx, ny, nz, D, D, D, *src, nt, 1e-3, 25.0, nb=nb, **kw)
    print(f"\n[gpu-wsweep] op={op} grid={nx}x{ny}x{nz} src={src} wz={wz} launch={launch} kw={kw}", flush=True)
    with Modeller(desc) as mdl:
        mdl.set_model(**m)
        mdl.set_launch(*launch)
        mdl.set_w_sweep(wz)
        if strip is not None:
            mdl.set_strip_net(strip)
        assert mdl.w_sweep() == wz
        out, snaps, _ = mdl.run(nsnap_max=nsnap)
    return m, desc, out, snaps


@pytest.mark.parametrize("op", ["iso", "vti"])
@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("storage", [0, 1, 2])
def test_w_sweep_bitwise(oracle, op, scheme, storage):
    """Both schemes and all three precision instances; several x and y tiles, a partial
    last x tile; the velocity sweep in 3 z segments."""
    src = (30, 17, 5)
    runs = [_run(op, 70, 37, 28, 150, src, wz, (4, 3, 4, 1), scheme=scheme, storage=storage, snap_interval=50)
            for wz in (0, 1)]
    (m, desc, a, sa), (_, _, b, sb) = runs
    print(f"[gpu-wsweep] {op} scheme {scheme} storage {storage}: max|d| {np.abs(a - b).max():.3e}")
    assert np.abs(a).max() > 0
    assert np.array_equal(a, b) and np.array_equal(sa, sb)
    if storage == 0 and scheme == 0:
        args = (op, 70, 37, 28, D, D, D, *src, 150, 1e-3, 25.0)
        ref, rsn, _ = oracle.run(oracle.make_shot(*args, nb=10, snap_interval=50), 32, nsnap_max=3, **m)
        assert np.array_equal(b, ref) and np.array_equal(sb, rsn)


@pytest.mark.parametrize("op", ["iso", "vti"])
@pytest.mark.parametrize("src", [(1, 1, 1), (26, 22, 24), (1, 22, 12), (13, 11, 2)])
def test_w_sweep_sources_near_faces(op, src):
    """Sources on the faces and next to the top: the stress sweep's window adds the source
    to P / Q on the source column at the planes the velocity sweep would, including the
    warm-up planes of the first segment."""
    runs = [_run(op, 26, 22, 24, 140, src, wz, ty_launch, snap_interval=139, nsnap=2)
            for wz, ty_launch in ((0, (8, 2, 8, 2)), (1, (2, 1, 8, 1)))]
    (_, _, a, sa), (_, _, b, sb) = runs
    assert np.abs(a).max() > 0
    assert np.array_equal(a, b) and np.array_equal(sa, sb), src


@pytest.mark.parametrize("ty", [2, 4, 8])
def test_w_sweep_tile_rows(ty):
    """Every tile-row count of the stress sweep that updates w; the requested stress z
    segments are replaced by one."""
    runs = [_run("vti", 45, 29, 33, 120, (20, 14, 6), wz, (ty, 1, ty, 5), scheme=1, snap_interval=119, nsnap=2)
            for wz in (0, 1)]
    (_, _, a, sa), (_, _, b, sb) = runs
    assert np.array_equal(a, b) and np.array_equal(sa, sb)


def test_w_sweep_with_learned_strip():
    """The learned strip corrects u, v, w after both sweeps: unchanged by which sweep made w."""
    from a3dfd import STRIP_NB, strip
    params, amp = strip.load_params()
    runs = [_run("vti", 32, 30, 28, 160, (16, 15, 6), wz, (4, 2, 4, 1), nb=STRIP_NB, strip=params,
                 snap_interval=80, nsnap=2) for wz in (0, 1)]
    (_, _, a, sa), (_, _, b, sb) = runs
    print(f"[gpu-wsweep] strip: max|d| {np.abs(a - b).max():.3e}, rel {rel_l2(a, b):.2e}")
    assert np.array_equal(a, b) and np.array_equal(sa, sb)


def test_w_sweep_not_for_tti_or_one_pass():
    """TTI's w needs x-y differences of Q, and the one-pass step has its own w: mode 1 is
    ignored there (the reported mode stays 0)."""
    from a3dfd import Modeller, make_desc
    for op, scheme in (("tti", 1), ("vti", 2)):
        with Modeller(make_desc(op, 20, 20, 20, D, D, D, 10, 10, 3, 10, 1e-3, 25.0, nb=10, scheme=scheme)) as mdl:
            mdl.set_w_sweep(1)
            assert mdl.w_sweep() == 0


def test_vti_model_update_reaches_stress_sweep():
    """The stress sweep reads an interleaved copy of the VTI model built at the first run
    after a3d_set_model: a second set_model on the same handle must be picked up (the
    record equals a fresh handle's), with w in either sweep."""
    from a3dfd import Modeller, make_desc
    from a3dfd.models import small_model
    nx, ny, nz = 40, 30, 26
    m1 = small_model("vti", nx, ny, nz, seed=51)
    m2 = small_model("vti", nx, ny, nz, seed=52)
    desc = make_desc("vti", nx, ny, nz, D, D, D, 20, 15, 5, 120, 1e-3, 25.0, nb=10, scheme=1)
    for wz in (0, 1):
        with Modeller(desc) as mdl:
            mdl.set_launch(4, 2, 4, 1)
            mdl.set_w_sweep(wz)
            mdl.set_model(**m1)
            first, _, _ = mdl.run()
            mdl.set_model(**m2)
            second, _, _ = mdl.run()
        with Modeller(desc) as fresh:
            fresh.set_launch(4, 2, 4, 1)
            fresh.set_w_sweep(wz)
            fresh.set_model(**m2)
            ref, _, _ = fresh.run()
        assert not np.array_equal(first, second)
        assert np.array_equal(second, ref), wz
