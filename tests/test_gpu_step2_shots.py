"""Several 2nd-order shots on one model stepped together (a3d_run_shots): one launch per time
step over every shot's tiles. Each shot's record must equal its own single-shot run bit for
bit (the reference runs one shot per cuda_3dfd call, Toa_3dfd_cuda.cu:366-508) — every
storage precision, several launch shapes, 1 to 16 shots, sources near the faces — and the
FP32 records the oracle's; bad arguments are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = 10.0
NX, NY, NZ, NT = 70, 41, 30, 90


def _desc(storage=0, sx=33, sy=19, sz=6, **kw):
    from a3dfd import make_desc
    return make_desc("2nd", NX, NY, NZ, D, D, D, sx, sy, sz, NT, 1e-3, 25.0, nb=20, storage=storage, **kw)


def _sources(n, seed=5):
    """Source cells (0-based, as the 2nd-order entry point takes them)."""
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(0, NX)), int(rng.integers(0, NY)), int(rng.integers(0, NZ))) for _ in range(n)]


@pytest.mark.parametrize("storage", [0, 1, 2])
def test_shots_together_equal_single_runs(storage):
    from a3dfd import Modeller
    from a3dfd.models import small_model
    m = small_model("2nd", NX, NY, NZ, seed=61)
    srcs = _sources(5) + [(0, 0, 0), (NX - 1, NY - 1, NZ - 1)]
    with Modeller(_desc(storage)) as mdl:
        mdl.set_model(**m)
        short = -(-(mdl.grid_info()["nnz"] - 8) // 4)   # segments of <= 4 planes: all loads at once
        for launch in ((4, 3, 4, 3), (2, 1, 2, 1), (8, 2, 8, 2), (4, short, 4, short)):
            mdl.set_launch(*launch)
            print(f"\n[gpu-step2-shots] storage={storage} launch={launch} shots={len(srcs)}", flush=True)
            recs, _ = mdl.run_shots(srcs)
            for k, (sx, sy, sz) in enumerate(srcs):
                one, _, _ = mdl.run(_desc(storage, sx, sy, sz))
                assert np.abs(one).max() > 0
                assert np.array_equal(recs[k], one), (launch, k)
            solo, _ = mdl.run_shots(srcs[-1:])   # one shot: the launch shape as given
            assert np.array_equal(solo[0], recs[-1]), launch


def test_sixteen_shots_and_one():
    from a3dfd import Modeller
    from a3dfd.models import small_model
    m = small_model("2nd", NX, NY, NZ, seed=63)
    srcs = _sources(16, seed=7)
    with Modeller(_desc()) as mdl:
        mdl.set_model(**m)
        recs, _ = mdl.run_shots(srcs)
        for k in (0, 7, 15):
            one, _, _ = mdl.run(_desc(0, *srcs[k]))
            assert np.array_equal(recs[k], one), k
        single, _ = mdl.run_shots(srcs[3:4])
        assert np.array_equal(single[0], recs[3])


def test_shots_together_equal_oracle(oracle):
    from a3dfd import Modeller
    from a3dfd.models import small_model
    m = small_model("2nd", NX, NY, NZ, seed=65)
    srcs = [(20, 10, 4), (50, 30, 12)]
    with Modeller(_desc()) as mdl:
        mdl.set_model(**m)
        recs, _ = mdl.run_shots(srcs)
    for k, (sx, sy, sz) in enumerate(srcs):
        ref, _, _ = oracle.run(oracle.make_shot("2nd", NX, NY, NZ, D, D, D, sx, sy, sz, NT, 1e-3, 25.0, nb=20), 32, **m)
        assert np.array_equal(recs[k], ref), k


def test_shots_together_refuse_bad_arguments():
    from a3dfd import A3DError, Modeller, make_desc
    from a3dfd.models import small_model
    m = small_model("2nd", NX, NY, NZ, seed=67)
    with Modeller(_desc()) as mdl:
        mdl.set_model(**m)
        for bad in ([], _sources(17), [(-1, 5, 5)], [(NX, 5, 5)]):
            with pytest.raises((A3DError, ValueError)):
                mdl.run_shots(bad)
        with pytest.raises(A3DError):
            mdl.run_shots(_sources(2), desc=_desc(snap_interval=30))
    vti = make_desc("vti", 30, 26, 24, D, D, D, 12, 10, 3, 20, 1e-3, 25.0, nb=10)
    with Modeller(vti) as mdl:
        mdl.set_model(**small_model("vti", 30, 26, 24, seed=3))
        with pytest.raises(A3DError):
            mdl.run_shots([(12, 10, 3)])


def test_descriptor_source_not_used():
    """The shots' sources come from the list; the descriptor's own (here outside the model)
    is not used, by the run or by the launch-shape tuning of a fresh context."""
    from a3dfd import Modeller
    from a3dfd.models import small_model
    m = small_model("2nd", NX, NY, NZ, seed=69)
    srcs = _sources(3, seed=11)
    with Modeller(_desc(sx=-5)) as mdl:
        mdl.set_model(**m)
        a, _ = mdl.run_shots(srcs)
    with Modeller(_desc()) as mdl:
        mdl.set_model(**m)
        b, _ = mdl.run_shots(srcs)
    assert np.abs(a).max() > 0
    assert np.array_equal(a, b)
