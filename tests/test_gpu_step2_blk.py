"""2nd-order step over short z segments (at most 4 planes) with every load of a segment issued
at once (k_step2_blk, chosen for segments of <= 4 planes): records and snapshots equal the
plane-streaming kernel's bit for bit for every precision, tile rows 2/4/8 and segments of 1
to 4 planes (a last, shorter segment included), sources near the faces, and the FP32
oracle's (reference cuda_step_fd3d, Toa_3dfd_cuda.cu:189-298)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = 10.0
NX, NY, NZ, NT = 70, 41, 30, 100


def _desc(storage=0, src=(33, 19, 6)):
    from a3dfd import make_desc
    return make_desc("2nd", NX, NY, NZ, D, D, D, *src, NT, 1e-3, 25.0, nb=20, storage=storage, snap_interval=45)


@pytest.mark.parametrize("storage", [0, 1, 2])
def test_short_segments_equal_streaming(storage):
    from a3dfd import Modeller
    from a3dfd.models import small_model
    m = small_model("2nd", NX, NY, NZ, seed=71)
    for src in ((33, 19, 6), (0, 0, 0), (NX - 1, NY - 1, NZ - 1)):
        with Modeller(_desc(storage, src)) as mdl:
            mdl.set_model(**m)
            nnz = mdl.grid_info()["nnz"]
            nzi = nnz - 8
            mdl.set_launch(4, 1, 4, 1)
            ref, rsn, _ = mdl.run(nsnap_max=3)
            assert np.abs(ref).max() > 0
            for ty in (2, 4, 8):
                for zl in (4, 3, 2, 1):
                    zseg = -(-nzi // zl)
                    print(f"\n[gpu-step2-blk] storage={storage} src={src} ty={ty} planes/segment<={zl} "
                          f"segments={zseg}", flush=True)
                    mdl.set_launch(ty, zseg, ty, zseg)
                    got, gsn, _ = mdl.run(nsnap_max=3)
                    assert np.array_equal(got, ref) and np.array_equal(gsn, rsn), (ty, zl)


def test_short_segments_equal_oracle(oracle):
    from a3dfd import Modeller
    from a3dfd.models import small_model
    m = small_model("2nd", NX, NY, NZ, seed=73)
    with Modeller(_desc()) as mdl:
        mdl.set_model(**m)
        nzi = mdl.grid_info()["nnz"] - 8
        mdl.set_launch(4, -(-nzi // 4), 4, -(-nzi // 4))
        got, gsn, _ = mdl.run(nsnap_max=3)
    ref, rsn, _ = oracle.run(oracle.make_shot("2nd", NX, NY, NZ, D, D, D, 33, 19, 6, NT, 1e-3, 25.0, nb=20,
                                              snap_interval=45), 32, nsnap_max=3, **m)
    assert np.array_equal(got, ref) and np.array_equal(gsn, rsn)
