"""The single-shot time loop captured in hipGraphs (a3d_set_graph): bitwise the records and
snapshots of host-launched steps for every operator, with graph chunks of 1 step, of a
length that does not divide the loop, and longer than the loop, cut at snapshot steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = {
    "2nd": dict(nb=8, src=(12, 10, 5)),
    "iso": dict(nb=10, src=(12, 10, 3)),
    "vti": dict(nb=10, src=(12, 10, 3), scheme=3),
    "tti": dict(nb=12, src=(12, 10, 3), scheme=1),
}


@pytest.mark.parametrize("op", ["2nd", "iso", "vti", "tti"])
def test_graph_loop_bitwise(op):
    from a3dfd import Modeller, make_desc
    from a3dfd.models import small_model
    kw = dict(CASES[op])
    src = kw.pop("src")
    nx, ny, nz, nt = 30, 26, 24, 97
    m = small_model(op, nx, ny, nz, seed=83)
    desc = make_desc(op, nx, ny, nz, 10.0, 10.0, 10.0, *src, nt, 1e-3, 25.0, snap_interval=30, **kw)
    print(f"\n[gpu-graph] {op}", flush=True)
    with Modeller(desc) as mdl:
        mdl.set_model(**m)
        mdl.set_launch(4, 2, 4, 2)
        ref, rsn, _ = mdl.run(nsnap_max=4)
        for steps in (1, 13, 1000):
            mdl.set_graph(steps)
            got, gsn, ms = mdl.run(nsnap_max=4)
            assert np.array_equal(got, ref) and np.array_equal(gsn, rsn), steps
        mdl.set_graph(0)
        again, _, _ = mdl.run(nsnap_max=4)
    assert np.abs(ref).max() > 0
    assert np.array_equal(again, ref)


def test_graph_loop_with_one_pass_step_and_strip():
    """The one-pass VTI step (scheme 2) and the learned strip's passes in the captured loop."""
    from a3dfd import STRIP_NB, Modeller, make_desc, strip
    from a3dfd.models import small_model
    nx, ny, nz, nt = 30, 26, 24, 80
    m = small_model("vti", nx, ny, nz, seed=85)
    for with_strip, kw in ((False, dict(nb=10, scheme=2)), (True, dict(nb=STRIP_NB, scheme=3))):
        desc = make_desc("vti", nx, ny, nz, 10.0, 10.0, 10.0, 15, 13, 4, nt, 1e-3, 25.0, **kw)
        with Modeller(desc) as mdl:
            mdl.set_model(**m)
            if with_strip:
                mdl.set_strip_net(strip.load_params()[0])
            ref, _, _ = mdl.run()
            mdl.set_graph(16)
            got, _, _ = mdl.run()
        assert np.abs(ref).max() > 0
        assert np.array_equal(got, ref), kw
