"""The drop-in path end to end on the GPU: readmodel.py's operator classes call
the four legacy entry points with the reference's ctypes arguments; the shot
files (decimated [iy][ix][it]) and snapshot files (names and contents) must
equal the CPU oracle run with the reference's per-operator constants."""
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

NX, NY, NZ, D = 18, 16, 14, 10.0


def _files(tmp_path, op, seed=21):
    from a3dfd.models import small_model
    import readmodel as rm
    m = small_model(op, NX, NY, NZ, seed=seed)
    paths = {}
    names = {"vel": "vel", "rho": "rho", "eps": "epsilon", "delta": "delta", "theta": "theta", "phi": "psi"}
    for k, a in m.items():
        p = str(tmp_path / f"{names[k]}.bin")
        np.ascontiguousarray(a).tofile(p)          # (ny, nx, nz) C order = model-file order
        paths[names[k]] = p
    model = rm.Model(NX, NY, NZ, D, D, D, vel_file=paths["vel"], rho_file=paths.get("rho"),
                     epsilon_file=paths.get("epsilon"), delta_file=paths.get("delta"),
                     theta_file=paths.get("theta"), psi_file=paths.get("psi"))
    return m, model


CASES = {
    # op: (class, oracle kwargs, source (ishot), snapshot name pattern, wtype, mute, fm, nt)
    "2nd": ("Acoustic3d2order", dict(nb=50, rec_z=1, snap_interval=500), (9, 8, 3), "ishot_0_it_{}.bin", 1, False,
            25.0, 120),
    "iso": ("Acoustic3d1order", dict(nb=20, pfac=10000.0, rec_z=0, snap_interval=50), (9, 8, 3),
            "ishot_0it_{}.bin", 2, True, 40.0, 200),
    "vti": ("Acoustic3dvti", dict(nb=20, pfac=10.0, sv=1, snap_interval=50), (9, 8, 2), "ishot_0_it_{}.bin", 1,
            True, 40.0, 200),
    "tti": ("Acoustic3dtti", dict(nb=40, pfac=10.0, sv=1, snap_interval=100, snap_skip0=1), (9, 8, 2),
            "ishot_0_it_{}.bin", 1, False, 25.0, 120),
}


@pytest.mark.parametrize("op", ["2nd", "iso", "vti", "tti"])
def test_legacy_entry_point_files(tmp_path, oracle, op):
    import readmodel as rm
    cls, okw, ishot, snapfmt, wtype, mute, fm, nt = CASES[op]
    m, model = _files(tmp_path, op)
    shots = tmp_path / "shots"
    snaps = tmp_path / "snaps"
    shots.mkdir()
    snaps.mkdir()
    sg = rm.ShotReceiver(ishot[0], ishot[1], ishot[2], 1, 1, 1, ishot[0], ishot[1], ishot[2],
                         dgx=20.0, dgy=10.0, dgt=0.002, shotfolder=str(shots) + "/")
    fd = getattr(rm, cls)(model=model, sg=sg, nt=nt, dt=0.001, fpeak=fm, snap_folder=str(snaps) + "/",
                          snap_interval=50, wtype=wtype, savesnap=True, cut_directwave=mute)
    print(f"\n[gpu-legacy] {op} via readmodel.{cls} nt={nt} source={ishot}", flush=True)
    fd.run(sg.shotlist[0], 0)
    got = np.fromfile(fd.shot_name(sg.shotlist[0]), dtype=np.float32)

    kw = dict(okw, wtype=wtype, mute=int(mute))
    ref, rsn, _ = oracle.run(oracle.make_shot(op, NX, NY, NZ, D, D, D, *ishot, nt, 0.001, fm, **kw), 32,
                             nsnap_max=8, **m)
    expect = ref[::1, ::2, ::2]
    assert got.size == expect.size
    assert np.abs(expect).max() > 0
    assert rel_l2(got, expect.ravel()) <= 1e-5
    if op != "tti":
        assert np.array_equal(got, expect.ravel())
    its = [it for it in range(nt) if it % kw["snap_interval"] == 0 and not (kw.get("snap_skip0") and it == 0)]
    assert len(its) == len(rsn)
    for k, it in enumerate(its):
        p = snaps / snapfmt.format(it)
        assert p.exists(), sorted(os.listdir(snaps))
        s = np.fromfile(p, dtype=np.float32).reshape(NY, NX, NZ)
        assert rel_l2(s, rsn[k]) <= 1e-5


def test_skip_existing_shot(tmp_path):
    import readmodel as rm
    m, model = _files(tmp_path, "vti")
    sg = rm.ShotReceiver(9, 8, 2, 1, 1, 1, 9, 8, 2, dgx=10.0, dgy=10.0, dgt=0.001, shotfolder=str(tmp_path) + "/")
    fd = rm.Acoustic3dvti(model=model, sg=sg, nt=40, dt=0.001, fpeak=25.0, snap_folder=str(tmp_path) + "/",
                          snap_interval=10)
    fd.run(sg.shotlist[0], 0)
    p = fd.shot_name(sg.shotlist[0])
    assert os.path.getsize(p) == 4 * NX * NY * 40
    mtime = os.path.getmtime(p)
    assert fd.run(sg.shotlist[0], 0, skip_existing=True) == 0
    assert os.path.getmtime(p) == mtime


def test_parallel_run_on_streams(tmp_path, oracle):
    """Several shots of one node run concurrently (one HIP stream each) and each
    equals its own single run."""
    import readmodel as rm
    m, model = _files(tmp_path, "vti", seed=23)
    sg = rm.ShotReceiver(5, 5, 2, 4, 3, 1, 13, 11, 2, dgx=10.0, dgy=10.0, dgt=0.001, shotfolder=str(tmp_path) + "/")
    fd = rm.Acoustic3dvti(model=model, sg=sg, nt=60, dt=0.001, fpeak=25.0, snap_folder=str(tmp_path) + "/",
                          snap_interval=10)
    print(f"\n[gpu-legacy] parallel_run of {sg.nshot} VTI shots on 3 streams", flush=True)
    fd.parallel_run(streams=3)
    for ishot in sg.shotlist:
        got = np.fromfile(fd.shot_name(ishot), dtype=np.float32).reshape(NY, NX, 60)
        ref, _, _ = oracle.run(oracle.make_shot("vti", NX, NY, NZ, D, D, D, ishot[0], ishot[1], ishot[2], 60,
                                                0.001, 25.0, nb=20), 32, **m)
        assert np.array_equal(got, ref)


def test_legacy_scheme_environment(tmp_path, oracle, monkeypatch):
    """A3D_SCHEME selects the stress update of the legacy entry points: 1 (split fields in PML
    cells only) within the parity bar of the reference arithmetic, 2 (one-pass VTI step) bitwise
    equal to 1."""
    import readmodel as rm
    m, model = _files(tmp_path, "vti", seed=29)
    got = {}
    for scheme in ("0", "1", "2"):
        monkeypatch.setenv("A3D_SCHEME", scheme)
        out = tmp_path / f"s{scheme}"
        out.mkdir()
        sg = rm.ShotReceiver(9, 8, 2, 1, 1, 1, 9, 8, 2, dgx=10.0, dgy=10.0, dgt=0.001, shotfolder=str(out) + "/")
        fd = rm.Acoustic3dvti(model=model, sg=sg, nt=150, dt=0.001, fpeak=25.0, snap_folder=str(out) + "/",
                              snap_interval=1000)
        fd.run(sg.shotlist[0], 0)
        got[scheme] = np.fromfile(fd.shot_name(sg.shotlist[0]), dtype=np.float32)
    ref, _, _ = oracle.run(oracle.make_shot("vti", NX, NY, NZ, D, D, D, 9, 8, 2, 150, 0.001, 25.0, nb=20), 32, **m)
    assert np.array_equal(got["0"], ref.ravel())
    assert rel_l2(got["1"], ref.ravel()) <= 1e-5
    assert np.array_equal(got["1"], got["2"])
