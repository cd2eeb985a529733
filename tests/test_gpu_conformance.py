"""The reference caller's own call shapes at the C boundary.

The reference's readmodel.py loads each operator with
cdll.LoadLibrary('./lib/<name>.so') relative to the working directory, sets no
argtypes and no restype, and wraps every argument itself
(readmodel.py:228-244, 278-296, 324-339, 366-386). Its 2nd-order call passes
kt = 300 and hands snap_interval to the bool slot and cut_directwave to the int
slot of cuda_3dfd (Toa_3dfd_cuda.cu:366-370). These tests make exactly those
calls, in that argument sequence, against this package's lib/ directory and
compare the written shot and snapshot files with the CPU oracle run with the
reference's per-operator constants."""
import os
from ctypes import c_bool, c_char_p, c_float, c_int, cdll

import numpy as np
import pytest

from conftest import PKG, rel_l2

pytestmark = pytest.mark.gpu

NX, NY, NZ, D = 18, 16, 14, 10.0
NAMES = {"vel": "vel", "rho": "rho", "eps": "epsilon", "delta": "delta", "theta": "theta", "phi": "psi"}


def _model_files(tmp_path, op, seed):
    from a3dfd.models import small_model
    m = small_model(op, NX, NY, NZ, seed=seed)
    paths = {}
    for k, a in m.items():
        p = str(tmp_path / f"{NAMES[k]}.bin")
        np.ascontiguousarray(a).tofile(p)
        paths[NAMES[k]] = p
    # the reference passes a rho file name to TTI even though it is not read
    paths.setdefault("rho", paths["vel"])
    return m, paths


def _b(s):
    return c_char_p(bytes(s, "utf-8"))


def _call(op, kernel, p, snap_folder, shotname, ishot, dg, nt, dt, fm, wtype, savesnap, cut, snap_interval):
    """The reference readmodel.py's ctypes call for `op`, argument for argument."""
    geo = [c_int(NX), c_int(NY), c_int(NZ), c_float(D), c_float(D), c_float(D),
           c_int(ishot[0]), c_int(ishot[1]), c_int(ishot[2]), c_int(ishot[3]), c_int(ishot[4]), c_int(ishot[5]),
           c_float(dg[0]), c_float(dg[1]), c_float(dg[2])]
    if op == "2nd":   # readmodel.py:324-339
        return kernel.cuda_3dfd(_b(p["vel"]), _b(snap_folder), _b(shotname), c_int(0), c_int(1), *geo,
                                c_int(nt), c_int(300), c_float(dt), c_float(fm), c_bool(savesnap),
                                c_int(snap_interval), c_bool(cut), c_int(0))
    if op == "iso":   # readmodel.py:228-244
        return kernel.cuda_3dfd_1order(_b(p["vel"]), _b(p["rho"]), _b(snap_folder), _b(shotname), c_int(0),
                                       c_int(1), *geo, c_int(nt), c_float(dt), c_int(wtype), c_float(fm),
                                       c_bool(savesnap), c_bool(cut), c_int(snap_interval), c_int(0))
    if op == "vti":   # readmodel.py:278-296
        return kernel.cuda_3dfd_vti(_b(p["vel"]), _b(p["rho"]), _b(p["epsilon"]), _b(p["delta"]),
                                    _b(snap_folder), _b(shotname), c_int(0), c_int(1), *geo, c_int(nt),
                                    c_float(dt), c_float(fm), c_bool(savesnap), c_bool(cut),
                                    c_int(snap_interval), c_int(0))
    # readmodel.py:366-386
    return kernel.cuda_3dfd_tti(_b(p["vel"]), _b(p["rho"]), _b(p["epsilon"]), _b(p["delta"]), _b(p["theta"]),
                                _b(p["psi"]), _b(snap_folder), _b(shotname), c_int(0), c_int(1), *geo,
                                c_int(nt), c_float(dt), c_float(fm), c_bool(savesnap), c_bool(cut),
                                c_int(snap_interval), c_int(0))


# op: (library name, oracle constants, (sx, sy, sz), snapshot steps, snapshot name, wtype, mute, fm, nt)
CASES = {
    "2nd": ("acoustic2order", dict(nb=50, rec_z=1, snap_interval=500), (9, 8, 3), [0], "ishot_0_it_{}.bin", 1,
            True, 25.0, 130),
    "iso": ("acoustic1order", dict(nb=20, pfac=10000.0, rec_z=0, snap_interval=40), (9, 8, 3), [0, 40, 80, 120],
            "ishot_0it_{}.bin", 3, True, 40.0, 150),
    "vti": ("acoustic1vti", dict(nb=20, pfac=10.0, sv=1, snap_interval=40), (9, 8, 2), [0, 40, 80, 120],
            "ishot_0_it_{}.bin", 1, True, 40.0, 150),
    "tti": ("acoustic1tti", dict(nb=40, pfac=10.0, sv=1, snap_interval=100, snap_skip0=1), (9, 8, 2), [100],
            "ishot_0_it_{}.bin", 1, False, 25.0, 130),
}


@pytest.mark.parametrize("op", ["2nd", "iso", "vti", "tti"])
def test_reference_caller_call_shape(tmp_path, monkeypatch, oracle, op):
    libname, okw, src, snap_its, snapfmt, wtype, cut, fm, nt = CASES[op]
    m, paths = _model_files(tmp_path, op, seed=41)
    snaps = tmp_path / "snaps"
    snaps.mkdir()
    shotname = str(tmp_path / f"{src[0]}.{src[1]}.bin")    # readmodel.py:223
    # the reference resolves ./lib/<name>.so against the working directory
    monkeypatch.chdir(PKG)
    kernel = cdll.LoadLibrary(f"./lib/{libname}.so")
    ishot = (src[0], src[1], src[2], 1, 1, 1)
    dg = (20.0, 10.0, 0.002)
    print(f"\n[gpu-conformance] {op}: cdll.LoadLibrary('./lib/{libname}.so'), reference argument sequence",
          flush=True)
    rc = _call(op, kernel, paths, str(snaps) + "/", shotname, ishot, dg, nt, 1e-3, fm, wtype, True, cut, 40)
    if op == "tti":
        assert rc == 0                                       # default restype int: 0 ok, 1 shot file
    got = np.fromfile(shotname, dtype=np.float32)

    # the 2nd-order entry ignores cut_directwave and snap_interval (the reference's are dead
    # there: Toa_3dfd_cuda.cu:456, :474-477), whichever slot the caller put them in
    kw = dict(okw, wtype=wtype, mute=int(cut) if op != "2nd" else 0)
    ref, rsn, _ = oracle.run(oracle.make_shot(op, NX, NY, NZ, D, D, D, *src, nt, 1e-3, fm, **kw), 32,
                             nsnap_max=8, **m)
    expect = ref[::1, ::2, ::2]                              # dgy/dy = 1, dgx/dx = 2, dgt/dt = 2
    assert np.abs(expect).max() > 0
    assert got.size == expect.size
    e = rel_l2(got, expect.ravel())
    print(f"[gpu-conformance] {op}: shot file rel-L2 vs oracle {e:.2e}")
    assert e <= 1e-5
    if op != "tti":
        assert np.array_equal(got, expect.ravel())
    written = sorted(os.listdir(snaps))
    assert written == sorted(snapfmt.format(it) for it in snap_its)
    assert len(rsn) == len(snap_its)
    for k, it in enumerate(snap_its):
        s = np.fromfile(snaps / snapfmt.format(it), dtype=np.float32).reshape(NY, NX, NZ)
        assert rel_l2(s, rsn[k]) <= 1e-5

