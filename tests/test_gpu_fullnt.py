"""Parity over whole time axes at the configurations' full grids, against the golden
fixtures tests/golden/make_fullsize_golden.py wrote from the pinned CPU oracle:

  2nd order 100^3 (sponge 20), nt 500, receivers at z=5, snapshots every 100 steps;
  ISO 256^3 (npml 20), nt 2000, direct-wave mute with a 2000 m/s first-arrival window;
  VTI 400x400x300 (npml 20), nt 3000, the benchmarked shot, FP32 and FP64 oracle runs.

The FP32 reference arithmetic (scheme 0) must be bitwise equal to the oracle on every
sampled receiver, the last receiver plane and the sampled snapshots; the faster schemes
within the 1e-5 bar; binary16 storage and the FP64 instance against the FP64 oracle (the
FP64 instance bitwise after the fixture's rounding to float)."""
import ast
import hashlib
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODEL_KEYS = {"cfg1": ("vel",), "cfg2": ("vel", "rho"), "cfg3": ("vel", "rho", "eps", "delta")}
FILES = {"cfg1": "cfg1_2nd_100_nt500.npz", "cfg2": "cfg2_iso_256_nt2000.npz",
         "cfg3": "cfg3_vti_400x400x300_nt3000.npz"}


_CACHE = {}


def _fixture(name):
    """(meta, model, arrays) of a fixture; the model is generated once per test session."""
    if name in _CACHE:
        return _CACHE[name]
    from a3dfd import models
    with np.load(os.path.join(HERE, FILES[name]), allow_pickle=False) as f:
        data = {k: f[k] for k in f.files}
    meta = ast.literal_eval(str(data.pop("meta")))
    c = {"cfg1": models.cfg1, "cfg2": models.cfg2, "cfg3": models.cfg3}[name]()
    for k in MODEL_KEYS[name]:
        digest = hashlib.sha256(np.ascontiguousarray(c[k]).tobytes()).hexdigest()
        assert digest == str(data[f"sha256_{k}"]), f"{name}: the model generator's {k} changed"
    _CACHE.clear()   # one configuration's model in memory at a time
    _CACHE[name] = (meta, c, data)
    return meta, c, data


def _run(name, meta, c, nsnap, **kw):
    """The configuration's shot on the GPU: (traces, last plane, snapshots), decimated as the fixture."""
    from a3dfd import Modeller, make_desc
    skip = ("config", "op", "nx", "ny", "nz", "dx", "dt", "nt", "fm", "sx", "sy", "sz", "decimation")
    shot_kw = {k: v for k, v in meta.items() if k not in skip}
    shot_kw.update(kw)
    d = make_desc(meta["op"], c["nx"], c["ny"], c["nz"], c["dx"], c["dy"], c["dz"], c["sx"], c["sy"], c["sz"],
                  c["nt"], c["dt"], c["fm"], **shot_kw)
    print(f"\n[gpu-fullnt] {name} {meta['op']} {c['nx']}x{c['ny']}x{c['nz']} nt {c['nt']} {kw}", flush=True)
    with Modeller(d) as mdl:
        mdl.set_model(**{k: c[k] for k in MODEL_KEYS[name]})
        shot, snaps, ms = mdl.run(nsnap_max=nsnap)
        launch = mdl.launch_info()
    print(f"[gpu-fullnt] {name} {kw}: loop {ms:.0f} ms, launch {launch}", flush=True)
    r0, rs, ts, ss = meta["decimation"]
    out = (shot[r0::rs, r0::rs, ::ts], shot[:, :, -1], snaps[:, ::ss, ::ss, ::ss] if nsnap else None)
    del shot
    return out


def _compare(tag, got, data, prec, bitwise, bound):
    tr, pl, sn = got
    errs = [rel_l2(tr, data[f"traces{prec}"]), rel_l2(pl, data[f"plane{prec}"])]
    if sn is not None:
        errs += [rel_l2(sn[k], data[f"snaps{prec}"][k]) for k in range(len(sn))]
    print(f"[gpu-fullnt] {tag} vs FP{prec} oracle: traces {errs[0]:.2e}, last plane {errs[1]:.2e}"
          + (f", snapshots {', '.join(f'{e:.2e}' for e in errs[2:])}" if sn is not None else ""), flush=True)
    assert np.abs(data[f"traces{prec}"]).max() > 0
    if bitwise:
        assert np.array_equal(tr, data[f"traces{prec}"]) and np.array_equal(pl, data[f"plane{prec}"])
        if sn is not None:
            assert np.array_equal(sn, data[f"snaps{prec}"])
    assert max(errs) <= bound, errs
    return errs


def test_cfg1_full_length():
    meta, c, data = _fixture("cfg1")
    got = _run("cfg1", meta, c, nsnap=5)
    _compare("cfg1 2nd order", got, data, 32, True, 0.0)


@pytest.mark.parametrize("scheme", [0, 3])
def test_cfg2_full_length_with_mute(scheme):
    meta, c, data = _fixture("cfg2")
    got = _run("cfg2", meta, c, nsnap=0, scheme=scheme)
    _compare(f"cfg2 ISO scheme {scheme}", got, data, 32, scheme == 0, 1e-5)
    # the record is muted before the 2000 m/s first arrival plus the wavelet and 60 samples:
    # the first samples of every sampled trace are zero
    assert np.all(got[0][:, :, :20] == 0.0)


# VTI 400x400x300, nt 3000: the benchmarked shot. The reference arithmetic (scheme 0) is
# bitwise equal to the FP32 oracle over the whole record and both snapshots. The benchmark
# path (stress scheme 4, autotuned launch shapes and w placement) and scheme 3 meet the 1e-5
# bar on the record and the snapshots. The receiver plane of the last step alone carries 1/165 of the
# record's peak amplitude, so its relative error is rounding-dominated: the FP32 oracle is
# itself 4.7e-5 from the FP64 oracle there; the GPU plane must stay within 3x that distance.
@pytest.mark.parametrize("scheme", [0, 3, 4])
def test_cfg3_full_length(scheme):
    meta, c, data = _fixture("cfg3")
    got = _run("cfg3", meta, c, nsnap=2, scheme=scheme)
    tr, pl, sn = got
    e32 = _compare(f"cfg3 VTI scheme {scheme}", got, data, 32, scheme == 0, 1.0)
    e64 = _compare(f"cfg3 VTI scheme {scheme}", got, data, 64, False, 1.0)
    own = rel_l2(data["plane32"], data["plane64"])   # the FP32 oracle's own distance on the last plane
    print(f"[gpu-fullnt] cfg3 scheme {scheme}: record {e32[0]:.2e}, snapshots {max(e32[2:]):.2e} vs FP32 oracle; "
          f"last plane {e64[1]:.2e} vs FP64 oracle (FP32 oracle: {own:.2e})", flush=True)
    assert e32[0] <= 1e-5 and max(e32[2:]) <= 1e-5, e32
    assert e64[1] <= 3 * own, (e64[1], own)
    assert e64[0] <= 1e-4 and max(e64[2:]) <= 1e-4, e64


# FP64 instance against the FP64 oracle: bitwise equal after the fixture's rounding to float.
def test_cfg3_full_length_fp64_instance():
    meta, c, data = _fixture("cfg3")
    got = _run("cfg3", meta, c, nsnap=2, storage=2)
    got = tuple(None if g is None else g.astype(np.float32) for g in got)
    _compare("cfg3 VTI FP64 instance", got, data, 64, True, 0.0)


# binary16 wavefield storage (per-field power-of-two scales) against the FP64 oracle, at
# twice the errors measured on MI355X: sampled record 7.1e-3, snapshots 5.5e-3 (it 1000) and
# 1.9e-2 (it 2000), last receiver plane 0.235. The last plane's amplitude is 1/165 of the
# record's peak, far below the field scale, where binary16's 11-bit significand leaves
# little relative precision; the record and snapshots carry the accuracy statement.
FP16_BOUNDS = {"traces": 1.4e-2, "plane": 0.47, "snaps": (1.1e-2, 3.8e-2)}


@pytest.mark.parametrize("scheme", [1, 4])
def test_cfg3_full_length_fp16_storage(scheme):
    meta, c, data = _fixture("cfg3")
    tr, pl, sn = _run("cfg3", meta, c, nsnap=2, storage=1, scheme=scheme)
    et, ep = rel_l2(tr, data["traces64"]), rel_l2(pl, data["plane64"])
    es = [rel_l2(sn[k], data["snaps64"][k]) for k in range(len(sn))]
    print(f"[gpu-fullnt] cfg3 FP16 storage, scheme {scheme}, vs FP64 oracle: traces {et:.2e}, last plane {ep:.2e}, "
          f"snapshots {', '.join(f'{e:.2e}' for e in es)}", flush=True)
    assert et <= FP16_BOUNDS["traces"] and ep <= FP16_BOUNDS["plane"], (et, ep)
    assert all(e <= b for e, b in zip(es, FP16_BOUNDS["snaps"])), es
