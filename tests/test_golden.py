"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the
pinned CPU oracle): the oracle must still reproduce them bit for bit, and the HIP path
must match them (bitwise for 2nd order, ISO and VTI; within the 1e-5 bar for TTI, whose
per-point cos/sin rounding may differ in rare double-rounding cases). FP32 against the
FP64 record: within 1e-4 (the VTI fixture's record is muted, so its norm is small)."""
import ast
import os

import numpy as np
import pytest

from conftest import rel_l2

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OPS = ["2nd", "iso", "vti", "tti"]
MODEL_KEYS = ("vel", "rho", "eps", "delta", "theta", "phi")


def _load(op):
    with np.load(os.path.join(HERE, f"{op}.npz"), allow_pickle=False) as f:
        meta = ast.literal_eval(str(f["meta"]))
        model = {k: f[f"model_{k}"] for k in MODEL_KEYS if f"model_{k}" in f}
        return meta, model, f["shot32"], f["snap32"], f["shot64"]


def _shot_kw(meta):
    kw = {k: v for k, v in meta.items() if k not in ("op", "nx", "ny", "nz", "d", "dt", "sx", "sy", "sz", "nt", "fm")}
    return kw


@pytest.mark.parametrize("op", OPS)
def test_oracle_reproduces_golden(oracle, op):
    meta, model, shot32, snap32, shot64 = _load(op)
    d = meta["d"]
    shot = oracle.make_shot(op, meta["nx"], meta["ny"], meta["nz"], d, d, d, meta["sx"], meta["sy"], meta["sz"],
                            meta["nt"], meta["dt"], meta["fm"], **_shot_kw(meta))
    s32, n32, _ = oracle.run(shot, 32, nsnap_max=4, **model)
    s64, _, _ = oracle.run(shot, 64, **model)
    assert np.array_equal(s32[::2, ::2], shot32) and np.array_equal(n32[-1], snap32)
    assert np.array_equal(s64[::2, ::2], shot64)
    assert rel_l2(shot32, shot64) <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("op", OPS)
def test_gpu_matches_golden(op):
    from a3dfd import Modeller, make_desc
    meta, model, shot32, snap32, shot64 = _load(op)
    d = meta["d"]
    desc = make_desc(op, meta["nx"], meta["ny"], meta["nz"], d, d, d, meta["sx"], meta["sy"], meta["sz"], meta["nt"],
                     meta["dt"], meta["fm"], **_shot_kw(meta))
    with Modeller(desc) as mdl:
        mdl.set_model(**model)
        shot, snaps, _ = mdl.run(nsnap_max=4)
    got, gsnap = shot[::2, ::2], snaps[-1]
    print(f"\n[gpu-golden] {op}: vs FP32 fixture {rel_l2(got, shot32):.2e}, snapshot {rel_l2(gsnap, snap32):.2e}, "
          f"vs FP64 fixture {rel_l2(got, shot64):.2e}")
    if op != "tti":
        assert np.array_equal(got, shot32) and np.array_equal(gsnap, snap32)
    assert rel_l2(got, shot32) <= 1e-5 and rel_l2(gsnap, snap32) <= 1e-5
    assert rel_l2(got, shot64) <= 1e-4
