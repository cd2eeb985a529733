#!/usr/bin/env python3
"""Writes full-length golden fixtures of three benchmark configurations from the pinned
CPU oracle, each over the configuration's whole time axis:

  cfg1_2nd_100_nt500.npz       2nd order 100^3, sponge 20, nt 500, receivers at z=5,
                               snapshots every 100 steps (FP32)
  cfg2_iso_256_nt2000.npz      ISO 256^3, npml 20, nt 2000, direct-wave mute with a
                               2000 m/s first-arrival window (FP32)
  cfg3_vti_400x400x300_nt3000.npz
                               VTI 400x400x300, npml 20, nt 3000, snapshots at 1000 and
                               2000 (FP32 and FP64)

Each file holds a decimated shot record (every DEC_R-th receiver in x and y from R0,
every DEC_T-th sample), the full receiver plane at the last step, decimated snapshots
(every 8th point per axis for cfg3, every 4th for cfg1), and SHA-256 digests of the
model arrays so a test can tell a model-generator change from a parity failure.

Usage: python tests/golden/make_fullsize_golden.py [--threads N] [--only cfg1,cfg2,cfg3] [--parts DIR]
(cfg3 takes about 40 min in FP32 and 80 min in FP64 on 6 threads of a 2.7 GHz Xeon.) With
--parts, each finished (configuration, precision) run is kept in DIR and not run again, so
an interrupted generation resumes where it stopped.
"""
import argparse
import hashlib
import importlib.util
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "acoustic3d-fd_amd"))

# name: (receiver start, receiver stride, time stride, snapshot stride)
DECIMATION = {"cfg1": (2, 5, 1, 4), "cfg2": (8, 16, 2, 8), "cfg3": (10, 20, 3, 8)}
FILES = {"cfg1": "cfg1_2nd_100_nt500.npz", "cfg2": "cfg2_iso_256_nt2000.npz",
         "cfg3": "cfg3_vti_400x400x300_nt3000.npz"}
MODEL_KEYS = {"cfg1": ("vel",), "cfg2": ("vel", "rho"), "cfg3": ("vel", "rho", "eps", "delta")}


def oracle():
    spec = importlib.util.spec_from_file_location("oracle_binding", os.path.join(ROOT, "oracle", "binding.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.load()
    return mod


def config(name):
    """(model dict, operator, shot keyword arguments, precisions) of a configuration."""
    from a3dfd import models
    if name == "cfg1":
        c = models.cfg1()
        return c, "2nd", dict(nb=20, rec_z=5, snap_interval=100), (32,)
    if name == "cfg2":
        c = models.cfg2()
        return c, "iso", dict(nb=20, mute=1, v_mute=2000.0, snap_interval=0), (32,)
    c = models.cfg3()
    return c, "vti", dict(nb=20, snap_interval=1000, snap_skip0=1), (32, 64)


def digests(c, keys):
    return {k: hashlib.sha256(np.ascontiguousarray(c[k]).tobytes()).hexdigest() for k in keys}


def decimate(name, shot, snaps):
    r0, rs, ts, ss = DECIMATION[name]
    out = {"traces": np.ascontiguousarray(shot[r0::rs, r0::rs, ::ts]).astype(np.float32),
           "plane": np.ascontiguousarray(shot[:, :, -1]).astype(np.float32)}
    if snaps is not None and len(snaps):
        out["snaps"] = np.ascontiguousarray(snaps[:, ::ss, ::ss, ::ss]).astype(np.float32)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--only", type=str, default="cfg1,cfg2,cfg3")
    ap.add_argument("--parts", type=str, default="")
    args = ap.parse_args()
    if args.parts:
        os.makedirs(args.parts, exist_ok=True)
    orc = oracle()
    for name in args.only.split(","):
        c, op, kw, precs = config(name)
        keys = MODEL_KEYS[name]
        m = {k: c[k] for k in keys}
        meta = dict(config=name, op=op, nx=c["nx"], ny=c["ny"], nz=c["nz"], dx=c["dx"], dt=c["dt"],
                    nt=c["nt"], fm=c["fm"], sx=c["sx"], sy=c["sy"], sz=c["sz"], decimation=DECIMATION[name], **kw)
        out = {"meta": np.array(repr(meta))}
        for k, v in digests(c, keys).items():
            out[f"sha256_{k}"] = np.array(v)
        for prec in precs:
            part = os.path.join(args.parts, f"{name}_fp{prec}.npz") if args.parts else ""
            if part and os.path.exists(part):
                with np.load(part, allow_pickle=False) as f:
                    out.update({k: f[k] for k in f.files})
                print(f"{name} FP{prec}: kept from {part}", flush=True)
                continue
            shot = orc.make_shot(op, c["nx"], c["ny"], c["nz"], c["dx"], c["dy"], c["dz"], c["sx"], c["sy"], c["sz"],
                                 c["nt"], c["dt"], c["fm"], nthreads=args.threads, **kw)
            t0 = time.time()
            nsnap = c["nt"] // kw["snap_interval"] + 1 if kw.get("snap_interval") else 0
            rec, snaps, secs = orc.run(shot, prec, nsnap_max=nsnap, **m)
            print(f"{name} FP{prec}: {time.time() - t0:.0f} s, max |record| {float(np.abs(rec).max()):.3e}", flush=True)
            got = {f"{k}{prec}": v for k, v in decimate(name, rec, snaps).items()}
            del rec, snaps
            if part:
                np.savez(part, **got)
            out.update(got)
        path = os.path.join(HERE, FILES[name])
        np.savez_compressed(path, **out)
        print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)", flush=True)


if __name__ == "__main__":
    main()
