#!/usr/bin/env python3
"""Benchmark: 1st-order VTI pseudo-acoustic shot, 400 x 400 x 300 cells,
dx = 12.5 m, dt = 1 ms, nt = 3000, 20-cell split PML, FP32, on one MI355X.
Default stress update: scheme 4 (one PML field per damped direction of a PML
cell, of which the reference's split p and q fields are per-point multiples;
within 1e-5 relative L2 of the reference arithmetic: 4.1e-7 on this shot's
record); --scheme 3 keeps the (p, q) pair per damped direction, --scheme 1 all
six in every PML cell, --scheme 0 runs the reference's split-everywhere
arithmetic bit for bit.

A "step" is one full shot (nt time steps of the velocity and stress sweeps)
over the synthetic seeded model, with the model and wavefields resident in
HBM when the timed region starts. value = grid-point updates per second
(updated points x nt x shots / seconds). The roofline entry is the stress
sweep (the dominant kernel): algorithmic bytes per launch / its average
launch time from HIP events recorded around the launches of every 8th time
step (--time-stride) on the context stream. cpu_baseline: the FP32 CPU oracle (OpenMP over all host cores) on a
bounded number of time steps of the same model, scaled to the same unit.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--nt NT]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "acoustic3d-fd_amd"))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per updated point per time step of the VTI formulation
# implemented (reference layout: split PML fields stored everywhere, FP32).
# velocity sweep: R P,Q,rho,u,v,w; W u,v,w            -> 9 x 4 = 36 B
# stress sweep:   R u,v,w, 6 split, vp,rho,eps,delta;
#                 W 6 split, P, Q                      -> 21 x 4 = 84 B
A_VEL, A_STRESS = 36, 84


def algorithmic_bytes(cfg, storage, scheme, npts, npml=20, wz=0):
    """Algorithmic bytes per updated point of the VTI velocity and stress sweeps for the
    formulation run: wavefields of ws bytes (4 FP32, 2 binary16, 8 FP64), split PML fields
    of ss bytes (FP32 under binary16), model of ms bytes.
    velocity: R P,Q,rho,u,v,w; W u,v,w. stress, split cell: R u,v,w, 6 split, vp,rho,eps,delta;
    W 6 split, P, Q. stress, unsplit cell (scheme 1 outside the PML): R u,v,w, P,Q, 4 model; W P,Q.
    wz (w updated by the stress sweep): velocity R P,rho,u,v; W u,v. stress reads u,v, the old Q
    one plane ahead (also the unsplit cell's Q), the old w, and writes w; otherwise as above.
    scheme 2 (one pass, FP32): (0, R P,Q,u,v,w, 4 model; W P,Q,u,v,w; + R/W 6 split in PML cells).
    scheme 3: the unsplit cell plus R/W of 2 split fields per damped direction; a cell damped in
    all three directions is the split cell (it reads no old P). scheme 4: as 3 with one field
    per damped direction (R/W 1 field; the all-damped cell R/W 3 fields)."""
    ws = {0: 4, 1: 2, 2: 8}[storage]
    ms = 8 if storage == 2 else 4
    ss = max(ws, 4)                     # split PML fields are FP32 under binary16 storage
    if wz:
        a_vel = 5 * ws + ms
        split = 7 * ws + 12 * ss + 4 * ms
        unsplit = 8 * ws + 4 * ms
    else:
        a_vel = 8 * ws + ms
        split = 5 * ws + 12 * ss + 4 * ms
        unsplit = 7 * ws + 4 * ms
    inner = (cfg["nx"] + 1) * (cfg["ny"] + 1) * (cfg["nz"] + 1)   # cells whose stress PML factors are all 1
    f = inner / npts
    if scheme == 3:
        return a_vel, dir_split_bytes(cfg, npml, unsplit, split, 4 * ss)
    if scheme == 4:
        return a_vel, dir_split_bytes(cfg, npml, unsplit, split - 6 * ss, 2 * ss)
    if scheme == 2:
        one = 10 * ws + 4 * ms
        return 0, f * one + (1 - f) * (one + 12 * ss)
    if scheme == 0 and storage != 1:
        return a_vel, float(split)
    return a_vel, f * unsplit + (1 - f) * split


def damped_counts(cfg, npml):
    """Updated cells by the number of directions (0..3) whose stress PML factor is not 1: per
    axis of n cells the padded extent n + 2 npml less the 4-cell ring on each side, of which
    n + 1 cells (x = npml .. npml + n) have factor 1."""
    c = [1]
    for k in ("nx", "ny", "nz"):
        inner = cfg[k] + 1
        damped = cfg[k] + 2 * npml - 8 - inner
        c = [(c[j] if j < len(c) else 0) * inner + (c[j - 1] * damped if j >= 1 else 0) for j in range(len(c) + 1)]
    return c


def dir_split_bytes(cfg, npml, unsplit, split, per_dir):
    """Mean stress bytes per updated point of scheme 3: unsplit + per_dir per damped direction,
    the split cell where all three are damped."""
    c = damped_counts(cfg, npml)
    tot = c[0] * unsplit + c[1] * (unsplit + per_dir) + c[2] * (unsplit + 2 * per_dir) + c[3] * split
    return tot / sum(c)


def sampled_steps(nt, stride):
    """Time steps of one shot whose kernels a3d_set_kernel_timing(ctx, stride) times."""
    return (nt + stride - 1) // stride if stride > 1 else nt


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(cfg, target_s=15.0):
    """Times the CPU oracle's FP32 time loop on a bounded number of steps."""
    spec = importlib.util.spec_from_file_location("oracle_binding", os.path.join(ROOT, "oracle", "binding.py"))
    orc = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(orc)
    if not os.path.exists(orc.LIB_PATH):
        return None
    orc.load()
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    shot = orc.make_shot("vti", cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["dy"], cfg["dz"], cfg["sx"],
                         cfg["sy"], cfg["sz"], cfg["nt"], cfg["dt"], cfg["fm"], nthreads=cores)
    m = dict(vel=cfg["vel"], rho=cfg["rho"], eps=cfg["eps"], delta=cfg["delta"])
    t2 = orc.time_steps(shot, 2, **m)
    nsteps = int(max(2, min(200, target_s / max(t2 / 2, 1e-3))))
    secs = orc.time_steps(shot, nsteps, **m)
    npts = (cfg["nx"] + 2 * 20 - 8) * (cfg["ny"] + 2 * 20 - 8) * (cfg["nz"] + 2 * 20 - 8)
    return dict(value=npts * nsteps / secs / 1e9, unit="Gpt/s", cores=cores, kind="port",
                sample=f"{nsteps} time steps of the same VTI shot (FP32 oracle, OpenMP), {secs:.1f} s")


def precision_study(cfg):
    """The 400x400x300 VTI shot in FP32, FP16 wavefield storage and FP64; errors vs FP64."""
    from a3dfd import Modeller, make_desc
    nz = cfg["nz"]
    out = {}
    for name, storage, scheme in (("fp64", 2, 0), ("fp32", 0, 0), ("fp32_unsplit", 0, 1), ("fp32_dirsplit", 0, 3),
                                  ("fp32_onefield", 0, 4), ("fp16", 1, 1), ("fp16_dirsplit", 1, 3),
                                  ("fp16_onefield", 1, 4)):
        d = make_desc("vti", cfg["nx"], cfg["ny"], nz, cfg["dx"], cfg["dy"], cfg["dz"], cfg["sx"], cfg["sy"],
                      cfg["sz"], cfg["nt"], cfg["dt"], cfg["fm"], storage=storage, scheme=scheme,
                      snap_interval=1000, snap_skip0=1)
        with Modeller(d) as mdl:
            mdl.set_model(vel=cfg["vel"], rho=cfg["rho"], eps=cfg["eps"], delta=cfg["delta"])
            mdl.run(shot=False)                                  # warm-up (and autotune)
            shot, snaps, ms = mdl.run(nsnap_max=3)               # it = 1000, 2000 (and timing)
        npts = (cfg["nx"] + 32) * (cfg["ny"] + 32) * (nz + 32)
        out[name] = dict(shot=shot, snap=snaps[0], snap2=snaps[1], ms=ms, gpts=npts * cfg["nt"] / ms / 1e6)
        log(f"bench: {name}: {ms:.1f} ms, {out[name]['gpts']:.2f} Gpt/s")
    ref = out["fp64"]
    res = {"workload": "vti_400x400x300_nt%d" % cfg["nt"], "reference": "GPU FP64 instance"}
    rl = lambda a, b: float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b.astype(np.float64)))  # noqa
    for name in ("fp32", "fp32_unsplit", "fp32_dirsplit", "fp32_onefield", "fp16", "fp16_dirsplit", "fp16_onefield"):
        o = out[name]
        res[name] = {"shot_rel_l2_vs_fp64": rl(o["shot"], ref["shot"]),
                     "snapshot_it1000_rel_l2_vs_fp64": rl(o["snap"], ref["snap"]),
                     "snapshot_it2000_rel_l2_vs_fp64": rl(o["snap2"], ref["snap2"]),
                     "loop_ms": round(o["ms"], 1), "Gpt_s": round(o["gpts"], 2)}
    a, b = out["fp32_unsplit"], out["fp32"]
    res["fp32_unsplit_vs_fp32"] = {"shot_rel_l2": rl(a["shot"], b["shot"]), "snapshot_it1000": rl(a["snap"], b["snap"]),
                                   "snapshot_it2000": rl(a["snap2"], b["snap2"])}
    for name in ("fp32_dirsplit", "fp32_onefield"):
        a = out[name]
        res[name + "_vs_fp32"] = {"shot_rel_l2": rl(a["shot"], b["shot"]), "snapshot_it1000": rl(a["snap"], b["snap"]),
                                  "snapshot_it2000": rl(a["snap2"], b["snap2"])}
    res["fp64"] = {"loop_ms": round(ref["ms"], 1), "Gpt_s": round(ref["gpts"], 2)}
    print(json.dumps(res), flush=True)


def accuracy_check(cfg, mdl, desc):
    """Errors of the benchmarked shot, measured after the timed region: its record and the
    snapshots at it = 1000 and 2000 against the same shot run by the GPU FP64 instance, and
    (for schemes other than 0) against scheme 0, which is bitwise the FP32 oracle restatement."""
    from a3dfd import Modeller, make_desc
    m = dict(vel=cfg["vel"], rho=cfg["rho"], eps=cfg["eps"], delta=cfg["delta"])
    d = type(desc).from_buffer_copy(desc)
    d.snap_interval, d.snap_skip0 = 1000, 1
    t0 = time.time()
    shot, snaps, _ = mdl.run(desc=d, nsnap_max=2)
    runs = {"this": (shot, snaps)}
    for name, storage, scheme in (("fp64", 2, 0), ("scheme0", 0, 0)):
        if name == "scheme0" and desc.scheme == 0:
            continue
        r = make_desc("vti", cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["dy"], cfg["dz"], cfg["sx"], cfg["sy"],
                      cfg["sz"], cfg["nt"], cfg["dt"], cfg["fm"], storage=storage, scheme=scheme, snap_interval=1000,
                      snap_skip0=1)
        with Modeller(r) as other:
            other.set_model(**m)
            runs[name] = other.run(nsnap_max=2)[:2]
    rl = lambda a, b: float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b.astype(np.float64)))  # noqa
    out = {}
    for ref in ("fp64", "scheme0"):
        if ref not in runs:
            continue
        (a, sa), (b, sb) = runs["this"], runs[ref]
        out[f"vs_{ref}"] = {"shot_rel_l2": rl(a, b), "snapshot_it1000_rel_l2": rl(sa[0], sb[0]),
                            "snapshot_it2000_rel_l2": rl(sa[1], sb[1])}
    log(f"bench: accuracy check {time.time() - t0:.1f} s: {out}")
    return out


def batched(args):
    """64 TTI shots with uneven aperture windows cut from one 600x600x300 model, nt=2500,
    40-cell PML: 64 sequential per-window runs (mode 0) against the batched sweeps with the
    time loop captured in hipGraphs (mode 2) and the persistent kernel (mode 3, --persist-ty
    tile rows, 250 steps per launch, launches captured in one hipGraph). Prints shots/hour."""
    from a3dfd import Batch, Window, make_desc, models
    os.environ.setdefault("A3D_PROGRESS", "1")
    t0 = time.time()
    cfg = models.cfg4_model()
    nt = args.nt if args.nt != 3000 else cfg["nt"]
    wins = models.uneven_windows(args.nwin, cfg["nx"], cfg["ny"], cfg["nz"])
    log(f"bench: TTI model and {len(wins)} windows generated in {time.time() - t0:.1f} s")
    desc = make_desc("tti", cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["dy"], cfg["dz"], 1, 1, 1, nt,
                     cfg["dt"], cfg["fm"], scheme=args.scheme)
    res = {}
    with Batch(desc, [Window(*w) for w in wins]) as b:
        b.set_model(vel=cfg["vel"], eps=cfg["eps"], delta=cfg["delta"], theta=cfg["theta"], phi=cfg["phi"])
        b.set_tti_angles(args.tti_angles)
        t1 = time.time()
        ncls = b.tune()
        info = b.info()
        log(f"bench: tuned {ncls} size classes in {time.time() - t1:.1f} s; updated points/step {info['updated_points']}")
        b.set_graph_chunk(250)
        modes = [int(x) for x in args.batch_modes.split(",")]
        for mode in modes:
            if mode == 3:
                b.set_persist(args.persist_ty, 250, args.persist_group)
                if args.persist_zseg:
                    b.set_zseg(*[int(v) for v in args.persist_zseg.split(",")])
            t1 = time.perf_counter()
            _, ms = b.run(mode=mode, shots=False)
            wall = time.perf_counter() - t1
            res[mode] = dict(loop_s=round(ms / 1e3, 2), wall_s=round(wall, 2),
                             shots_per_hour=round(len(wins) * 3600e3 / ms, 1),
                             Gpt_s=round(info["updated_points"] * nt / ms / 1e6, 2))
            log(f"bench: mode {mode}: {res[mode]}")
    line = {
        "metric": "shots/hour, 64 uneven-window TTI shots (batched hipGraph vs sequential per-window runs)",
        "value": None, "unit": "shots/h", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic seeded TTI model (layered + noise, smooth eps/delta/tilt/azimuth)",
        "config": {"workload": f"tti_600x600x300_{len(wins)}win_nt{nt}", "npml": 40, "nt": nt, "scheme": args.scheme,
                   "tti_angles": args.tti_angles,
                   "windows": "uniform in [160..320]^2 x [200..300] cells, seed 4",
                   "updated_points_per_step": info["updated_points"], "size_classes": ncls,
                   "z_segments": list(info["zseg"])},
        "sequential": res.get(0), "batched_graph": res.get(2), "persistent": res.get(3),
        "persistent_tile_rows": args.persist_ty if 3 in res else None,
    }
    if 0 in res:
        for mode, key in ((2, "speedup"), (3, "speedup_persistent")):
            if mode in res:
                line[key] = round(res[0]["loop_s"] / res[mode]["loop_s"], 3)
    best = max((m for m in (2, 3) if m in res), key=lambda m: res[m]["shots_per_hour"], default=None)
    line["value"] = res[best]["shots_per_hour"] if best is not None else None
    print(json.dumps(line), flush=True)


def learned_boundary(scheme=4, nt=2000, n=256, seed=95000, heldout=12):
    """The learned-boundary half of the metric, measured in the bench run: the VTI 256^3
    configuration (d = 10 m, dt = 0.8 ms, nt 2000, seeded study model and source) as per-shot
    runs with the 20-cell PML, with the 6-cell strip with its head closed (plain ramp) and with
    the committed learned strip (its head mode and network period), all under the same stress
    scheme. shots/hour from each run's time loop (best of two, no snapshots); reflected
    energy and shot-record error against an extended-domain FP64 run of the same shot
    (snapshots every 100 steps)."""
    from a3dfd import STRIP_NB, models, strip
    t0 = time.time()
    m = models.strip_model(seed, n)
    src = models.strip_source(seed, n)
    params, amp = strip.load_params()
    head, period = strip.load_head(), strip.load_period()
    every = 100
    ref_shot, ref_snaps, _ = strip.run_case(m, src, nt, every, "extended", margin=64)
    out = {}
    for name, b in (("pml20", "pml"), ("strip_plain", "strip0"), ("strip_learned", "strip")):
        kw = dict(amp=amp, scheme=scheme, head=head, period=period)
        shot, snaps, _ = strip.run_case(m, src, nt, every, b, params, **kw)
        ms = min(strip.run_case(m, src, nt, nt + 1, b, params, **kw)[2] for _ in range(2))
        out[name] = {"shots_per_hour": round(3600e3 / ms, 1), "loop_ms": round(ms, 1),
                     "reflected_db": round(float(strip.energy_db(snaps, ref_snaps)), 2),
                     "shot_rel_l2_vs_fp64": round(strip.rel_l2(shot, ref_shot), 6)}
    judged = lambda k: {"reflected_db": out[k]["reflected_db"], "shot_rel_l2": out[k]["shot_rel_l2_vs_fp64"],   # noqa
                        "loop_ms": out[k]["loop_ms"]}
    res = {"workload": f"vti_{n}^3_nt{nt}", "scheme": scheme, "network": {"head": head, "period": period,
                                                                           "ramp_amplitude": amp},
           "runs": out,
           "speedup_learned_vs_pml20": round(out["strip_learned"]["shots_per_hour"] / out["pml20"]["shots_per_hour"], 3),
           # updated points per step, 20-cell PML over the 6-cell strip: the ratio a boundary that costs
           # nothing per step could reach
           "volume_bound": round(((n + 2 * 20 - 8) / (n + 2 * STRIP_NB - 8)) ** 3, 3),
           "gain_db_over_plain_strip": round(out["strip_plain"]["reflected_db"] - out["strip_learned"]["reflected_db"], 2),
           "choice": strip.choose([{"strip": judged("strip_learned"), "pml": judged("pml20")}]),
           "judge": "extended domain (+64 cells per face, 40-cell PML) in FP64"}
    # the same comparison on held-out study models (seeds 90000 + k, 128^3, nt 800): the one
    # model above says little about the network's gain on its own
    seeds = [90000 + k for k in range(heldout)]
    rows = strip.judge(seeds, params, n=128, nt=800, snap_every=20, amp=amp, head=head, period=period,
                           scheme=scheme, log=log) if heldout else []
    if rows:
        res["held_out_in_this_run"] = heldout_summary(rows, seeds, 128, 800, scheme, head, period, amp, with_rows=False)
    log(f"bench: learned boundary {time.time() - t0:.1f} s: {res}")
    return res


def heldout_summary(rows, seeds, n, nt, scheme, head, period, amp, with_rows=True):
    """Mean gain of the learned strip over its closed head (with standard error), reflected
    energies, share of models at or below -40 dB and loop times over strip.judge rows."""
    g = np.array([r["strip0"]["reflected_db"] - r["strip"]["reflected_db"] for r in rows])
    db = {k: np.array([r[k]["reflected_db"] for r in rows]) for k in ("pml", "strip0", "strip")}
    ms = {k: np.array([r[k]["loop_ms"] for r in rows]) for k in ("pml", "strip0", "strip")}
    out = {"workload": f"vti_{n}^3_nt{nt}", "models": len(rows), "seeds": [seeds[0], seeds[-1]], "scheme": scheme,
           "network": {"head": head, "period": period, "ramp_amplitude": amp},
           "gain_db_over_closed_head": {"mean": round(float(g.mean()), 3),
                                        "stderr": round(float(g.std(ddof=1) / np.sqrt(len(g))), 3) if len(g) > 1 else None,
                                        "median": round(float(np.median(g)), 3), "min": round(float(g.min()), 2),
                                        "max": round(float(g.max()), 2), "share_positive": round(float((g > 0).mean()), 3)},
           "mean_reflected_db": {k: round(float(v.mean()), 2) for k, v in db.items()},
           "share_at_or_below_-40dB": {k: round(float((v <= -40.0).mean()), 3) for k, v in db.items()},
           "loop_ms_mean": {k: round(float(v.mean()), 1) for k, v in ms.items()},
           "speedup_learned_vs_pml20": round(float(ms["pml"].mean() / ms["strip"].mean()), 3),
           "judge": "extended domain (+64 cells per face, 40-cell PML) in FP64"}
    if with_rows:
        out["rows"] = [{"seed": r["seed"], "fm": r["fm"], "src": r["src"], "pml_db": r["pml"]["reflected_db"],
                        "closed_db": r["strip0"]["reflected_db"], "learned_db": r["strip"]["reflected_db"]}
                       for r in rows]
    return out


def strip_heldout(args):
    """The committed learned strip on --strip-heldout held-out study models (seeds 90000 + k,
    --strip-n cubed, nt --strip-nt, one source each, scheme --scheme): reflected energy of the
    20-cell PML, the strip with its head closed and the learned strip against extended-domain
    FP64 runs, loop times; mean gain over the closed head with its standard error, and the
    share of models at or below -40 dB."""
    from a3dfd import strip
    params, amp = strip.load_params()
    head, period = strip.load_head(), strip.load_period()
    n, nt = args.strip_n, args.strip_nt
    every = max(1, nt // 40)
    seeds = [90000 + k for k in range(args.strip_heldout)]
    rows = strip.judge(seeds, params, n=n, nt=nt, snap_every=every, amp=amp, head=head, period=period,
                       scheme=args.scheme, log=log)
    print(json.dumps(heldout_summary(rows, seeds, n, nt, args.scheme, head, period, amp)), flush=True)


def strip_coef(args):
    """The learned strip's coefficient head (network file --strip-coef, head 2) on the VTI 256^3
    study model (seed 95000, nt --strip-nt), scheme 3: loop time (best of two runs without
    snapshots) and reflected energy / shot error against an extended-domain FP64 run, for the
    network run every k steps (--strip-periods), the same head with a closed network (m = 1),
    the plain strip and the 20-cell PML; then the held-out 64^3 model of the gain test."""
    from a3dfd import models, strip
    with np.load(args.strip_coef, allow_pickle=False) as f:
        net = f["params"].astype(np.float32)
    n, nt, seed = args.strip_n, args.strip_nt, 95000
    m = models.strip_model(seed, n)
    src = models.strip_source(seed, n)
    ref_shot, ref_snaps, _ = strip.run_case(m, src, nt, 100, "extended", margin=64)
    rows = {}

    def one(name, b, params=None, head=0, period=1):
        kw = dict(amp=10.0, scheme=args.scheme, head=head, period=period)
        shot, snaps, _ = strip.run_case(m, src, nt, 100, b, params, **kw)
        ms = min(strip.run_case(m, src, nt, nt + 1, b, params, **kw)[2] for _ in range(2))
        rows[name] = {"loop_ms": round(ms, 1), "reflected_db": round(float(strip.energy_db(snaps, ref_snaps)), 2),
                      "shot_rel_l2_vs_fp64": round(strip.rel_l2(shot, ref_shot), 5)}
        log(f"bench: {name}: {rows[name]}")

    one("pml20", "pml")
    one("strip_plain_damping_head_shut", "strip0")
    for k in [int(x) for x in args.strip_periods.split(",")]:
        one(f"coefficient_head_period{k}", "strip", net, head=strip.COEF_HEAD, period=k)
    closed = strip.coef_params(strip.coef_closed_theta(), np.zeros(strip.NPARAMS, np.float32))
    one("coefficient_head_closed_period8", "strip", closed, head=strip.COEF_HEAD, period=8)
    n2, nt2, every, s2 = 64, 400, 20, 90100
    m2, src2 = models.strip_model(s2, n2), models.strip_source(s2, n2)
    _, ref2, _ = strip.run_case(m2, src2, nt2, every, "extended", margin=32)
    held = {"plain_strip_db": round(float(strip.energy_db(strip.run_case(m2, src2, nt2, every, "strip0", amp=10.0)[1],
                                                           ref2)), 2)}
    for k in [int(x) for x in args.strip_periods.split(",")]:
        sn = strip.run_case(m2, src2, nt2, every, "strip", net, amp=10.0, head=strip.COEF_HEAD, period=k)[1]
        held[f"coefficient_head_period{k}_db"] = round(float(strip.energy_db(sn, ref2)), 2)
    print(json.dumps({"workload": f"vti_{n}^3_nt{nt}", "network": os.path.basename(args.strip_coef), "rows": rows,
                      "held_out_64^3_seed90100_nt400": held,
                      "judge": "extended domain (+64 cells per face, 40-cell PML) in FP64"}), flush=True)


def strip_bench(args):
    """Learned absorbing strip on the 256^3 VTI study model (d = 10 m, dt = 0.8 ms, nt = 2000):
    shots/hour of per-shot runs with the 20-cell PML against the 6-cell learned strip, each under
    stress schemes 0 (reference formulation), 1 and 3, and each boundary's reflected energy and
    shot-record error against an extended-domain FP64 run. value: the learned strip's best
    shots/hour over the three schemes."""
    from a3dfd import models, strip
    t0 = time.time()
    n, nt, every = args.strip_n, args.strip_nt, 50
    m = models.strip_model(95000, n)
    src = models.strip_source(95000, n)
    params, amp = strip.load_params()
    head = strip.load_head()
    if params is None:
        log("bench: no trained strip network found; the strip runs with its damping head shut")
        params = strip.closed_params()
    log(f"bench: strip study model {n}^3 generated in {time.time() - t0:.1f} s; strip amplitude {amp}")
    ref_shot, ref_snaps, ext_ms = strip.run_case(m, src, nt, every, "extended", margin=64)
    log(f"bench: extended FP64 reference run {ext_ms:.0f} ms")
    out = {}
    for name, b, sch in (("pml20", "pml", 0), ("strip_closed", "strip0", 0), ("strip_learned", "strip", 0),
                         ("pml20_scheme1", "pml", 1), ("strip_learned_scheme1", "strip", 1),
                         ("pml20_scheme3", "pml", 3), ("strip_learned_scheme3", "strip", 3)):
        shot, snaps, ms = strip.run_case(m, src, nt, every, b, params, amp=amp, scheme=sch, head=head)
        _, _, ms2 = strip.run_case(m, src, nt, every, b, params, amp=amp, scheme=sch, head=head)   # steady timing
        out[name] = {"loop_ms": round(min(ms, ms2), 1), "shots_per_hour": round(3600e3 / min(ms, ms2), 1),
                     "reflected_db": round(float(strip.energy_db(snaps, ref_snaps)), 2),
                     "shot_rel_l2_vs_fp64": strip.rel_l2(shot, ref_shot)}
        log(f"bench: {name}: {out[name]}")
    base = out["pml20"]["shots_per_hour"]
    judged = lambda k: {"reflected_db": out[k]["reflected_db"], "shot_rel_l2": out[k]["shot_rel_l2_vs_fp64"],   # noqa
                        "loop_ms": out[k]["loop_ms"]}
    choice = strip.choose([{"strip": judged("strip_learned_scheme3"), "pml": judged("pml20_scheme3")}])
    sph = lambda k: out[k]["shots_per_hour"]   # noqa: E731
    same = {str(k): round(sph(f"strip_learned{sfx}") / sph(f"pml20{sfx}"), 3)
            for k, sfx in ((0, ""), (1, "_scheme1"), (3, "_scheme3"))}
    best_strip = max(sph(k) for k in ("strip_learned", "strip_learned_scheme1", "strip_learned_scheme3"))
    best_pml = max(sph(k) for k in ("pml20", "pml20_scheme1", "pml20_scheme3"))
    line = {
        "metric": "shots/hour, learned 6-cell absorbing strip vs 20-cell PML (VTI 256^3, per-shot runs)",
        "value": best_strip, "unit": "shots/h", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic seeded VTI model (dipping layers, smooth eps/delta)",
        "config": {"workload": f"vti_{n}^3_nt{nt}_strip", "d": 10.0, "dt": 8e-4, "strip_amp": amp,
                   "judge": "extended domain (+64 cells per face, 40-cell PML) in FP64"},
        "runs": out, "speedup_vs_pml20": round(out["strip_learned"]["shots_per_hour"] / base, 3),
        "speedup_scheme1_vs_pml20": round(out["strip_learned_scheme1"]["shots_per_hour"] / base, 3),
        "speedup_same_scheme": same,
        "speedup_best_strip_vs_pml20_reference": round(best_strip / base, 3),
        "speedup_best_strip_vs_best_pml20": round(best_strip / best_pml, 3),
        "choice": choice,
    }
    print(json.dumps(line), flush=True)


def batched_strip(args):
    """Shots/hour of the learned-strip batched form against 20-cell-PML per-shot runs: --nwin
    shots (default 8) on the 256^3 VTI study model (d = 10 m, dt = 0.8 ms, nt = --strip-nt),
    sources at seeded positions, every window the full model. Runs (stress scheme --scheme
    unless noted): the strip with the committed network (its head and period) batched with
    the time loop in hipGraphs (mode 2) and as per-shot runs (mode 0); the 20-cell PML as
    per-shot runs under --scheme and scheme 0 (reference formulation) and batched in graphs."""
    from a3dfd import STRIP_NB, Batch, Window, make_desc, models, strip
    os.environ.setdefault("A3D_PROGRESS", "1")
    n, nt = args.strip_n, args.strip_nt
    nshot = 8 if args.nwin == 64 else min(args.nwin, 16)   # --nwin defaults to the TTI batch's 64
    m = models.strip_model(95000, n)
    fm = models.strip_source(95000, n)["fm"]
    srcs = [models.strip_source(95000 + k, n) for k in range(nshot)]
    wins = [Window(0, 0, 0, n, n, n, s["sx"], s["sy"], s["sz"]) for s in srcs]
    params, amp = strip.load_params()
    head, period = strip.load_head(), strip.load_period()
    if params is None:
        params, head, period = strip.closed_params(), 0, 1
    model = dict(vel=m["vel"], rho=m["rho"], eps=m["eps"], delta=m["delta"])
    res = {}

    def run(name, nb, scheme, mode, net):
        d = make_desc("vti", n, n, n, m["dx"], m["dy"], m["dz"], 1, 1, 1, nt, m["dt"], fm, nb=nb, scheme=scheme)
        with Batch(d, wins) as b:
            b.set_model(**model)
            if net is not None:
                b.set_strip_net(net, amp=amp, head=head, period=period)
            b.tune()
            b.set_graph_chunk(250)
            _, ms = b.run(mode=mode, shots=False)
        res[name] = {"loop_s": round(ms / 1e3, 2), "shots_per_hour": round(nshot * 3600e3 / ms, 1)}
        log(f"bench: {name}: {res[name]}")

    sch = args.scheme
    run("strip_batched_graph", STRIP_NB, sch, 2, params)
    run("strip_per_shot", STRIP_NB, sch, 0, params)
    run("pml20_per_shot", 20, sch, 0, None)
    run("pml20_per_shot_scheme0", 20, 0, 0, None)
    run("pml20_batched_graph", 20, sch, 2, None)
    sph = lambda k: res[k]["shots_per_hour"]   # noqa: E731
    line = {
        "metric": "shots/hour, learned-strip batched form vs 20-cell-PML per-shot runs (VTI 256^3)",
        "value": sph("strip_batched_graph"), "unit": "shots/h", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic seeded VTI study model, sources at seeded positions",
        "config": {"workload": f"vti_{n}^3_nt{nt}_{nshot}shots_strip_batched", "strip_amp": amp,
                   "fm": round(fm, 2), "scheme": args.scheme, "network": {"head": head, "period": period}},
        "runs": res,
        "speedup_vs_pml20_per_shot": round(sph("strip_batched_graph") / sph("pml20_per_shot"), 3),
        "speedup_vs_pml20_per_shot_reference_formulation": round(sph("strip_batched_graph")
                                                                 / sph("pml20_per_shot_scheme0"), 3),
        "speedup_vs_pml20_batched": round(sph("strip_batched_graph") / sph("pml20_batched_graph"), 3),
    }
    print(json.dumps(line), flush=True)


STRIP_PROFILES = {
    "none": (0.0,) * 6,
    "uniform_0.002": (0.002,) * 6, "uniform_0.005": (0.005,) * 6, "uniform_0.01": (0.01,) * 6,
    "uniform_0.02": (0.02,) * 6,
    "outer1": (0.02, 0.0, 0.0, 0.0, 0.0, 0.0), "outer2": (0.03, 0.01, 0.0, 0.0, 0.0, 0.0),
    "outer3": (0.05, 0.02, 0.01, 0.0, 0.0, 0.0), "ramp": (0.03, 0.025, 0.02, 0.015, 0.01, 0.005),
    "inner": (0.0, 0.0, 0.0, 0.005, 0.01, 0.01),
    # shaped like the extra per-step damping that doubling a 5-amplitude ramp adds
    # (1 - a2(2D)/a2(D), D = dt d0 (k/6)^2, capped below the head's 0.25), and fractions of it
    "ramp_shaped": (0.2499, 0.2499, 0.2499, 0.22, 0.10, 0.027),
    "ramp_shaped_half": (0.125, 0.125, 0.125, 0.11, 0.05, 0.0135),
    "ramp_shaped_quarter": (0.0625, 0.0625, 0.0625, 0.055, 0.025, 0.00675),
}


def strip_profiles(args):
    """Reflected energy (dB, against the extended-domain FP64 run) of the 6-cell strip with fixed
    damping-head profiles (strip.profile_params: factor per depth, 1 = outermost cell) on top of
    its PML ramp (amplitude --strip-amp, default 10), beside the 20-cell PML, on seeded 128^3
    models (nt 800, every 20th step), for the head mode --strip-head (0 every field, 1 the
    normal-axis PML components); also the ramp alone at amplitudes 5, 7, 10, 15, 20, 30."""
    from a3dfd import models, strip
    n, nt, every = 128, 800, 20
    head, base = args.strip_head, args.strip_amp
    rows = []
    for seed in (10000, 10001, 10002, 90000, 90001, 90002):
        m = models.strip_model(seed, n)
        src = models.strip_source(seed, n)
        _, ref, _ = strip.run_case(m, src, nt, every, "extended", margin=64)
        _, sp, _ = strip.run_case(m, src, nt, every, "pml")
        row = {"seed": seed, "pml20": round(float(strip.energy_db(sp, ref)), 2)}
        for name, gains in STRIP_PROFILES.items():
            _, ss, _ = strip.run_case(m, src, nt, every, "strip", strip.profile_params(gains), amp=base, head=head)
            row[name] = round(float(strip.energy_db(ss, ref)), 2)
        for amp in (5.0, 7.0, 10.0, 15.0, 20.0, 30.0):
            _, ss, _ = strip.run_case(m, src, nt, every, "strip0", amp=amp)
            row[f"ramp_amp{amp:g}"] = round(float(strip.energy_db(ss, ref)), 2)
        rows.append(row)
        log(f"bench: strip profiles seed {seed}: {row}")
    line = {"metric": "reflected energy (dB) of the 6-cell strip with fixed damping-head profiles",
            "unit": "dB", "higher_is_better": False, "n_gpus": 1, "dtype": "f32",
            "config": {"workload": f"vti_{n}^3_nt{nt}_strip_profiles", "strip_amp": base, "head": head,
                       "profiles": {k: list(v) for k, v in STRIP_PROFILES.items()},
                       "judge": "extended domain (+64 cells per face, 40-cell PML) in FP64, every 20th step"},
            "rows": rows}
    print(json.dumps(line), flush=True)


# Algorithmic bytes per updated point (FP32): (velocity sweep or 2nd-order step, stress sweep with
# split fields, stress sweep without them) -- reads + writes of each sweep; OP_BYTES_WZ with w
# updated by the stress sweep (ISO: stress reads P one plane ahead and the old w, writes w).
OP_BYTES = {"2nd": (16, 0, 0), "iso": (32, 48, 28), "vti": (36, 84, 44), "tti": (48, 96, 56)}
OP_BYTES_WZ = {"iso": (24, 56, 32), "vti": (24, 92, 48)}


def operators(args):
    """Every operator at its configuration's grid: 2nd order 100^3 (nt 500, sponge 20), ISO 256^3
    (nt 2000, npml 20), VTI 400x400x300 (nt 3000), TTI 600x600x300 (one full-model shot, nt 500,
    npml 40); the 1st-order ones with both stress schemes. One JSON line per run."""
    from a3dfd import Modeller, make_desc, models
    runs = []
    c1 = models.cfg1()
    runs.append(("2nd", c1, dict(nb=20), 0, c1["nt"]))
    c2 = models.cfg2()
    runs += [("iso", c2, dict(v_mute=2000.0), s, c2["nt"]) for s in (0, 1, 3)]
    c3 = models.cfg3()
    runs += [("vti", c3, {}, s, c3["nt"]) for s in (0, 1, 3, 4)]
    c4 = models.cfg4_model()
    c4.update(sx=301, sy=301, sz=1)
    runs += [("tti", c4, {}, s, 500) for s in (0, 1, 3, 4)]
    runs.append(("tti", c4, {"tti_angles": 1}, 4, 500))   # rotation terms from the angles in the sweeps
    if args.ops:
        runs = [r for r in runs if r[0] in args.ops.split(",")]
    for op, c, kw, scheme, nt in runs:
        keys = {"2nd": ("vel",), "iso": ("vel", "rho"), "vti": ("vel", "rho", "eps", "delta"),
                "tti": ("vel", "eps", "delta", "theta", "phi")}[op]
        kw = dict(kw)
        angles = kw.pop("tti_angles", 0)
        d = make_desc(op, c["nx"], c["ny"], c["nz"], c["dx"], c["dy"], c["dz"], c["sx"], c["sy"], c["sz"], nt,
                      c["dt"], c["fm"], scheme=scheme, **kw)
        with Modeller(d) as mdl:
            mdl.set_model(**{k: c[k] for k in keys})
            err_terms = None
            if angles:
                # the record against the run that reads the terms (same scheme), then the angles
                ref_rec, _, _ = mdl.run()
                mdl.set_tti_angles(1)
                rec, _, _ = mdl.run()
                err_terms = float(np.sqrt(((rec.astype(np.float64) - ref_rec) ** 2).sum()
                                          / max((ref_rec.astype(np.float64) ** 2).sum(), 1e-300)))
            mdl.run(shot=False)                     # warm-up (and autotune)
            mdl.set_kernel_timing(True, stride=8)
            mdl.kernel_times(reset=True)
            _, _, ms_timed = mdl.run(shot=False)
            kms, kn = mdl.kernel_times(reset=True)
            mdl.set_kernel_timing(False)
            # loop times without per-kernel events: host-launched steps, then the loop in hipGraphs
            _, _, ms = mdl.run(shot=False)
            mdl.set_graph(args.graph_steps)
            mdl.run(shot=False)
            _, _, ms_graph = mdl.run(shot=False)
            mdl.set_graph(0)
            together = {}
            if op == "2nd":
                # several shots on the model stepped together (one launch per step over all of them)
                for k in (4, 8):
                    srcs = [(c["sx"] - 20 + 5 * j, c["sy"], c["sz"]) for j in range(k)]
                    mdl.run_shots(srcs, shot=False)
                    together[k] = min(mdl.run_shots(srcs, shot=False)[1] for _ in range(2))
            gi = mdl.grid_info()
            launch = mdl.launch_info()
        npts = gi["updated_points"]
        a0, a_split, a_unsplit = OP_BYTES_WZ[op] if launch["w_sweep"] else OP_BYTES[op]
        nbb = d.nb
        inner = (c["nx"] + 1) * (c["ny"] + 1) * (c["nz"] + 1) if op != "2nd" else npts
        f = min(1.0, inner / npts)
        a1 = a_split if scheme == 0 else f * a_unsplit + (1 - f) * a_split
        if scheme in (3, 4):
            # old P (Q) reads a split cell skips: those the sweep does not take from its w window
            drop = {"iso": 0 if launch["w_sweep"] else 4, "vti": 4 if launch["w_sweep"] else 8, "tti": 8}[op]
            per_dir = (a_split - a_unsplit + drop) / 3
            a_all = a_split
            if scheme == 4 and op != "iso":   # one field per damped direction instead of the (p, q) pair
                per_dir -= 8
                a_all -= 24
            a1 = dir_split_bytes(c, nbb, a_unsplit, a_all, per_dir)
        if angles:   # the two angles (8 B) instead of the four rotation terms (16 B) per sweep
            a0 -= 8
            a1 -= 8
        k0 = kms[0] / max(kn[0], 1)
        k1 = kms[1] / max(kn[1], 1) if kn[1] else 0.0
        g0 = a0 * npts / (k0 * 1e-3) / 1e9
        g1 = a1 * npts / (k1 * 1e-3) / 1e9 if k1 else 0.0
        gpts = npts * nt / ms / 1e6
        line = {"operator": op, "workload": f"{op}_{c['nx']}x{c['ny']}x{c['nz']}_nt{nt}", "boundary": nbb,
                "scheme": scheme, "Gpt_s": round(gpts, 2), "loop_ms": round(ms, 1),
                "graph": {"steps_per_graph": args.graph_steps, "loop_ms": round(ms_graph, 2),
                          "Gpt_s": round(npts * nt / ms_graph / 1e6, 2)},
                "loop_ms_kernel_events": round(ms_timed, 1), "updated_points": npts,
                "launch": launch,
                "kernels": {("k_step2" if op == "2nd" else "k_vel"): {"avg_ms": round(k0, 4), "bytes_per_point": a0,
                                                                       "GBs": round(g0, 1),
                                                                       "frac": round(g0 / HBM_PEAK_GBS, 4)}},
                "whole_step_GBs": round((a0 + a1) * gpts, 1),
                "whole_step_frac": round((a0 + a1) * gpts / HBM_PEAK_GBS, 4)}
        if together:
            line["shots_together"] = {str(k): {"loop_ms": round(v, 2), "Gpt_s": round(k * npts * nt / v / 1e6, 2),
                                               "GBs": round(a0 * k * npts * nt / v / 1e6, 1),
                                               "frac": round(a0 * k * npts * nt / v / 1e6 / HBM_PEAK_GBS, 4)}
                                      for k, v in together.items()}
        if angles:
            # 8 bytes of angles instead of the 16-byte rotation terms in each sweep
            line["tti_angles"] = {"record_rel_l2_vs_terms_read": err_terms, "bytes_per_point_saved_per_sweep": 8}
        if op != "2nd":
            line["kernels"]["k_stress"] = {"avg_ms": round(k1, 4), "bytes_per_point": round(a1, 2),
                                           "GBs": round(g1, 1), "frac": round(g1 / HBM_PEAK_GBS, 4)}
        log(f"bench: {line['workload']} scheme {scheme}: {gpts:.2f} Gpt/s")
        print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed shots")
    ap.add_argument("--warmup", type=int, default=1, help="untimed shots")
    ap.add_argument("--nt", type=int, default=3000, help="time steps per shot")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-strip", action="store_true",
                    help="skip the learned-boundary block (strip vs 20-cell PML at VTI 256^3) of the default line")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the accuracy check (the shot against the FP64 instance and scheme 0) after the timing")
    ap.add_argument("--time-stride", type=int, default=8,
                    help="HIP events around the kernels of every N-th time step of the timed shots")
    ap.add_argument("--launch", type=str, default="", help="ty_vel,zseg_vel,ty_stress,zseg_stress (default: autotuned)")
    ap.add_argument("--w-sweep", type=int, default=None, choices=(-1, 0, 1),
                    help="sweep that updates w: 0 velocity, 1 stress, -1 autotuned (default)")
    ap.add_argument("--zchunk", type=int, default=0, help="two-sweep step in z chunks of this many planes (0: off)")
    ap.add_argument("--precision-study", action="store_true",
                    help="run the shot in FP32, FP16-storage and FP64 and print the errors vs FP64")
    ap.add_argument("--storage", type=int, default=0, help="0 FP32 (default), 1 FP16 wavefield storage, 2 FP64")
    ap.add_argument("--scheme", type=int, default=4,
                    help="0: split PML fields everywhere (bitwise reference arithmetic); 1: split only in PML cells; "
                         "2: scheme 1 in one pass per step (bitwise equal to 1); 3: split fields only for the "
                         "damped directions of PML cells; 4: one field per damped direction")
    ap.add_argument("--probe-stress", action="store_true",
                    help="time the VTI stress sweep's memory access pattern replayed without arithmetic")
    ap.add_argument("--probe-copy", action="store_true",
                    help="only run the calibration streams (2 GiB float4 copy and read) and print GB/s")
    ap.add_argument("--batched", action="store_true",
                    help="64 uneven-window TTI shots: sequential per-window runs vs the batched hipGraph form")
    ap.add_argument("--nwin", type=int, default=64, help="windows of the batched run")
    ap.add_argument("--batch-modes", type=str, default="0,2,3",
                    help="with --batched: modes run (0 sequential, 2 batched graphs, 3 persistent kernel)")
    ap.add_argument("--tti-angles", type=int, default=1, choices=(0, 1),
                    help="with --batched: the TTI sweeps evaluate the rotation terms from the angles (1) or read "
                         "them (0)")
    ap.add_argument("--persist-ty", type=int, default=4, choices=(4, 8),
                    help="with --batched: tile rows of the persistent kernel")
    ap.add_argument("--persist-group", type=int, default=1,
                    help="with --batched: consecutive tiles per ticket of the persistent kernel")
    ap.add_argument("--persist-zseg", type=str, default="",
                    help="with --batched: z segments vel,stress for the persistent run (default: the tuned ones)")
    ap.add_argument("--operators", action="store_true",
                    help="every operator at its configuration's grid, both stress schemes; one JSON line each")
    ap.add_argument("--strip", action="store_true",
                    help="learned absorbing strip against the 20-cell PML: shots/hour and reflected energy")
    ap.add_argument("--ops", type=str, default="", help="with --operators: comma list of 2nd,iso,vti,tti (default all)")
    ap.add_argument("--graph-steps", type=int, default=100,
                    help="with --operators: time steps per captured hipGraph of the graph-mode loop")
    ap.add_argument("--batched-strip", action="store_true",
                    help="learned-strip batched form vs 20-cell-PML per-shot runs (shots/hour, VTI 256^3)")
    ap.add_argument("--strip-profiles", action="store_true",
                    help="reflected energy of the strip with fixed damping profiles on top of its PML ramp")
    ap.add_argument("--strip-head", type=int, default=0, choices=(0, 1),
                    help="with --strip-profiles: 0 damping of every field, 1 of the normal-axis PML components")
    ap.add_argument("--strip-amp", type=float, default=10.0,
                    help="with --strip-profiles: PML amplitude of the ramp under the profiles")
    ap.add_argument("--strip-coef", type=str, default="",
                    help="coefficient-head study of the given network file (speed and reflection vs network period)")
    ap.add_argument("--strip-periods", type=str, default="1,4,8,16")
    ap.add_argument("--strip-heldout", type=int, default=0,
                    help="judge the committed learned strip on this many held-out study models (--strip-n, --strip-nt)")
    ap.add_argument("--strip-n", type=int, default=256)
    ap.add_argument("--strip-nt", type=int, default=2000)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs every shot on a single MI355X "
              "(multi-GPU shot sharding is a design note, not part of this build).", file=sys.stderr)
        sys.exit(2)

    from a3dfd import Modeller, make_desc
    from a3dfd import models
    if args.probe_copy:
        from a3dfd._lib import probe_bandwidth
        c, r = probe_bandwidth(0, 2 << 30, 10)
        print(json.dumps({"probe": "2 GiB buffers, 10 reps: copy one float4 per thread (256 per workgroup), read four "
                                   "per thread (1024 per workgroup), non-temporal loads and stores; before them the "
                                   "grid-stride float4 copy (1024 x 256 threads) the byte counters are calibrated on",
                          "copy_GBs_read_plus_write": round(c, 1), "read_only_GBs": round(r, 1)}), flush=True)
        return

    if args.probe_stress:
        from a3dfd._lib import probe_stress_pattern
        res = probe_stress_pattern(440, 440, 340)
        print(json.dumps({"probe": "VTI stress sweep access pattern replayed without arithmetic, padded "
                          "440x440x340 (400x400x300 + 20-cell PML + 4-cell ring), 64x4 tiles, 20 launches each",
                          "patterns": {k: {"ms": round(v[0], 4), "algorithmic_GBs": round(v[1], 1)}
                                       for k, v in res.items()}}), flush=True)
        return

    if args.batched:
        return batched(args)
    if args.strip:
        return strip_bench(args)
    if args.strip_coef:
        return strip_coef(args)
    if args.strip_profiles:
        return strip_profiles(args)
    if args.strip_heldout:
        return strip_heldout(args)
    if args.batched_strip:
        return batched_strip(args)
    if args.operators:
        return operators(args)
    t0 = time.time()
    cfg = models.cfg3()
    cfg["nt"] = args.nt
    log(f"bench: model generated in {time.time() - t0:.1f} s")
    if args.precision_study:
        return precision_study(cfg)
    desc = make_desc("vti", cfg["nx"], cfg["ny"], cfg["nz"], cfg["dx"], cfg["dy"], cfg["dz"], cfg["sx"],
                     cfg["sy"], cfg["sz"], cfg["nt"], cfg["dt"], cfg["fm"], storage=args.storage,
                     scheme=args.scheme)
    mdl = Modeller(desc, device=0)
    mdl.set_model(vel=cfg["vel"], rho=cfg["rho"], eps=cfg["eps"], delta=cfg["delta"])
    if args.zchunk:
        mdl.set_zchunk(args.zchunk)
    if args.launch:
        mdl.set_launch(*[int(x) for x in args.launch.split(",")])
    if args.w_sweep is not None:
        mdl.set_w_sweep(args.w_sweep)
    gi = mdl.grid_info()
    npts = gi["updated_points"]
    log(f"bench: padded {gi['nnx']}x{gi['nny']}x{gi['nnz']}, updated points {npts}, launch {mdl.launch_info()}")

    for w in range(args.warmup):
        _, _, ms = mdl.run(shot=False)
        log(f"bench: warmup shot {w}: {ms:.1f} ms loop, launch {mdl.launch_info()}")
    launch = mdl.launch_info()

    mdl.set_kernel_timing(True, stride=args.time_stride)
    mdl.kernel_times(reset=True)
    loop_ms = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        _, _, ms = mdl.run(shot=False)
        loop_ms.append(ms)
        log(f"bench: shot {k}: {ms:.1f} ms loop, {npts * args.nt / ms / 1e6:.2f} Gpt/s")
    t_wall = time.perf_counter() - t_start
    kms, kn = mdl.kernel_times(reset=True)
    mdl.set_kernel_timing(False)
    acc = None
    if args.nt == 3000 and args.storage != 2 and not args.no_check:
        acc = accuracy_check(cfg, mdl, desc)
    mdl.close()
    lb = None
    if not args.no_strip:
        try:
            lb = learned_boundary()
        except Exception as e:   # informative; never fail the headline line on it
            log(f"bench: learned boundary failed: {e}")

    total_loop_s = sum(loop_ms) / 1e3
    gpts = npts * args.nt * args.steps / total_loop_s / 1e9
    steps_timed = args.steps * sampled_steps(args.nt, args.time_stride)
    avg_vel = kms[0] / steps_timed   # per time step (a chunked step launches each sweep once per chunk)
    avg_str = kms[1] / steps_timed
    main_kernel = "k_fused_vti" if args.scheme == 2 else "k_stress<VTI>"
    wz = launch["w_sweep"]
    a_vel, a_str = algorithmic_bytes(cfg, args.storage, args.scheme, npts, wz=wz)
    ach_str = a_str * npts / (avg_str * 1e-3) / 1e9
    ach_vel = a_vel * npts / (avg_vel * 1e-3) / 1e9 if kn[0] else 0.0
    step_gbs = (a_vel + a_str) * gpts
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_vti_400x400x300.json")
    if os.path.exists(tpath) and args.storage == 0:
        with open(tpath) as f:
            traffic = json.load(f).get(f"scheme{args.scheme}" + ("_wz" if wz else ""), {}).get("stress_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (FP32 oracle) ...")
        try:
            cpu = cpu_baseline(cfg)
        except Exception as e:  # the baseline is informative; never fail the GPU line on it
            log(f"bench: CPU baseline failed: {e}")

    ref_form = None
    if args.storage != 1:
        rv, rs = algorithmic_bytes(cfg, args.storage, 0, npts, wz=0)
        ref_form = {"bytes_per_point": rv + rs, "equivalent_GBs": round(gpts * (rv + rs), 1),
                    "frac": round(gpts * (rv + rs) / HBM_PEAK_GBS, 4)}
    line = {
        "metric": "Gpt-updates/s (VTI 400x400x300 FP32 shot) & % HBM roofline",
        "value": round(gpts, 3),
        "unit": "Gpt/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_wall / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {0: "f32", 1: "f32 (f16 wavefield storage)", 2: "f64"}[args.storage],
        "data": "synthetic seeded VTI model (layered 1500-5000 m/s + salt body, Gardner density, "
                "smooth epsilon/delta); inputs resident in HBM",
        "config": {"workload": "vti_400x400x300_nt%d" % args.nt, "nx": cfg["nx"], "ny": cfg["ny"],
                   "nz": cfg["nz"], "npml": 20, "nt": args.nt, "dx": cfg["dx"], "dt": cfg["dt"],
                   "updated_points": npts, "step": "one shot", "parallelism": "single GPU",
                   "launch": launch, "scheme": args.scheme,
                   "parity": {0: "bitwise equal to the FP32 oracle (the uncontracted restatement of the reference's "
                                 "arithmetic; its build with contracted multiply-adds, the reference compiler's "
                                 "default, stays within 1e-5 on records)",
                              1: "split PML fields in PML cells only: same equations, other rounding; error vs "
                                 "scheme 0 (bitwise the FP32 oracle) measured below",
                              2: "scheme 1 in one pass: bitwise equal to scheme 1",
                              3: "split PML fields for damped directions only: same equations, other rounding; "
                                 "error vs scheme 0 (bitwise the FP32 oracle) measured below",
                              4: "one PML field per damped direction (the reference's split p and q fields are "
                                 "per-point multiples of it): same equations, other rounding; error vs scheme 0 "
                                 "(bitwise the FP32 oracle) measured below"}[args.scheme],
                   "errors_measured_in_this_run": acc},
        "roofline": {"bound": "hbm", "kernel": main_kernel, "achieved": round(ach_str, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach_str / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "bytes_per_point": round(a_str, 2),
                     "avg_launch_ms": round(avg_str, 4), "launches": kn[1], "zchunk": args.zchunk,
                     "timed_every": args.time_stride},
        "kernels": {"k_vel<VTI>": {"avg_launch_ms": round(avg_vel, 4), "achieved_GBs": round(ach_vel, 1),
                                   "bytes_per_point": a_vel, "launches": kn[0]},
                    "whole_step": {"achieved_GBs": round(step_gbs, 1),
                                   "frac": round(step_gbs / HBM_PEAK_GBS, 4), "bytes_per_point": round(a_vel + a_str, 2)},
                    # the same updates counted at the bytes of the reference's own formulation (split PML
                    # fields everywhere, w in the velocity sweep): the bandwidth it would need for this rate
                    "whole_step_reference_formulation": ref_form},
        "cpu_baseline": cpu,
        "learned_boundary": lb,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
