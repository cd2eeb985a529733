#!/usr/bin/env python3
"""Trains the learned absorbing strip on the GPU and judges it.

1. Teacher-forced samples from seeded random VTI models (a 40-cell-PML teacher).
2. Training with the library's GPU kernels (Adam, cosine-decayed learning rate).
3. Judge on held-out seeded models: reflected energy and shot-record error of the
   learned strip, the same strip without the network and the 20-cell PML, all
   against an extended-domain FP64 run; the fallback rule picks 'strip' only at
   <= -40 dB on every held-out model.

Writes the parameters (and the training/judge summary) to --out.

Usage: python train_strip.py [--models 48] [--n 128] [--nt 600] [--epochs 30] [--out a3dfd/strip_net.npz]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from a3dfd import strip  # noqa: E402


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--models", type=int, default=48, help="training models (seeds 10000+k)")
    ap.add_argument("--n", type=int, default=128, help="training model size (n^3)")
    ap.add_argument("--nt", type=int, default=600, help="time steps per training shot")
    ap.add_argument("--every", type=int, default=3, help="sample every k-th step")
    ap.add_argument("--keep", type=float, default=0.003, help="fraction of strip cells sampled per sampled step")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--rounds", type=int, default=3,
                    help="data-aggregation rounds: round r >= 1 re-samples with the strip grid running the "
                         "current network for --horizon steps between synchronisations")
    ap.add_argument("--horizon", type=int, default=24)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--heldout", type=int, default=3, help="held-out judge models (seeds 90000+k)")
    ap.add_argument("--judge-nt", type=int, default=800)
    ap.add_argument("--amp", type=float, default=10.0,
                    help="PML amplitude of the strip's 6-cell ramp (10: the reference's constant)")
    ap.add_argument("--tune-amp", action="store_true",
                    help="pick the amplitude with the best worst-case reflection over 8 training models")
    ap.add_argument("--adaptive", action="store_true",
                    help="normal head (PML-like damping) on a weaker ramp: each training model's best uniform "
                         "extra damping (scanned against its extended-domain run) is the network's target")
    ap.add_argument("--coef", action="store_true",
                    help="coefficient head: each (model, source) case's best constant multiplier of the strip's "
                         "PML damping (scanned against the 40-cell-PML teacher's interior) labels that case's "
                         "strip-cell samples; the network learns the multiplier from the local state")
    ap.add_argument("--coef-profile", action="store_true",
                    help="coefficient head, depth profile: per-depth multipliers found by coordinate search on "
                         "the training cases' reflected energy (against their 40-cell-PML teachers), then the "
                         "network trained on the GPU on teacher-forced samples labelled with their depth's "
                         "multiplier")
    ap.add_argument("--period", type=int, default=1, help="coefficient head: network every k steps")
    ap.add_argument("--labels-from", default="",
                    help="coefficient head: take the (model, source) labels m* and their scans from the summary "
                         "of an earlier run's network file instead of scanning again")
    ap.add_argument("--bias", choices=("mean", "pooled"), default="mean",
                    help="coefficient head: output bias at start: the mean label, or the constant multiplier "
                         "with the lowest reflected energy pooled over the labelled cases (from their scans)")
    ap.add_argument("--scheme", type=int, default=3, help="stress scheme of the strip runs")
    ap.add_argument("--out", default=None,
                    help="network file (default: the committed coefficient-head network for --coef, the "
                         "damping-head network otherwise)")
    args = ap.parse_args()
    if args.out is None:
        args.out = strip.NET_PATH if args.coef or args.coef_profile else strip.DAMP_NET_PATH
    if args.adaptive:
        return adaptive(args)
    if args.coef_profile:
        return coef_profile(args)
    if args.coef:
        return coef(args)

    seeds = [10000 + k for k in range(args.models)]
    t0 = time.time()
    amp, amp_db = args.amp, {}
    if args.tune_amp:
        amp, amp_db = strip.tune_amp(seeds[:8], n=args.n, log=log)
    log(f"train_strip: strip PML amplitude {amp} ({time.time() - t0:.1f} s)")
    t0 = time.time()
    X, T, S = strip.collect(seeds, n=args.n, nt=args.nt, every=args.every, keep=args.keep, amp=amp, log=log)
    log(f"train_strip: {len(X)} samples in {time.time() - t0:.1f} s; target rms {np.sqrt((T ** 2).mean(0))}")
    params, hist = strip.train(X, T, epochs=args.epochs, batch=args.batch, lr=args.lr)
    log(f"train_strip: round 0: loss {hist[0]:.5f} -> {hist[-1]:.5f} (target mean square "
        f"{float((T ** 2).sum(1).mean()):.5f})")
    losses = [hist]
    for r in range(1, args.rounds):
        t0 = time.time()
        Xr, Tr, Sr = strip.collect(seeds, n=args.n, nt=args.nt, every=args.every, keep=args.keep,
                                   horizon=args.horizon, params=params, amp=amp, log=None)
        X, T = np.concatenate([X, Xr]), np.concatenate([T, Tr])
        params, hist = strip.train(X, T, epochs=max(5, args.epochs // 2), batch=args.batch, lr=args.lr / 2,
                                   init=params, seed=r)
        losses.append(hist)
        log(f"train_strip: round {r}: +{len(Xr)} samples ({time.time() - t0:.1f} s), loss {hist[0]:.5f} -> "
            f"{hist[-1]:.5f}")
    rows = strip.judge([90000 + k for k in range(args.heldout)], params, n=args.n, nt=args.judge_nt, amp=amp,
                       log=log)
    choice = strip.choose(rows)
    summary = {"samples": int(len(X)), "models": args.models, "n": args.n, "nt": args.nt, "amp": amp,
               "amp_worst_db": {str(k): v for k, v in amp_db.items()},
               "rounds": args.rounds, "horizon": args.horizon,
               "loss_first_epoch": losses[0][0], "loss_last_epoch": losses[-1][-1],
               "target_mean_square": float((T ** 2).sum(1).mean()), "judge": rows, "choice": choice}
    strip.save(params, args.out, amp=np.float32(amp), loss=np.concatenate(losses), summary=json.dumps(summary))
    print(json.dumps(summary), flush=True)


def adaptive(args):
    """Normal head (damping of the normal-axis PML components: a local increase of the PML's own
    damping) on the strip's ramp at amplitude --amp. Per training model, the uniform damping
    factor g* with the lowest reflected energy against the model's extended-domain FP64 run
    (strip.best_uniform_gain) labels that model's teacher-forced feature samples; the network,
    trained on the GPU, learns g* from the local wavefield and model features, so it can damp
    more where the ramp alone lets the wall's reflection through and less where the ramp's own
    onset reflects. Judged on held-out models against the ramp alone at --amp and at 10."""
    seeds = [10000 + k for k in range(args.models)]
    Xs, Ts, labels = [], [], []
    t0 = time.time()
    for seed in seeds:
        gstar, tab = strip.best_uniform_gain(seed, n=args.n, nt=args.judge_nt, amp=args.amp, head=1)
        X, _, _ = strip.collect([seed], n=args.n, nt=args.nt, every=args.every, keep=args.keep, amp=args.amp)
        Xs.append(X)
        Ts.append(strip.adaptive_targets(X, gstar))
        labels.append({"seed": seed, "g_star": gstar, "db": {str(k): v for k, v in tab.items()}, "samples": len(X)})
        log(f"train_strip: seed {seed}: g* {gstar} {tab} ({len(X)} samples, {time.time() - t0:.1f} s)")
    X, T = np.concatenate(Xs), np.concatenate(Ts)
    g_mean = float(np.mean([r["g_star"] for r in labels]))
    r = min(max(g_mean, 1e-4) / 0.25, 0.5)   # the head's factor is 0.25 sigmoid(y)
    init = strip.init_params(0, out_bias=float(np.log(r / (1 - r))))
    params, hist = strip.train(X, T, epochs=args.epochs, batch=args.batch, lr=args.lr, init=init)
    log(f"train_strip: {len(X)} samples, loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    rows = strip.judge([90000 + k for k in range(args.heldout)], params, n=args.n, nt=args.judge_nt, amp=args.amp,
                       head=1, extra_amps=(10.0,), log=log)
    choice = strip.choose(rows)
    summary = {"mode": "adaptive", "head": 1, "samples": int(len(X)), "models": args.models, "n": args.n,
               "nt": args.nt, "amp": args.amp, "labels": labels, "loss_first_epoch": hist[0],
               "loss_last_epoch": hist[-1], "judge": rows, "choice": choice}
    strip.save(params, args.out, amp=np.float32(args.amp), head=np.int32(1), loss=np.asarray(hist),
               summary=json.dumps(summary))
    print(json.dumps(summary), flush=True)


def coef(args):
    """Coefficient head (mode 2). Cases: every source (1-4) of every training model (seeds
    10000 + k). Per case, the best constant damping multiplier m* (strip.coef_label, against
    the case's 40-cell-PML teacher run) labels the case's teacher-forced strip-cell samples
    (features as the head reads them); the network is trained on the GPU (a3d_mlp kernels) to
    give m* from the local state, and judged on held-out models against the plain strip (m = 1)
    and the 20-cell PML, all against extended-domain FP64 runs."""
    seeds = [10000 + k for k in range(args.models)]
    Xs, Ts, labels = [], [], []
    known = {}
    if args.labels_from:
        with np.load(args.labels_from, allow_pickle=False) as f:
            for r in json.loads(str(f["summary"]))["labels"]:
                known[(r["seed"], r["source"])] = (r["m_star"], {float(k): v for k, v in r["db"].items()})
    t0 = time.time()
    for seed in seeds:
        for j, src in enumerate(strip.model_sources(seed, args.n)):
            if (seed, j) in known:
                mstar, tab = known[(seed, j)]
            else:
                case = strip.CoefCase(seed, j, n=args.n, nt=args.nt, snap_every=30, amp=args.amp, scheme=args.scheme)
                try:
                    mstar, tab = strip.coef_label(case)
                finally:
                    case.close()
            X, _, _ = strip.collect([seed], n=args.n, nt=args.nt, every=args.every, keep=args.keep, amp=args.amp,
                                    head=strip.COEF_HEAD, sources=[src], params=strip.coef_params(
                                        strip.coef_closed_theta(), np.zeros(strip.NPARAMS, np.float32)))
            Xs.append(X)
            Ts.append(strip.coef_targets(X, mstar))
            labels.append({"seed": seed, "source": j, "fm": round(src["fm"], 2), "m_star": round(mstar, 4),
                           "db": {str(k): v for k, v in tab.items()}, "samples": len(X)})
        log(f"train_strip: model {seed}: m* {[r['m_star'] for r in labels if r['seed'] == seed]} "
            f"({sum(len(x) for x in Xs)} samples, {time.time() - t0:.1f} s)")
    X, T = np.concatenate(Xs), np.concatenate(Ts)
    mbar = float(np.mean([r["m_star"] for r in labels]))
    m0 = mbar if args.bias == "mean" else strip.pooled_multiplier([r["db"] for r in labels])
    r = m0 / strip.COEF_MMAX
    init = strip.init_params(0, out_bias=float(np.log(r / (1 - r))))
    params, hist = strip.train(X, T, epochs=args.epochs, batch=args.batch, lr=args.lr, init=init)
    log(f"train_strip: {len(X)} samples, mean m* {mbar:.3f}, start bias m {m0:.3f}, "
        f"loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    rows = strip.judge([90000 + k for k in range(args.heldout)], params, n=args.n, nt=args.judge_nt, amp=args.amp,
                       head=strip.COEF_HEAD, log=log, period=args.period, scheme=args.scheme)
    choice = strip.choose(rows)
    gains = [round(r["strip0"]["reflected_db"] - r["strip"]["reflected_db"], 2) for r in rows]
    log(f"train_strip: held-out gain over the closed head {gains} dB, mean {np.mean(gains):.2f} dB")
    summary = {"mode": "coefficient", "head": strip.COEF_HEAD, "samples": int(len(X)), "models": args.models,
               "cases": len(labels), "n": args.n, "nt": args.nt, "amp": args.amp, "period": args.period,
               "scheme": args.scheme, "epochs": args.epochs, "lr": args.lr, "bias": args.bias,
               "start_multiplier": m0, "labels": labels, "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1],
               "judge": rows, "heldout_gain_db": gains, "heldout_mean_gain_db": float(np.mean(gains)),
               "choice": choice}
    strip.save(params, args.out, amp=np.float32(args.amp), head=np.int32(strip.COEF_HEAD),
               period=np.int32(args.period), loss=np.asarray(hist), summary=json.dumps(summary))
    print(json.dumps(summary), flush=True)


def coef_profile(args):
    """Coefficient head (mode 2) with a learned depth profile. (1) Training cases: every source
    (1-4, 10-30 Hz) of training models 10000 + k, each with its 40-cell-PML teacher run; (2) the
    per-depth multipliers of the strip's PML damping that minimise the cases' mean reflected
    energy, by coordinate search (strip.search_depth_profile); (3) teacher-forced samples of
    the same models (and a share of them with the wavefield zeroed, the state ahead of the
    wave) labelled with their depth's multiplier, on which the network is trained with the GPU
    training kernels on the multiplier itself (training mode 2); (4) the judge on held-out models (the trained network, the
    exact profile and the closed head m = 1, all against extended-domain FP64 runs)."""
    seeds = [10000 + k for k in range(args.models)]
    t0 = time.time()
    pool, srcs = [], []
    for seed in seeds:
        for j, src in enumerate(strip.model_sources(seed, args.n)):
            pool.append(strip.CoefCase(seed, j, n=args.n, nt=args.nt, snap_every=30, amp=args.amp,
                                       period=0, scheme=args.scheme))
    log(f"train_strip: {len(pool)} training cases from {len(seeds)} models ({time.time() - t0:.1f} s)")
    try:
        closed_db = float(np.mean([c.db(strip.depth_coef_params(np.ones(6))) for c in pool]))
        mult, best, hist = strip.search_depth_profile(pool, log=log)
    finally:
        for c in pool:
            c.close()
    log(f"train_strip: depth profile {mult.round(3).tolist()} (1 = outermost): {best:.3f} dB against "
        f"{closed_db:.3f} dB closed ({time.time() - t0:.1f} s)")
    Xs = []
    for seed in seeds:
        X, _, _ = strip.collect([seed], n=args.n, nt=args.nt, every=args.every, keep=args.keep, amp=args.amp,
                                head=strip.COEF_HEAD, params=strip.depth_coef_params(np.ones(6)))
        Xs.append(X)
    X, T = strip.depth_targets(np.concatenate(Xs), mult)
    m0 = float(np.mean(mult))
    r = m0 / strip.COEF_MMAX
    init = strip.init_params(0, out_bias=float(np.log(r / (1 - r))))
    params, loss = strip.train(X, T, epochs=args.epochs, batch=args.batch, lr=args.lr, init=init,
                               head=strip.COEF_HEAD)
    log(f"train_strip: {len(X)} samples, loss {loss[0]:.3e} -> {loss[-1]:.3e} ({time.time() - t0:.1f} s)")
    held = [90000 + k for k in range(args.heldout)]
    rows = strip.judge(held, params, n=args.n, nt=args.judge_nt, amp=args.amp, head=strip.COEF_HEAD, log=log,
                       period=args.period, scheme=args.scheme)
    exact = strip.judge(held, strip.depth_coef_params(mult), n=args.n, nt=args.judge_nt, amp=args.amp,
                        head=strip.COEF_HEAD, period=0, scheme=args.scheme)
    gains = [round(r["strip0"]["reflected_db"] - r["strip"]["reflected_db"], 2) for r in rows]
    gains_exact = [round(r["strip0"]["reflected_db"] - r["strip"]["reflected_db"], 2) for r in exact]
    log(f"train_strip: held-out gain over the closed head {gains} dB (mean {np.mean(gains):.2f}); "
        f"exact profile {np.mean(gains_exact):.2f} dB")
    summary = {"mode": "coefficient, learned depth profile", "head": strip.COEF_HEAD, "models": args.models,
               "cases": int(sum(len(strip.model_sources(s, args.n)) for s in seeds)), "n": args.n, "nt": args.nt,
               "amp": args.amp, "period": args.period, "scheme": args.scheme, "epochs": args.epochs, "lr": args.lr,
               "depth_multipliers": [round(float(v), 4) for v in mult], "train_db": best, "train_closed_db": closed_db,
               "search": hist, "samples": int(len(X)), "loss_first_epoch": loss[0], "loss_last_epoch": loss[-1],
               "judge": rows, "heldout_gain_db": gains, "heldout_mean_gain_db": float(np.mean(gains)),
               "heldout_gain_db_exact_profile": gains_exact, "choice": strip.choose(rows)}
    strip.save(params, args.out, amp=np.float32(args.amp), head=np.int32(strip.COEF_HEAD),
               period=np.int32(args.period), loss=np.asarray(loss), summary=json.dumps(summary))
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
