"""Learned absorbing strip of the VTI operator: training data, training, judge and fallback.

The strip replaces the 20-cell PML by the 6 outer updated cells of every face
(grid padding nb = STRIP_NB = 10 instead of 20), run with the reference's
split-PML update at that width plus a correction from a small network after
every stress sweep (the library's a3d_set_strip_net). This module:

- `collect`: teacher-forced samples. A teacher shot with a wide PML (nb = 40)
  runs on seeded random models; at sampled steps its state is copied onto the
  strip grid, the strip grid's own stress sweep is applied, and the network's
  target is what the teacher's next P, Q differ from it by, over the strip
  cells (a3d_strip_collect).
- `train`: fits the network with the library's GPU training kernels (a3d_mlp).
- `judge`: reflected energy. Reflected energy of a boundary = FP64 norm of
  (P_run - P_extended) over the interior and all snapshot steps, over the norm
  of P_extended, in dB; the extended run has the model padded by `margin`
  cells per face and a 40-cell PML, computed in FP64 on the GPU.
- `choose`: the fallback rule: the learned strip is used only when, on every
  held-out model, its reflected energy is at most -40 dB, its shot record is
  within 1e-3 relative L2 of the extended FP64 run and it runs faster than the
  20-cell PML; otherwise the plain 20-cell PML.
"""
import os
import time

import numpy as np

from . import models
from ._lib import STRIP_NB, A3DError, Mlp, Modeller, make_desc, strip_collect

HERE = os.path.dirname(os.path.abspath(__file__))
# the committed trained network: the coefficient head (mode 2); the damping-head network
# trained by teacher forcing is kept for the head studies
NET_PATH = os.path.join(HERE, "strip_coef_net.npz")
DAMP_NET_PATH = os.path.join(HERE, "strip_net.npz")
TEACHER_NB = 40
PML_NB = 20
THRESHOLD_DB = -40.0
MAX_TRACE_REL_L2 = 1e-3   # shot record against the extended-domain FP64 run
NPARAMS = 64 * 16 + 64 + 64 * 64 + 64 + 2 * 64 + 2   # W1 b1 W2 b2 W3 b3


def _desc(m, src, nt, nb, storage=0, **kw):
    return make_desc("vti", m["nx"], m["ny"], m["nz"], m["dx"], m["dy"], m["dz"], src["sx"], src["sy"], src["sz"], nt,
                     m["dt"], src["fm"], nb=nb, storage=storage, **kw)


def _set(mdl, m):
    mdl.set_model(vel=m["vel"], rho=m["rho"], eps=m["eps"], delta=m["delta"])


def collect(seeds, n=128, nt=600, every=3, keep=0.02, horizon=1, params=None, amp=None, device=0, log=None,
            head=0, sources=None):
    """Samples over seeded models (one source each): X (N, 16), T (N, 2), S (N,). Samples whose
    local amplitude s is below 1e-4 of the largest s of their shot are dropped. horizon > 1:
    the strip grid runs `horizon` steps on its own (with network `params`) between
    synchronisations with the teacher, so the samples cover the states the network itself
    produces."""
    Xs, Ts, Ss = [], [], []
    teacher = student = None
    try:
        for k, seed in enumerate(seeds):
            m = models.strip_model(seed, n)
            src = models.strip_source(seed, n) if sources is None else sources[k]
            dt_ = _desc(m, src, nt, TEACHER_NB)
            if teacher is None:
                teacher = Modeller(dt_, device)
                student = Modeller(_desc(m, src, nt, STRIP_NB), device)
                # builds the strip cell list; the network runs between synchronisations
                if amp is not None:
                    student.set_strip_profile(amp)
                student.set_strip_head(head)
                student.set_strip_net(closed_params() if params is None else params)
            _set(teacher, m)
            _set(student, m)
            X, T, S = strip_collect(teacher, student, dt_, every=every, start=0, keep=keep, seed=seed,
                                    horizon=horizon)
            ok = S > 1e-4 * S.max() if S.size else S.astype(bool)
            Xs.append(X[ok])
            Ts.append(T[ok])
            Ss.append(S[ok])
            if log:
                log(f"strip collect: model {k + 1}/{len(seeds)} seed {seed}: {int(ok.sum())} samples")
    finally:
        for mdl in (teacher, student):
            if mdl is not None:
                mdl.close()
    return np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ss)


def train(X, T, epochs=30, batch=4096, lr=2e-3, lr_end=2e-5, seed=0, init=None, device=0, head=0):
    """Trains the network on the GPU; returns (params, loss per epoch). head 2: the targets are
    the coefficient head's multipliers (T[:, 0]), fitted directly."""
    with Mlp(device) as net:
        net.set_params(init if init is not None else init_params(seed))
        net.set_head(head)
        hist = net.train(X, T, epochs, batch=batch, lr=lr, lr_end=lr_end, seed=seed)
        return net.params(), hist


def init_params(seed, out_bias=-4.0):
    """Glorot-uniform weights, zero biases, output layer scaled by 0.1 and output bias out_bias
    (damping head nearly closed: g = 0.25 * sigmoid(-4) = 0.0045)."""
    rng = np.random.default_rng(seed)
    shapes = [(64, 16), (64,), (64, 64), (64,), (2, 64), (2,)]
    parts = []
    for i, s in enumerate(shapes):
        if len(s) == 2:
            lim = np.sqrt(6.0 / (s[0] + s[1])) * (0.1 if i == 4 else 1.0)
            parts.append(rng.uniform(-lim, lim, size=s).ravel())
        else:
            parts.append(np.full(s, out_bias if i == 5 else 0.0))
    return np.concatenate(parts).astype(np.float32)


def closed_params():
    """A network whose damping head is shut (output bias -30: g = 0.25 * sigmoid(-30) ~ 2e-14,
    1 - g rounds to 1 in FP32): the strip runs as the plain 6-cell PML."""
    p = np.zeros(NPARAMS, np.float32)
    p[-2:] = -30.0
    return p


def profile_params(gains, steep=60.0):
    """A network whose damping head gives the fixed factor gains[d - 1] (P and Q alike) at strip
    depth d = 1..6 (1 = outermost), whatever the local state: hidden unit j of layer 1 is a step
    tanh(steep * (depth/6 - (j + 0.5)/6)) (+1 for d >= j + 1, -1 below), layer 2 passes each step
    on as tanh(steps / 2), and the output weights make y(d) = logit(gains[d - 1] / 0.25), from the
    differences y(d) - y(d - 1) = 2 c_(d-1), c_j = W3_j tanh(1/2). A zero gain is taken as 1e-14
    (1 - g rounds to 1 in FP32). Used to measure what fixed damping profiles on top of the strip's
    PML ramp do to the reflected energy."""
    gains = np.asarray(gains, np.float64)
    if gains.shape != (6,) or np.any(gains < 0) or np.any(gains >= 0.25):
        raise ValueError("six damping factors in [0, 0.25) expected")
    r = np.maximum(gains, 1e-14) / 0.25
    y = np.log(r / (1.0 - r))
    c = np.zeros(6)
    c[1:] = (y[1:] - y[:-1]) / 2.0
    s = np.tanh(0.5)
    w1, b1 = np.zeros((64, 16)), np.zeros(64)
    w2, b2 = np.zeros((64, 64)), np.zeros(64)
    w3, b3 = np.zeros((2, 64)), np.zeros(2)
    for j in range(6):
        w1[j, 14] = steep
        b1[j] = -steep * (j + 0.5) / 6.0
        w2[j, j] = 0.5
        w3[:, j] = c[j] / s
    b3[:] = y[0] + c[1:].sum()
    return np.concatenate([a.ravel() for a in (w1, b1, w2, b2, w3, b3)]).astype(np.float32)


def save(params, path=NET_PATH, **meta):
    np.savez(path, params=np.asarray(params, np.float32), **{k: np.asarray(v) for k, v in meta.items()})


def load_params(path=NET_PATH):
    """The committed trained network and its strip PML amplitude, or (None, None) when absent."""
    if not os.path.exists(path):
        return None, None
    with np.load(path, allow_pickle=False) as f:
        return f["params"].astype(np.float32), float(f["amp"]) if "amp" in f else None


def load_head(path=NET_PATH):
    """The head mode the committed network was trained for (0 when absent or not recorded)."""
    if not os.path.exists(path):
        return 0
    with np.load(path, allow_pickle=False) as f:
        return int(f["head"]) if "head" in f else 0


def load_period(path=NET_PATH):
    """The network period (steps between runs) recorded with the committed network (1 when absent)."""
    if not os.path.exists(path):
        return 1
    with np.load(path, allow_pickle=False) as f:
        return int(f["period"]) if "period" in f else 1


def extended(m, src, margin):
    """Model and source of the extended-domain run: every face padded by `margin` cells
    (edge replication, as the boundary padding itself does)."""
    pad = lambda a: np.ascontiguousarray(np.pad(a, margin, mode="edge"))   # noqa: E731
    me = dict(m, nx=m["nx"] + 2 * margin, ny=m["ny"] + 2 * margin, nz=m["nz"] + 2 * margin,
              vel=pad(m["vel"]), rho=pad(m["rho"]), eps=pad(m["eps"]), delta=pad(m["delta"]))
    se = dict(src, sx=src["sx"] + margin, sy=src["sy"] + margin, sz=src["sz"] + margin)
    return me, se


def run_case(m, src, nt, snap_every, boundary, params=None, margin=64, amp=None, scheme=0, device=0, head=0,
             period=1):
    """One shot under a boundary: 'strip' (learned, needs params), 'strip0' (the 6-cell strip
    with its damping head shut: its own PML only), 'pml' (20-cell PML), 'extended' (FP64, model
    padded by margin, 40-cell PML). amp: the strip's PML amplitude (None: the library default).
    scheme: the stress update of the non-extended runs (0 reference, 1 split fields in PML cells only).
    head: what the network does (0 damp every field, 1 damp the normal-axis PML components,
    2 scale each strip cell's PML damping inside the sweeps); period: steps between network
    runs (head 2; 0: once, after the first step, with the network evaluated in FP32).
    Returns (shot (ny, nx, nt), snapshots (k, ny, nx, nz), loop_ms) on the original interior."""
    if boundary == "extended":
        me, se = extended(m, src, margin)
        d = _desc(me, se, nt, TEACHER_NB, storage=2, rec_z=margin, snap_interval=snap_every, snap_skip0=1)
        with Modeller(d, device) as mdl:
            _set(mdl, me)
            shot, snaps, ms = mdl.run(nsnap_max=nt // snap_every)
        c = slice(margin, margin + m["nx"])
        return shot[margin:margin + m["ny"], c], snaps[:, margin:margin + m["ny"], c, margin:margin + m["nz"]], ms
    nb = PML_NB if boundary == "pml" else STRIP_NB
    d = _desc(m, src, nt, nb, snap_interval=snap_every, snap_skip0=1, scheme=scheme)
    with Modeller(d, device) as mdl:
        _set(mdl, m)
        if boundary in ("strip", "strip0"):
            if amp is not None:
                mdl.set_strip_profile(amp)
            if boundary == "strip":
                mdl.set_strip_head(head)
                mdl.set_strip_period(period)
                if head == COEF_HEAD and period == 0:
                    mdl.set_strip_precision(0)   # run once per shot: FP32 evaluation costs nothing
                mdl.set_strip_net(params)
            elif head == COEF_HEAD:
                # strip0 beside a coefficient head: the plain ramp through the same sweeps, every
                # multiplier 1 (the closed head, run once)
                mdl.set_strip_head(COEF_HEAD)
                mdl.set_strip_period(0)
                mdl.set_strip_net(coef_params(coef_closed_theta(), np.zeros(NPARAMS, np.float32)))
            else:
                # strip0: the plain ramp (a damping head shut)
                mdl.set_strip_head(0)
                mdl.set_strip_net(closed_params())
        return mdl.run(nsnap_max=nt // snap_every)


def energy_db(snaps, ref):
    """10 log10 of sum (snaps - ref)^2 / sum ref^2, accumulated in FP64."""
    num = den = 0.0
    for a, b in zip(snaps, ref):
        b64 = b.astype(np.float64)
        num += float(np.sum((a.astype(np.float64) - b64) ** 2))
        den += float(np.sum(b64 ** 2))
    return 10.0 * np.log10(max(num, 1e-300) / den)


def rel_l2(a, b):
    b = b.astype(np.float64)
    return float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))


ADAPT_GAINS = (0.0, 0.001, 0.002, 0.004, 0.007, 0.01, 0.015, 0.02)


def best_uniform_gain(seed, n=128, nt=800, snap_every=20, amp=5.0, head=1, gains=ADAPT_GAINS, margin=64, device=0):
    """The per-model optimum of a uniform damping factor on top of the strip's ramp: reflected
    energy of the strip with profile_params((g,) * 6) for every g in gains, against the
    extended-domain FP64 run of seeded model `seed`. Returns (g*, {g: dB})."""
    m = models.strip_model(seed, n)
    src = models.strip_source(seed, n)
    _, ref, _ = run_case(m, src, nt, snap_every, "extended", margin=margin, device=device)
    tab = {}
    for g in gains:
        _, snaps, _ = run_case(m, src, nt, snap_every, "strip", profile_params((g,) * 6), amp=amp, head=head,
                               device=device)
        tab[g] = round(float(energy_db(snaps, ref)), 2)
    return min(tab, key=tab.get), tab


def adaptive_targets(X, gstar):
    """Regression targets that make the damping head give g* (the head's prediction is
    (-x_0 g_P, -x_7 g_Q), x_0 and x_7 the cell's own P/s and Q/s)."""
    return (-X[:, [0, 7]] * np.float32(gstar)).astype(np.float32)


def judge(seeds, params, n=128, nt=800, snap_every=20, margin=64, amp=None, device=0, log=None, head=0,
          extra_amps=(), period=1, scheme=0):
    """Reflected energy (dB) and shot-record rel. L2 against the extended FP64 run for the
    learned strip, the strip without network and the 20-cell PML, on held-out models; also
    the strip without network at each ramp amplitude in extra_amps (key strip0_amp<a>)."""
    rows = []
    for seed in seeds:
        m = models.strip_model(seed, n)
        src = models.strip_source(seed, n)
        t0 = time.time()
        ref_shot, ref_snaps, _ = run_case(m, src, nt, snap_every, "extended", margin=margin, device=device)
        row = {"seed": seed, "fm": round(src["fm"], 2), "src": [src["sx"], src["sy"], src["sz"]]}
        cases = [("pml", "pml", amp), ("strip0", "strip0", amp), ("strip", "strip", amp)]
        cases += [(f"strip0_amp{a:g}", "strip0", a) for a in extra_amps]
        for key, b, a in cases:
            if b == "strip" and params is None:
                continue
            try:
                shot, snaps, ms = run_case(m, src, nt, snap_every, b, params, amp=a, device=device, head=head,
                                           period=period, scheme=scheme)
            except A3DError as e:   # an unstable boundary is judged as failed
                row[key] = {"reflected_db": float("inf"), "shot_rel_l2": float("inf"), "loop_ms": None, "error": str(e)}
                continue
            row[key] = {"reflected_db": round(float(energy_db(snaps, ref_snaps)), 2),
                        "shot_rel_l2": rel_l2(shot, ref_shot), "loop_ms": round(ms, 2)}
        rows.append(row)
        if log:
            log(f"strip judge: seed {seed} ({time.time() - t0:.1f} s): {row}")
    return rows


def tune_amp(seeds, amps=(4.0, 5.0, 6.0, 7.0, 8.5, 10.0), n=128, nt=800, snap_every=20, device=0, log=None):
    """The strip's PML amplitude with the lowest worst-case reflected energy (damping head shut)
    over the given training models (the judge asks every model to pass); returns
    (amp, {amp: worst dB})."""
    db = {a: [] for a in amps}
    for seed in seeds:
        m = models.strip_model(seed, n)
        src = models.strip_source(seed, n)
        _, ref, _ = run_case(m, src, nt, snap_every, "extended", device=device)
        for a in amps:
            _, snaps, _ = run_case(m, src, nt, snap_every, "strip0", amp=a, device=device)
            db[a].append(energy_db(snaps, ref))
    worst = {a: round(float(np.max(v)), 2) for a, v in db.items()}
    if log:
        log(f"strip profile: worst reflected dB by amplitude {worst}")
    return min(worst, key=worst.get), worst


def choose(rows, threshold_db=THRESHOLD_DB, max_rel_l2=MAX_TRACE_REL_L2):
    """The fallback rule: 'strip' only when, on every held-out model, the learned strip's
    reflected energy is <= threshold_db, its shot record is within max_rel_l2 (relative L2)
    of the extended-domain FP64 run, and its time loop is shorter than the 20-cell PML's;
    otherwise 'pml'. Rows need "strip" and "pml" entries (strip.judge's rows)."""
    if not rows or any("strip" not in r or "pml" not in r for r in rows):
        return "pml"
    for r in rows:
        s, p = r["strip"], r["pml"]
        if not (s["reflected_db"] <= threshold_db and s["shot_rel_l2"] <= max_rel_l2):
            return "pml"
        if s.get("loop_ms") is None or p.get("loop_ms") is None or not s["loop_ms"] < p["loop_ms"]:
            return "pml"
    return "strip"


# ---------------------------------------------------------------------------
# Coefficient head (head mode 2): the network's first output y_0 sets each strip cell's
# multiplier m = COEF_MMAX sigmoid(y_0) of its PML damping, inside the sweeps' update.
# ---------------------------------------------------------------------------
COEF_MMAX = 3.0
COEF_HEAD = 2
O_W1, O_B1, O_W2, O_B2, O_W3, O_B3 = 0, 1024, 1088, 5184, 5248, 5376
COEF_CLOSED_BIAS = float(np.log(1.0 / (COEF_MMAX - 1.0)))   # m = 1: the plain ramp
COURANT_STEPS = (0.14, 0.18, 0.22, 0.26, 0.30, 0.34)       # vp dt / d of the study models: 0.12-0.36
RATIO_STEPS = (0.5, 0.7, 0.85, 1.0, 1.2, 1.5, 2.0)         # vp over the ramp's d0 sample: 1/3 - 3
N_DESIGNED = 6 + len(COURANT_STEPS) + len(RATIO_STEPS)


def coef_hidden(base, steep=60.0):
    """Hidden layers of the coefficient head: units 0-5 are steps in the strip depth (input 14,
    +1 at depth >= j + 1), units 6-11 steps in the local Courant number vp dt / d (input 11),
    units 12-18 steps in the velocity over the ramp's d0 sample (input 15, 8 / vp scale), all
    passed on by layer 2 as tanh(step / 2); the other units are those of `base` (a network
    trained on the strip's wavefields), cut off from the designed ones. The output layer is
    left to the caller."""
    p = np.array(base, np.float32)
    w1 = p[O_W1:O_B1].reshape(64, 16)
    b1 = p[O_B1:O_W2]
    w2 = p[O_W2:O_B2].reshape(64, 64)
    b2 = p[O_B2:O_W3]
    w1[:, 15] = 0.0                      # input 15 is new to the base network's units
    for j in range(N_DESIGNED):
        w1[j] = 0.0
        w2[j] = 0.0
        b2[j] = 0.0
        w2[j, j] = 0.5
    for j in range(6):
        w1[j, 14] = steep
        b1[j] = -steep * (j + 0.5) / 6.0
        w1[6 + j, 11] = steep
        b1[6 + j] = -steep * COURANT_STEPS[j]
    for j, r in enumerate(RATIO_STEPS):
        k = 12 + j
        w1[k, 15] = 8.0 / r
        b1[k] = -8.0
    w2[N_DESIGNED:, :N_DESIGNED] = 0.0
    return p


def coef_params(theta, hidden):
    """Coefficient-head network: hidden layers from `hidden`, output 0 with weights theta[:64]
    and bias theta[64] (output 1 unused: zero)."""
    p = np.array(hidden, np.float32)
    p[O_W3:O_B3] = 0.0
    p[O_W3:O_W3 + 64] = np.asarray(theta[:64], np.float32)
    p[O_B3] = np.float32(theta[64])
    p[O_B3 + 1] = 0.0
    return p


def coef_closed_theta():
    th = np.zeros(65, np.float64)
    th[64] = COEF_CLOSED_BIAS
    return th


def model_sources(seed, n):
    """The 1-4 Ricker sources (10-30 Hz, anywhere in the volume) of seeded study model `seed`."""
    rng = np.random.default_rng(seed + 17)
    k = int(rng.integers(1, 5))
    return [models.strip_source(seed * 8 + j, n) for j in range(k)]


class CoefCase:
    """One (model, source) shot of the coefficient head's training: the 40-cell-PML teacher's
    interior snapshots (FP32) and a strip-grid context (head 2) to run candidate networks."""

    def __init__(self, seed, src_index, n=128, nt=600, snap_every=30, amp=10.0, period=1, scheme=3, device=0):
        self.m = models.strip_model(seed, n)
        srcs = model_sources(seed, n)
        self.src = srcs[src_index % len(srcs)]
        self.nt, self.every = nt, snap_every
        d = _desc(self.m, self.src, nt, TEACHER_NB, snap_interval=snap_every, snap_skip0=1, scheme=scheme)
        with Modeller(d, device) as t:
            _set(t, self.m)
            _, self.ref, _ = t.run(shot=False, nsnap_max=nt // snap_every)
        d = _desc(self.m, self.src, nt, STRIP_NB, snap_interval=snap_every, snap_skip0=1, scheme=scheme)
        self.mdl = Modeller(d, device)
        _set(self.mdl, self.m)
        self.mdl.set_strip_profile(amp)
        self.mdl.set_strip_head(COEF_HEAD)
        self.mdl.set_strip_period(period)
        if period == 0:
            self.mdl.set_strip_precision(0)

    def db(self, params):
        """Reflected energy (dB) of the strip run with network `params` against the teacher."""
        self.mdl.set_strip_net(params)
        try:
            _, snaps, _ = self.mdl.run(shot=False, nsnap_max=self.nt // self.every)
        except A3DError:
            return 0.0   # unstable: as bad as no absorption
        return float(energy_db(snaps, self.ref))

    def close(self):
        self.mdl.close()


def coef_pool(seeds, n=128, nt=600, snap_every=30, amp=10.0, period=1, device=0):
    """Training cases: for each seed, one of its model's 1-4 sources (the k-th seed of a model
    takes source k), with the teacher's snapshots computed once."""
    visits, pool = {}, []
    for s in seeds:
        visits[s] = visits.get(s, -1) + 1
        pool.append(CoefCase(s, visits[s], n=n, nt=nt, snap_every=snap_every, amp=amp, period=period, device=device))
    return pool


def train_coef(pool, hidden, iters=100, active=None, batch=None, a=0.02, c=0.1, alpha=0.602, gamma=0.101,
               npert=2, seed=0, theta0=None, log=None):
    """Fits the coefficient head's output layer against the teacher's interior by
    simultaneous-perturbation stochastic approximation (SPSA): each iteration averages `npert`
    two-sided estimates (J(theta + c_k D) - J(theta - c_k D)) / (2 c_k) D over the active
    coordinates (D random +-1), on a minibatch of `batch` cases of `pool` (all when None),
    and steps theta -= a_k times the estimate, with a_k = a / (k + 1 + A)^alpha (A = iters /
    10) and c_k = c / (k + 1)^gamma. J = mean reflected energy in dB against the teacher.
    active: indices of theta (64 output weights, then the bias) to fit (default all).
    Returns (theta, history of (iteration, mean J over the two-sided evaluations))."""
    rng = np.random.default_rng(seed)
    theta = coef_closed_theta() if theta0 is None else np.array(theta0, np.float64)
    act = np.arange(theta.size) if active is None else np.asarray(active)
    A = max(1, iters // 10)
    hist = []
    for k in range(iters):
        ak = a / (k + 1 + A) ** alpha
        ck = c / (k + 1) ** gamma
        cases = pool if batch is None else [pool[i] for i in rng.choice(len(pool), batch, replace=False)]
        g = np.zeros(theta.size)
        js = []
        for _ in range(npert):
            d = np.zeros(theta.size)
            d[act] = rng.choice((-1.0, 1.0), size=act.size)
            jp = float(np.mean([cs.db(coef_params(theta + ck * d, hidden)) for cs in cases]))
            jm = float(np.mean([cs.db(coef_params(theta - ck * d, hidden)) for cs in cases]))
            g += (jp - jm) / (2.0 * ck) * d / npert
            js += [jp, jm]
        theta -= ak * g
        hist.append((k, float(np.mean(js))))
        if log:
            log(f"coef train: iteration {k}: J {np.mean(js):.3f} dB, |step| {ak * np.abs(g).max():.4f}, "
                f"bias {theta[64]:.3f}")
    return theta, hist


COEF_SCAN = (0.25, 0.35, 0.5, 0.7, 0.85, 1.0, 1.3)


def coef_label(case, hidden=None, mults=COEF_SCAN):
    """The best constant damping multiplier of a (model, source) case: the reflected energy
    against the teacher for every multiplier in `mults`, refined by a parabola in log m through
    the best grid point and its neighbours. Returns (m*, {m: dB})."""
    hid = np.zeros(NPARAMS, np.float32) if hidden is None else hidden
    tab = {}
    for mult in mults:
        th = np.zeros(65)
        th[64] = np.log((mult / COEF_MMAX) / (1.0 - mult / COEF_MMAX))
        tab[mult] = round(case.db(coef_params(th, hid)), 3)
    ms = list(mults)
    k = int(np.argmin([tab[x] for x in ms]))
    best = ms[k]
    if 0 < k < len(ms) - 1:
        x = np.log([ms[k - 1], ms[k], ms[k + 1]])
        y = np.array([tab[ms[k - 1]], tab[ms[k]], tab[ms[k + 1]]])
        a2, a1, _ = np.polyfit(x, y, 2)
        if a2 > 0:
            best = float(np.clip(np.exp(-a1 / (2 * a2)), ms[k - 1], ms[k + 1]))
    return float(best), tab


def _layer2(params, x):
    """Layer-2 activations of network `params` for input x (FP64, numpy)."""
    p = np.asarray(params, np.float64)
    w1, b1 = p[O_W1:O_B1].reshape(64, 16), p[O_B1:O_W2]
    w2, b2 = p[O_W2:O_B2].reshape(64, 64), p[O_B2:O_W3]
    return np.tanh(w2 @ np.tanh(w1 @ x + b1) + b2)


def depth_coef_params(mult):
    """Coefficient-head network whose multiplier at strip depth d = 1..6 (1 = outermost) is
    mult[d - 1], whatever the local state: the depth units of coef_hidden (steps at d >= j + 1)
    with output weights solving y(d) = logit(mult[d - 1] / COEF_MMAX) for d = 1..6 (minimum
    norm; the Courant and velocity-ratio units get no weight)."""
    mult = np.asarray(mult, np.float64)
    if mult.shape != (6,) or np.any(mult <= 0) or np.any(mult >= COEF_MMAX):
        raise ValueError(f"six multipliers in (0, {COEF_MMAX}) expected")
    hidden = coef_hidden(np.zeros(NPARAMS, np.float32))
    rows = []
    for d in range(1, 7):
        x = np.zeros(16)
        x[14] = d / 6.0
        rows.append(np.append(_layer2(hidden, x)[:6], 1.0))
    r = mult / COEF_MMAX
    sol, *_ = np.linalg.lstsq(np.array(rows), np.log(r / (1.0 - r)), rcond=None)
    th = np.zeros(65)
    th[:6], th[64] = sol[:6], sol[6]
    return coef_params(th, hidden)


def search_depth_profile(pool, start=0.7, steps=(0.4, 0.2, 0.1), sweeps=2, log=None):
    """Per-depth multipliers of the coefficient head found by coordinate search on the mean
    reflected energy of the training cases in `pool` (CoefCase, judged against their teachers):
    from the constant `start`, each depth's log multiplier is moved by +-step while that lowers
    the mean dB by more than 0.02, for each step size in turn. Returns (multipliers, mean dB,
    history of (stage, mean dB))."""
    def J(logm):
        p = depth_coef_params(np.exp(logm))
        v = float(np.mean([c.db(p) for c in pool]))
        if log:
            log(f"depth profile: candidate {np.exp(logm).round(3).tolist()}: {v:.3f} dB")
        return v

    cur = np.log(np.full(6, float(start)))
    best = J(cur)
    hist = [("start", best)]
    for step in steps:
        for _ in range(sweeps):
            improved = False
            for k in range(6):
                for sgn in (1.0, -1.0):
                    trial = cur.copy()
                    trial[k] += sgn * step
                    if not 0.05 < np.exp(trial[k]) < 0.95 * COEF_MMAX:
                        continue
                    v = J(trial)
                    if v < best - 0.02:
                        best, cur, improved = v, trial, True
                        break
            hist.append((f"step {step}", best))
            if not improved:
                break
    return np.exp(cur), best, hist


def depth_targets(X, mult, quiet=0.05, seed=0):
    """Samples and targets for training the coefficient head on its multiplier (strip.train with
    head=2): each sample's target is the multiplier of its strip depth (input 14 = depth / 6).
    A fraction `quiet` of the samples is added again with a zero wavefield (inputs 0-10), the
    state of a cell the wave has not reached, whose multiplier holds when the wave arrives
    (the network runs every few steps). Returns (X, T)."""
    rng = np.random.default_rng(seed)
    q = X[rng.choice(len(X), int(quiet * len(X)), replace=False)].copy()
    q[:, :11] = 0.0
    X = np.concatenate([X, q])
    d = np.clip(np.rint(X[:, 14] * 6.0).astype(int), 1, 6)
    T = np.zeros((len(X), 2), np.float32)
    T[:, 0] = np.asarray(mult, np.float32)[d - 1]
    return X, T


def pooled_multiplier(tabs, grid=np.exp(np.linspace(np.log(0.25), np.log(1.3), 200))):
    """The constant multiplier with the lowest reflected energy averaged over cases, from each
    case's scan {m: dB} interpolated linearly in log m."""
    tot = np.zeros(grid.size)
    for tab in tabs:
        ms = np.array(sorted(float(k) for k in tab))
        db = np.array([tab[k] for k in sorted(tab, key=float)], np.float64)
        tot += np.interp(np.log(grid), np.log(ms), db)
    return float(grid[int(np.argmin(tot))])


def coef_targets(X, mstar):
    """Regression targets that make the damping head's factor g_0 = 0.25 sigmoid(y_0) equal
    m* / 12, i.e. the coefficient head's multiplier 3 sigmoid(y_0) = m* (the training kernels
    fit the head's prediction (-x_0 g_0, -x_7 g_1)); output 1 is unused and pulled to 0."""
    T = np.zeros((len(X), 2), np.float32)
    T[:, 0] = -X[:, 0] * np.float32(mstar / (4.0 * COEF_MMAX))
    return T
