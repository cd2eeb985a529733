"""CPU pinning of the learned-strip checkers: the FP64 network restatement (finite
differences, single-path closed forms) and the oracle's strip correction (one
feature at a time against the strip's definition, zero network = plain PML)."""
import importlib.util
import os

import numpy as np
import pytest

from conftest import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _mlp():
    spec = importlib.util.spec_from_file_location("oracle_mlp", os.path.join(ROOT, "oracle", "mlp.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


mlp = _mlp()


def test_parameter_count_matches_library_constant():
    from a3dfd import strip
    assert mlp.NPAR == strip.NPARAMS == 5378


def test_fp64_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    p = mlp.init(1, scale=1.0).astype(np.float64)
    p[mlp.NPAR - 200:] += rng.normal(0, 0.3, 200)   # non-trivial output layer and biases
    X = rng.normal(0, 1, (48, mlp.NIN))
    T = rng.normal(0, 1, (48, mlp.NOUT))
    _, g = mlp.loss_grad(p, X, T)
    idx = np.concatenate([rng.choice(mlp.NPAR, 60, replace=False), np.arange(mlp.NPAR - 130, mlp.NPAR, 7)])
    h = 1e-6
    for i in idx:
        e = np.zeros(mlp.NPAR)
        e[i] = h
        fd = (mlp.loss(p + e, X, T) - mlp.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_fp64_coefficient_head_gradient_matches_central_differences():
    """Training mode 2 (the coefficient head's multiplier m = 3 sigmoid(y_0) as the target):
    the hand-written gradient against central differences; output 1 carries no loss."""
    rng = np.random.default_rng(4)
    p = mlp.init(2, scale=1.0).astype(np.float64)
    p[mlp.NPAR - 200:] += rng.normal(0, 0.3, 200)
    X = rng.normal(0, 1, (40, mlp.NIN))
    T = rng.uniform(0.2, 2.5, (40, mlp.NOUT))
    _, g = mlp.loss_grad(p, X, T, head=2)
    idx = np.concatenate([rng.choice(mlp.NPAR, 60, replace=False), np.arange(mlp.NPAR - 130, mlp.NPAR, 7)])
    h = 1e-6
    for i in idx:
        e = np.zeros(mlp.NPAR)
        e[i] = h
        fd = (mlp.loss(p + e, X, T, head=2) - mlp.loss(p - e, X, T, head=2)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])
    w3_1 = slice(mlp.NPAR - 2 - 64, mlp.NPAR - 2)   # W3 row 1 and b3[1]: no loss through output 1
    assert not np.any(g[w3_1]) and g[-1] == 0.0
    # the multiplier of a network whose output 0 is the constant b: 3 sigmoid(b)
    q = np.zeros(mlp.NPAR)
    q[-2] = 0.4
    pred, _ = mlp.forward_coef(q, X)
    np.testing.assert_allclose(pred[:, 0], 3.0 / (1.0 + np.exp(-0.4)), rtol=1e-15)


def test_depth_profile_network_gives_its_multipliers():
    """strip.depth_coef_params: the FP64 network gives each strip depth its multiplier, whatever
    the wavefield and model inputs."""
    from a3dfd import strip
    mult = np.array([0.9, 0.75, 0.6, 0.5, 0.8, 1.4])
    p = strip.depth_coef_params(mult)
    rng = np.random.default_rng(6)
    X = rng.uniform(-1, 1, (600, mlp.NIN))
    d = rng.integers(1, 7, 600)
    X[:, 14] = d / 6.0
    X[:, 11] = rng.uniform(0.12, 0.36, 600)
    X[:, 15] = rng.uniform(0.4, 2.5, 600)
    pred, _ = mlp.forward_coef(p, X)
    np.testing.assert_allclose(pred[:, 0], mult[d - 1], rtol=2e-4)


def _single_path(k, a=0.7, b=1.3, c=-0.9, c2=0.4, bias=(0.05, -0.02)):
    """Network whose raw outputs depend on input k only: y = (c, c2) * tanh(b * tanh(a * x_k)) + bias."""
    w = {n: np.zeros(s) for n, s in mlp.SHAPES}
    w["W1"][3, k] = a
    w["W2"][5, 3] = b
    w["W3"][0, 5] = c
    w["W3"][1, 5] = c2
    w["b3"][:] = bias
    return np.concatenate([w[n].ravel() for n, _ in mlp.SHAPES]).astype(np.float32)


def _sig(v):
    return 1.0 / (1.0 + np.exp(-v))


def test_fp64_forward_single_path_closed_form():
    rng = np.random.default_rng(1)
    X = rng.normal(0, 1, (10, mlp.NIN))
    for k in (0, 7, 15):
        pred, _ = mlp.forward(_single_path(k), X)
        f = np.tanh(np.float32(1.3) * np.tanh(np.float32(0.7) * X[:, k]))
        g0 = 0.25 * _sig(np.float32(-0.9) * f + np.float32(0.05))
        g1 = 0.25 * _sig(np.float32(0.4) * f + np.float32(-0.02))
        np.testing.assert_allclose(pred[:, 0], -X[:, 0] * g0, rtol=1e-6, atol=1e-9)
        np.testing.assert_allclose(pred[:, 1], -X[:, 7] * g1, rtol=1e-6, atol=1e-9)


# --- the oracle's strip correction against the strip's definition -------------------------

MM, CELLS = 4, 6


def _state(shape, seed):
    rng = np.random.default_rng(seed)
    f = {k: rng.normal(0, 1, shape).astype(np.float32)
         for k in ("P", "Q", "u", "v", "w", "px", "py", "pz", "qx", "qy", "qz")}
    f["vp"] = rng.uniform(1500, 4500, shape).astype(np.float32)
    f["rho"] = rng.uniform(1900, 2600, shape).astype(np.float32)
    f["eps"] = rng.uniform(0, 0.25, shape).astype(np.float32)
    f["del"] = rng.uniform(0, 0.1, shape).astype(np.float32)
    return f


def _owner(shape):
    """Per padded cell (iy, ix, iz): depth (1 = outermost updated cell; > CELLS: not a strip
    cell), owning axis (0 x, 1 y, 2 z) and side (1 = high face); ties go to z, then y, then x."""
    nny, nnx, nnz = shape
    best = np.full(shape, 1 << 30)
    axis = np.zeros(shape, int)
    side = np.zeros(shape, int)
    grids = np.meshgrid(np.arange(nny), np.arange(nnx), np.arange(nnz), indexing="ij")
    pos = {0: grids[1], 1: grids[0], 2: grids[2]}
    ext = {0: nnx, 1: nny, 2: nnz}
    for a in (2, 1, 0):
        for sd, dist in ((0, pos[a] - MM + 1), (1, ext[a] - MM - pos[a])):
            upd = dist < best
            best = np.where(upd, dist, best)
            axis = np.where(upd, a, axis)
            side = np.where(upd, sd, side)
    inside = ((grids[0] >= MM) & (grids[0] < nny - MM) & (grids[1] >= MM) & (grids[1] < nnx - MM)
              & (grids[2] >= MM) & (grids[2] < nnz - MM))
    return best, axis, side, inside & (best <= CELLS)


# numpy axis of each physical axis in (y, x, z) arrays
NP_AXIS = {0: 1, 1: 0, 2: 2}


def _expected_feature(f, k, idx, axis, side, depth, cdt):
    """Feature k of the strip cell idx = (iy, ix, iz), from the strip's definition."""
    a = NP_AXIS[axis]
    out = 1 if side else -1
    t1, t2 = {0: (1, 2), 1: (0, 2), 2: (0, 1)}[axis]
    t1, t2 = NP_AXIS[t1], NP_AXIS[t2]

    def at(name, ax, dlt):
        j = list(idx)
        j[ax] += dlt
        return f[name][tuple(j)]

    vel = f[{0: "u", 1: "v", 2: "w"}[axis]]
    un = np.float32(0.5) * (vel[idx] + at({0: "u", 1: "v", 2: "w"}[axis], a, -1))
    local = [f["P"][idx], at("P", a, -out), at("P", a, out), at("P", t1, -1), at("P", t1, 1), at("P", t2, -1),
             at("P", t2, 1), f["Q"][idx], at("Q", a, -out), at("Q", a, out),
             np.float32(un if side else -un) * (f["rho"][idx] * f["vp"][idx])]
    s = max(np.float32(1e-30), max(abs(np.float32(v)) for v in local))
    if k < 11:
        return np.float32(local[k]) / s, s
    return [f["vp"][idx] * np.float32(cdt[axis]), f["eps"][idx], f["del"][idx],
            np.float32(depth) * np.float32(1.0 / 6.0), np.float32(0.0)][k - 11], s


def test_round_half_matches_numpy_binary16(oracle):
    rng = np.random.default_rng(7)
    v = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e-4, 2000), rng.normal(0, 1e-6, 1000),
                        rng.uniform(-60000, 60000, 1000)]).astype(np.float32)
    # exact ties between binary16 neighbours (round to even)
    ties = (np.float16(np.arange(1, 200)).astype(np.float32) + np.float32(2.0 ** -11)).astype(np.float32)
    for x in np.concatenate([v, ties, -ties]):
        assert oracle.round_half(x) == float(np.float16(x)), x


@pytest.mark.parametrize("k", list(range(15)))
def test_oracle_strip_feature(oracle, k):
    shape = (21, 19, 23)   # (nny, nnx, nnz): updated 13 x 11 x 15, every kind of strip cell
    dt, dx, dy, dz = 8e-4, 10.0, 12.0, 9.0
    cdt = (np.float32(dt) / np.float32(dx), np.float32(dt) / np.float32(dy), np.float32(dt) / np.float32(dz))
    f = _state(shape, 100 + k)
    before = {n: a.copy() for n, a in f.items()}
    oracle.set_strip_net(_single_path(k))
    oracle.set_strip_half(0)
    try:
        oracle.strip_apply(dt, dx, dy, dz, f)
    finally:
        oracle.set_strip_net(None)
        oracle.set_strip_half(1)
    depth, axis, side, strip = _owner(shape)
    for n in ("P", "Q", "u", "v", "w", "px", "qz"):
        assert np.array_equal(f[n][~strip], before[n][~strip]), n
    cells = np.argwhere(strip)
    rng = np.random.default_rng(k)
    for idx in cells[rng.choice(len(cells), 300, replace=False)]:
        idx = tuple(int(i) for i in idx)
        x, s = _expected_feature(before, k, idx, int(axis[idx]), int(side[idx]), int(depth[idx]), cdt)
        t = np.tanh(1.3 * np.tanh(0.7 * float(x)))
        fp = 1.0 - 0.25 * _sig(-0.9 * t + 0.05)
        fq = 1.0 - 0.25 * _sig(0.4 * t - 0.02)
        # P, its split components, u, v by one factor; Q, its split components, w by the other
        for names, fac in ((("P", "px", "py", "pz", "u", "v"), fp), (("Q", "qx", "qy", "qz", "w"), fq)):
            for n in names:
                exp = float(before[n][idx]) * fac
                assert abs(float(f[n][idx]) - exp) <= 1e-6 * abs(exp) + 1e-30, (k, idx, n, float(f[n][idx]), exp)


@pytest.mark.parametrize("k", [0, 2, 7, 10, 11, 14])
def test_oracle_strip_binary16_network(oracle, k):
    """The matrix-core form of the network: inputs, hidden weights and activations rounded to
    binary16, the last layer in float; the single-path network against numpy's float16."""
    shape = (21, 19, 23)
    dt, dx, dy, dz = 8e-4, 10.0, 12.0, 9.0
    cdt = (np.float32(dt) / np.float32(dx), np.float32(dt) / np.float32(dy), np.float32(dt) / np.float32(dz))
    f = _state(shape, 300 + k)
    before = {n: a.copy() for n, a in f.items()}
    oracle.set_strip_net(_single_path(k))
    try:
        oracle.strip_apply(dt, dx, dy, dz, f)
    finally:
        oracle.set_strip_net(None)
    depth, axis, side, strip = _owner(shape)
    h = lambda v: float(np.float16(np.float32(v)))   # noqa: E731
    cells = np.argwhere(strip)
    rng = np.random.default_rng(k)
    for idx in cells[rng.choice(len(cells), 200, replace=False)]:
        idx = tuple(int(i) for i in idx)
        x, _ = _expected_feature(before, k, idx, int(axis[idx]), int(side[idx]), int(depth[idx]), cdt)
        h1 = h(np.tanh(h(0.7) * h(x)))
        h2 = h(np.tanh(h(1.3) * h1))
        fp = 1.0 - 0.25 * _sig(np.float32(-0.9) * h2 + np.float32(0.05))
        got = float(f["P"][idx]) / float(before["P"][idx]) if before["P"][idx] != 0 else fp
        # one binary16 step of h2 (a rounding decided differently near a tie) moves the factor by < 1e-4
        assert abs(got - fp) <= 1e-4, (k, idx, got, fp)


def test_oracle_strip_closed_network(oracle):
    """A shut damping head leaves the strip's own 6-cell PML: any network with output
    biases -30 gives the same run bit for bit, and it differs from the 10-cell reference PML."""
    from a3dfd import models, strip
    m = models.small_model("vti", 20, 18, 16, seed=5)
    shot = oracle.make_shot("vti", 20, 18, 16, 10.0, 10.0, 10.0, 10, 9, 2, 120, 1e-3, 25.0, nb=10,
                            snap_interval=40, mute_tt=0)
    kw = dict(vel=m["vel"], rho=m["rho"], eps=m["eps"], delta=m["delta"])
    a, sa, _ = oracle.run(shot, 32, nsnap_max=4, **kw)
    other = mlp.init(3, out_bias=-30.0)
    other[-130:-2] *= 0.01
    outs = []
    for net in (strip.closed_params(), other):
        oracle.set_strip_net(net)
        try:
            outs.append(oracle.run(shot, 32, nsnap_max=4, **kw)[:2])
        finally:
            oracle.set_strip_net(None)
    (b, sb), (c, sc) = outs
    assert np.array_equal(b, c) and np.array_equal(sb, sc)
    assert not np.array_equal(a, b)
    assert np.abs(a).max() > 0


@pytest.mark.parametrize("gains", [(0.05, 0.02, 0.01, 0.0, 0.0, 0.0), (0.002,) * 6, (0.0, 0.0, 0.0, 0.005, 0.01, 0.01),
                                   (0.03, 0.025, 0.02, 0.015, 0.01, 0.005)])
def test_profile_network_gives_its_depth_profile(gains):
    """strip.profile_params: the FP64 network's damping factor at every depth is the requested
    one, whatever the 11 local-state and 3 model features are."""
    from a3dfd import strip
    p = strip.profile_params(gains)
    rng = np.random.default_rng(5)
    for d in range(1, 7):
        X = np.zeros((200, mlp.NIN))
        X[:, :11] = rng.uniform(-1, 1, (200, 11))
        X[:, 11:14] = rng.uniform(0, 0.5, (200, 3))
        X[:, 14] = d / 6.0
        _, (_, _, _, g, _) = mlp.forward(p, X)
        want = max(gains[d - 1], 1e-14)
        assert np.allclose(g, want, rtol=2e-3, atol=0), (d, g.min(), g.max(), want)


def test_oracle_normal_head_damps_only_the_normal_pml_components(oracle):
    """Head mode 1 with a network of known per-depth factors: in a strip cell whose stress
    PML factor along its normal axis a differs from 1, p_a and q_a lose the fraction g (P and
    Q lose the same amount), the normal velocity is scaled by 1 - g, and every other field is
    untouched; all other cells are untouched."""
    from a3dfd import strip
    gains = (0.05, 0.02, 0.01, 0.004, 0.003, 0.002)
    shape = (26, 28, 24)   # (nny, nnx, nnz): interior 6 x 8 x 4 plus nb = 10 per side
    nb = CELLS + MM
    f = _state(shape, 21)
    before = {k: v.copy() for k, v in f.items()}
    oracle.set_strip_net(strip.profile_params(gains))
    oracle.set_strip_head(1)
    try:
        oracle.strip_apply(1e-3, 10.0, 10.0, 10.0, f)
    finally:
        oracle.set_strip_head(0)
        oracle.set_strip_net(None)
    depth, axis, side, is_strip = _owner(shape)
    grids = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), np.arange(shape[2]), indexing="ij")
    pos = {0: grids[1], 1: grids[0], 2: grids[2]}
    ext = {0: shape[1], 1: shape[0], 2: shape[2]}
    split = {0: ("px", "qx", "u"), 1: ("py", "qy", "v"), 2: ("pz", "qz", "w")}
    n_damped = 0
    for idx in zip(*np.nonzero(np.ones(shape, bool))):
        a = int(axis[idx])
        damped = bool(is_strip[idx]) and (pos[a][idx] < nb or pos[a][idx] > ext[a] - nb)
        if not damped:
            for k in before:
                assert f[k][idx] == before[k][idx], (k, idx)
            continue
        n_damped += 1
        g = gains[int(depth[idx]) - 1]
        pa, qa, va = split[a]
        for k in ("px", "py", "pz", "qx", "qy", "qz", "u", "v", "w"):
            if k not in (pa, qa, va):
                assert f[k][idx] == before[k][idx], (k, idx)
        for name, tot in ((pa, "P"), (qa, "Q")):
            b, n = float(before[name][idx]), float(f[name][idx])
            assert abs(n - b * (1 - g)) <= 3e-3 * g * abs(b) + 1e-6 * abs(b), (name, idx, b, n, g)
            # the total loses what the component loses
            lost = b - n
            assert abs((float(before[tot][idx]) - float(f[tot][idx])) - lost) <= 1e-6 * (abs(b) + abs(float(before[tot][idx]))), (tot, idx)
        b, n = float(before[va][idx]), float(f[va][idx])
        assert abs(n - b * (1 - g)) <= 3e-3 * g * abs(b) + 1e-6 * abs(b), (va, idx)
    assert n_damped > 0.8 * int(is_strip.sum())


def coef_const_params(mult):
    """Coefficient-head network whose multiplier is the constant mult everywhere: zero weights,
    output-0 bias logit(mult / 3)."""
    p = np.zeros(mlp.NPAR, np.float32)
    r = mult / 3.0
    p[-2] = np.float32(np.log(r / (1.0 - r)))
    return p


@pytest.mark.parametrize("mult", [0.5, 2.0])
def test_oracle_coefficient_head_scales_the_ramp(oracle, mult):
    """Head 2 ("coefficient") with a constant multiplier m is the strip's PML ramp with its
    damping scaled by m in both the velocity and the stress update: it equals the plain strip
    (shut damping head) at ramp amplitude 10 m within FP32 rounding of the coefficients, and
    differs from the amplitude-10 ramp. A period of 3 steps between network runs, or one run
    after the first step (period 0), changes nothing for a constant network (bitwise)."""
    from a3dfd import models, strip
    m = models.small_model("vti", 18, 16, 16, seed=9)
    kw = dict(vel=m["vel"], rho=m["rho"], eps=m["eps"], delta=m["delta"])
    shot = oracle.make_shot("vti", 18, 16, 16, 10.0, 10.0, 10.0, 9, 8, 3, 150, 1e-3, 25.0, nb=10,
                            snap_interval=75, snap_skip0=1, mute_tt=0)
    p = coef_const_params(mult)
    b = np.float32(p[-2])
    mf = float(np.float32(3.0) / (np.float32(1.0) + np.exp(-b)))     # the multiplier as the oracle computes it
    runs = {}
    try:
        oracle.set_strip_net(strip.closed_params())
        for amp in (10.0, 10.0 * mf):
            oracle.set_strip_profile(amp)
            runs[amp] = oracle.run(shot, 32, nsnap_max=2, **kw)[:2]
        oracle.set_strip_profile(10.0)
        oracle.set_strip_head(2)
        oracle.set_strip_net(p)
        head2 = oracle.run(shot, 32, nsnap_max=2, **kw)[:2]
        oracle.set_strip_period(3)
        head2p = oracle.run(shot, 32, nsnap_max=2, **kw)[:2]
        oracle.set_strip_period(0)          # the network after the first step only
        head2o = oracle.run(shot, 32, nsnap_max=2, **kw)[:2]
    finally:
        oracle.set_strip_period(1)
        oracle.set_strip_head(0)
        oracle.set_strip_profile(10.0)
        oracle.set_strip_net(None)
    (a, sa), (c, sc) = runs[10.0 * mf], runs[10.0]
    g, sg = head2
    assert np.abs(g).max() > 0
    assert rel_l2(g, a) < 1e-5 and rel_l2(sg, sa) < 1e-5
    assert rel_l2(g, c) > 100 * rel_l2(g, a) and rel_l2(sg, sc) > 100 * rel_l2(sg, sa)
    assert np.array_equal(head2p[0], g) and np.array_equal(head2p[1], sg)
    assert np.array_equal(head2o[0], g) and np.array_equal(head2o[1], sg)
