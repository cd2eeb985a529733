"""CPU pinning of the learned-strip checkers: the FP64 network restatement (finite
differences, single-path closed forms) and the oracle's strip correction (one
feature at a time against the strip's definition, zero network = plain PML)."""
import importlib.util
import os

import numpy as np
import pytest

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


def _single_path(k, a=0.7, b=1.3, c=-0.9, c2=0.4, bias=(0.05, -0.02)):
    """Network whose outputs depend on input k only: y = (c, c2) * tanh(b * tanh(a * x_k)) + bias."""
    w = {n: np.zeros(s) for n, s in mlp.SHAPES}
    w["W1"][3, k] = a
    w["W2"][5, 3] = b
    w["W3"][0, 5] = c
    w["W3"][1, 5] = c2
    w["b3"][:] = bias
    return np.concatenate([w[n].ravel() for n, _ in mlp.SHAPES]).astype(np.float32)


def test_fp64_forward_single_path_closed_form():
    rng = np.random.default_rng(1)
    X = rng.normal(0, 1, (10, mlp.NIN))
    for k in (0, 7, 15):
        y, _ = mlp.forward(_single_path(k), X)
        f = np.tanh(np.float32(1.3) * np.tanh(np.float32(0.7) * X[:, k]))
        np.testing.assert_allclose(y[:, 0], np.float32(-0.9) * f + np.float32(0.05), rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(y[:, 1], np.float32(0.4) * f + np.float32(-0.02), rtol=1e-6, atol=1e-7)


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


@pytest.mark.parametrize("k", list(range(15)))
def test_oracle_strip_feature(oracle, k):
    shape = (21, 19, 23)   # (nny, nnx, nnz): updated 13 x 11 x 15, every kind of strip cell
    dt, dx, dy, dz = 8e-4, 10.0, 12.0, 9.0
    cdt = (np.float32(dt) / np.float32(dx), np.float32(dt) / np.float32(dy), np.float32(dt) / np.float32(dz))
    f = _state(shape, 100 + k)
    before = {n: a.copy() for n, a in f.items()}
    oracle.set_strip_net(_single_path(k))
    try:
        oracle.strip_apply(dt, dx, dy, dz, f)
    finally:
        oracle.set_strip_net(None)
    depth, axis, side, strip = _owner(shape)
    dP = (f["P"].astype(np.float64) - before["P"])
    dQ = (f["Q"].astype(np.float64) - before["Q"])
    assert np.array_equal(f["P"][~strip], before["P"][~strip]) and np.array_equal(f["Q"][~strip], before["Q"][~strip])
    cells = np.argwhere(strip)
    rng = np.random.default_rng(k)
    for idx in cells[rng.choice(len(cells), 300, replace=False)]:
        idx = tuple(int(i) for i in idx)
        x, s = _expected_feature(before, k, idx, int(axis[idx]), int(side[idx]), int(depth[idx]), cdt)
        g = np.tanh(1.3 * np.tanh(0.7 * float(x)))
        eP = float(s) * (-0.9 * g + 0.05)
        eQ = float(s) * (0.4 * g - 0.02)
        # FP32 rounding of the network's terms (before their partial cancellation) and of the sum
        tolP = 4e-6 * float(s) * (0.9 * abs(g) + 0.05) + 2e-6 * abs(before["P"][idx])
        tolQ = 4e-6 * float(s) * (0.4 * abs(g) + 0.02) + 2e-6 * abs(before["Q"][idx])
        assert abs(dP[idx] - eP) <= tolP, (k, idx, dP[idx], eP)
        assert abs(dQ[idx] - eQ) <= tolQ, (k, idx, dQ[idx], eQ)
        # the owning axis' split components take the same correction, the others none
        names = ("px", "py", "pz")
        for a in range(3):
            dsp = float(f[names[a]][idx]) - float(before[names[a]][idx])
            dsq = float(f["q" + names[a][1]][idx]) - float(before["q" + names[a][1]][idx])
            if a == int(axis[idx]):
                assert abs(dsp - eP) <= tolP + 2e-6 * abs(before[names[a]][idx])
                assert abs(dsq - eQ) <= tolQ + 2e-6 * abs(before["q" + names[a][1]][idx])
            else:
                assert dsp == 0.0 and dsq == 0.0


def test_oracle_strip_zero_network_is_plain_pml(oracle):
    from a3dfd import models
    m = models.small_model("vti", 20, 18, 16, seed=5)
    shot = oracle.make_shot("vti", 20, 18, 16, 10.0, 10.0, 10.0, 10, 9, 2, 120, 1e-3, 25.0, nb=10,
                            snap_interval=40, mute_tt=0)
    kw = dict(vel=m["vel"], rho=m["rho"], eps=m["eps"], delta=m["delta"])
    a, sa, _ = oracle.run(shot, 32, nsnap_max=4, **kw)
    oracle.set_strip_net(np.zeros(mlp.NPAR, np.float32))
    try:
        b, sb, _ = oracle.run(shot, 32, nsnap_max=4, **kw)
    finally:
        oracle.set_strip_net(None)
    assert np.array_equal(a, b) and np.array_equal(sa, sb)
    assert np.abs(a).max() > 0
