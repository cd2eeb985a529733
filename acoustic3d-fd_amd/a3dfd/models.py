"""Seeded synthetic earth models for the benchmark configurations, and the
stability (Courant) check of the explicit schemes.

All arrays are float32 in the reference file order: shape (ny, nx, nz),
C-ordered, so z is the fastest axis (reference iso:249-255).
"""
import numpy as np
from scipy.ndimage import gaussian_filter, zoom

# Sum of |staggered coefficients| (reference iso:37) and of the 2nd-derivative
# stencil weights' symbol at Nyquist (Toa_3dfd_cuda.cu:32).
STAG_SUM = 1.196289 + 0.0797526 + 0.009570313 + 0.0006975447
_S = (-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0)
LAP_NYQ = -(_S[0] - 2 * _S[1] + 2 * _S[2] - 2 * _S[3] + 2 * _S[4])


def gardner(v):
    """Gardner's relation rho = 310 * v**0.25 (kg/m^3)."""
    return (310.0 * np.power(v, 0.25)).astype(np.float32)


def smooth_field(shape, seed, sigma_cells=8.0, coarse=8):
    """A smooth random field in [0, 1]: Gaussian-filtered noise on a grid
    `coarse` times coarser, trilinearly upsampled (cheap at 10^8 points)."""
    rng = np.random.default_rng(seed)
    cs = tuple(max(4, int(np.ceil(s / coarse)) + 1) for s in shape)
    f = gaussian_filter(rng.standard_normal(cs), sigma=max(sigma_cells / coarse, 0.5), mode="reflect")
    f = zoom(f, [s / c for s, c in zip(shape, cs)], order=1, grid_mode=False)[: shape[0], : shape[1], : shape[2]]
    lo, hi = f.min(), f.max()
    return ((f - lo) / (hi - lo + 1e-30)).astype(np.float32)


def courant_1st(vmax, dt, dmin, eps=0.0):
    """Stability number of the staggered 8th-order leapfrog (ISO, VTI with epsilon >= delta):
    v dt/d * sum|c| * sqrt(3 + 4 eps); the scheme is stable when it is <= 1."""
    return vmax * dt / dmin * STAG_SUM * np.sqrt(3.0 + 4.0 * eps)


def courant_2nd(vmax, dt, dmin):
    """Stability number of the 8th-order leapfrog: v dt/d * sqrt(3 * Nyquist symbol) / 2 (<= 1 stable)."""
    return vmax * dt / dmin * np.sqrt(3.0 * LAP_NYQ) / 2.0


def cap_epsilon(v, eps, dt, dmin, margin=0.98):
    """Caps epsilon where v*sqrt(3+4eps) would exceed the staggered-scheme limit."""
    lim = (margin * dmin / (np.maximum(v, 1.0) * dt * STAG_SUM)) ** 2
    emax = (lim - 3.0) / 4.0
    return np.minimum(eps, np.maximum(emax, 0.0)).astype(np.float32)


# ---------------------------------------------------------------------------
# Benchmark configurations
# ---------------------------------------------------------------------------

def cfg1(seed=1, n=100):
    """2nd-order isotropic, 100^3, dx=10: 2000 m/s above z=50, 3000 below, times
    (1 + 0.05 * smoothed N(0,1) noise / max|.|) with sigma = 5 cells."""
    rng = np.random.default_rng(seed)
    z = np.arange(n)
    v = np.where(z < n // 2, 2000.0, 3000.0)[None, None, :] * np.ones((n, n, 1))
    noise = gaussian_filter(rng.standard_normal((n, n, n)), sigma=5.0, mode="reflect")
    noise /= np.abs(noise).max()
    vel = (v * (1.0 + 0.05 * noise)).astype(np.float32)
    return dict(op="2nd", nx=n, ny=n, nz=n, dx=10.0, dy=10.0, dz=10.0, dt=1e-3, nt=500, fm=20.0,
                sx=50, sy=50, sz=5, vel=vel)


def dipping_layers(shape, d, seed, nlayer, vmin, vmax, max_dip_deg):
    """Flat-to-dipping layers: interface k at depth z_k + tan(dip_k)(x cos a_k + y sin a_k)."""
    ny, nx, nz = shape
    rng = np.random.default_rng(seed)
    vs = np.sort(rng.uniform(vmin, vmax, nlayer))
    z0 = np.sort(rng.uniform(0.1, 0.95, nlayer - 1)) * nz
    dips = np.deg2rad(rng.uniform(0.0, max_dip_deg, nlayer - 1))
    az = rng.uniform(0.0, 2 * np.pi, nlayer - 1)
    y = (np.arange(ny) - ny / 2)[:, None]
    x = (np.arange(nx) - nx / 2)[None, :]
    z = np.arange(nz)[None, None, :]
    layer = np.zeros(shape, np.int8)
    for k in range(nlayer - 1):
        depth = z0[k] + np.tan(dips[k]) * (x * np.cos(az[k]) + y * np.sin(az[k]))
        layer += (z >= depth[:, :, None]).astype(np.int8)
    return vs[layer].astype(np.float32)


def cfg2(seed=2, n=256):
    """1st-order isotropic with density, 256^3, dx=10, dt=0.8 ms: 8 seeded dipping layers
    1500-4500 m/s (dips 0-15 deg), Gardner density, Ricker 25 Hz at the surface centre."""
    vel = dipping_layers((n, n, n), 10.0, seed, 8, 1500.0, 4500.0, 15.0)
    return dict(op="iso", nx=n, ny=n, nz=n, dx=10.0, dy=10.0, dz=10.0, dt=8e-4, nt=2000, fm=25.0,
                sx=n // 2 + 1, sy=n // 2 + 1, sz=1, vel=vel, rho=gardner(vel), v_mute=2000.0)


def cfg3(seed=3, nx=400, ny=400, nz=300, d=12.5, dt=1e-3):
    """1st-order VTI, 400x400x300, dx=12.5, dt=1 ms: seeded layered model 1500-5000 m/s with a
    Gaussian salt-like body (4500 m/s, radius 60 cells), Gardner density, smooth random
    epsilon in [0, 0.25] and delta in [0, 0.1] with epsilon >= delta, epsilon capped where
    the Courant condition would fail. Ricker 15 Hz at the surface centre."""
    shape = (ny, nx, nz)
    vel = dipping_layers(shape, d, seed, 10, 1500.0, 5000.0, 8.0)
    # salt body: Gaussian weight blends toward 4500 m/s within radius 60 cells
    cy, cx, cz = ny / 2, nx / 2, nz / 2
    y = (np.arange(ny) - cy)[:, None, None] ** 2
    x = (np.arange(nx) - cx)[None, :, None] ** 2
    z = (np.arange(nz) - cz)[None, None, :] ** 2
    wgt = np.exp(-(x + y + z) / (2.0 * (60.0 / 2.0) ** 2)).astype(np.float32)
    vel = (vel * (1.0 - wgt) + 4500.0 * wgt).astype(np.float32)
    eps = 0.25 * smooth_field(shape, seed + 100)
    dl = np.minimum(0.1 * smooth_field(shape, seed + 200), eps)
    eps = cap_epsilon(vel, eps, dt, d)
    dl = np.minimum(dl, eps).astype(np.float32)
    return dict(op="vti", nx=nx, ny=ny, nz=nz, dx=d, dy=d, dz=d, dt=dt, nt=3000, fm=15.0,
                sx=nx // 2 + 1, sy=ny // 2 + 1, sz=1, vel=vel, rho=gardner(vel), eps=eps.astype(np.float32),
                delta=dl)


def cfg4_model(seed=4, nx=600, ny=600, nz=300, d=12.5, dt=1e-3):
    """TTI 600x600x300: layered 1500-5000 m/s plus smooth noise, epsilon 0-0.25, delta 0-0.1,
    tilt 0-45 deg and azimuth 0-360 deg as smooth random fields (radians)."""
    shape = (ny, nx, nz)
    vel = dipping_layers(shape, d, seed, 10, 1500.0, 5000.0, 8.0)
    vel = (vel * (1.0 + 0.03 * (smooth_field(shape, seed + 50) - 0.5))).astype(np.float32)
    eps = 0.25 * smooth_field(shape, seed + 100)
    dl = np.minimum(0.1 * smooth_field(shape, seed + 200), eps)
    eps = cap_epsilon(vel, eps, dt, d, margin=0.9)
    dl = np.minimum(dl, eps).astype(np.float32)
    theta = (np.deg2rad(45.0) * smooth_field(shape, seed + 300)).astype(np.float32)
    phi = (2 * np.pi * smooth_field(shape, seed + 400)).astype(np.float32)
    return dict(op="tti", nx=nx, ny=ny, nz=nz, dx=d, dy=d, dz=d, dt=dt, nt=2500, fm=15.0, vel=vel,
                eps=eps.astype(np.float32), delta=dl, theta=theta, phi=phi)


def small_model(op, nx=24, ny=20, nz=22, seed=0, d=10.0):
    """A small random smooth model for parity tests: velocity 1800-3200 m/s, Gardner
    density, epsilon 0-0.2, delta 0-0.08 (<= epsilon), tilt 0-0.6 rad, azimuth 0-2pi."""
    shape = (ny, nx, nz)
    rng = np.random.default_rng(seed)
    vel = (1800.0 + 1400.0 * gaussian_filter(rng.random(shape), 2.0, mode="reflect")).astype(np.float32)
    vel = np.clip(vel, 1800.0, 3200.0).astype(np.float32)
    m = dict(vel=vel)
    if op in ("iso", "vti"):
        m["rho"] = gardner(vel)
    if op in ("vti", "tti"):
        e = (0.2 * gaussian_filter(rng.random(shape), 2.0, mode="reflect")).astype(np.float32)
        m["eps"] = e
        m["delta"] = np.minimum(0.4 * e, (0.08 * rng.random(shape)).astype(np.float32)).astype(np.float32)
    if op == "tti":
        m["theta"] = (0.6 * gaussian_filter(rng.random(shape), 2.0, mode="reflect")).astype(np.float32)
        m["phi"] = (2 * np.pi * gaussian_filter(rng.random(shape), 2.0, mode="reflect")).astype(np.float32)
    return m


def uneven_windows(nshot, nx, ny, nz, seed=4, lo=(160, 160, 200), hi=(320, 320, 300)):
    """Seeded shots on the surface, each with an aperture window of uneven size drawn
    uniformly in [lo, hi] cells (x, y, z), centred on its shot and clipped to the model.
    Returns a3d_window tuples (ox, oy, oz, wx, wy, wz, sx, sy, sz), sources 1-based
    inside the window at depth 1."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nshot):
        wx = int(rng.integers(lo[0], hi[0] + 1))
        wy = int(rng.integers(lo[1], hi[1] + 1))
        wz = int(min(rng.integers(lo[2], hi[2] + 1), nz))
        gx = int(rng.integers(0, nx))
        gy = int(rng.integers(0, ny))
        ox = int(np.clip(gx - wx // 2, 0, nx - wx))
        oy = int(np.clip(gy - wy // 2, 0, ny - wy))
        out.append((ox, oy, 0, wx, wy, wz, gx - ox + 1, gy - oy + 1, 1))
    return out


def strip_model(seed, n=128, d=10.0, dt=8e-4, vmin=1500.0, vmax=4500.0):
    """Random VTI model of the learned-strip study (n^3, d = 10 m, dt = 0.8 ms): 4-10 dipping
    layers in [vmin, vmax] (dips 0-30 deg) times smooth +-5 % noise, Gardner density,
    smooth epsilon in [0, 0.25] and delta in [0, 0.1] with delta <= epsilon, epsilon capped
    for stability."""
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    nl = int(rng.integers(4, 11))
    vel = dipping_layers(shape, d, seed, nl, vmin, vmax, 30.0)
    vel = np.clip(vel * (1.0 + 0.1 * (smooth_field(shape, seed + 50) - 0.5)), vmin, vmax).astype(np.float32)
    eps = 0.25 * smooth_field(shape, seed + 100)
    dl = np.minimum(0.1 * smooth_field(shape, seed + 200), eps)
    eps = cap_epsilon(vel, eps, dt, d, margin=0.95)
    dl = np.minimum(dl, eps).astype(np.float32)
    return dict(op="vti", nx=n, ny=n, nz=n, dx=d, dy=d, dz=d, dt=dt, vel=vel, rho=gardner(vel),
                eps=eps.astype(np.float32), delta=dl)


def strip_source(seed, n):
    """Seeded source of the learned-strip study: 1-based interior position anywhere in the
    volume (so waves reach the faces at all angles) and a Ricker peak frequency in 10-30 Hz."""
    rng = np.random.default_rng(seed + 7)
    sx, sy, sz = (int(v) for v in rng.integers(1, n + 1, 3))
    return dict(sx=sx, sy=sy, sz=sz, fm=float(rng.uniform(10.0, 30.0)))
