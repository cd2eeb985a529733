"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load this module, and only as the checker or the timed CPU
baseline. The product (acoustic3d-fd_amd/) never imports it.
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
# the same source with a*b+c contracted into fused multiply-adds (the reference's nvcc
# default): bounds what contraction changes
FMA_PATH = os.path.join(HERE, "_build", "liboracle_fma.so")

OPS = {"2nd": 0, "iso": 1, "vti": 2, "tti": 3}


class OracleShot(ctypes.Structure):
    _fields_ = [
        ("op", c_int), ("nx", c_int), ("ny", c_int), ("nz", c_int),
        ("dx", c_float), ("dy", c_float), ("dz", c_float),
        ("nb", c_int), ("sx", c_int), ("sy", c_int), ("sz", c_int),
        ("nt", c_int), ("dt", c_float), ("fm", c_float),
        ("wtype", c_int), ("pfac", c_float), ("rec_z", c_int), ("sv", c_int),
        ("snap_interval", c_int), ("snap_skip0", c_int),
        ("mute", c_int), ("v_mute", c_float), ("mute_tt", c_int),
        ("nthreads", c_int),
    ]


_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_fma = None


def load_contracted():
    """The oracle built with contracted multiply-adds (oracle/_build/liboracle_fma.so)."""
    global _fma
    if _fma is None:
        if not os.path.exists(FMA_PATH):
            raise RuntimeError(f"{FMA_PATH} missing: run make -C oracle")
        _fma = _declare(ctypes.CDLL(FMA_PATH))
    return _fma


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
    _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


def _declare(lib):
    fp = POINTER(c_float)
    lib.oracle_run.argtypes = [c_int, POINTER(OracleShot), fp, fp, fp, fp, fp, fp, c_void_p, c_void_p, c_int,
                               POINTER(c_double)]
    lib.oracle_run.restype = c_int
    lib.oracle_time_steps.argtypes = [POINTER(OracleShot), fp, fp, fp, fp, fp, fp, c_int]
    lib.oracle_time_steps.restype = c_double
    lib.oracle_ricker_table.argtypes = [c_int, c_float, c_float, c_int, c_void_p]
    lib.oracle_wavelet.argtypes = [c_int, c_int, c_float, c_float, c_int, c_void_p]
    lib.oracle_pml_tables.argtypes = [c_int, c_int, c_int, c_float, c_float, c_void_p, c_void_p, c_void_p,
                                      c_void_p]
    lib.oracle_d0.argtypes = [fp, c_int, c_int, c_int, c_int, c_float, c_float, c_float]
    lib.oracle_d0.restype = c_double
    lib.oracle_pad.argtypes = [fp, c_int, c_int, c_int, c_int, fp]
    lib.oracle_sponge_factor.argtypes = [c_int, c_int, POINTER(c_double)]
    lib.oracle_taper.argtypes = [c_int, POINTER(c_double)]
    lib.oracle_staggered_dx.argtypes = [fp, c_int, c_int, fp]
    lib.oracle_laplacian_z.argtypes = [fp, c_int, fp]
    lib.oracle_set_strip_net.argtypes = [fp, c_int]
    lib.oracle_set_strip_net.restype = c_int
    lib.oracle_strip_apply.argtypes = [c_int, c_int, c_int, c_float, c_float, c_float, c_float] + [fp] * 15
    lib.oracle_strip_apply.restype = c_int
    lib.oracle_set_strip_profile.argtypes = [c_float]
    lib.oracle_set_strip_head.argtypes = [c_int]
    lib.oracle_set_strip_head.restype = c_int
    lib.oracle_set_strip_half.argtypes = [c_int]
    lib.oracle_set_strip_half.restype = c_int
    lib.oracle_round_half.argtypes = [c_float]
    lib.oracle_round_half.restype = c_float
    lib.oracle_set_strip_profile.restype = c_int
    lib.oracle_set_strip_period.argtypes = [c_int]
    lib.oracle_set_strip_period.restype = c_int
    lib.oracle_set_tti_fault.argtypes = [c_int]
    lib.oracle_set_tti_fault.restype = c_int
    return lib


def _fp(a):
    return None if a is None else a.ctypes.data_as(POINTER(c_float))


def _arr(a, shape):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    assert a.shape == shape, (a.shape, shape)
    return a


# Reference per-operator constants (see the reference files cited in oracle.cpp).
DEFAULTS = {
    "2nd": dict(nb=50, pfac=1.0, rec_z=1, sv=0, wtype=1),
    "iso": dict(nb=20, pfac=10000.0, rec_z=0, sv=0, wtype=1),
    "vti": dict(nb=20, pfac=10.0, rec_z=0, sv=1, wtype=1),
    "tti": dict(nb=40, pfac=10.0, rec_z=0, sv=1, wtype=1),
}


def make_shot(op, nx, ny, nz, dx, dy, dz, sx, sy, sz, nt, dt, fm, nthreads=0, **kw):
    d = dict(DEFAULTS[op])
    d.update(dict(snap_interval=0, snap_skip0=0, mute=0, v_mute=0.0, mute_tt=60))
    d.update(kw)
    return OracleShot(OPS[op], nx, ny, nz, dx, dy, dz, d["nb"], sx, sy, sz, nt, dt, fm, d["wtype"], d["pfac"],
                      d["rec_z"], d["sv"], d["snap_interval"], d["snap_skip0"], d["mute"], d["v_mute"],
                      d["mute_tt"], nthreads)


def run(shot, prec=32, vel=None, rho=None, eps=None, delta=None, theta=None, phi=None, nsnap_max=0,
        contracted=False):
    """Runs the oracle; returns (shot (ny,nx,nt), snaps (k,ny,nx,nz) or None, loop seconds).
    contracted=True runs the build with fused multiply-adds."""
    lib = load_contracted() if contracted else load()
    shape = (shot.ny, shot.nx, shot.nz)
    arrs = [_arr(a, shape) for a in (vel, rho, eps, delta, theta, phi)]
    dt = np.float32 if prec == 32 else np.float64
    out = np.zeros((shot.ny, shot.nx, shot.nt), dt)
    snaps = np.zeros((max(nsnap_max, 1), shot.ny, shot.nx, shot.nz), dt) if nsnap_max > 0 else None
    secs = c_double(0.0)
    n = lib.oracle_run(prec, ctypes.byref(shot), *[_fp(a) for a in arrs], out.ctypes.data_as(c_void_p),
                       None if snaps is None else snaps.ctypes.data_as(c_void_p), nsnap_max,
                       ctypes.byref(secs))
    if n < 0:
        raise RuntimeError(f"oracle_run failed with {n}")
    if snaps is not None:
        snaps = snaps[: min(n, nsnap_max)]
    return out, snaps, secs.value


def set_strip_net(params):
    """Learned absorbing strip of the FP32 VTI instance (nb = 10), or None to switch it off."""
    lib = load()
    if params is None:
        lib.oracle_set_strip_net(None, 0)
        return
    p = np.ascontiguousarray(params, dtype=np.float32).ravel()
    if lib.oracle_set_strip_net(_fp(p), p.size) != 0:
        raise ValueError(f"strip network: wrong parameter count {p.size}")


def set_strip_half(on):
    load().oracle_set_strip_half(int(on))


def round_half(f):
    return load().oracle_round_half(float(f))


def set_strip_profile(amp):
    if load().oracle_set_strip_profile(amp) != 0:
        raise ValueError("bad strip profile amplitude")


def strip_apply(dt, dx, dy, dz, f):
    """One strip correction in place on the padded float32 arrays of dict f (shape (nny, nnx, nnz),
    z fastest): P, Q, u, v, w, px, py, pz, qx, qy, qz, vp, rho, eps, del."""
    lib = load()
    names = ("P", "Q", "u", "v", "w", "px", "py", "pz", "qx", "qy", "qz", "vp", "rho", "eps", "del")
    for k in names:
        if f[k].dtype != np.float32 or not f[k].flags.c_contiguous:
            raise ValueError(f"{k}: contiguous float32 expected")
    nny, nnx, nnz = f["P"].shape
    if lib.oracle_strip_apply(nnx, nny, nnz, dt, dx, dy, dz, *[_fp(f[k]) for k in names]) != 0:
        raise RuntimeError("no strip network set")


def time_steps(shot, nsteps, vel=None, rho=None, eps=None, delta=None, theta=None, phi=None):
    lib = load()
    shape = (shot.ny, shot.nx, shot.nz)
    arrs = [_arr(a, shape) for a in (vel, rho, eps, delta, theta, phi)]
    return lib.oracle_time_steps(ctypes.byref(shot), *[_fp(a) for a in arrs], nsteps)


def set_strip_head(mode):
    """0: the strip's damping factors scale every field of a cell; 1: the split components
    along the normal axis and the normal velocity; 2: the coefficient head (the network's
    multiplier of each strip cell's PML damping, inside the update)."""
    if load().oracle_set_strip_head(int(mode)) != 0:
        raise ValueError("bad strip head mode")


def set_tti_fault(code):
    """Plants a mistake in the TTI rotation (0: none). Codes 1..8 flip the sign of a
    velocity-sweep rotation term, 9..16 of a stress-sweep term, 17 and 18 swap an
    operand (oracle.cpp lists them); used only by tests that must show they catch one."""
    if load().oracle_set_tti_fault(int(code)) != 0:
        raise ValueError(f"bad TTI fault code {code}")


def set_strip_period(steps):
    """Coefficient head: the network runs after the stress update of every `steps`-th step."""
    if load().oracle_set_strip_period(int(steps)) != 0:
        raise ValueError("bad strip period")
