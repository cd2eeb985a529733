"""ctypes binding of lib/libacoustic3d.so (the C ABI declared in include/acoustic3d.h).

The library is the only compute path: there is no CPU fallback. If it is
missing, importing the binding raises, naming the build step.
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_longlong, c_void_p

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libacoustic3d.so")

OP_2ND, OP_ISO, OP_VTI, OP_TTI = 0, 1, 2, 3
OP_NAMES = {"2nd": OP_2ND, "iso": OP_ISO, "vti": OP_VTI, "tti": OP_TTI}

A3D_OK = 0
A3D_EUNSTABLE = -5

# Legacy library file names of the reference (Makefile:10-13); each is a
# symlink to libacoustic3d.so.
LEGACY_NAMES = {
    OP_2ND: "acoustic2order.so",
    OP_ISO: "acoustic1order.so",
    OP_VTI: "acoustic1vti.so",
    OP_TTI: "acoustic1tti.so",
}


class ShotDesc(ctypes.Structure):
    """Mirror of a3d_shot_desc (include/acoustic3d.h)."""

    _fields_ = [
        ("op", c_int),
        ("nx", c_int), ("ny", c_int), ("nz", c_int),
        ("dx", c_float), ("dy", c_float), ("dz", c_float),
        ("nb", c_int),
        ("sx", c_int), ("sy", c_int), ("sz", c_int),
        ("nt", c_int),
        ("dt", c_float), ("fm", c_float),
        ("wtype", c_int),
        ("pfac", c_float),
        ("rec_z", c_int),
        ("sv", c_int),
        ("snap_interval", c_int),
        ("snap_skip0", c_int),
        ("mute", c_int),
        ("v_mute", c_float),
        ("mute_tt", c_int),
        ("storage", c_int),
        ("scheme", c_int),
    ]


class Window(ctypes.Structure):
    """Mirror of a3d_window: origin and size of an aperture window in the full model,
    and the source inside it (1-based, operator convention)."""

    _fields_ = [("ox", c_int), ("oy", c_int), ("oz", c_int), ("nx", c_int), ("ny", c_int), ("nz", c_int),
                ("sx", c_int), ("sy", c_int), ("sz", c_int)]


# Reference per-operator constants: boundary width (Toa_3dfd_cuda.cu:19; iso:395,
# vti:414, tti:418), source factor (iso:407, vti:431, tti:437), receiver depth
# (Toa_3dfd_cuda.cu:159 interior z=1; iso:302 interior z=0), SV taper (vti:567).
REFERENCE_DEFAULTS = {
    OP_2ND: dict(nb=50, pfac=1.0, rec_z=1, sv=0, wtype=1),
    OP_ISO: dict(nb=20, pfac=10000.0, rec_z=0, sv=0, wtype=1),
    OP_VTI: dict(nb=20, pfac=10.0, rec_z=0, sv=1, wtype=1),
    OP_TTI: dict(nb=40, pfac=10.0, rec_z=0, sv=1, wtype=1),
}


def make_desc(op, nx, ny, nz, dx, dy, dz, sx, sy, sz, nt, dt, fm, **kw):
    """Builds a ShotDesc with the reference defaults of `op`, overridden by kw."""
    op = OP_NAMES.get(op, op)
    d = dict(REFERENCE_DEFAULTS[op])
    d.update(dict(snap_interval=0, snap_skip0=0, mute=0, v_mute=0.0, mute_tt=60, storage=0, scheme=0))
    d.update(kw)
    return ShotDesc(op, nx, ny, nz, dx, dy, dz, d["nb"], sx, sy, sz, nt, dt, fm, d["wtype"], d["pfac"],
                    d["rec_z"], d["sv"], d["snap_interval"], d["snap_skip0"], d["mute"], d["v_mute"],
                    d["mute_tt"], d["storage"], d["scheme"])


def desc_dict(d):
    return {name: getattr(d, name) for name, _ in ShotDesc._fields_}


_lib = None


def load(path=None):
    """Loads libacoustic3d.so and declares the signatures. Raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()'"
            " or make -C acoustic3d-fd_amd)")
    lib = ctypes.CDLL(p)
    vp = c_void_p
    fp = POINTER(c_float)
    lib.a3d_create.argtypes = [POINTER(ShotDesc), c_int, POINTER(vp)]
    lib.a3d_destroy.argtypes = [vp]
    lib.a3d_set_model.argtypes = [vp, fp, fp, fp, fp, fp, fp]
    lib.a3d_run.argtypes = [vp, POINTER(ShotDesc), fp, fp, c_int, POINTER(c_int), POINTER(c_double)]
    lib.a3d_set_kernel_timing.argtypes = [vp, c_int]
    lib.a3d_kernel_times.argtypes = [vp, POINTER(c_double), POINTER(c_longlong), c_int]
    lib.a3d_set_launch.argtypes = [vp, c_int, c_int, c_int, c_int]
    lib.a3d_get_launch.argtypes = [vp, POINTER(c_int)]
    lib.a3d_autotune.argtypes = [vp, POINTER(ShotDesc), c_int, POINTER(c_double)]
    lib.a3d_grid_info.argtypes = [vp, POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(c_longlong)]
    lib.a3d_batch_create.argtypes = [POINTER(ShotDesc), c_int, POINTER(Window), c_int, POINTER(vp)]
    lib.a3d_batch_destroy.argtypes = [vp]
    lib.a3d_batch_set_model.argtypes = [vp, fp, fp, fp, fp, fp, fp]
    lib.a3d_batch_tune.argtypes = [vp, c_int, POINTER(c_int)]
    lib.a3d_batch_run.argtypes = [vp, c_int, POINTER(fp), POINTER(c_double)]
    lib.a3d_batch_info.argtypes = [vp, POINTER(c_longlong), POINTER(c_int), POINTER(c_int)]
    lib.a3d_batch_set_graph_chunk.argtypes = [vp, c_int]
    lib.a3d_batch_set_persist.argtypes = [vp, c_int, c_int, c_int]
    lib.a3d_batch_set_tti_angles.argtypes = [vp, c_int]
    lib.a3d_batch_set_zseg.argtypes = [vp, c_int, c_int]
    lib.a3d_batch_set_strip_net.argtypes = [vp, fp, c_int]
    lib.a3d_batch_set_w_sweep.argtypes = [vp, c_int]
    lib.a3d_batch_set_strip_profile.argtypes = [vp, c_float]
    lib.a3d_batch_get_zseg.argtypes = [vp, POINTER(c_int), POINTER(c_int)]
    for name in ("a3d_batch_set_zseg", "a3d_batch_get_zseg", "a3d_batch_set_strip_net", "a3d_batch_set_strip_profile",
                 "a3d_batch_set_w_sweep"):
        getattr(lib, name).restype = c_int
    for name in ("a3d_batch_create", "a3d_batch_destroy", "a3d_batch_set_model", "a3d_batch_tune", "a3d_batch_run",
                 "a3d_batch_info", "a3d_batch_set_graph_chunk"):
        getattr(lib, name).restype = c_int
    lib.a3d_set_strip_net.argtypes = [vp, fp, c_int]
    lib.a3d_strip_info.argtypes = [vp, POINTER(c_int)]
    lib.a3d_set_strip_profile.argtypes = [vp, c_float]
    lib.a3d_set_strip_precision.argtypes = [vp, c_int]
    lib.a3d_set_strip_head.argtypes = [vp, c_int]
    lib.a3d_set_strip_period.argtypes = [vp, c_int]
    lib.a3d_set_strip_period.restype = c_int
    lib.a3d_batch_set_strip_head.argtypes = [vp, c_int]
    lib.a3d_batch_set_strip_head.restype = c_int
    lib.a3d_batch_set_strip_period.argtypes = [vp, c_int]
    lib.a3d_batch_set_strip_period.restype = c_int
    lib.a3d_set_zchunk.argtypes = [vp, c_int]
    lib.a3d_set_graph.argtypes = [vp, c_int]
    lib.a3d_run_shots.argtypes = [vp, POINTER(ShotDesc), c_int, POINTER(c_int), fp, POINTER(c_double)]
    lib.a3d_run_shots.restype = c_int
    lib.a3d_set_graph.restype = c_int
    lib.a3d_set_w_sweep.argtypes = [vp, c_int]
    lib.a3d_set_tti_angles.argtypes = [vp, c_int]
    lib.a3d_get_w_sweep.argtypes = [vp, POINTER(ShotDesc), POINTER(c_int)]
    lib.a3d_strip_collect.argtypes = [vp, vp, POINTER(ShotDesc), c_int, c_int, c_int, c_float, ctypes.c_uint, fp, fp,
                                      fp, c_int, POINTER(c_int)]
    lib.a3d_mlp_create.argtypes = [c_int, POINTER(vp)]
    lib.a3d_mlp_destroy.argtypes = [vp]
    lib.a3d_mlp_set_params.argtypes = [vp, fp, c_int]
    lib.a3d_mlp_set_head.argtypes = [vp, c_int]
    lib.a3d_mlp_get_params.argtypes = [vp, fp, c_int]
    lib.a3d_mlp_predict.argtypes = [vp, fp, c_int, fp]
    lib.a3d_mlp_loss_grad.argtypes = [vp, fp, fp, c_int, POINTER(c_double), fp]
    lib.a3d_mlp_train.argtypes = [vp, fp, fp, c_int, c_int, c_int, c_float, c_float, ctypes.c_uint, POINTER(c_double)]
    for name in ("a3d_set_strip_net", "a3d_strip_info", "a3d_set_strip_profile", "a3d_set_strip_precision",
                 "a3d_set_strip_head",
                 "a3d_set_zchunk", "a3d_set_w_sweep", "a3d_get_w_sweep", "a3d_strip_collect", "a3d_mlp_nparams", "a3d_mlp_create",
                 "a3d_mlp_destroy", "a3d_mlp_set_params", "a3d_mlp_get_params", "a3d_mlp_predict",
                 "a3d_mlp_loss_grad", "a3d_mlp_train"):
        getattr(lib, name).restype = c_int
    # allocation-free shot: model arrays and the record are device addresses (c_void_p)
    lib.a3d_run_static.argtypes = [POINTER(ShotDesc), c_int, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.a3d_run_static.restype = c_int
    lib.a3d_probe_bandwidth.argtypes = [c_int, ctypes.c_size_t, c_int, POINTER(c_double), POINTER(c_double)]
    lib.a3d_probe_bandwidth.restype = c_int
    lib.a3d_probe_stress_pattern.argtypes = [c_int, c_int, c_int, c_int, c_int, POINTER(c_double), POINTER(c_double)]
    lib.a3d_probe_stress_pattern.restype = c_int
    lib.a3d_probe_rot_sincos.argtypes = [c_int, fp, fp, fp, c_int]
    lib.a3d_probe_rot_sincos.restype = c_int
    lib.a3d_last_error.restype = c_char_p
    lib.a3d_version.restype = c_char_p
    for name in ("a3d_create", "a3d_destroy", "a3d_set_model", "a3d_run", "a3d_set_kernel_timing",
                 "a3d_kernel_times", "a3d_set_launch", "a3d_get_launch", "a3d_autotune", "a3d_grid_info"):
        getattr(lib, name).restype = c_int
    if path is None:
        _lib = lib
    return lib


def probe_bandwidth(device=0, nbytes=2 << 30, reps=10):
    """Returns (copy GB/s counting read+write, read-only GB/s) of the calibration streams."""
    lib = load()
    c, r = c_double(), c_double()
    _check(lib, lib.a3d_probe_bandwidth(device, nbytes, reps, ctypes.byref(c), ctypes.byref(r)))
    return c.value, r.value


STRESS_PATTERNS = ("stream_mix_every_lane", "interior_in_place_with_leads", "plus_z_pml", "plus_y_pml",
                   "plus_x_pml_whole_sweep")


def probe_stress_pattern(nnx, nny, nnz, device=0, reps=20):
    """{pattern: (ms per launch, algorithmic GB/s)} of the stress sweep's replayed access pattern."""
    lib = load()
    ms, nb = (c_double * 5)(), (c_double * 5)()
    _check(lib, lib.a3d_probe_stress_pattern(device, nnx, nny, nnz, reps, ms, nb))
    return {k: (ms[i], nb[i] / (ms[i] * 1e-3) / 1e9) for i, k in enumerate(STRESS_PATTERNS)}


def probe_rot_sincos(t, device=0):
    """(sin t, cos t) as the TTI sweeps evaluate them from the angles (float32 arrays)."""
    lib = load()
    t = np.ascontiguousarray(t, dtype=np.float32)
    s, c = np.empty_like(t), np.empty_like(t)
    f = lambda a: a.ctypes.data_as(POINTER(c_float))
    _check(lib, lib.a3d_probe_rot_sincos(device, f(t), f(s), f(c), int(t.size)))
    return s, c


class A3DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"acoustic3d error {code}: {msg}")
        self.code = code


def _check(lib, rc):
    if rc != A3D_OK:
        raise A3DError(rc, lib.a3d_last_error().decode())
    return rc


def _fptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(POINTER(c_float))


def _model(a, shape):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.shape != shape:
        raise ValueError(f"model array shape {a.shape} != {shape} (ny, nx, nz)")
    return a


class Modeller:
    """One device-resident model and its wavefields (an a3d_ctx).

    Arrays use the reference file order: models are (ny, nx, nz) C-ordered,
    z fastest; shot records are (ny, nx, nt), time fastest.
    """

    def __init__(self, desc, device=0):
        self.lib = load()
        self.desc = desc
        self._h = c_void_p()
        _check(self.lib, self.lib.a3d_create(ctypes.byref(desc), device, ctypes.byref(self._h)))
        self._keep = ()

    @property
    def shape(self):
        return (self.desc.ny, self.desc.nx, self.desc.nz)

    def set_model(self, vel, rho=None, eps=None, delta=None, theta=None, phi=None):
        arrs = [_model(a, self.shape) for a in (vel, rho, eps, delta, theta, phi)]
        _check(self.lib, self.lib.a3d_set_model(self._h, *[_fptr(a) for a in arrs]))

    def run(self, desc=None, shot=True, nsnap_max=0):
        """Runs one shot; returns (shot or None, snapshots (k, ny, nx, nz) or None, loop_ms)."""
        d = desc or self.desc
        out = np.empty((d.ny, d.nx, d.nt), np.float32) if shot else None
        snaps = np.empty((nsnap_max, d.ny, d.nx, d.nz), np.float32) if nsnap_max > 0 else None
        ns = c_int(0)
        ms = c_double(0.0)
        rc = self.lib.a3d_run(self._h, ctypes.byref(d), _fptr(out), _fptr(snaps), nsnap_max, ctypes.byref(ns),
                              ctypes.byref(ms))
        _check(self.lib, rc)
        if snaps is not None:
            snaps = snaps[: min(ns.value, nsnap_max)]
        return out, snaps, ms.value

    def run_shots(self, sources, desc=None, shot=True):
        """2nd order: the shots with sources [(sx, sy, sz), ...] (1..16) stepped together, one
        launch per time step; returns (records (n, ny, nx, nt) or None, loop_ms)."""
        d = desc or self.desc
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int32).reshape(-1, 3))
        n = src.shape[0]
        out = np.empty((n, d.ny, d.nx, d.nt), np.float32) if shot else None
        ms = c_double(0.0)
        _check(self.lib, self.lib.a3d_run_shots(self._h, ctypes.byref(d), n, src.ctypes.data_as(POINTER(c_int)),
                                                _fptr(out), ctypes.byref(ms)))
        return out, ms.value

    def set_kernel_timing(self, on=True, stride=1):
        """Time the hot kernels with HIP events; stride N > 1 times every N-th time step."""
        _check(self.lib, self.lib.a3d_set_kernel_timing(self._h, int(stride) if on else 0))

    def kernel_times(self, reset=True):
        ms = (c_double * 2)()
        n = (c_longlong * 2)()
        _check(self.lib, self.lib.a3d_kernel_times(self._h, ms, n, int(reset)))
        return [ms[0], ms[1]], [n[0], n[1]]

    def set_launch(self, ty_vel, zseg_vel, ty_stress, zseg_stress):
        _check(self.lib, self.lib.a3d_set_launch(self._h, ty_vel, zseg_vel, ty_stress, zseg_stress))

    def set_zchunk(self, planes):
        """Two-sweep step in z chunks of `planes` planes (0: whole grid); bit-identical results."""
        _check(self.lib, self.lib.a3d_set_zchunk(self._h, planes))

    def set_graph(self, steps):
        """Time loop as hipGraphs of at most `steps` steps (0: host launches)."""
        _check(self.lib, self.lib.a3d_set_graph(self._h, int(steps)))

    def set_tti_angles(self, on):
        """TTI (FP32/FP16 storage): the sweeps read the tilt and azimuth (8 bytes per point)
        and evaluate the rotation terms in FP32 (1), or read the terms built once per model
        (0, default; bitwise the oracle's arithmetic)."""
        _check(self.lib, self.lib.a3d_set_tti_angles(self._h, int(on)))

    def set_w_sweep(self, mode):
        """Which sweep updates w (ISO, VTI): 0 velocity sweep, 1 stress sweep (bit-identical,
        fewer bytes per step), -1 the autotuner's choice (default)."""
        _check(self.lib, self.lib.a3d_set_w_sweep(self._h, mode))

    def w_sweep(self, desc=None):
        """1 if a shot of desc (default: the handle's) runs with w updated by the stress sweep."""
        m = c_int(0)
        _check(self.lib, self.lib.a3d_get_w_sweep(self._h, ctypes.byref(desc or self.desc), ctypes.byref(m)))
        return m.value

    def launch_info(self):
        v = (c_int * 6)()
        _check(self.lib, self.lib.a3d_get_launch(self._h, v))
        return dict(ty_vel=v[0], zseg_vel=v[1], blocks_vel=v[2], ty_stress=v[3], zseg_stress=v[4],
                    blocks_stress=v[5], w_sweep=self.w_sweep())

    def autotune(self, desc=None, steps=10):
        best = (c_double * 2)()
        _check(self.lib, self.lib.a3d_autotune(self._h, ctypes.byref(desc or self.desc), steps, best))
        return [best[0], best[1]]

    def set_strip_net(self, params):
        """Learned absorbing strip (VTI, FP32, nb = STRIP_NB): network parameters, or None
        for the plain PML of width nb."""
        if params is None:
            _check(self.lib, self.lib.a3d_set_strip_net(self._h, None, 0))
            return
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        _check(self.lib, self.lib.a3d_set_strip_net(self._h, _fptr(p), p.size))

    def set_strip_profile(self, amp):
        """PML amplitude of the learned strip's 6-cell ramp (default 10)."""
        _check(self.lib, self.lib.a3d_set_strip_profile(self._h, amp))

    def set_strip_head(self, mode):
        """What the strip's network does: 0 damp every field (sponge), 1 damp the split
        components along the outward normal axis and the normal velocity (PML-like), 2 scale
        each strip cell's PML damping inside the sweeps (coefficient head)."""
        _check(self.lib, self.lib.a3d_set_strip_head(self._h, int(mode)))

    def set_strip_period(self, steps):
        """Coefficient head: the network runs after the stress sweep of every `steps`-th step."""
        _check(self.lib, self.lib.a3d_set_strip_period(self._h, int(steps)))

    def set_strip_precision(self, mfma):
        """1: strip network on the matrix cores (binary16 operands, FP32 sums); 0: FP32."""
        _check(self.lib, self.lib.a3d_set_strip_precision(self._h, int(mfma)))

    def strip_cells(self):
        n = c_int()
        _check(self.lib, self.lib.a3d_strip_info(self._h, ctypes.byref(n)))
        return n.value

    def grid_info(self):
        a, b, c = c_int(), c_int(), c_int()
        n = c_longlong()
        _check(self.lib, self.lib.a3d_grid_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                                 ctypes.byref(n)))
        return dict(nnx=a.value, nny=b.value, nnz=c.value, updated_points=n.value)

    def close(self):
        if self._h:
            self.lib.a3d_destroy(self._h)
            self._h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


STRIP_NB = 10
MLP_NIN, MLP_NOUT = 16, 2


def strip_collect(teacher, student, desc, every=4, start=0, keep=0.05, seed=0, cap=1 << 20, horizon=1):
    """Teacher-forced strip samples (a3d_strip_collect): returns X (n, 16), T (n, 2), S (n,).
    horizon > 1: the student runs freely (with its network) between synchronisations."""
    lib = load()
    X = np.empty((cap, MLP_NIN), np.float32)
    T = np.empty((cap, MLP_NOUT), np.float32)
    S = np.empty(cap, np.float32)
    n = c_int()
    _check(lib, lib.a3d_strip_collect(teacher._h, student._h, ctypes.byref(desc), every, start, horizon, keep, seed,
                                      _fptr(X), _fptr(T), _fptr(S), cap, ctypes.byref(n)))
    k = n.value
    return X[:k].copy(), T[:k].copy(), S[:k].copy()


class Mlp:
    """The strip network on the GPU (a3d_mlp): 16 -> 64 -> 64 -> 2, tanh, FP32."""

    def __init__(self, device=0):
        self.lib = load()
        self.nparams = self.lib.a3d_mlp_nparams()
        self._h = c_void_p()
        _check(self.lib, self.lib.a3d_mlp_create(device, ctypes.byref(self._h)))

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32).ravel()
        _check(self.lib, self.lib.a3d_mlp_set_params(self._h, _fptr(p), p.size))

    def set_head(self, head):
        """What predict / loss_grad / train fit: 0 the damping head's change, 2 the coefficient
        head's multiplier (targets (m, anything), output 1 without loss)."""
        _check(self.lib, self.lib.a3d_mlp_set_head(self._h, int(head)))

    def params(self):
        p = np.empty(self.nparams, np.float32)
        _check(self.lib, self.lib.a3d_mlp_get_params(self._h, _fptr(p), p.size))
        return p

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty((X.shape[0], MLP_NOUT), np.float32)
        _check(self.lib, self.lib.a3d_mlp_predict(self._h, _fptr(X), X.shape[0], _fptr(Y)))
        return Y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(self.nparams, np.float32)
        loss = c_double()
        _check(self.lib, self.lib.a3d_mlp_loss_grad(self._h, _fptr(X), _fptr(T), X.shape[0], ctypes.byref(loss),
                                                    _fptr(g)))
        return loss.value, g

    def train(self, X, T, epochs, batch=4096, lr=2e-3, lr_end=1e-5, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        hist = (c_double * epochs)()
        _check(self.lib, self.lib.a3d_mlp_train(self._h, _fptr(X), _fptr(T), X.shape[0], epochs, batch, lr, lr_end,
                                                seed, hist))
        return list(hist)

    def close(self):
        if self._h:
            self.lib.a3d_mlp_destroy(self._h)
            self._h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many shots, each an aperture window of one device-resident model (a3d_batch)."""

    def __init__(self, desc, windows, device=0):
        self.lib = load()
        self.desc = desc
        self.windows = (Window * len(windows))(*windows)
        self._h = c_void_p()
        _check(self.lib, self.lib.a3d_batch_create(ctypes.byref(desc), len(windows), self.windows, device,
                                                   ctypes.byref(self._h)))

    def set_model(self, vel, rho=None, eps=None, delta=None, theta=None, phi=None):
        shape = (self.desc.ny, self.desc.nx, self.desc.nz)
        arrs = [_model(a, shape) for a in (vel, rho, eps, delta, theta, phi)]
        _check(self.lib, self.lib.a3d_batch_set_model(self._h, *[_fptr(a) for a in arrs]))

    def tune(self, steps=3):
        n = c_int()
        _check(self.lib, self.lib.a3d_batch_tune(self._h, steps, ctypes.byref(n)))
        return n.value

    def run(self, mode=2, shots=True, which=None):
        """mode 0: sequential plain runs; 1: batched; 2: batched, one hipGraph; 3: persistent kernel
        (device-side time loop, launches captured in one hipGraph). Returns (shots, loop_ms)."""
        nw = len(self.windows)
        outs = [None] * nw
        ptrs = (POINTER(c_float) * nw)()
        for w, win in enumerate(self.windows):
            if shots and (which is None or w in which):
                outs[w] = np.empty((win.ny, win.nx, self.desc.nt), np.float32)
                ptrs[w] = outs[w].ctypes.data_as(POINTER(c_float))
        ms = c_double()
        _check(self.lib, self.lib.a3d_batch_run(self._h, mode, ptrs, ctypes.byref(ms)))
        return outs, ms.value

    def set_graph_chunk(self, steps):
        _check(self.lib, self.lib.a3d_batch_set_graph_chunk(self._h, steps))

    def set_persist(self, tile_rows=4, steps=100, group=1):
        """Mode 3: tile rows of the persistent kernel (4 or 8), time steps per launch and
        consecutive tiles per ticket."""
        _check(self.lib, self.lib.a3d_batch_set_persist(self._h, int(tile_rows), int(steps), int(group)))

    def set_tti_angles(self, on):
        """TTI: every window's sweeps evaluate the rotation terms from the angles (1) or
        read them (0, default)."""
        _check(self.lib, self.lib.a3d_batch_set_tti_angles(self._h, int(on)))

    def info(self):
        nw = len(self.windows)
        pts, n = c_longlong(), c_int()
        tr = (c_int * (2 * nw))()
        _check(self.lib, self.lib.a3d_batch_info(self._h, ctypes.byref(pts), ctypes.byref(n), tr))
        zv, zs = c_int(), c_int()
        _check(self.lib, self.lib.a3d_batch_get_zseg(self._h, ctypes.byref(zv), ctypes.byref(zs)))
        return dict(updated_points=pts.value, nwin=n.value, tile_rows=[(tr[2 * w], tr[2 * w + 1]) for w in range(nw)],
                    zseg=(zv.value, zs.value))

    def set_strip_net(self, params, amp=None, head=None, period=None):
        """The learned strip on every window (VTI, FP32, nb = STRIP_NB); None: the plain PML.
        head / period: as Modeller.set_strip_head / set_strip_period."""
        if amp is not None:
            _check(self.lib, self.lib.a3d_batch_set_strip_profile(self._h, float(amp)))
        if head is not None:
            _check(self.lib, self.lib.a3d_batch_set_strip_head(self._h, int(head)))
        if period is not None:
            _check(self.lib, self.lib.a3d_batch_set_strip_period(self._h, int(period)))
        if params is None:
            _check(self.lib, self.lib.a3d_batch_set_strip_net(self._h, None, 0))
            return
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        _check(self.lib, self.lib.a3d_batch_set_strip_net(self._h, _fptr(p), p.size))

    def set_w_sweep(self, mode):
        """Batched modes (ISO, VTI): w updated by 0 the velocity sweep, 1 the stress sweep, -1 tuned."""
        _check(self.lib, self.lib.a3d_batch_set_w_sweep(self._h, int(mode)))

    def set_zseg(self, vel, stress):
        """z segments per tile column of every window, per sweep (1..64); tune() picks them."""
        _check(self.lib, self.lib.a3d_batch_set_zseg(self._h, int(vel), int(stress)))

    def close(self):
        if self._h:
            self.lib.a3d_batch_destroy(self._h)
            self._h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
