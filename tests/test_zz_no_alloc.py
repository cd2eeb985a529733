"""The allocation-free entry point a3d_run_static (no counterpart in the reference): called
through the C ABI with device buffers the test allocates itself (HIP runtime via ctypes),
its record must equal a3d_run's bit for bit, on two calls in a row on the same device, the
second with a smaller grid and time axis (the scratch array keeps the first call's
contents, which the second must not see)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HIP_H2D, HIP_D2H = 1, 2


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipSetDevice.argtypes = [ctypes.c_int]
    for f in (hip.hipMalloc, hip.hipFree, hip.hipMemcpy, hip.hipSetDevice):
        f.restype = ctypes.c_int
    return hip


class DeviceBuffers:
    """Device copies of host float32 arrays, freed on exit."""

    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, HIP_H2D) == 0
        return p

    def empty(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.ptrs.append(p)
        return p

    def get(self, p, shape):
        out = np.empty(shape, np.float32)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, HIP_D2H) == 0
        return out

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        for p in self.ptrs:
            self.hip.hipFree(p)


KEYS = {"2nd": ("vel",), "iso": ("vel", "rho"), "vti": ("vel", "rho", "eps", "delta"),
        "tti": ("vel", "eps", "delta", "theta", "phi")}
# (operator, grid, source, nt, extra descriptor fields): the second case of each pair is smaller
CASES = {
    "vti": [((40, 36, 30), (20, 18, 4), 200, dict(nb=10, scheme=4, mute=1, mute_tt=0)),
            ((24, 20, 22), (12, 10, 3), 120, dict(nb=10, scheme=0))],
    "2nd": [((48, 40, 36), (24, 20, 5), 160, dict(nb=20, rec_z=5)),
            ((24, 20, 22), (12, 10, 3), 100, dict(nb=8))],
    "tti": [((36, 30, 28), (18, 15, 4), 140, dict(nb=12, scheme=3)),
            ((24, 20, 22), (12, 10, 3), 90, dict(nb=12, scheme=4))],
}


@pytest.mark.parametrize("op", ["vti", "2nd", "tti"])
def test_static_entry_equals_a3d_run_twice(op):
    from a3dfd import Modeller, load, make_desc
    from a3dfd.models import small_model
    lib = load()
    hip = _hip()
    assert hip.hipSetDevice(0) == 0
    for k, ((nx, ny, nz), src, nt, kw) in enumerate(CASES[op]):
        m = small_model(op, nx, ny, nz, seed=71 + k)
        desc = make_desc(op, nx, ny, nz, 10.0, 10.0, 10.0, *src, nt, 1e-3, 25.0, **kw)
        print(f"\n[gpu-static] {op} call {k}: grid {nx}x{ny}x{nz} nt {nt} {kw}", flush=True)
        with Modeller(desc) as mdl:
            mdl.set_model(**m)
            ref, _, _ = mdl.run()
        with DeviceBuffers(hip) as buf:
            args = [buf.put(m[key]) if key in m else None for key in ("vel", "rho", "eps", "delta", "theta", "phi")]
            shot = buf.empty(4 * nx * ny * nt)
            rc = lib.a3d_run_static(ctypes.byref(desc), 0, *args, shot, None)
            assert rc == 0, lib.a3d_last_error().decode()
            got = buf.get(shot, (ny, nx, nt))
        print(f"[gpu-static] {op} call {k}: max |record| {np.abs(ref).max():.3e}, "
              f"max |diff| {np.abs(got - ref).max():.3e}", flush=True)
        assert np.abs(ref).max() > 0
        assert np.array_equal(got, ref)


def test_static_entry_rejects_what_it_does_not_run():
    from a3dfd import load, make_desc
    lib = load()
    hip = _hip()
    assert hip.hipSetDevice(0) == 0
    with DeviceBuffers(hip) as buf:
        v = buf.empty(4 * 600 * 600 * 300)
        shot = buf.empty(4)
        big = make_desc("vti", 600, 600, 300, 10.0, 10.0, 10.0, 300, 300, 1, 100, 1e-3, 25.0)
        assert lib.a3d_run_static(ctypes.byref(big), 0, v, v, v, v, None, None, shot, None) == -1
        assert b"MiB of scratch" in lib.a3d_last_error()
        snap = make_desc("vti", 20, 20, 20, 10.0, 10.0, 10.0, 10, 10, 1, 10, 1e-3, 25.0, snap_interval=5)
        assert lib.a3d_run_static(ctypes.byref(snap), 0, v, v, v, v, None, None, shot, None) == -1
        fp16 = make_desc("vti", 20, 20, 20, 10.0, 10.0, 10.0, 10, 10, 1, 10, 1e-3, 25.0, storage=1)
        assert lib.a3d_run_static(ctypes.byref(fp16), 0, v, v, v, v, None, None, shot, None) == -1
        assert lib.a3d_run_static(ctypes.byref(fp16), 0, None, v, v, v, None, None, shot, None) == -1
