"""CPU checks of bench.py's contract pieces that need no GPU: the multi-GPU refusal and
the algorithmic-bytes accounting behind roofline.achieved."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_algorithmic_bytes():
    b = _bench()
    cfg = dict(nx=400, ny=400, nz=300)
    npts = 61959168                      # (440-8) * (440-8) * (340-8)
    assert b.algorithmic_bytes(cfg, 0, 0, npts) == (36, 84.0)          # reference formulation, FP32
    assert b.algorithmic_bytes(cfg, 2, 0, npts) == (72, 168.0)         # FP64
    a_vel, a_str = b.algorithmic_bytes(cfg, 0, 1, npts)                # split fields in PML cells only
    f = 401 * 401 * 301 / npts
    assert a_vel == 36 and abs(a_str - (f * 44 + (1 - f) * 84)) < 1e-9
    assert abs(a_str - 52.75) < 0.01
    a_vel, a_str = b.algorithmic_bytes(cfg, 1, 0, npts)                # binary16 wavefields: always scheme 1
    assert a_vel == 8 * 2 + 4 and abs(a_str - (f * (7 * 2 + 16) + (1 - f) * (5 * 2 + 12 * 4 + 16))) < 1e-9
    assert b.OP_BYTES["vti"] == (36, 84, 44) and b.OP_BYTES["2nd"][0] == 16
    # w updated by the stress sweep: velocity R P,rho,u,v W u,v; the stress sweep reads the
    # old Q one plane ahead (its own Q too) and the old w, and writes w
    a_vel, a_str = b.algorithmic_bytes(cfg, 0, 1, npts, wz=1)
    assert a_vel == 24 and abs(a_str - (f * 48 + (1 - f) * 92)) < 1e-9
    assert b.algorithmic_bytes(cfg, 0, 0, npts, wz=1) == (24, 92.0)
    assert b.OP_BYTES_WZ["vti"] == (24, 92, 48) and b.OP_BYTES_WZ["iso"] == (24, 56, 32)


def test_sampled_steps():
    """Time steps timed per shot at a3d_set_kernel_timing(ctx, N): it % N == 0 for it < nt."""
    b = _bench()
    for nt in (1, 7, 8, 9, 3000):
        for stride in (1, 2, 8, 40):
            assert b.sampled_steps(nt, stride) == sum(1 for it in range(nt) if it % stride == 0)


def test_damped_counts_and_dir_split_bytes():
    """Cells by number of damped directions, against an enumeration of the updated cells with
    the PML profile's rule (factor 1 where the ramp argument (npml - id) or (id - n - npml) is
    <= 0), and the scheme-3 mean bytes at its extremes."""
    b = _bench()
    cfg = dict(nx=7, ny=5, nz=6)
    npml, MM = 6, 4

    def damped(i, n):
        ra = (npml - i) if i < npml else (i - n - npml if i >= npml + n else 0)
        return ra > 0

    cnt = [0, 0, 0, 0]
    for x in range(MM, cfg["nx"] + 2 * npml - MM):
        for y in range(MM, cfg["ny"] + 2 * npml - MM):
            for z in range(MM, cfg["nz"] + 2 * npml - MM):
                cnt[damped(x, cfg["nx"]) + damped(y, cfg["ny"]) + damped(z, cfg["nz"])] += 1
    assert b.damped_counts(cfg, npml) == cnt
    big = dict(nx=400, ny=400, nz=300)
    assert sum(b.damped_counts(big, 20)) == 61959168
    # VTI with w made by the stress sweep: 48 / 64 / 80 / 92 B by damped directions
    c = b.damped_counts(big, 20)
    want = (c[0] * 48 + c[1] * 64 + c[2] * 80 + c[3] * 92) / sum(c)
    assert abs(b.algorithmic_bytes(big, 0, 3, 61959168, wz=1)[1] - want) < 1e-9
    c1 = b.algorithmic_bytes(big, 0, 1, 61959168, wz=1)[1]
    assert want < c1 and abs(want - 51.79) < 0.01
