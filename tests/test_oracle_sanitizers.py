"""CPU: the oracle built with the host address and undefined-behaviour sanitizers
(oracle/Makefile target `sanitize`, driver oracle/sanitize_check.cpp) runs small cases of
every operator in FP32 and FP64 (2nd order, ISO with all six wavelets and the mute, VTI
and TTI with the SV taper, the mute and snapshots) and the learned strip under each head:
an out-of-bounds access or undefined operation in the restatement aborts the run."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_address_and_ub_sanitizers():
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-C", odir, "sanitize"], check=True, capture_output=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(odir, "_build", "sanitize_check")], capture_output=True, text=True,
                       timeout=600, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "sanitize check: all cases ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
