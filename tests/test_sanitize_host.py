"""Host sanitizer build (CPU): the .cw loader and writer, scw_cw_unpack and
the planner's host half compiled with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/sanitize/Makefile) and driven by
tests/sanitize/host_driver.cpp on valid structures, every truncation of a
valid file, corrupted headers, random files, lattice / dense / scattered /
empty regions and invalid tables.  Any memory error, leak or undefined
behaviour aborts the driver; a wrong status fails it."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")
LIB = os.path.join(os.path.dirname(HERE), "..", "scwatter_amd", "libscwatter_amd.so")


@pytest.mark.timeout(600)
def test_host_paths_under_asan_and_ubsan(tmp_path):
    if not os.path.exists("/opt/rocm/bin/hipcc") or shutil.which("make") is None:
        pytest.skip("hipcc or make not available")
    if not os.path.exists(LIB):
        pytest.skip("libscwatter_amd.so not built")
    b = subprocess.run(["make", "-s", "-C", HERE, "-j4"], capture_output=True, text=True)
    assert b.returncode == 0, b.stdout[-2000:] + b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([os.path.join(HERE, "_build", "host_driver"), str(tmp_path)], capture_output=True, text=True,
                       env=env, timeout=500)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out and "LeakSanitizer" not in out, out[-4000:]
    assert r.returncode == 0 and "all checks passed" in r.stdout, out[-4000:]
