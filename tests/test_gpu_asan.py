"""Host-side AddressSanitizer / UndefinedBehaviorSanitizer run of libfdfd.

fdfd.jl_amd/lib/asan_driver is the library's sources with the host code
instrumented (device code is not) linked into tests/asan_driver.cpp, which
drives every host path of the C ABI on small problems.  A heap overflow,
use-after-free, double free or undefined behaviour in the host
runtime aborts the program with a sanitizer report; the test requires a clean
exit and the driver's own PASSED line.  build_lib.build_asan() (called by
__graft_entry__.build()) makes the program.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fdfd.jl_amd", "lib", "asan_driver")

pytestmark = pytest.mark.gpu


def test_host_runtime_under_asan_ubsan():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing: run build_lib.build_asan() (or __graft_entry__.build())")
    env = dict(os.environ)
    # leak detection needs ptrace, which the runtime's own threads defeat; the
    # other ASan checks and every UBSan check abort the run
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:verify_asan_link_order=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    print("running", EXE, flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    print(out[-4000:])
    assert "ERROR: AddressSanitizer" not in out
    assert "runtime error:" not in out
    assert r.returncode == 0, out[-2000:]
    assert "PASSED: 0 failures" in r.stdout
