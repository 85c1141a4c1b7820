"""The host-only set-up logic of libsfem_hip.so (element geometry, the
element-block partition, the CSR tables; solidfem_amd/csrc/sfem_host.cpp)
built for the CPU with g++ and run under AddressSanitizer and
UndefinedBehaviorSanitizer by the invariant checks of tests/host_check.cpp:
any out-of-bounds access, use of freed memory, leak, signed overflow or
misaligned access aborts the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "solidfem_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_host_setup_logic_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "host_check")
    build = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=all", f"-I{CSRC}", os.path.join(ROOT, "tests", "host_check.cpp"),
             os.path.join(CSRC, "sfem_host.cpp"), "-o", exe]
    r = subprocess.run(build, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "all host checks passed" in r.stdout
