"""The library's host planning (line-list preparation, request plans,
workspace layout, work items, job tables, re-plans after LDM growth) built
with a host compiler under AddressSanitizer and UndefinedBehaviorSanitizer
and run on the CPU (tests/host/plan_check.cpp): any write past a host
buffer the library sizes aborts the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM_INC = "/opt/rocm/include"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(os.path.join(ROCM_INC, "hip")),
                    reason="needs g++ and the HIP headers")
def test_host_planning_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "plan_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fno-sanitize-recover=undefined", "-D__HIP_PLATFORM_AMD__", "-I", ROCM_INC,
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "radis_amd", "csrc"),
           os.path.join(ROOT, "tests", "host", "plan_check.cpp"), "-o", exe, "-lpthread"]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert r.stdout.strip().endswith("ok")
