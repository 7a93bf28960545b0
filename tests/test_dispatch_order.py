"""Host-side dispatch order of multi-view renders (tvk::group_costs,
tvk::dispatch_order), built from the library's host sources with hipcc
(host code only) and run on the CPU: group costs against a brute-force sum over
the bricks' pixel rectangles; for single groups and rows, a permutation of the
launch's groups in 8 contiguous, balanced ranges in decreasing cost."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="no hipcc")
def test_dispatch_order_host_check(tmp_path):
    exe = tmp_path / "dispatch_order_check"
    cmd = [HIPCC, "-x", "hip", "--offload-host-only", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "tuvok_amd", "csrc"),
           "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "dispatch_order_check.cpp"),
           os.path.join(ROOT, "tuvok_amd", "csrc", "host", "host.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
