"""The library's host-side bookkeeping (updes_amd/csrc/host_logic.cpp:
workspace carving, the GMRES least-squares problem and stopping decisions,
the panel leaves' launch geometry) built with AddressSanitizer and UBSan
and driven by tests/host_logic_check.cpp on the CPU."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_logic_under_sanitizers():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "host_logic_check")
        cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
               "-fno-sanitize-recover=all", os.path.join(ROOT, "updes_amd", "csrc", "host_logic.cpp"),
               os.path.join(ROOT, "tests", "host_logic_check.cpp"), "-o", exe]
        subprocess.run(cmd, check=True, capture_output=True, text=True)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
        res = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
        assert res.returncode == 0, res.stdout + res.stderr
        assert res.stdout.strip() == "ok", res.stdout + res.stderr
