"""The request loop's host logic (imageworsener_amd/csrc/iw_server.h: the
admission, flag polling and flight completion behind iwamd_server_*) under
ThreadSanitizer with a mock executor on the CPU (tests/server_mock.cpp).

Built with the ROCm LLVM clang++: its ThreadSanitizer runtime understands
std::condition_variable::wait_for (gcc 11's reports a false "double lock"
there)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "imageworsener_amd", "csrc")


def _clang():
    for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++")):
        if c and os.path.exists(c):
            return c
    return None


def test_request_loop_under_tsan(tmp_path):
    cxx = _clang()
    if cxx is None:
        pytest.skip("no clang++ with ThreadSanitizer")
    exe = str(tmp_path / "server_mock")
    subprocess.run([cxx, "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-Wall", "-I" + CSRC,
                    os.path.join(HERE, "server_mock.cpp"), "-o", exe, "-lpthread"], check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([exe], env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "all checks passed" in r.stdout
