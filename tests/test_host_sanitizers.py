"""The host layer under the sanitizers: tests/sanitize/host_checks.cpp with
fs2_amd/csrc/fof.cpp and host_numerics.cpp (no GPU code), built by g++ with
AddressSanitizer + UndefinedBehaviorSanitizer and, separately, with
ThreadSanitizer for the multithreaded kd-tree build and union-find, and run
with several host threads."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fs2_amd", "csrc")
SOURCES = [os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp"),
           os.path.join(CSRC, "fof.cpp"), os.path.join(CSRC, "host_numerics.cpp")]
TABLE = os.path.join(ROOT, "tests", "golden", "planck_matterpower.dat")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.mark.parametrize("flags,env", [
    (["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
     {"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1", "UBSAN_OPTIONS": "print_stacktrace=1"}),
    (["-fsanitize=thread"], {"TSAN_OPTIONS": "halt_on_error=1"}),
], ids=["asan_ubsan", "tsan"])
def test_host_layer_clean_under_sanitizer(tmp_path, flags, env):
    exe = os.path.join(tmp_path, "host_checks")
    cmd = (["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-pthread"] + flags +
           ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I", os.path.join(ROOT, "include"),
            "-I", CSRC] + SOURCES + ["-o", exe])
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-3000:]
    run_env = dict(os.environ, OMP_NUM_THREADS="8", **env)
    r = subprocess.run([exe, TABLE], capture_output=True, text=True, env=run_env,
                       cwd=tmp_path, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "host checks passed" in r.stdout
    for marker in ("ERROR: AddressSanitizer", "runtime error:", "WARNING: ThreadSanitizer",
                   "LeakSanitizer"):
        assert marker not in r.stderr, r.stderr[-4000:]
