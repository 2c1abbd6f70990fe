"""The host topology build under ThreadSanitizer and AddressSanitizer/UBSan
(CPU only).  Its passes run on several threads with atomic maxima and
increments; tests/native/topo_check.cpp builds whole-scaffold and copy
schedules with the product's topology.cpp, compiled by g++ with each
sanitizer, and the digests must not depend on the thread count or the
sanitizer build.  Marked slow: the sanitized builds take about a minute."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "pose-memory-wiping_amd", "csrc")
DRIVER = os.path.join(REPO, "tests", "native", "topo_check.cpp")
SHAPES = ["3000", "64", "1", "4096", "64", "1", "4096", "64", "0", "5000", "128", "1", "32768", "64", "1"]

pytestmark = pytest.mark.slow


def _build(tmp_path, name, flags):
    exe = str(tmp_path / name)
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fno-omit-frame-pointer", *flags, "-I", CSRC,
           os.path.join(CSRC, "topology.cpp"), DRIVER, "-o", exe, "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=600)
    return exe


def _run(exe, threads, env_extra):
    env = dict(os.environ, POSE_GPU_BUILD_THREADS=str(threads), **env_extra)
    r = subprocess.run([exe, *SHAPES], capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr
    assert "runtime error" not in r.stderr, r.stderr
    return r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_topology_build_sanitized(tmp_path):
    plain = _build(tmp_path, "plain", ["-O2"])
    ref = _run(plain, 1, {})
    tsan = _build(tmp_path, "tsan", ["-fsanitize=thread"])
    assert _run(tsan, 7, {"TSAN_OPTIONS": "halt_on_error=1"}) == ref
    asan = _build(tmp_path, "asan", ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"])
    assert _run(asan, 5, {"ASAN_OPTIONS": "detect_leaks=1"}) == ref
