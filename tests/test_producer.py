"""The producer CLI (pose_wipe_bench): argument handling on the CPU, and a real
wipe of a temporary file on the GPU checked byte for byte against the oracle
and by the producer's own --verify (reference CPU API)."""
import os
import re
import subprocess
import tempfile

import pytest

from pyoracle import REF_DIR, ref_available

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "pose-memory-wiping_amd", "pose_wipe_bench")
SEED = bytes([0x44, 0x69, 0x73, 0x6b, 0x57, 0x69, 0x70, 0x65, 0x01, 0x23, 0x45, 0x67, 0x89,
              0xab, 0xcd, 0xef, 0xfe, 0xed, 0xfa, 0xce, 0x0b, 0xad, 0xf0, 0x0d, 0xaa, 0xbb,
              0xcc, 0xdd, 0xee, 0xff, 0x00, 0x11])


def run(*args, timeout=300):
    return subprocess.run([BIN, *args], capture_output=True, text=True, timeout=timeout)


def test_cli_rejects_bad_arguments():
    assert run().returncode == 2
    assert run("--disk", "/nonexistent", "--chunk-blocks", "1000").returncode == 2
    assert run("--disk", "/nonexistent", "--chunk-blocks", str(1 << 21)).returncode == 2
    assert run("--bogus").returncode == 2
    # One process drives one GPU; the reference's team flags are validated as
    # it validates them (disk_wipe_bench.cu:767-806).
    r = run("--disk", "/nonexistent", "--gpus", "0,1")
    assert r.returncode == 2 and "one GPU per process" in r.stderr
    assert run("--disk", "/nonexistent", "--gpus", "0,0").returncode == 2
    assert run("--disk", "/nonexistent", "--gpus", "x").returncode == 2
    assert run("--disk", "/nonexistent", "--mode", "bogus").returncode == 2


def test_cli_refuses_to_write_without_confirm():
    with tempfile.NamedTemporaryFile() as f, tempfile.NamedTemporaryFile() as g:
        f.truncate(1 << 20)
        g.truncate(1 << 19)
        r = run("--disk", f.name, "--disk", g.name, "--chunk-blocks", "1024", "--mode", "gpus-per-disk")
        assert r.returncode == 1
        assert "REFUSING" in r.stderr
        assert r.stdout.count("target      :") == 2
        f.seek(0)
        g.seek(0)
        assert f.read() == bytes(1 << 20) and g.read() == bytes(1 << 19)


def test_cli_bad_second_target_writes_nothing():
    """Every target is opened and sized before the first write: a second
    target that cannot be opened ends the run with the first untouched."""
    with tempfile.NamedTemporaryFile() as f:
        f.truncate(1 << 20)
        r = run("--disk", f.name, "--disk", "/nonexistent/target", "--chunk-blocks", "1024", "--confirm")
        assert r.returncode == 1 and "/nonexistent/target" in r.stderr
        f.seek(0)
        assert f.read() == bytes(1 << 20)


def test_verify_library_checked_before_writing():
    """--verify resolves the reference CPU library (--cpu-lib, else
    $POSE_COMMON_LIB, else libpose_common.so) before the target is opened: a
    missing library, or one without the graph.h symbols, ends the run with
    status 2 and the target untouched."""
    with tempfile.NamedTemporaryFile() as f:
        f.truncate(1 << 20)
        env = dict(os.environ, POSE_COMMON_LIB="/nonexistent/libpose_common.so")
        for extra in ([], ["--cpu-lib", "libc.so.6"]):
            r = subprocess.run([BIN, "--disk", f.name, "--chunk-blocks", "1024", "--confirm", "--verify", "3",
                                *extra], capture_output=True, text=True, timeout=60, env=env)
            assert r.returncode == 2, r.stderr
            assert "--verify needs the reference CPU library" in r.stderr
        f.seek(0)
        assert f.read() == bytes(1 << 20)


@pytest.mark.gpu
def test_wipe_file_and_verify(oracle):
    cb, n = 1 << 10, 24
    with tempfile.NamedTemporaryFile(dir=os.environ.get("TMPDIR", "/tmp")) as f:
        f.truncate(n * cb * 32 + 1000)          # tail smaller than a super-chunk stays
        args = ["--disk", f.name, "--chunk-blocks", str(cb), "--confirm", "--hbm-window-mib", "2048",
                "--stage-mib", "64", "--challenge", "50"]
        env = None
        if ref_available(16):
            # The library found by default: $POSE_COMMON_LIB (no --cpu-lib flag).
            args += ["--verify", "5", "--cpu-threads", "16"]
            env = dict(os.environ, POSE_COMMON_LIB=os.path.join(REF_DIR, "libposeref_t16.so"))
        r = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=300, env=env)
        print(r.stdout, r.stderr, flush=True)
        assert r.returncode == 0, r.stderr
        assert "GPU wipe:" in r.stdout
        assert "50/50 random blocks match GPU answers" in r.stdout
        if ref_available(16):
            assert "5/5 sampled super-chunks match" in r.stdout
        f.seek(0)
        data = f.read()
    assert data[:n * cb * 32] == oracle.label_region(SEED, cb, n, 0)
    assert data[n * cb * 32:] == bytes(1000)
    # Verifier spot checks without a CPU library: intact, then with the second
    # half of the labels overwritten (random challenges must catch it).
    with tempfile.NamedTemporaryFile(dir=os.environ.get("TMPDIR", "/tmp")) as g:
        g.write(data[:n * cb * 32])
        g.flush()
        r = run("--disk", g.name, "--chunk-blocks", str(cb), "--verify-only", "--challenge", "40")
        assert r.returncode == 0 and "40/40 challenged blocks match" in r.stdout, r.stdout + r.stderr
        g.seek(n * cb * 16)
        g.write(bytes(n * cb * 16))
        g.flush()
        r = run("--disk", g.name, "--chunk-blocks", str(cb), "--verify-only", "--challenge", "40")
        assert r.returncode == 1 and "MISMATCH" in r.stderr and "VERIFY FAILED" in r.stdout
    if ref_available(16):
        with tempfile.NamedTemporaryFile(dir=os.environ.get("TMPDIR", "/tmp")) as g:
            g.write(data[:n * cb * 32])
            g.flush()
            r = run("--disk", g.name, "--chunk-blocks", str(cb), "--verify-only", "--verify", "4",
                    "--cpu-lib", os.path.join(REF_DIR, "libposeref_t16.so"), "--cpu-threads", "16")
            assert r.returncode == 0 and "VERIFY OK" in r.stdout


@pytest.mark.gpu
def test_dry_run_labels_without_writing():
    with tempfile.NamedTemporaryFile(dir=os.environ.get("TMPDIR", "/tmp")) as f:
        f.truncate(8 << 20)
        r = run("--disk", f.name, "--chunk-blocks", "4096", "--dry-run", "--hbm-window-mib", "2048",
                "--stage-mib", "64")
        assert r.returncode == 0, r.stderr
        assert "GPU wipe: 8 MiB" in r.stdout
        f.seek(0)
        assert f.read() == bytes(8 << 20)


@pytest.mark.gpu
def test_wipe_two_targets_in_turn(oracle):
    """Two targets in one run (the reference's repeatable --disk with one GPU):
    each gets the labels of its own absolute positions, the larger first; the
    reference's team flags are accepted (--gpus-per-disk beyond one GPU warns)."""
    cb = 1 << 10
    sizes = [10, 24]
    tmp = os.environ.get("TMPDIR", "/tmp")
    with tempfile.NamedTemporaryFile(dir=tmp) as a, tempfile.NamedTemporaryFile(dir=tmp) as b:
        a.truncate(sizes[0] * cb * 32)
        b.truncate(sizes[1] * cb * 32)
        js = os.path.join(tmp, f"wipe_{os.getpid()}.json")
        r = run("--disk", a.name, "--disk", b.name, "--chunk-blocks", str(cb), "--confirm", "--gpus", "0",
                "--mode", "disk-per-gpu", "--gpus-per-disk", "2", "--hbm-window-mib", "2048", "--stage-mib", "64",
                "--challenge", "20", "--json", js)
        print(r.stdout, r.stderr, flush=True)
        assert r.returncode == 0, r.stderr
        assert "exceeds 1 selected GPU" in r.stderr
        assert r.stdout.count("GPU wipe:") == 2
        agg = re.search(r"aggregate +: ([0-9.]+) MiB/s", r.stdout)
        assert agg and float(agg.group(1)) > 0, r.stdout
        assert r.stdout.index(b.name + ": topology") < r.stdout.index(a.name + ": topology")
        assert r.stdout.count("20/20 random blocks match GPU answers") == 2
        import json
        with open(js) as jf:
            res = json.load(jf)
        os.unlink(js)
        assert res["ok"] and res["chunk_blocks"] == cb and res["gpu"] == 0
        assert res["total_bytes"] == sum(sizes) * cb * 32
        by_path = {t["path"]: t for t in res["targets"]}
        for f, n in ((a, sizes[0]), (b, sizes[1])):
            t = by_path[f.name]
            assert t["bytes_done"] == n * cb * 32 and t["super_chunks"] == n
            assert t["challenge"] == {"checked": 20, "mismatches": 0}
        for f, n in ((a, sizes[0]), (b, sizes[1])):
            f.seek(0)
            assert f.read() == oracle.label_region(SEED, cb, n, 0), f.name
