"""GPU tests of the batched (hipGraph) and serving paths: every replay and every
served request must be byte-identical to the CPU labeler, and a request's labels
must be the same whether it runs alone or mixed with concurrent requests."""
import os
import sys
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import pose_gpu as pg  # noqa: E402
CPU_THREADS = min(16, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    return torch.device("cuda:0")


def test_plan_graph_replays(dev, oracle):
    """A captured plan replayed several times, with eager device work and
    synchronisations between replays; every replay checked."""
    topo = {cb: pg.Topology(cb) for cb in (1 << 10, 1 << 12)}
    meta = [(b"\x01" * 32, 1 << 10, 5, 0), (b"\x02" * 32, 1 << 12, 3, 9 << 12),
            (b"\x03" * 32, 1 << 10, 2, 77 << 10)]
    outs = [torch.zeros(n * cb * 32, dtype=torch.uint8, device=dev) for _, cb, n, _ in meta]
    sessions = [(topo[cb], k, rbo, n, o.data_ptr()) for (k, cb, n, rbo), o in zip(meta, outs)]
    exp = [oracle.label_superchunks(k, cb, list(range(n)), rbo, CPU_THREADS) for k, cb, n, rbo in meta]
    plan = pg.Plan(sessions, graph=True)
    assert plan.info.uses_graph == 1
    for rep in range(3):
        say(f"replay {rep}")
        for o in outs:
            o.fill_(rep + 1)           # eager work between replays
        torch.cuda.synchronize()
        plan.run()
        for o, e in zip(outs, exp):
            assert o.cpu().numpy().tobytes() == e, rep
        x = torch.ones(1 << 20, device=dev).sum()   # unrelated eager work
        torch.cuda.synchronize()
        assert float(x) == float(1 << 20)
    plan.close()


def test_plan_without_graph_matches(dev, oracle):
    T = pg.Topology(1 << 11)
    o = torch.zeros(4 * (1 << 11) * 32, dtype=torch.uint8, device=dev)
    plan = pg.Plan([(T, b"k" * 32, 1 << 13, 4, o.data_ptr())], slots=2, graph=False)
    assert plan.info.uses_graph == 0 and plan.info.slots == 2
    plan.run()
    assert o.cpu().numpy().tobytes() == oracle.label_superchunks(b"k" * 32, 1 << 11, [0, 1, 2, 3], 1 << 13, CPU_THREADS)
    plan.close()


def test_plan_groups_parallel_branches(dev, oracle):
    """Sessions of different launch shapes (64- and 128-lane topologies, one or
    four lanes per node) become separate persistent launches, run as parallel
    branches of one hipGraph (or forked streams without a graph); every replay
    of every session is byte-identical to the oracle."""
    t64 = pg.Topology(1 << 10, 64, 1)
    t128 = pg.Topology(1 << 12, 128, 0)
    t64b = pg.Topology(1 << 11, 64, 1)
    meta = [(b"\x01" * 32, t64, 3, 0, (0, 2, 1)),
            (b"\x02" * 32, t128, 2, 5 << 12, (0, 2, 1)),
            (b"\x03" * 19, t64b, 2, 7 << 11, (0, 1, 4)),
            (b"\x04" * 32, t64, 1, 9 << 10, (0, 1, 4))]
    outs = [torch.zeros(n * T.info.chunk_blocks * 32, dtype=torch.uint8, device=dev) for _, T, n, _, _ in meta]
    exp = [oracle.label_superchunks(k, T.info.chunk_blocks, list(range(n)), rbo, CPU_THREADS)
           for k, T, n, rbo, _ in meta]
    sessions = [(T, k, rbo, n, o.data_ptr()) for (k, T, n, rbo, _), o in zip(meta, outs)]
    for graph in (True, False):
        plan = pg.Plan(sessions, graph=graph, launches=[m[4] for m in meta])
        say(f"groups graph={graph}: {plan.groups()}")
        assert plan.info.groups == 3 and plan.info.uses_graph == int(graph)
        shapes = sorted((g["lanes"], g["lanes_per_node"], g["super_chunks"]) for g in plan.groups())
        assert shapes == [(64, 1, 3), (64, 4, 3), (128, 1, 2)]
        assert plan.info.super_chunks == 8
        for rep in range(2):
            for o in outs:
                o.fill_(0xA5 + rep)
            torch.cuda.synchronize()
            plan.run()
            for o, e in zip(outs, exp):
                assert o.cpu().numpy().tobytes() == e, (graph, rep)
        plan.close()
    for T in (t64, t128, t64b):
        T.close()


def _requests(seed, n_req):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_req):
        cb = int(1 << rng.integers(8, 15))
        n = int(rng.integers(1, 6))
        key = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        rbo = int(rng.integers(0, 1 << 30)) * cb
        out.append((key, cb, n, rbo))
    return out


def test_server_alone_and_in_mix_identical(dev, oracle):
    reqs = _requests(5, 10)
    # Alone: one request at a time through a single-stream server.
    alone = []
    srv = pg.Server(hbm_budget_bytes=8 << 30, pinned_bytes=1 << 30, max_inflight=1)
    for key, cb, n, rbo in reqs:
        buf = bytearray(n * cb * 32)
        rid = srv.submit(key, cb, rbo, n, host_out=buf)
        assert srv.wait(rid) == 0
        alone.append(bytes(buf))
    srv.close()
    # Mixed: all submitted while others run, four streams, small launches so
    # requests are split across launches and interleave on the device.
    srv = pg.Server(hbm_budget_bytes=8 << 30, pinned_bytes=1 << 30, max_inflight=4,
                    max_super_per_launch=3)
    bufs, ids = [], []
    for i, (key, cb, n, rbo) in enumerate(reqs):
        buf = bytearray(n * cb * 32)
        bufs.append(buf)
        ids.append(srv.submit(key, cb, rbo, n, host_out=buf))
    assert srv.drain() == 0
    st = srv.stats()
    srv.close()
    say(f"mix: launches={st['launches']} peak_inflight={st['peak_inflight']}")
    assert st["super_chunks_done"] == sum(n for _, _, n, _ in reqs)
    assert st["peak_inflight"] >= 2, "requests did not overlap on the device"
    for (key, cb, n, rbo), a, b in zip(reqs, alone, bufs):
        assert a == bytes(b)
        assert a == oracle.label_superchunks(key, cb, list(range(n)), rbo, CPU_THREADS)


def test_server_file_sink_offsets(dev, oracle):
    reqs = [(b"\x05" * 32, 1 << 10, 3, 0, 0), (b"\x06" * 32, 1 << 12, 2, 5 << 12, 3 << 15)]
    size = max(off + n * cb * 32 for _, cb, n, _, off in reqs)
    with tempfile.NamedTemporaryFile(dir=os.environ.get("TMPDIR", "/tmp")) as f:
        f.truncate(size)
        srv = pg.Server(hbm_budget_bytes=4 << 30, pinned_bytes=256 << 20, sink_fd=f.fileno())
        for key, cb, n, rbo, off in reqs:
            srv.submit(key, cb, rbo, n, sink_offset=off)
        assert srv.drain() == 0
        srv.close()
        f.seek(0)
        data = f.read()
    for key, cb, n, rbo, off in reqs:
        assert data[off:off + n * cb * 32] == oracle.label_superchunks(key, cb, list(range(n)), rbo, CPU_THREADS)


def test_server_rejects_bad_requests(dev):
    # 8 MiB: below one 2^17 super-chunk's 8 MiB scratch pool + 4 MiB output.
    srv = pg.Server(hbm_budget_bytes=8 << 20, pinned_bytes=16 << 20)
    with pytest.raises(RuntimeError):
        srv.submit(b"x" * 32, 1000, 0, 1)          # not a power of two
    with pytest.raises(RuntimeError):
        srv.submit(b"x" * 32, 1 << 21, 0, 1)       # beyond 2^20
    with pytest.raises(RuntimeError):
        srv.submit(b"x" * 32, 1 << 17, 0, 1)       # one super-chunk exceeds the budget
    assert srv.drain() == 0
    srv.close()


def test_autotune_choice_is_valid_and_exact(dev, oracle):
    cb = 1 << 12
    best, trials = pg.autotune(cb, 48, levels=True, force=True)
    say(f"autotune: {best}")
    assert trials and best["lanes"] in (64, 128, 256) and best["mode"] in (0, 1)
    assert best["batch_mib_per_s"] == max(t["batch_mib_per_s"] for t in trials)
    again, _ = pg.autotune(cb, 40)            # same bucket: cached answer
    assert again == best
    T = pg.Topology(cb, best["lanes"], best["mode"])
    o = torch.zeros(3 * cb * 32, dtype=torch.uint8, device=dev)
    pg.label_batch([(T, b"t" * 32, 0, 3, o.data_ptr())], best["slots"] and min(3, best["slots"]))
    assert o.cpu().numpy().tobytes() == oracle.label_superchunks(b"t" * 32, cb, [0, 1, 2], 0, CPU_THREADS)
    T.close()


def test_server_split_launch_at_2_17(dev, golden):
    """A request of one 2^17 super-chunk on an idle server (labeled as two
    copy jobs in the four-lane form): the streamed labels, copied into the
    caller's buffer, equal the reference build's super-chunk, and the
    statistics account for the kernel, the device-to-pinned copy and the
    writer."""
    import hashlib
    cb = 1 << 17
    r = [r for r in golden["regions"] if r["cb"] == cb][0]
    srv = pg.Server(hbm_budget_bytes=4 << 30, pinned_bytes=64 << 20, preload=[cb])
    buf = bytearray(cb * 32)
    say("server: one 2^17 super-chunk")
    rid = srv.submit(bytes.fromhex(r["seed"]), cb, r["block_offset"], 1, host_out=buf)
    assert srv.wait(rid) == 0
    st = srv.stats()
    srv.close()
    assert bytes(buf[:32]).hex() == r["rank0"][0]
    assert hashlib.sha256(bytes(buf)).hexdigest() == r["super_sha256"][0]
    assert st["launches"] == 1 and st["super_chunks_done"] == 1
    assert st["kernel_ms_total"] > 0 and st["copy_ms_total"] > 0 and st["sink_ms_total"] > 0
    say(f"kernel {st['kernel_ms_total']:.0f} ms (split: about 400), copy {st['copy_ms_total']:.1f} ms, "
        f"writer {st['sink_ms_total']:.1f} ms")


def _close_within(obj, seconds):
    """Close a Server or Plan on a helper thread; fail if it does not return."""
    import threading
    t = threading.Thread(target=obj.close, daemon=True)
    t.start()
    t.join(seconds)
    assert not t.is_alive(), f"close() did not return within {seconds} s"


def test_server_failed_launch_ends_cleanly(dev, oracle, monkeypatch, capfd):
    """A launch whose retirement reports a failure (host-side injection: the
    second launch to retire; no device fault) fails every live request once:
    the request finished before it still waits to 0, the others to -1, drain
    returns -1, later submits are refused, the error is reported once and
    destroy returns."""
    monkeypatch.setenv("POSE_GPU_SERVER_FAIL_RETIRE", "2")
    srv = pg.Server(hbm_budget_bytes=1 << 30, pinned_bytes=64 << 20, max_inflight=1,
                    max_super_per_launch=1)
    monkeypatch.delenv("POSE_GPU_SERVER_FAIL_RETIRE")
    cb = 1 << 10
    first = bytearray(cb * 32)
    ids = [srv.submit(b"\x31" * 32, cb, 0, 1, host_out=first)]
    ids += [srv.submit(bytes([0x40 + i]) * 32, cb, i * cb, 3) for i in range(3)]
    say("failure injected at the second retirement")
    assert srv.drain() == -1
    rcs = [srv.wait(i) for i in ids]
    assert rcs == [0, -1, -1, -1], rcs
    assert bytes(first) == oracle.label_superchunks(b"\x31" * 32, cb, [0], 0, CPU_THREADS)
    with pytest.raises(RuntimeError):
        srv.submit(b"\x32" * 32, cb, 0, 1)
    st = srv.stats()
    assert st["launches"] == 2, st
    _close_within(srv, 60)
    err = capfd.readouterr().err
    assert err.count("pose_gpu_server: ") == 1, err


def test_server_sink_write_error(dev, tmp_path):
    """A sink that refuses writes (a read-only descriptor): the request fails,
    wait and drain return -1, nothing hangs, and destroy returns."""
    path = tmp_path / "sink.bin"
    path.write_bytes(b"\0" * (1 << 20))
    fd = os.open(str(path), os.O_RDONLY)
    try:
        srv = pg.Server(hbm_budget_bytes=1 << 30, pinned_bytes=64 << 20, sink_fd=fd)
        rid = srv.submit(b"\x55" * 32, 1 << 10, 0, 2)
        assert srv.wait(rid) == -1
        assert srv.drain() == -1
        _close_within(srv, 60)
    finally:
        os.close(fd)


def test_server_failure_no_writes_after_wait(dev, oracle, monkeypatch):
    """After a failure, wait and drain return only when no launch of the
    request can still write its host buffer: launches of four streams are
    running when the third retirement reports a failure (host-side injection),
    and every buffer stays as it was once drain has returned.  Requests
    reported complete hold the CPU labeler's bytes."""
    import time
    monkeypatch.setenv("POSE_GPU_SERVER_FAIL_RETIRE", "3")
    srv = pg.Server(hbm_budget_bytes=1 << 30, pinned_bytes=64 << 20, max_inflight=4,
                    max_super_per_launch=1)
    monkeypatch.delenv("POSE_GPU_SERVER_FAIL_RETIRE")
    cb, n = 1 << 10, 4
    seeds = [bytes([0x60 + i]) * 32 for i in range(6)]
    bufs = [bytearray(b"\xa5" * (n * cb * 32)) for _ in seeds]
    ids = [srv.submit(s, cb, i * n * cb, n, host_out=b) for i, (s, b) in enumerate(zip(seeds, bufs))]
    say("failure injected at the third retirement, four streams")
    assert srv.drain() == -1
    rcs = [srv.wait(i) for i in ids]
    assert -1 in rcs, rcs
    snap = [bytes(b) for b in bufs]
    time.sleep(0.5)
    assert [bytes(b) for b in bufs] == snap, "a host buffer changed after drain returned"
    for i, (rc, s, b) in enumerate(zip(rcs, seeds, bufs)):
        if rc == 0:
            assert bytes(b) == oracle.label_superchunks(s, cb, list(range(n)), i * n * cb, CPU_THREADS), i
    _close_within(srv, 60)


def test_server_sink_error_stops_admission(dev, tmp_path):
    """A request whose sink write fails is not admitted further: of its 16
    one-super-chunk launches only those started before the failure run."""
    path = tmp_path / "sink.bin"
    path.write_bytes(b"\0" * (1 << 20))
    fd = os.open(str(path), os.O_RDONLY)
    try:
        srv = pg.Server(hbm_budget_bytes=1 << 30, pinned_bytes=64 << 20, sink_fd=fd, max_inflight=1,
                        max_super_per_launch=1)
        rid = srv.submit(b"\x56" * 32, 1 << 10, 0, 16)
        assert srv.wait(rid) == -1
        assert srv.drain() == -1
        st = srv.stats()
        say(f"launches after the failed write: {st['launches']} of 16")
        assert st["launches"] < 16 and st["super_chunks_done"] == 0, st
        _close_within(srv, 60)
    finally:
        os.close(fd)


def test_plan_scratch_halving(dev, oracle, monkeypatch):
    """The plan's slot count halves until its scratch fits the budget
    (POSE_GPU_SCRATCH_LIMIT caps it here; the reference halves on allocation
    failure, common/src/cuda/label_gpu_warp.cu:163-175), and one slot is
    still tried when not even that fits; the labels stay byte-identical."""
    cb, n = 1 << 10, 8
    T = pg.Topology(cb, 64, 1)
    region = (T.info.pool_slots * 32 + 255) // 256 * 256 + cb * 32
    seed = b"\x66" * 32
    exp = oracle.label_superchunks(seed, cb, list(range(n)), 13 * cb, CPU_THREADS)
    for limit, want in ((3 * region, 2), (region // 2, 1), (0, 8)):
        if limit:
            monkeypatch.setenv("POSE_GPU_SCRATCH_LIMIT", str(limit))
        else:
            monkeypatch.delenv("POSE_GPU_SCRATCH_LIMIT", raising=False)
        buf = torch.zeros(n * cb * 32, dtype=torch.uint8, device=dev)
        plan = pg.Plan([(T, seed, 13 * cb, n, buf.data_ptr())], slots=n, graph=True, lanes_per_node=1,
                       split=pg.SPLIT_OFF)
        say(f"scratch limit {limit}: {plan.info.slots} slots")
        assert plan.info.slots == want and plan.info.scratch_bytes == want * region
        plan.run()
        torch.cuda.synchronize()
        assert buf.cpu().numpy().tobytes() == exp, limit
        plan.close()
    T.close()
