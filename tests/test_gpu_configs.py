"""GPU tests at the benchmark configurations' own shapes.

- Streamed batch: 64 super-chunks at chunk_blocks = 2^17 from a random
  aligned start (SplitMix64 stream of seed 2), through the label server into
  the caller's host buffer; 16 random super-chunks against the CPU labeler.
- Uneven sessions: 48 sessions from the stream of seed 3 (chunk_blocks in
  {2^12, 2^14, 2^15, 2^17}, 1..32 super-chunks each), launch shape autotuned
  per chunk_blocks, all launch groups in one hipGraph replayed twice; one
  sampled super-chunk of every session against the CPU labeler each time.
- Serving: a Poisson request stream (seed 4; chunk_blocks 2^12..2^17,
  1..128 super-chunks) into a server whose HBM budget is small enough that
  requests wait for memory; sampled requests against the same request run
  alone and against the CPU labeler.
- Headline shape: 8,192 super-chunks at 2^17 in one captured plan on 4,096
  one-lane runners; sampled super-chunks against the CPU labeler, replay
  checksums, and random super-chunks against the four-lane form alone.

Integer/byte work: every compared label is bit-exact.
"""
import time

import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import pose_gpu as pg  # noqa: E402
import workloads as wl  # noqa: E402


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    return torch.device("cuda:0")


def test_streamed_64_superchunks_at_2_17(dev, cpu_superchunks):
    key, cb, n, rbo, sm = wl.streamed_batch()
    sb = cb * 32
    host = bytearray(n * sb)
    srv = pg.Server(hbm_budget_bytes=16 << 30, pinned_bytes=1 << 30, preload=[cb])
    say(f"streamed: {n} super-chunks at 2^17 from block {rbo}")
    rid = srv.submit(key, cb, rbo, n, host_out=host)
    assert srv.wait(rid) == 0
    st = srv.stats()
    srv.close()
    assert st["super_chunks_done"] == n
    picks = set()
    while len(picks) < 16:
        picks.add(sm.below(n))
    picks = sorted(picks)
    exp = cpu_superchunks([(key, cb, rbo + j * cb) for j in picks])
    for j, e in zip(picks, exp):
        assert bytes(host[j * sb:(j + 1) * sb]) == e, j


def test_uneven_sessions_autotuned_graph(dev, cpu_superchunks):
    sessions = wl.uneven_sessions()
    cbs = sorted({cb for _, cb, _, _ in sessions})
    tuned, topos = {}, {}
    for cb in cbs:
        n_cb = sum(n for _, c, n, _ in sessions if c == cb)
        t0 = time.perf_counter()
        best, _ = pg.autotune(cb, n_cb)
        say(f"autotune cb={cb} n={n_cb} ({time.perf_counter() - t0:.1f} s): {best}")
        tuned[cb] = best
        topos[cb] = pg.Topology(cb, best["lanes"], best["mode"])
    outs = [torch.empty(n * cb * 32, dtype=torch.uint8, device=dev) for _, cb, n, _ in sessions]
    launches = [(tuned[cb]["slots"], tuned[cb]["waves_per_simd"], tuned[cb]["lanes_per_node"],
                 tuned[cb].get("split", pg.SPLIT_AUTO)) for _, cb, _, _ in sessions]
    plan = pg.Plan([(topos[cb], key, rbo, n, o.data_ptr()) for (key, cb, n, rbo), o in zip(sessions, outs)],
                   graph=True, launches=launches)
    say(f"plan: {plan.info.groups} groups {plan.groups()}")
    assert plan.info.uses_graph == 1 and plan.info.super_chunks == sum(n for _, _, n, _ in sessions)
    # One sampled super-chunk per session (a different one on each replay).
    sm = wl.SplitMix(33)
    for rep in range(2):
        picks = [sm.below(n) for _, _, n, _ in sessions]
        exp = cpu_superchunks([(key, cb, rbo + j * cb) for (key, cb, n, rbo), j in zip(sessions, picks)])
        for o in outs:
            o.fill_(0xC3 + rep)
        torch.cuda.synchronize()
        ms = plan.run()
        torch.cuda.synchronize()
        say(f"replay {rep}: {ms:.1f} ms")
        for s, ((key, cb, n, rbo), o, j, e) in enumerate(zip(sessions, outs, picks, exp)):
            sb = cb * 32
            assert o[j * sb:(j + 1) * sb].cpu().numpy().tobytes() == e, (rep, s, cb, n, j)
    plan.close()
    for t in topos.values():
        t.close()


def test_poisson_serving_under_binding_budget(dev, cpu_superchunks):
    budget = 2 << 30
    reqs, _ = wl.poisson_requests(16.0, 2.0)
    total = sum(n for _, _, _, n, _ in reqs)
    say(f"serving: {len(reqs)} requests, {total} super-chunks, budget {budget >> 20} MiB")
    srv = pg.Server(hbm_budget_bytes=budget, pinned_bytes=512 << 20,
                    preload=[1 << k for k in range(12, 18)])
    kept = []
    t0 = time.perf_counter()
    for t, key, cb, n, rbo in reqs:
        now = time.perf_counter() - t0
        if now < t:
            time.sleep(t - now)
        buf = None
        if n <= 6 and len(kept) < 6:
            buf = bytearray(n * cb * 32)
            kept.append((key, cb, n, rbo, buf))
        srv.submit(key, cb, rbo, n, host_out=buf)
    assert srv.drain() == 0
    st = srv.stats()
    srv.close()
    say(f"served: {st['super_chunks_done']} super-chunks, p50 {st['p50_ms']:.0f} ms, p99 {st['p99_ms']:.0f} ms, "
        f"{st['mib_per_s']:.0f} MiB/s, peak inflight {st['peak_inflight']}, peak HBM {st['peak_hbm_bytes'] >> 20} MiB, "
        f"memory waits {st['memory_waits']}, launches cut by memory {st['memory_cut_launches']}")
    assert st["super_chunks_done"] == total and st["requests_done"] == len(reqs)
    assert st["peak_hbm_bytes"] <= budget
    assert st["peak_inflight"] >= 2, "requests did not overlap on the device"
    assert st["memory_waits"] + st["memory_cut_launches"] >= 1, "the budget never held a request back"
    assert kept, "no request small enough to sample"
    # Each sampled request alone, on a fresh single-stream server.
    alone = pg.Server(hbm_budget_bytes=budget, pinned_bytes=512 << 20, max_inflight=1)
    for key, cb, n, rbo, buf in kept:
        solo = bytearray(n * cb * 32)
        assert alone.wait(alone.submit(key, cb, rbo, n, host_out=solo)) == 0
        assert bytes(solo) == bytes(buf), (cb, n)
    alone.close()
    items = [(key, cb, rbo + j * cb) for key, cb, n, rbo, _ in kept for j in sorted({0, n - 1})]
    exp = iter(cpu_superchunks(items))
    for key, cb, n, rbo, buf in kept:
        for j in sorted({0, n - 1}):
            assert bytes(buf[j * cb * 32:(j + 1) * cb * 32]) == next(exp), (cb, n, j)


def test_headline_shape_8192_superchunks_at_2_17(dev, cpu_superchunks):
    """The batched benchmark's own shape: 8,192 super-chunks at 2^17 from a
    random aligned start (the stream of seed 2), one captured plan on 4,096
    one-lane job runners, 32 GiB of labels.  At this size the checks are
    size-independent ones: the first, a middle and the last super-chunk
    against the CPU labeler; two replays give the same per-super-chunk
    checksums (the output is cleared between them); and four random
    super-chunks equal the same super-chunks labeled alone in the four-lane
    form, split into copies."""
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    key, cb, n, rbo, sm = wl.streamed_batch(n_super=32 * cus)
    sb = cb * 32
    T = pg.Topology(cb)
    out = torch.empty(n * sb, dtype=torch.uint8, device=dev)
    plan = pg.Plan([(T, key, rbo, n, out.data_ptr())], 0, 4, graph=True)
    info = plan.info
    say(f"headline shape: {n} super-chunks, {info.slots} runners, {info.lanes_per_node} lane(s) per node")
    assert info.lanes_per_node == 1 and info.slots == 16 * cus

    def checksums():
        # Two per-super-chunk sums of the labels as int64 words (the second
        # weighted), 512 super-chunks at a time to bound the temporaries.
        w = out.view(torch.int64).view(n, sb // 8)
        parts = []
        for r in range(0, n, 512):
            blk = w[r:r + 512]
            parts.append(torch.stack([blk.sum(dim=1), (blk * 0x9E3779B1).sum(dim=1)]))
        return torch.cat(parts, dim=1).cpu()

    ms = plan.run()
    torch.cuda.synchronize()
    say(f"replay 1: {ms:.0f} ms")
    first = checksums()
    js = [0, n // 2, n - 1]
    exp = cpu_superchunks([(key, cb, rbo + j * cb) for j in js])
    for j, e in zip(js, exp):
        assert out[j * sb:(j + 1) * sb].cpu().numpy().tobytes() == e, j
    out.zero_()
    ms = plan.run()
    torch.cuda.synchronize()
    say(f"replay 2: {ms:.0f} ms")
    assert torch.equal(checksums(), first)
    plan.close()
    alone = torch.empty(sb, dtype=torch.uint8, device=dev)
    for _ in range(4):
        j = sm.below(n)
        alone.fill_(0x5A)
        pg.label_batch([(T, key, rbo + j * cb, 1, alone.data_ptr())], 0, 0, lanes_per_node=4, split=pg.SPLIT_ON)
        torch.cuda.synchronize()
        assert torch.equal(alone, out[j * sb:(j + 1) * sb]), j
    T.close()
    del out
