"""GPU parity: the HIP labeler (through its C ABI) against the oracle and the
golden fixtures written by the reference build.  Integer/byte work: the bar is
bit-exact equality of every label byte."""
import hashlib
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import pose_gpu as pg  # noqa: E402  (raises if the HIP library is not built)
from pyoracle import RefLib, ref_available, splitmix_key  # noqa: E402

CPU_THREADS = min(16, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    return torch.device("cuda:0")


def dbuf(dev, nbytes, fill=None):
    t = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    if fill is not None:
        t.fill_(fill)
    return t


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().tobytes()


def cpu_labels(oracle, seed, cb, js, rbo):
    """Expected super-chunks js (indices within the region)."""
    if ref_available(16) and list(js) == list(range(len(js))):
        ref = RefLib(16)
        try:
            return ref.label_region(seed, cb, len(js), rbo)
        finally:
            ref.close()
    return oracle.label_superchunks(seed, cb, list(js), rbo, CPU_THREADS)


def test_node_single_matches_fixture(dev, golden):
    ks, ki = golden["keys"]["src"], golden["keys"]["int"]
    for c in golden["label_node"]:
        p0 = bytes.fromhex(c["p0"]) if c["p0"] else None
        p1 = bytes.fromhex(c["p1"]) if c["p1"] else None
        got = pg.label_node_single(c["is_source"], int(c["id"]), p0, p1, bytes.fromhex(c["seed"]),
                                   bytes.fromhex(c["desc"]), ks, ki)
        assert got.hex() == c["out"]


def test_cb1024_four_superchunks_prng_seed0(dev, golden, oracle):
    cb, n = 1 << 10, 4
    seed = splitmix_key(0)
    say("label_hbm cb=1024 n=4 seed=SM64(0)")
    buf = dbuf(dev, n * cb * 32, 0)
    assert pg.label_hbm(seed, cb, buf.data_ptr(), n * cb * 32, 0) == 0
    got = host(buf)
    fx = [r for r in golden["regions"] if r["cb"] == cb and r["n_super"] == 4 and r["seed"] == seed.hex()][0]
    for i in range(n):
        assert hashlib.sha256(got[i * cb * 32:(i + 1) * cb * 32]).hexdigest() == fx["super_sha256"][i]
    assert got == oracle.label_region(seed, cb, n, 0)


def test_golden_regions_small(dev, golden):
    """Every fixture super-chunk up to 2^15 (rank order, chunk seeds from offsets)."""
    topo = {}
    for r in golden["regions"]:
        cb = r["cb"]
        if cb > (1 << 15):
            continue
        say(f"golden region cb={cb} n={r['n_super']} off={r['block_offset']}")
        if cb not in topo:
            topo[cb] = pg.Topology(cb)
        n = r["n_super"]
        buf = dbuf(dev, n * cb * 32, 0)
        assert topo[cb].label(bytes.fromhex(r["seed"]), buf.data_ptr(), n * cb * 32, r["block_offset"]) == 0
        got = host(buf)
        for i in range(n):
            assert hashlib.sha256(got[i * cb * 32:(i + 1) * cb * 32]).hexdigest() == r["super_sha256"][i], (cb, i)
    for t in topo.values():
        t.close()


def test_topology_reuse_seeds_offsets(dev, oracle):
    rng = np.random.default_rng(11)
    for k in [0, 1, 3, 6, 9, 12]:
        cb = 1 << k
        T = pg.Topology(cb)
        for trial in range(2):
            seed = rng.integers(0, 256, 32 if trial == 0 else 19, dtype=np.uint8).tobytes()
            rbo = int(rng.integers(0, 1 << 35)) * cb
            n = 3
            say(f"reuse cb={cb} trial={trial} rbo={rbo}")
            buf = dbuf(dev, n * cb * 32, 0)
            assert T.label(seed, buf.data_ptr(), n * cb * 32, rbo) == 0
            assert host(buf) == oracle.label_superchunks(seed, cb, [0, 1, 2], rbo, CPU_THREADS)
        T.close()


def test_unaligned_and_wrapping_offsets(dev, oracle):
    """An offset inside a chunk and one where offset + j*cb wraps past 2^64:
    super-chunk j takes chunk index (offset + j*cb) / cb in 64-bit arithmetic
    (common/src/graph.c:1646-1649), through the topology call, the batch call
    and the label server alike."""
    rng = np.random.default_rng(23)
    cases = [(1 << 4, 7), (1 << 10, int(rng.integers(0, 1 << 40)) * (1 << 10) + 513),
             (1 << 8, (1 << 64) - (1 << 8) - 5), (1 << 12, (1 << 64) - 3 * (1 << 12) + 1)]
    srv = pg.Server(hbm_budget_bytes=1 << 30, pinned_bytes=64 << 20)
    try:
        for cb, rbo in cases:
            n = 3
            seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
            exp = oracle.label_region(seed, cb, n, rbo, CPU_THREADS)
            T = pg.Topology(cb)
            say(f"unaligned offset cb={cb} rbo={rbo}: topology call")
            buf = dbuf(dev, n * cb * 32, 0)
            assert T.label(seed, buf.data_ptr(), n * cb * 32, rbo) == 0
            assert host(buf) == exp, (cb, rbo)
            for lpn in (1, 4):
                say(f"unaligned offset cb={cb} rbo={rbo}: batch call, {lpn} lanes per node")
                buf2 = dbuf(dev, n * cb * 32, 0)
                pg.label_batch([(T, seed, rbo, n, buf2.data_ptr())], lanes_per_node=lpn)
                assert host(buf2) == exp, (cb, rbo, lpn)
            say(f"unaligned offset cb={cb} rbo={rbo}: label server")
            out = bytearray(n * cb * 32)
            assert srv.wait(srv.submit(seed, cb, rbo, n, host_out=out)) == 0
            assert bytes(out) == exp, (cb, rbo)
            T.close()
    finally:
        srv.close()


def test_tail_and_short_buffers_untouched(dev, oracle):
    cb = 256
    sb = cb * 32
    seed = b"\x42" * 32
    say("tail cb=256")
    buf = dbuf(dev, 2 * sb + sb // 2, 0xAB)
    assert pg.label_hbm(seed, cb, buf.data_ptr(), 2 * sb + sb // 2, 0) == 0
    got = host(buf)
    assert got[:2 * sb] == oracle.label_region(seed, cb, 2, 0)
    assert got[2 * sb:] == b"\xab" * (sb // 2)
    small = dbuf(dev, sb - 32, 0xCD)
    assert pg.label_hbm(seed, cb, small.data_ptr(), sb - 32, 0) == 0
    assert host(small) == b"\xcd" * (sb - 32)


def test_error_paths(dev):
    buf = dbuf(dev, 1 << 16, 0)
    assert pg.label_hbm(b"\0" * 32, 1024, buf.data_ptr(), 1 << 16, 0, algo=1) == -1
    assert pg.label_hbm(b"\0" * 32, 1 << 21, buf.data_ptr(), 1 << 26, 0) == -1
    assert pg.lib.pose_graph_hbm_topo_build(1 << 21, 0) is None


@pytest.mark.parametrize("lanes,mode", [(64, 0), (128, 1), (256, 1), (256, 0), (512, 1)])
def test_lane_widths_and_modes(dev, oracle, lanes, mode):
    cb, n = 1 << 12, 3
    seed = hashlib.sha256(f"{lanes}/{mode}".encode()).digest()
    say(f"lanes={lanes} mode={mode} cb={cb}")
    T = pg.Topology(cb, lanes, mode)
    buf = dbuf(dev, n * cb * 32, 0)
    assert T.label(seed, buf.data_ptr(), n * cb * 32, 7 * cb) == 0
    assert host(buf) == oracle.label_superchunks(seed, cb, [0, 1, 2], 7 * cb, CPU_THREADS)
    T.close()


@pytest.mark.parametrize("lanes", [64, 128, 256, 512])
def test_one_lane_form_at_low_and_high_occupancy(dev, oracle, lanes):
    """A one-lane launch with few resident waves per SIMD runs the
    software-pipelined kernel, a full one the plain kernel: both byte-identical
    to the oracle over more super-chunks than eight waves per SIMD hold."""
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    cb, waves = 1 << 6, lanes // 64
    n = cus * 32 // waves + 7
    seed = hashlib.sha256(f"occupancy/{lanes}".encode()).digest()
    T = pg.Topology(cb, lanes, 1)
    exp = oracle.label_superchunks(seed, cb, list(range(n)), 3 * cb, CPU_THREADS)
    for wps in (8, 1):
        say(f"lanes={lanes} cb={cb} n={n} waves_per_simd={wps}")
        buf = dbuf(dev, n * cb * 32, 0)
        pg.label_batch([(T, seed, 3 * cb, n, buf.data_ptr())], 0, wps, lanes_per_node=1)
        assert host(buf) == exp, wps
    T.close()


def test_non_power_of_two_chunk_blocks(dev, oracle):
    for cb in [3, 100, 1000]:
        say(f"cb={cb}")
        seed = b"\x07" * 32
        buf = dbuf(dev, 2 * cb * 32, 0)
        assert pg.label_hbm(seed, cb, buf.data_ptr(), 2 * cb * 32, cb) == 0
        assert host(buf) == oracle.label_superchunks(seed, cb, [0, 1], cb, CPU_THREADS)


def test_cb16384_256_superchunks_one_per_launch_full_compare(dev, oracle):
    cb, n = 1 << 14, 256
    seed = splitmix_key(1)
    say("cb=16384 n=256 full compare")
    T = pg.Topology(cb)
    buf = dbuf(dev, n * cb * 32, 0)
    assert T.label(seed, buf.data_ptr(), n * cb * 32, 0) == 0
    got = host(buf)
    T.close()
    exp = cpu_labels(oracle, seed, cb, range(n), 0)
    assert got == exp


def test_batch_uneven_sessions_one_launch(dev, oracle):
    rng = np.random.default_rng(3)
    topo = {cb: pg.Topology(cb) for cb in (1 << 10, 1 << 12, 1 << 14)}
    sessions, bufs, meta = [], [], []
    for s in range(9):
        cb = [1 << 10, 1 << 12, 1 << 14][s % 3]
        n = int(rng.integers(1, 5))
        seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        rbo = int(rng.integers(0, 1 << 30)) * cb
        b = dbuf(dev, n * cb * 32, 0)
        bufs.append(b)
        sessions.append((topo[cb], seed, rbo, n, b.data_ptr()))
        meta.append((seed, cb, n, rbo))
    say("batch of 9 uneven sessions")
    pg.label_batch(sessions)
    for b, (seed, cb, n, rbo) in zip(bufs, meta):
        assert host(b) == oracle.label_superchunks(seed, cb, list(range(n)), rbo, CPU_THREADS), (cb, n)
    for t in topo.values():
        t.close()


def test_cb131072_against_fixture(dev, golden):
    cb = 1 << 17
    say("cb=2^17 one super-chunk vs fixture")
    r = [r for r in golden["regions"] if r["cb"] == cb][0]
    buf = dbuf(dev, cb * 32, 0)
    assert pg.label_hbm(bytes.fromhex(r["seed"]), cb, buf.data_ptr(), cb * 32, r["block_offset"]) == 0
    got = host(buf)
    assert got[:32].hex() == r["rank0"][0]
    assert hashlib.sha256(got).hexdigest() == r["super_sha256"][0]


@pytest.mark.parametrize("lanes,cbs", [(64, (1, 16, 1000, 1 << 10, 1 << 12)), (128, (1 << 11, 1 << 14)),
                                       (256, (1 << 10, 1 << 13))])
def test_quad_latency_form(dev, oracle, lanes, cbs):
    """Four lanes per node (DPP column exchange, LDS-staged message words)."""
    for cb in cbs:
        say(f"quad lanes={lanes} cb={cb}")
        T = pg.Topology(cb, lanes, 1)
        seed = hashlib.sha256(f"quad/{lanes}/{cb}".encode()).digest()
        n = 3
        buf = dbuf(dev, n * cb * 32, 0)
        pg.label_batch([(T, seed, 5 * cb, n, buf.data_ptr())], 2, 0, lanes_per_node=4)
        assert host(buf) == oracle.label_superchunks(seed, cb, [0, 1, 2], 5 * cb, CPU_THREADS), cb
        T.close()


def test_quad_mixed_batch_graph(dev, oracle):
    topo = {cb: pg.Topology(cb, 64, 1) for cb in (1 << 9, 1 << 12)}
    meta = [(b"\x11" * 32, 1 << 9, 4, 0), (b"\x22" * 32, 1 << 12, 2, 3 << 12), (b"\x33" * 19, 1 << 9, 1, 9 << 9)]
    outs = [dbuf(dev, n * cb * 32, 0) for _, cb, n, _ in meta]
    plan = pg.Plan([(topo[cb], k, rbo, n, o.data_ptr()) for (k, cb, n, rbo), o in zip(meta, outs)],
                   graph=True, lanes_per_node=4)
    assert plan.info.lanes_per_node == 4
    for rep in range(2):
        plan.run()
        for (k, cb, n, rbo), o in zip(meta, outs):
            assert host(o) == oracle.label_superchunks(k, cb, list(range(n)), rbo, CPU_THREADS), (rep, cb)
    plan.close()


@pytest.mark.parametrize("lpn", [1, 4])
def test_split_copies_forced(dev, oracle, lpn):
    """Split on: every super-chunk labeled as its scaffold's two copies, each
    by its own workgroup job (one- and four-lane forms), byte-identical to the
    oracle; chunk_blocks include ones whose second copy has fewer outputs than
    the first (not a power of two) or none (1).  Guard bytes after the output
    stay untouched."""
    for cb in (1, 3, 100, 1000, 1 << 10, 1 << 12):
        say(f"split on lpn={lpn} cb={cb}")
        T = pg.Topology(cb, 64, 1)
        seed = hashlib.sha256(f"split/{lpn}/{cb}".encode()).digest()
        n = 5
        buf = dbuf(dev, n * cb * 32 + 64, 0xA5)
        pg.label_batch([(T, seed, 7 * cb, n, buf.data_ptr())], 3, 0, lanes_per_node=lpn, split=pg.SPLIT_ON)
        got = host(buf)
        assert got[:n * cb * 32] == oracle.label_superchunks(seed, cb, list(range(n)), 7 * cb, CPU_THREADS), cb
        assert got[n * cb * 32:] == b"\xa5" * 64
        T.close()


def test_split_auto_choice_and_replays(dev, oracle):
    """Automatic split: a batch of at most one super-chunk per two CUs in the
    four-lane form runs one job per copy, every job resident at once; a larger
    batch, or split off, runs whole super-chunks.  Each plan is replayed from
    its hipGraph twice, byte-identical to the oracle each time."""
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    cb = 1 << 9
    T = pg.Topology(cb, 64, 1)
    cases = [(3, pg.SPLIT_AUTO, 1), (cus // 2, pg.SPLIT_AUTO, 1), (cus // 2 + 1, pg.SPLIT_AUTO, 0),
             (3, pg.SPLIT_OFF, 0)]
    for n, split, want in cases:
        say(f"split auto n={n} split={split}")
        seed = hashlib.sha256(f"auto/{n}/{split}".encode()).digest()
        buf = dbuf(dev, n * cb * 32, 0)
        plan = pg.Plan([(T, seed, 11 * cb, n, buf.data_ptr())], graph=True, split=split)
        (g,) = plan.groups()
        assert g["lanes_per_node"] == 4 and g["super_chunks"] == n
        assert g["split"] == want and g["jobs"] == (2 * n if want else n)
        if want:
            assert g["slots"] == 2 * n
        exp = oracle.label_superchunks(seed, cb, list(range(n)), 11 * cb, CPU_THREADS)
        for rep in range(2):
            buf.zero_()
            plan.run()
            assert host(buf) == exp, (n, split, rep)
        plan.close()
    T.close()


def test_copy_schedules_on_device_match_host(dev, oracle):
    """The copy schedules stay on the host until the first launch that labels
    a copy on its own uploads them.  At a power-of-two chunk_blocks copy 1 is
    copy 0 shifted, so the device holds copy 0's records and near entries once
    (20 bytes per position) and a block-0 CV array per copy (32 bytes per
    position): the topology's device bytes grow by exactly that.  Read back
    from the device (byte offsets converted back to slots and ranks, copy 1's
    ids shifted) both copies equal the host build, and both launches give the
    oracle's labels."""
    cb = 1 << 10
    D = pg.Topology(cb, 64, 1)
    H = pg.Topology(cb, 64, 1, host_only=True)
    whole_bytes = D.device_bytes()
    assert whole_bytes == (D.info.steps + 1) * 64 * 52
    seed = b"\x77" * 32
    exp = oracle.label_superchunks(seed, cb, [0, 1, 2], 9 * cb, CPU_THREADS)
    for split in (pg.SPLIT_OFF, pg.SPLIT_ON):
        buf = dbuf(dev, 3 * cb * 32, 0)
        pg.label_batch([(D, seed, 9 * cb, 3, buf.data_ptr())], 0, 0, lanes_per_node=4, split=split)
        assert host(buf) == exp, split
        if split == pg.SPLIT_OFF:
            assert D.device_bytes() == whole_bytes     # whole jobs do not need the copies
    rows = [D.part(c)[0]["steps"] + 1 for c in (0, 1)]
    assert rows[0] == rows[1]
    assert D.device_bytes() == whole_bytes + rows[0] * 64 * 20 + (rows[0] + rows[1]) * 64 * 32
    for c in (0, 1):
        di, ds, dr, dn = D.part(c)
        hi, hs, hr, hn = H.part(c)
        assert di == hi
        assert np.array_equal(ds, hs) and np.array_equal(dr, hr) and np.array_equal(dn, hn)
    D.close()
    H.close()
    # Wider topologies carry no copy schedules, and split on is refused there.
    W = pg.Topology(cb, 128, 1)
    assert W.part(0) is None
    buf = dbuf(dev, cb * 32, 0)
    with pytest.raises(RuntimeError):
        pg.label_batch([(W, b"k" * 32, 0, 1, buf.data_ptr())], 0, 0, lanes_per_node=1, split=pg.SPLIT_ON)
    W.close()


def test_split_at_2_17_both_forms(dev, golden):
    """One super-chunk at 2^17 labeled as two copies in the four-lane form
    (the automatic choice) and in the one-lane form, against the reference
    build's fixture."""
    cb = 1 << 17
    r = [r for r in golden["regions"] if r["cb"] == cb][0]
    T = pg.Topology(cb)
    for lpn in (4, 1):
        say(f"cb=2^17 split lpn={lpn}")
        buf = dbuf(dev, cb * 32, 0)
        pg.label_batch([(T, bytes.fromhex(r["seed"]), r["block_offset"], 1, buf.data_ptr())], 0, 0,
                       lanes_per_node=lpn, split=pg.SPLIT_ON)
        got = host(buf)
        assert got[:32].hex() == r["rank0"][0]
        assert hashlib.sha256(got).hexdigest() == r["super_sha256"][0]
    T.close()


def test_max_chunk_blocks_2_20(dev, oracle):
    """The largest chunk_blocks the reference accepts (POSE_GRAPH_MAX_M = 2^20,
    common/include/graph.h:36): one super-chunk, about 403 M scaffold nodes and
    3.1 M dependency levels, labeled as two copies (the automatic choice) and
    as a whole scaffold, both byte-identical to the CPU labeler."""
    cb = 1 << 20
    seed = bytes(range(100, 132))
    rbo = 5 * cb
    say("cb=2^20: topology build")
    T = pg.Topology(cb)
    assert T.info.levels == 3145706
    buf = dbuf(dev, cb * 32, 0)
    got = {}
    for split in (pg.SPLIT_AUTO, pg.SPLIT_OFF):
        say(f"cb=2^20 split={split}")
        buf.zero_()
        pg.label_batch([(T, seed, rbo, 1, buf.data_ptr())], 0, 0, split=split)
        got[split] = host(buf)
    T.close()
    assert got[pg.SPLIT_AUTO] == got[pg.SPLIT_OFF]
    say("cb=2^20: CPU labeler")
    assert got[pg.SPLIT_AUTO] == cpu_labels(oracle, seed, cb, [0], rbo)


@pytest.mark.parametrize("k", [16, 18, 19])
def test_deep_chunk_blocks(dev, oracle, k):
    """The depths between the fixture sizes and the maximum (2^16, 2^18,
    2^19: ring pages, recycled slots and copy schedules at scale): one
    super-chunk labeled as two copies and as a whole scaffold, both equal to
    the CPU labeler."""
    cb = 1 << k
    seed = bytes(range(7, 39))
    rbo = 3 * cb + 11
    say(f"cb=2^{k}: topology build")
    T = pg.Topology(cb)
    buf = dbuf(dev, cb * 32, 0)
    got = {}
    for split in (pg.SPLIT_AUTO, pg.SPLIT_OFF):
        say(f"cb=2^{k} split={split}")
        buf.zero_()
        pg.label_batch([(T, seed, rbo, 1, buf.data_ptr())], 0, 0, split=split)
        got[split] = host(buf)
    T.close()
    assert got[pg.SPLIT_AUTO] == got[pg.SPLIT_OFF]
    say(f"cb=2^{k}: CPU labeler")
    assert got[pg.SPLIT_AUTO] == cpu_labels(oracle, seed, cb, [0], rbo)


def test_random_shapes_all_forms(dev, oracle):
    """Random chunk_blocks (mostly not powers of two), super-chunk counts,
    start positions and launch forms (one or four lanes per node, whole or
    split into copies, a few or many slots), each against the oracle."""
    rng = np.random.default_rng(2024)
    topo = {}
    for case in range(10):
        cb = int(rng.integers(1, 3000))
        n = int(rng.integers(1, 5))
        rbo = int(rng.integers(0, 1 << 20)) * cb
        lpn = [1, 4][int(rng.integers(0, 2))]
        split = [pg.SPLIT_AUTO, pg.SPLIT_OFF, pg.SPLIT_ON][int(rng.integers(0, 3))]
        slots = [0, 1, 3][int(rng.integers(0, 3))]
        seed = rng.integers(0, 256, int(rng.integers(1, 64)), dtype=np.uint8).tobytes()
        say(f"case {case}: cb={cb} n={n} lanes/node={lpn} split={split} slots={slots}")
        if cb not in topo:
            topo[cb] = pg.Topology(cb, 64, 1)
        buf = dbuf(dev, n * cb * 32, 0x5A)
        pg.label_batch([(topo[cb], seed, rbo, n, buf.data_ptr())], slots, 0, lanes_per_node=lpn, split=split)
        assert host(buf) == oracle.label_superchunks(seed, cb, list(range(n)), rbo, CPU_THREADS), case
    for t in topo.values():
        t.close()


@pytest.mark.parametrize("seed_len", [0, 1, 31, 33, 3000])
def test_session_seed_lengths(dev, oracle, seed_len):
    """The session seed passes to the chunk-seed derivation with its own
    length, any length (the reference's pose_chunk_seed, common/src/hash.c:218-233):
    empty, short, just under and over 32 bytes, and long."""
    cb, n = 1 << 9, 3
    seed = bytes((7 * i + seed_len) & 0xFF for i in range(seed_len))
    say(f"seed length {seed_len}")
    T = pg.Topology(cb)
    buf = dbuf(dev, n * cb * 32, 0)
    assert T.label(seed, buf.data_ptr(), n * cb * 32, 11 * cb) == 0
    assert host(buf) == oracle.label_superchunks(seed, cb, [0, 1, 2], 11 * cb, CPU_THREADS)
    T.close()
