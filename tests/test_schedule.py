"""CPU checks of the product's host-side schedule (no GPU): the exported step
records are replayed step by step with the oracle's node hash, with the GPU's
semantics (every read of a step happens before any write of that step), and
must give the oracle's super-chunk labels.  Also checks the schedule's
structural invariants directly."""
import ctypes as C
import re

import numpy as np
import pytest

import pose_gpu as pg

PERSIST = 0x80000000


def replay_rows(oracle, rec, n_steps, cb, cseed, persist):
    """Replays n_steps rows of step records into `persist` (rank -> label),
    with a scratch pool of its own; returns the number of nodes labeled."""
    desc = oracle.descriptor(cb)
    scratch = {0: bytes(32)}

    def read(loc):
        if loc & PERSIST:
            return persist[loc & ~PERSIST]
        return scratch[loc]

    n_done = 0
    for s in range(n_steps):
        out = []
        for t in range(rec.shape[1]):
            x, y, z, w = (int(v) for v in rec[s, t])
            np_ = w & 3
            if np_ == 3:
                continue
            nid = w >> 2
            p0 = read(y) if np_ >= 1 else None
            p1 = read(z) if np_ == 2 else None
            if np_ < 2:
                assert z == 0, "absent pred1 must point at the zero slot"
            if np_ == 0:
                assert y == 0
            out.append((x, oracle.label_node(np_ == 0, nid, p0, p1, cseed, desc)))
        for x, lab in out:
            if x & PERSIST:
                persist[x & ~PERSIST] = lab
            else:
                assert x >= 2, "slots 0 and 1 are reserved"
                scratch[x] = lab
        n_done += len(out)
    return n_done


def replay(oracle, T, cseed):
    steps, rec = T.export()
    info = T.info
    cb = info.chunk_blocks
    persist = {}
    assert replay_rows(oracle, rec, info.steps, cb, cseed, persist) == info.nodes
    assert sorted(persist) == list(range(cb))
    return b"".join(persist[r] for r in range(cb))


@pytest.mark.parametrize("cb,lanes,mode", [(1, 64, 1), (2, 64, 0), (5, 64, 1), (16, 64, 1),
                                           (64, 64, 0), (64, 128, 1), (256, 64, 1),
                                           (1024, 64, 1), (1024, 256, 0)])
def test_replayed_schedule_matches_oracle(oracle, cb, lanes, mode):
    T = pg.Topology(cb, lanes, mode, host_only=True)
    cseed = bytes(range(32, 64))
    assert replay(oracle, T, cseed) == oracle.label_chunk(cseed, cb)
    T.close()


# Depths the reference's own slot-map test checks (common/test/test_slot_map.c:139-146
# runs cb = 2^12..2^17); 2^14 and up are slow here (seconds to a minute each).
_DEEP = [pytest.param(1 << k, 1, marks=pytest.mark.slow, id=f"2^{k}") for k in (14, 15, 17)]


@pytest.mark.parametrize("cb,mode", [(1 << 10, 0), (1 << 10, 1), (1 << 12, 1)] + _DEEP)
def test_schedule_invariants(oracle, cb, mode):
    T = pg.Topology(cb, 64, mode, host_only=True)
    steps, rec = T.export()
    info = T.info
    npred, p0, p1, rank = oracle.edges(cb)
    n = len(npred)
    assert info.nodes == n
    body = rec[:info.steps]
    w = body[:, :, 3].astype(np.int64)
    live = (w & 3) != 3
    ids = (w >> 2)[live]
    # every node exactly once, and the final padding row is idle
    assert np.array_equal(np.sort(ids), np.arange(n))
    assert np.all((rec[info.steps, :, 3] & 3) == 3)
    step_of = np.empty(n, np.int64)
    step_of[ids] = np.nonzero(live)[0]
    # predecessors strictly earlier
    has0 = npred >= 1
    has1 = npred == 2
    assert np.all(step_of[p0[has0]] < step_of[np.nonzero(has0)[0]])
    assert np.all(step_of[p1[has1]] < step_of[np.nonzero(has1)[0]])
    # locations: outputs at their rank, internal nodes in slots >= 2
    loc = np.empty(n, np.int64)
    loc[ids] = body[:, :, 0].astype(np.int64)[live]
    out = rank >= 0
    assert np.array_equal(loc[out], PERSIST | rank[out].astype(np.int64))
    assert np.all(loc[~out] >= 2) and np.all(loc[~out] < info.pool_slots)
    # slot liveness: nodes sharing a slot have disjoint [birth, last read] ranges
    last = step_of.copy()
    np.maximum.at(last, p0[has0], step_of[np.nonzero(has0)[0]])
    np.maximum.at(last, p1[has1], step_of[np.nonzero(has1)[0]])
    inner = np.nonzero(~out)[0]
    order = np.lexsort((step_of[inner], loc[inner]))
    sl, b, e = loc[inner][order], step_of[inner][order], last[inner][order]
    same = sl[1:] == sl[:-1]
    assert np.all(b[1:][same] > e[:-1][same])
    if mode == 0:
        assert info.steps >= info.levels
    T.close()


@pytest.mark.parametrize("cb,lanes,mode", [(1 << 10, 64, 1), (1 << 12, 64, 1), (1 << 10, 128, 0),
                                           (1 << 11, 256, 1)])
def test_near_table(oracle, cb, lanes, mode):
    """The near table names, per predecessor, the lane of the immediately
    preceding step that wrote it (0xFFFF otherwise, and for absent
    predecessors and idle lanes)."""
    T = pg.Topology(cb, lanes, mode, host_only=True)
    steps, rec = T.export()
    near = T.export_near()
    info = T.info
    npred, p0, p1, rank = oracle.edges(cb)
    n = len(npred)
    w = rec[:, :, 3].astype(np.int64)
    live = (w & 3) != 3
    pos = np.nonzero(live)
    ids = (w >> 2)[live]
    step_of = np.empty(n, np.int64)
    lane_of = np.empty(n, np.int64)
    step_of[ids], lane_of[ids] = pos[0], pos[1]
    exp = np.full(rec.shape[:2], 0xFFFFFFFF, np.int64)
    s = pos[0]
    h0 = np.full(len(ids), 0xFFFF, np.int64)
    h1 = np.full(len(ids), 0xFFFF, np.int64)
    a0 = npred[ids] >= 1
    q0 = p0[ids].astype(np.int64)
    m0 = a0 & (step_of[q0] == s - 1)
    h0[m0] = lane_of[q0[m0]]
    a1 = npred[ids] == 2
    q1 = p1[ids].astype(np.int64)
    m1 = a1 & (step_of[q1] == s - 1)
    h1[m1] = lane_of[q1[m1]]
    exp[pos] = h0 | (h1 << 16)
    assert np.array_equal(near.astype(np.int64), exp)
    # Some predecessors do come from the previous step, and none at step 0.
    assert np.count_nonzero(m0) + np.count_nonzero(m1) > 0
    assert np.all(near[0] == 0xFFFFFFFF)
    # The kernels take bit 15 of a half, sign-extended, as the far mask: it
    # is set exactly for 0xFFFF (lane indexes stay below 2^15).
    halves = np.stack([near & 0xFFFF, near >> 16]).astype(np.int64)
    assert np.array_equal(((halves >> 15) & 1) == 1, halves == 0xFFFF)
    T.close()


def test_header_symbols_exported():
    """Every function declared in include/gpu_label.h is exported by the library."""
    import os
    hdr = open(os.path.join(os.path.dirname(pg.HERE), "include", "gpu_label.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(pose_\w+)\s*\(", hdr))
    assert len(names) >= 10
    for n in sorted(names):
        assert hasattr(pg.lib, n), n


def test_error_paths_cpu():
    # out-of-range chunk_blocks: host build refuses
    assert not pg.lib.pose_gpu_topo_build_host(1 << 21, 64, 1)
    assert not pg.lib.pose_gpu_topo_build_host(1024, 0, 1)
    assert not pg.lib.pose_gpu_topo_build_host(1024, 64, 7)
    # non-BLAKE3 algorithms are refused before touching a device
    assert pg.lib.pose_graph_hbm_topo_build(1024, 1) is None
    assert pg.lib.pose_graph_label_hbm(b"\0" * 32, 32, 1024, None, 1 << 20, 0, 2) == -1


def test_product_chunk_seed_matches_reference_fixtures(golden):
    """pose_gpu_chunk_seed (host side of the verifier) against the chunk seeds
    the reference build wrote (12 seed lengths x 7 indices, incl. 2^64-1)."""
    for c in golden["chunk_seeds"]:
        assert pg.chunk_seed(bytes.fromhex(c["seed"]), int(c["index"])).hex() == c["out"]
    assert pg.lib.pose_gpu_chunk_seed(None, 5, 0, C.create_string_buffer(32)) == -1


@pytest.mark.parametrize("cb", [1, 2, 3, 5, 16, 100, 256, 1024])
def test_copy_schedules_replay_to_oracle(oracle, cb):
    """The scaffold's two copies, each replayed on its own schedule with its
    own scratch pool, write exactly their output ranks; together they give the
    oracle's super-chunk (cb = 1: the second copy has no output)."""
    T = pg.Topology(cb, 64, pg.SCHED_LIST, host_only=True)
    cseed = bytes(range(7, 39))
    persist = {}
    n_total = 0
    for c in (0, 1):
        info, steps, rec, near = T.part(c)
        before = set(persist)
        n_total += replay_rows(oracle, rec, info["steps"], cb, cseed, persist)
        assert sorted(set(persist) - before) == list(range(info["first_rank"], info["end_rank"]))
    assert n_total == T.info.nodes
    assert b"".join(persist[r] for r in range(cb)) == oracle.label_chunk(cseed, cb)
    T.close()


@pytest.mark.parametrize("cb,mode", [(1 << 10, 1), (1 << 12, 1), (1 << 10, 0), (3000, 1)] + _DEEP)
def test_copy_schedule_invariants(oracle, cb, mode):
    """Copy c holds node ids [c*n/2, (c+1)*n/2), every one exactly once, with
    no predecessor outside the copy; predecessors sit in earlier steps; slots
    sharing a location have disjoint lifetimes; the near table names the
    previous step's lane; a copy needs at least the critical path's steps and
    no more than the whole scaffold."""
    T = pg.Topology(cb, 64, mode, host_only=True)
    npred, p0, p1, rank = oracle.edges(cb)
    n = len(npred)
    for c in (0, 1):
        info, steps, rec, near = T.part(c)
        lo, hi = info["first_node"], info["end_node"]
        assert (lo, hi) == ((0, n // 2) if c == 0 else (n // 2, n))
        S = info["steps"]
        assert T.info.levels <= S <= T.info.steps
        body = rec[:S]
        w = body[:, :, 3].astype(np.int64)
        live = (w & 3) != 3
        ids = (w >> 2)[live]
        assert np.array_equal(np.sort(ids), np.arange(lo, hi))
        assert np.all((rec[S, :, 3] & 3) == 3)
        # every predecessor inside the copy
        mine = np.arange(lo, hi)
        has0, has1 = npred[mine] >= 1, npred[mine] == 2
        assert np.all((p0[mine][has0] >= lo) & (p0[mine][has0] < hi))
        assert np.all((p1[mine][has1] >= lo) & (p1[mine][has1] < hi))
        pos = np.nonzero(live)
        step_of = np.full(n, -1, np.int64)
        lane_of = np.full(n, -1, np.int64)
        step_of[ids], lane_of[ids] = pos[0], pos[1]
        kids0, kids1 = mine[has0], mine[has1]
        assert np.all(step_of[p0[kids0]] < step_of[kids0])
        assert np.all(step_of[p1[kids1]] < step_of[kids1])
        # outputs at their absolute rank, the copy's ranks exactly
        loc = np.full(n, -1, np.int64)
        loc[ids] = body[:, :, 0].astype(np.int64)[live]
        out = rank[mine] >= 0
        assert np.array_equal(loc[mine][out], PERSIST | rank[mine][out].astype(np.int64))
        assert np.array_equal(np.sort(rank[mine][out]), np.arange(info["first_rank"], info["end_rank"]))
        inner = mine[~out]
        assert np.all(loc[inner] >= 2) and np.all(loc[inner] < info["pool_slots"])
        last = step_of.copy()
        np.maximum.at(last, p0[kids0], step_of[kids0])
        np.maximum.at(last, p1[kids1], step_of[kids1])
        order = np.lexsort((step_of[inner], loc[inner]))
        sl, b, e = loc[inner][order], step_of[inner][order], last[inner][order]
        same = sl[1:] == sl[:-1]
        assert np.all(b[1:][same] > e[:-1][same])
        # near table
        q0 = p0[ids].astype(np.int64)
        q1 = p1[ids].astype(np.int64)
        a0 = npred[ids] >= 1
        a1 = npred[ids] == 2
        h0 = np.where(a0 & (step_of[q0] == pos[0] - 1), lane_of[q0], 0xFFFF)
        h1 = np.where(a1 & (step_of[q1] == pos[0] - 1), lane_of[q1], 0xFFFF)
        exp = np.full(rec.shape[:2], 0xFFFFFFFF, np.int64)
        exp[pos] = h0 | (h1 << 16)
        assert np.array_equal(near.astype(np.int64), exp)
    T.close()


def test_copy_schedules_only_at_64_lanes():
    T = pg.Topology(1 << 10, 128, pg.SCHED_LIST, host_only=True)
    assert T.part(0) is None and T.part(1) is None
    T.close()
    T = pg.Topology(1 << 10, 64, pg.SCHED_LIST, host_only=True)
    assert T.part(2) is None
    T.close()


def test_copy_critical_path_at_2_17():
    """At chunk_blocks = 2^17, one copy at 64 nodes per step schedules in
    exactly the scaffold's 393,197 dependency levels, against 568,762 steps
    for the whole scaffold."""
    T = pg.Topology(1 << 17, 64, pg.SCHED_LIST, host_only=True)
    assert T.info.levels == 393197
    assert T.info.steps == 568762
    for c in (0, 1):
        info = pg.TopoPart()
        assert pg.lib.pose_gpu_topo_part(T.h, c, C.byref(info), None, None, None) == 0
        assert info.steps == 393197
        assert (info.first_rank, info.end_rank) == ((0, 1 << 16) if c == 0 else (1 << 16, 1 << 17))
    T.close()


_RNG_CASES = [(int(c), int(l), int(m)) for c, l, m in zip(
    np.random.default_rng(1234).integers(1, 600, 30),
    np.random.default_rng(4321).integers(1, 161, 30),
    np.random.default_rng(99).integers(0, 2, 30))]


@pytest.mark.parametrize("cb,lanes,mode", _RNG_CASES)
def test_random_shapes_replay_to_oracle(oracle, cb, lanes, mode):
    """Random chunk_blocks (mostly not powers of two), odd and even step
    widths (odd widths build the whole schedule directly, even ones from the
    mirrored copy order) and both schedule modes: the whole-scaffold schedule
    replays to the oracle's super-chunk."""
    T = pg.Topology(cb, lanes, mode, host_only=True)
    cseed = bytes((7 * cb + lanes + k) & 0xFF for k in range(32))
    assert replay(oracle, T, cseed) == oracle.label_chunk(cseed, cb)
    T.close()


@pytest.mark.parametrize("cb", [int(c) for c in np.random.default_rng(77).integers(1, 2000, 12)])
def test_random_chunk_blocks_copy_schedules(oracle, cb):
    """Random chunk_blocks at 64 lanes: the two copy schedules, each replayed
    with its own scratch pool, write their rank ranges and together give the
    oracle's super-chunk."""
    T = pg.Topology(cb, 64, pg.SCHED_LIST, host_only=True)
    cseed = bytes((cb + k) & 0xFF for k in range(32))
    persist = {}
    for c in (0, 1):
        info, steps, rec, near = T.part(c)
        replay_rows(oracle, rec, info["steps"], cb, cseed, persist)
    assert sorted(persist) == list(range(cb))
    assert b"".join(persist[r] for r in range(cb)) == oracle.label_chunk(cseed, cb)
    T.close()


def _schedule_digest(cb, lanes, mode):
    import hashlib
    T = pg.Topology(cb, lanes, mode, host_only=True)
    h = hashlib.sha256()
    steps, rec = T.export()
    h.update(steps.tobytes())
    h.update(rec.tobytes())
    h.update(T.export_near().tobytes())
    for c in (0, 1):
        p = T.part(c)
        if p is not None:
            info, steps, rec, near = p
            h.update(repr(sorted(info.items())).encode())
            h.update(steps.tobytes())
            h.update(rec.tobytes())
            h.update(near.tobytes())
    h.update(repr((T.info.steps, T.info.pool_slots, T.info.levels)).encode())
    T.close()
    return h.hexdigest()


@pytest.mark.parametrize("cb,lanes,mode", [(3000, 64, 1), (4096, 64, 1), (4096, 64, 0),
                                           (5000, 128, 1), (1 << 15, 64, 1)])
def test_schedule_independent_of_build_threads(cb, lanes, mode, monkeypatch):
    """The host build runs its passes on several threads (atomic maxima and
    increments among them); the whole-scaffold and copy schedules must not
    depend on how many: one thread, an odd count and the default agree."""
    got = []
    for threads in ("1", "7", None):
        if threads is None:
            monkeypatch.delenv("POSE_GPU_BUILD_THREADS", raising=False)
        else:
            monkeypatch.setenv("POSE_GPU_BUILD_THREADS", threads)
        got.append(_schedule_digest(cb, lanes, mode))
    assert got[0] == got[1] == got[2]
