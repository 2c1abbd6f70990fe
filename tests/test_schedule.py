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


def replay(oracle, T, cseed):
    steps, rec = T.export()
    info = T.info
    cb = info.chunk_blocks
    desc = oracle.descriptor(cb)
    scratch = {0: bytes(32)}
    persist = {}

    def read(loc):
        if loc & PERSIST:
            return persist[loc & ~PERSIST]
        return scratch[loc]

    n_done = 0
    for s in range(info.steps):
        out = []
        for t in range(info.lanes):
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
    assert n_done == info.nodes
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


@pytest.mark.parametrize("cb,mode", [(1 << 10, 0), (1 << 10, 1), (1 << 12, 1)])
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
