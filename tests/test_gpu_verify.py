"""GPU challenge answers (pose_gpu_challenge_batch) against the reference's
pose_graph_challenge fixtures, the oracle, and the labeler's own persisted
layout.  Integer/byte work: every answer must match bit for bit."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import pose_gpu as pg  # noqa: E402

CPU_THREADS = min(16, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    return torch.device("cuda:0")


def test_challenges_match_reference_fixtures(dev, golden):
    """All 24 pose_graph_challenge vectors of the reference build (chunk_blocks
    1..4097, including non-powers of two; first, middle and last rank)."""
    cs = golden["challenges"]
    say(f"{len(cs)} fixture challenges in one call")
    got = pg.challenge_batch([(bytes.fromhex(c["seed"]), c["m"], c["idx"]) for c in cs])
    for c, g in zip(cs, got):
        assert g.hex() == c["out"], (c["m"], c["idx"])


def test_challenges_mixed_batch_against_oracle(dev, oracle):
    """Random challenges over several chunk_blocks, with repeats of the same
    super-chunk and of the same rank, in one call."""
    rng = np.random.default_rng(21)
    seeds = [rng.integers(0, 256, 32, dtype=np.uint8).tobytes() for _ in range(5)]
    cbs = [1 << 8, 1000, 1 << 10, 1 << 12]
    chal = []
    for _ in range(60):
        s = seeds[int(rng.integers(0, len(seeds)))]
        cb = cbs[int(rng.integers(0, len(cbs)))]
        chal.append((s, cb, int(rng.integers(0, cb))))
    chal.append(chal[0])
    say(f"{len(chal)} random challenges")
    got = pg.challenge_batch(chal)
    full = {}
    for (s, cb, r), g in zip(chal, got):
        if (s, cb) not in full:
            full[(s, cb)] = oracle.label_chunk(s, cb)
        assert g == full[(s, cb)][32 * r:32 * (r + 1)], (cb, r)


def test_challenges_answer_the_persisted_layout(dev):
    """Absolute block b of a labeled session is challenge
    (chunk_seed(session, b // cb), cb, b % cb): the verifier's answers equal the
    bytes the labeler persisted, across super-chunk boundaries and offsets."""
    cb, n = 1 << 11, 6
    session = b"verifier session key, 35 bytes.."
    rbo = 123 * cb
    buf = torch.zeros(n * cb * 32, dtype=torch.uint8, device=dev)
    assert pg.label_hbm(session, cb, buf.data_ptr(), n * cb * 32, rbo) == 0
    host = buf.cpu().numpy().tobytes()
    rng = np.random.default_rng(5)
    blocks = [rbo, rbo + cb - 1, rbo + cb, rbo + n * cb - 1] + [rbo + int(rng.integers(0, n * cb)) for _ in range(40)]
    chal = [(pg.chunk_seed(session, b // cb), cb, b % cb) for b in blocks]
    say(f"{len(chal)} challenges on a labeled session, cb={cb}")
    got = pg.challenge_batch(chal)
    for b, g in zip(blocks, got):
        k = b - rbo
        assert g == host[32 * k:32 * (k + 1)], b


def test_challenge_at_depth_2_17(dev, golden):
    """One 2^17 super-chunk (fixture region): ranks 0, the middle and the last
    equal the fixture's first label and the labeler's own output."""
    cb = 1 << 17
    r = [r for r in golden["regions"] if r["cb"] == cb][0]
    seed = bytes.fromhex(r["seed"])
    idx = r["block_offset"] // cb
    cs = pg.chunk_seed(seed, idx)
    buf = torch.zeros(cb * 32, dtype=torch.uint8, device=dev)
    assert pg.label_hbm(seed, cb, buf.data_ptr(), cb * 32, r["block_offset"]) == 0
    host = buf.cpu().numpy().tobytes()
    say("challenges at 2^17")
    got, ms = pg.challenge_batch([(cs, cb, 0), (cs, cb, cb // 2), (cs, cb, cb - 1)], with_time=True)
    assert got[0].hex() == r["rank0"][0]
    assert got[0] == host[:32] and got[1] == host[32 * (cb // 2):32 * (cb // 2 + 1)] and got[2] == host[-32:]
    assert ms > 0


def test_challenge_errors(dev):
    s = bytes(32)
    with pytest.raises(RuntimeError):
        pg.challenge_batch([(s, 64, 64)])
    with pytest.raises(RuntimeError):
        pg.challenge_batch([(s, 0, 0)])
    with pytest.raises(RuntimeError):
        pg.challenge_batch([(s, (1 << 20) + 1, 0)])
    assert pg.challenge_batch([]) == []


def test_challenges_label_only_the_needed_copy(dev, oracle):
    """A super-chunk whose challenges all fall in one copy of its scaffold has
    only that copy labeled; answers from copy 0 alone, copy 1 alone and both
    copies agree with the oracle's super-chunk, for powers of two and not
    (chunk_blocks 1000: copy 1 holds fewer ranks) and for chunk_blocks 1 (copy
    1 holds none)."""
    rng = np.random.default_rng(8)
    for cb in (1, 3, 1000, 1 << 10, 1 << 12):
        s = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        full = oracle.label_chunk(s, cb)
        half_w = 1 << max(0, (cb - 1).bit_length() - 1) if cb > 1 else 1
        lo = [r for r in (0, half_w // 2, half_w - 1) if r < cb]
        hi = [r for r in (half_w, (half_w + cb) // 2, cb - 1) if half_w <= r < cb]
        for ranks in (lo, hi, lo + hi):
            if not ranks:
                continue
            say(f"cb={cb} ranks={ranks}")
            got = pg.challenge_batch([(s, cb, r) for r in ranks])
            for r, g in zip(ranks, got):
                assert g == full[32 * r:32 * (r + 1)], (cb, r)


def test_challenge_release_and_rebuild(dev, golden):
    """Releasing the verifier's cached topologies frees them; the next call
    builds them again and answers the same."""
    cs = golden["challenges"][:8]
    chal = [(bytes.fromhex(c["seed"]), c["m"], c["idx"]) for c in cs]
    first = pg.challenge_batch(chal)
    free0 = torch.cuda.mem_get_info()[0]
    pg.lib.pose_gpu_challenge_release()
    free1 = torch.cuda.mem_get_info()[0]
    say(f"released {(free1 - free0) / 2**20:.1f} MiB of cached topologies")
    assert free1 >= free0
    assert pg.challenge_batch(chal) == first
    assert [g.hex() for g in first] == [c["out"] for c in cs]
    pg.lib.pose_gpu_challenge_release()
    pg.lib.pose_gpu_challenge_release()        # nothing cached: a no-op
