"""The synthetic workload streams (pose-memory-wiping_amd/workloads.py) used by
bench.py and the GPU configuration tests: SplitMix64 keys equal the fixture
keys and the oracle's SplitMix64, draws stay in range, starts are aligned."""
import ctypes as C
import math

import workloads as wl


def test_splitmix_matches_fixtures_and_oracle(golden, oracle):
    for k in golden["splitmix_keys"]:
        assert wl.SplitMix(k["prng_seed"]).key().hex() == k["key"]
    assert wl.SplitMix(0).next() == 0xE220A8397B1DCDAF
    st = C.c_uint64(12345)
    sm = wl.SplitMix(12345)
    for _ in range(100):
        assert oracle.L.oracle_splitmix64_next(C.byref(st)) == sm.next()


def test_uneven_sessions_shape():
    s = wl.uneven_sessions()
    assert len(s) == 48
    assert s[0][0] == wl.SplitMix(3).key()
    for key, cb, n, rbo in s:
        assert len(key) == 32 and cb in (1 << 12, 1 << 14, 1 << 15, 1 << 17)
        assert 1 <= n <= 32 and rbo % cb == 0 and rbo * 32 < (1 << 40)
    assert len({cb for _, cb, _, _ in s}) == 4


def test_poisson_requests_shape():
    reqs, sm = wl.poisson_requests(16.0, 20.0)
    assert reqs[0][0] == 0.0 and reqs[0][1] == wl.SplitMix(4).key()
    times = [r[0] for r in reqs]
    assert times == sorted(times) and times[-1] < 20.0
    # about rate x seconds arrivals (Poisson: 320 +- 4 sigma)
    assert abs(len(reqs) - 320) < 4 * math.sqrt(320)
    for t, key, cb, n, rbo in reqs:
        assert 1 << 12 <= cb <= 1 << 17 and cb & (cb - 1) == 0
        assert 1 <= n <= 128 and rbo % cb == 0
    more, _ = wl.poisson_requests(16.0, 1.0, sm=sm)
    assert more[0][1] != reqs[0][1]


def test_streamed_batch():
    key, cb, n, rbo, _ = wl.streamed_batch()
    assert key == wl.SplitMix(2).key() and cb == 1 << 17 and n == 64 and rbo % cb == 0
