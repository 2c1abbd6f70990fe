"""The oracle (oracle/pose_oracle.c) pinned against the reference's own test
vectors, fixtures written by the reference build, closed forms and a second
independent formulation.  CPU only."""
import hashlib

import numpy as np
import pytest

from pyoracle import RefLib, ref_available, splitmix_key


def pat(n):
    return bytes(i % 251 for i in range(n))


def test_blake3_official_vectors(oracle):
    # common/test/test_hash.c:54-66 (official BLAKE3 vectors, unkeyed)
    assert oracle.blake3(b"").hex() == "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"
    assert oracle.blake3(b"\x00").hex() == "2d3adedff11b61f14c886e35afa036736dcd87a74d27b5c1510225d0f592e213"
    assert oracle.blake3(pat(1024)).hex() == "42214739f095a406f3fc83deb889744ac00df831c10daa55189b5d121c855af7"


def test_blake3_all_modes_vs_fixture(oracle, golden):
    b = golden["blake3"]
    key = b["key"].encode()
    for c in b["cases"]:
        data = pat(c["len"])
        assert oracle.blake3(data).hex() == c["hash"], c["len"]
        assert oracle.keyed(key, data).hex() == c["keyed"], c["len"]
        assert oracle.derive_key(b["context"], data).hex() == c["derive"], c["len"]


def test_keys(oracle, golden):
    s, i = oracle.keys()
    assert s == golden["keys"]["src"]
    assert i == golden["keys"]["int"]
    assert s != i


def test_descriptors(oracle, golden):
    for d in golden["descriptors"]:
        assert oracle.parameter_n(d["m"]) == d["n"]
        assert oracle.descriptor(d["m"]).hex() == d["desc"], d["m"]


def test_parameter_n(oracle, golden):
    # common/test/test_graph.c:29-41
    kat = {1: 0, 2: 0, 3: 1, 4: 1, 5: 2, 8: 2, 9: 3, 16: 3, 17: 4, 32: 4}
    for m, n in kat.items():
        assert oracle.parameter_n(m) == n
    for r in golden["param_n"]:
        assert oracle.parameter_n(r["m"]) == r["n"]


def test_chunk_seeds(oracle, golden):
    for c in golden["chunk_seeds"]:
        got = oracle.chunk_seed(bytes.fromhex(c["seed"]), int(c["index"]))
        assert got.hex() == c["out"], (len(c["seed"]) // 2, c["index"])


def test_label_node(oracle, golden):
    for c in golden["label_node"]:
        p0 = bytes.fromhex(c["p0"]) if c["p0"] else None
        p1 = bytes.fromhex(c["p1"]) if c["p1"] else None
        got = oracle.label_node(c["is_source"], int(c["id"]), p0, p1,
                                bytes.fromhex(c["seed"]), bytes.fromhex(c["desc"]))
        assert got.hex() == c["out"]


def test_label_node_source_ignores_preds(oracle):
    seed, desc = bytes(32), oracle.descriptor(1024)
    a = oracle.label_node(1, 7, None, None, seed, desc)
    b = oracle.label_node(1, 7, b"\x01" * 32, b"\x02" * 32, seed, desc)
    assert a == b
    # absent pred == 32 zero bytes for an internal node (common/include/hash.h:47-57)
    assert oracle.label_node(0, 7, None, None, seed, desc) == oracle.label_node(0, 7, bytes(32), bytes(32), seed, desc)


def closed_form_count(m):
    # 2*S(n+1), S/C/B recurrences; cross-checked against disk-calc's closed form
    def B(d): return (d + 1) << d
    def Cc(l): return 1 if l == 0 else 2 * Cc(l - 1) + 2 * B(l - 1)
    def S(l): return 1 if l == 0 else S(l - 1) + B(l - 1) + Cc(l - 1)
    n = max(0, (m - 1).bit_length() - 1) if m > 1 else 0
    return 2 * S(n + 1)


def test_counts_closed_form(oracle, golden):
    for t in golden["topology"]:
        assert oracle.scaffold_count(t["m"]) == t["count"] == closed_form_count(t["m"])
    # scaffold node counts per super-chunk (closed form 2*S(n+1))
    assert closed_form_count(1 << 17) == 36175870
    assert closed_form_count(1 << 14) == 3047422


def topo_digest(npred, p0, p1, rank):
    h = hashlib.sha256()
    for a in (npred, p0, p1, rank):
        h.update(a.tobytes())
    return h.hexdigest()


def test_topology_digests(oracle, golden):
    for t in golden["topology"]:
        if t["m"] > (1 << 15):
            continue
        npred, p0, p1, rank = oracle.edges(t["m"])
        assert topo_digest(npred, p0, p1, rank) == t["digest"], t["m"]
        assert int((npred == 0).sum()) == t["n_source"] == 2


def test_topology_invariants(oracle):
    for m in [1, 2, 3, 5, 16, 100, 1024]:
        npred, p0, p1, rank = oracle.edges(m)
        ids = np.arange(len(npred))
        assert np.all(p0[npred >= 1] < ids[npred >= 1])
        assert np.all(p1[npred == 2] < ids[npred == 2])
        assert np.all(p0[npred == 2] < p1[npred == 2])
        r = np.sort(rank[rank >= 0])
        assert np.array_equal(r, np.arange(m))


def test_two_formulations_agree(oracle):
    # full-scratch id-indexed labeling vs in-place arena labeling
    for m in [1, 2, 3, 4, 5, 7, 8, 9, 15, 16, 17, 31, 32, 100, 1000, 1024]:
        seed = hashlib.sha256(str(m).encode()).digest()
        assert oracle.label_chunk(seed, m, "inplace") == oracle.label_chunk(seed, m, "edges"), m


def test_region_fixtures(oracle, golden):
    for r in golden["regions"]:
        if r["cb"] > (1 << 14):
            continue
        out = oracle.label_region(bytes.fromhex(r["seed"]), r["cb"], r["n_super"], r["block_offset"])
        sb = r["cb"] * 32
        for i in range(r["n_super"]):
            assert hashlib.sha256(out[i * sb:(i + 1) * sb]).hexdigest() == r["super_sha256"][i], (r["cb"], i)


def test_challenges_match_inplace(oracle, golden):
    # pose_graph_challenge(i) == in-place rank i (common/test/test_inplace.c:114-128)
    for c in golden["challenges"]:
        out = oracle.label_chunk(bytes.fromhex(c["seed"]), c["m"], "edges")
        assert out[c["idx"] * 32:(c["idx"] + 1) * 32].hex() == c["out"]


def test_splitmix_keys(golden):
    for k in golden["splitmix_keys"]:
        assert splitmix_key(k["prng_seed"]).hex() == k["key"]
    # SplitMix64 seed 0 first output
    assert splitmix_key(0)[:8] == (0xE220A8397B1DCDAF).to_bytes(8, "little")


@pytest.mark.skipif(not ref_available(8), reason="reference build absent")
def test_oracle_vs_reference_build_random_regions(oracle):
    ref = RefLib(8)
    rng = np.random.default_rng(7)
    for cb in [1 << 6, 1 << 9, 1 << 11]:
        seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        off = int(rng.integers(0, 1 << 20)) * cb
        a = ref.label_region(seed, cb, 3, off)
        b = oracle.label_region(seed, cb, 3, off)
        assert a == b, cb
    ref.close()


@pytest.mark.skipif(not ref_available(8), reason="reference build absent")
def test_oracle_vs_reference_build_unaligned_offsets(oracle):
    """region_block_offset need not be a multiple of chunk_blocks: super-chunk j
    takes chunk index (offset + j*cb) / cb in 64-bit arithmetic
    (common/src/graph.c:1646-1649), so an offset inside a chunk, and one where
    offset + j*cb wraps past 2^64, both follow the reference build."""
    ref = RefLib(8)
    rng = np.random.default_rng(17)
    cases = []
    for cb in [1 << 4, 1 << 7, 1 << 10, 100]:
        cases.append((cb, int(rng.integers(0, 1 << 30)) * cb + int(rng.integers(1, cb))))
        cases.append((cb, cb - 1))
    cases.append((1 << 8, (1 << 64) - (1 << 8) - 5))   # the second super-chunk's index wraps
    for cb, off in cases:
        seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        a = ref.label_region(seed, cb, 3, off)
        b = oracle.label_region(seed, cb, 3, off)
        assert a == b, (cb, off)
    ref.close()
