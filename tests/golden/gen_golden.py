#!/usr/bin/env python3
"""Write tests/golden/golden.json from the reference CPU library.

The reference's common/ C library is compiled from its own sources by
``make -C oracle ref`` into oracle/_ref/ (this container only: it needs the
reference checkout).  This script calls it through ctypes and records inputs
and outputs as small JSON fixtures, which are committed; the tests then pin the
oracle (and, through it, the HIP labeler) against them without the reference.

What is recorded, and the reference symbol that produced it:
  blake3        blake3_hasher_init / _init_keyed / _init_derive_key, _update,
                _finalize (common/vendor/blake3/blake3.c) over i % 251 inputs
  keys          pose_label_node_keys          (common/src/hash.c:62-67)
  descriptors   pose_graph_descriptor         (common/src/hash.c:198-216)
  chunk_seeds   pose_chunk_seed               (common/src/hash.c:218-233)
  label_node    pose_label_node, BLAKE3       (common/src/hash.c:110-144)
  param_n       pose_graph_parameter_n        (common/src/graph.c:562-570)
  topology      pose_graph_edges              (common/src/graph.c:1044-1100),
                sha256 of the packed (np, p0, p1, rank) arrays + level stats
  slot_pool     pose_graph_slot_map pool size (common/src/graph.c:1102-1211)
  regions       pose_graph_label_region_pooled_inplace (common/src/graph.c:1604-1689)
  challenges    pose_graph_challenge          (common/src/graph.c:1019-1042)
"""
import ctypes as C
import hashlib
import json
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_SO = os.path.join(REPO, "oracle", "_ref", "libposeref_t8.so")
REF_THREADS = 8


def load():
    lib = C.CDLL(REF_SO)
    u8p = C.POINTER(C.c_uint8)
    lib.blake3_hasher_init.argtypes = [C.c_void_p]
    lib.blake3_hasher_init_keyed.argtypes = [C.c_void_p, C.c_char_p]
    lib.blake3_hasher_init_derive_key.argtypes = [C.c_void_p, C.c_char_p]
    lib.blake3_hasher_update.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
    lib.blake3_hasher_finalize.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
    lib.pose_label_node_keys.argtypes = [C.c_void_p, C.c_void_p]
    lib.pose_graph_descriptor.argtypes = [C.c_char_p, C.c_uint64, C.c_uint64]
    lib.pose_chunk_seed.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t, C.c_uint64]
    lib.pose_label_node.argtypes = [C.c_char_p, C.c_int, C.c_uint64, C.c_char_p,
                                    C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
    lib.pose_graph_parameter_n.argtypes = [C.c_uint64]
    lib.pose_graph_parameter_n.restype = C.c_int
    lib.pose_graph_edges.argtypes = [C.c_uint64, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
    lib.pose_graph_edges.restype = C.c_int
    lib.pose_graph_slot_map.argtypes = [C.c_uint32, C.c_void_p, C.POINTER(C.c_uint32)]
    lib.pose_graph_slot_map.restype = C.c_int
    lib.pose_graph_scaffold_node_count.argtypes = [C.c_uint32]
    lib.pose_graph_scaffold_node_count.restype = C.c_uint64
    lib.pose_graph_scratch_bytes_inplace.argtypes = [C.c_uint32]
    lib.pose_graph_scratch_bytes_inplace.restype = C.c_size_t
    lib.pose_graph_pool_create.restype = C.c_void_p
    lib.pose_graph_pool_destroy.argtypes = [C.c_void_p]
    lib.pose_graph_label_region_pooled_inplace.argtypes = [
        C.c_char_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_uint64,
        C.c_void_p, C.c_void_p, C.c_int, C.c_uint32]
    lib.pose_graph_label_region_pooled_inplace.restype = C.c_int
    lib.pose_graph_challenge.argtypes = [C.c_char_p, C.c_size_t, C.c_uint64,
                                         C.c_uint64, C.c_char_p, C.c_int]
    lib.pose_graph_challenge.restype = C.c_int
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    return lib, libc


def pattern(n):
    return bytes(i % 251 for i in range(n))


def b3(lib, mode, data, key_or_ctx=None):
    h = C.create_string_buffer(4096)
    if mode == "hash":
        lib.blake3_hasher_init(h)
    elif mode == "keyed":
        lib.blake3_hasher_init_keyed(h, key_or_ctx)
    else:
        lib.blake3_hasher_init_derive_key(h, key_or_ctx)
    lib.blake3_hasher_update(h, data, len(data))
    out = C.create_string_buffer(32)
    lib.blake3_hasher_finalize(h, out, 32)
    return out.raw.hex()


def splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def key_from_seed(seed):
    st = seed
    out = b""
    for _ in range(4):
        st, v = splitmix64(st)
        out += struct.pack("<Q", v)
    return out


class Node(C.Structure):
    _fields_ = [("num_preds", C.c_int), ("pred", C.c_uint64 * 2), ("challenge_rank", C.c_int)]


def topology(lib, libc, m):
    """Packed arrays + digest of the reference topology for m output blocks."""
    import numpy as np
    ptr = C.c_void_p()
    cnt = C.c_uint64()
    assert lib.pose_graph_edges(m, C.byref(ptr), C.byref(cnt)) == 0
    n = cnt.value
    assert C.sizeof(Node) == 32
    raw = (C.c_uint8 * (32 * n)).from_address(ptr.value)
    arr = np.frombuffer(bytes(raw), dtype=np.dtype([("np", "<i4"), ("pad", "<i4"),
                                                   ("p", "<u8", (2,)), ("rank", "<i4"),
                                                   ("pad2", "<i4")]))
    libc.free(ptr)
    npred = arr["np"].astype(np.uint8)
    p0 = np.where(npred >= 1, arr["p"][:, 0], 0).astype(np.uint32)
    p1 = np.where(npred >= 2, arr["p"][:, 1], 0).astype(np.uint32)
    rank = arr["rank"].astype(np.int32)
    return npred, p0, p1, rank


def topo_digest(npred, p0, p1, rank):
    h = hashlib.sha256()
    h.update(npred.tobytes())
    h.update(p0.tobytes())
    h.update(p1.tobytes())
    h.update(rank.tobytes())
    return h.hexdigest()


def level_stats(npred, p0, p1):
    import numpy as np
    n = len(npred)
    lv = np.zeros(n, dtype=np.int64)
    # ids are topological; vectorised passes are not possible, loop in chunks
    for i in range(n):
        k = npred[i]
        if k == 0:
            lv[i] = 0
        elif k == 1:
            lv[i] = lv[p0[i]] + 1
        else:
            a, b = lv[p0[i]], lv[p1[i]]
            lv[i] = (a if a > b else b) + 1
    L = int(lv.max()) + 1
    widths = np.bincount(lv, minlength=L)
    return {"levels": L, "max_width": int(widths.max()), "min_width": int(widths.min()),
            "levels_below_64": int((widths < 64).sum()),
            "sum_ceil_w64": int(((widths + 63) // 64).sum())}


def region(lib, seed, cb, n_super, block_off):
    pool = lib.pose_graph_pool_create()
    sb = lib.pose_graph_scratch_bytes_inplace(cb)
    bufs = [C.create_string_buffer(sb) for _ in range(REF_THREADS)]
    arr = (C.c_void_p * REF_THREADS)(*[C.cast(b, C.c_void_p) for b in bufs])
    out = C.create_string_buffer(n_super * cb * 32)
    rc = lib.pose_graph_label_region_pooled_inplace(seed, len(seed), out, n_super * cb * 32,
                                                    block_off, arr, pool, 0, cb)
    lib.pose_graph_pool_destroy(pool)
    assert rc == 0
    return out.raw


def main():
    lib, libc = load()
    g = {"source": "reference common/ CPU library compiled from source (gcc -O3)",
         "generator": "tests/golden/gen_golden.py"}

    # BLAKE3 in the three modes over the official-vector input pattern.
    key32 = b"whats the Elvish word for friend"
    ctx = b"BLAKE3 2019-12-27 16:29:52 test vectors context"
    lens = [0, 1, 63, 64, 65, 127, 128, 129, 1023, 1024, 1025, 2048, 2049, 3072,
            3073, 4096, 4097, 5120, 5121, 6144, 6145, 7168, 7169, 8192, 8193,
            16384, 31744, 102400]
    g["blake3"] = {"key": key32.decode(), "context": ctx.decode(), "cases": [
        {"len": n, "hash": b3(lib, "hash", pattern(n)),
         "keyed": b3(lib, "keyed", pattern(n), key32),
         "derive": b3(lib, "derive", pattern(n), ctx)} for n in lens]}

    ks = (C.c_uint32 * 8)()
    ki = (C.c_uint32 * 8)()
    lib.pose_label_node_keys(ks, ki)
    g["keys"] = {"src": [int(x) for x in ks], "int": [int(x) for x in ki]}

    ms = [1, 2, 3, 4, 5, 7, 8, 9, 16, 17, 100, 1000, 1024, 4095, 4096, 4097,
          1 << 14, 1 << 15, 1 << 17, 1 << 20]
    g["descriptors"] = [{"m": m, "n": lib.pose_graph_parameter_n(m),
                         "desc": (lambda o: (lib.pose_graph_descriptor(o, m, lib.pose_graph_parameter_n(m)), o.raw.hex())[1])(C.create_string_buffer(32))}
                        for m in ms]
    g["param_n"] = [{"m": m, "n": lib.pose_graph_parameter_n(m)}
                    for m in list(range(1, 40)) + [1000, 1024, 1025, 1 << 17, (1 << 20)]]

    seeds = [b"\x00" * 32, bytes(range(1, 33)), key_from_seed(0), key_from_seed(1),
             b"", b"\x07", pattern(55), pattern(56), pattern(100), pattern(1016),
             pattern(1017), pattern(3000)]
    idxs = [0, 1, 2, 255, 1 << 20, (1 << 40) + 3, 0xFFFFFFFFFFFFFFFF]
    cs = []
    for s in seeds:
        for i in idxs:
            o = C.create_string_buffer(32)
            lib.pose_chunk_seed(o, s, len(s), i)
            cs.append({"seed": s.hex(), "index": str(i), "out": o.raw.hex()})
    g["chunk_seeds"] = cs

    d10 = C.create_string_buffer(32)
    lib.pose_graph_descriptor(d10, 1024, 9)
    ln = []
    import random
    rnd = random.Random(12345)
    for case in range(24):
        is_src = case % 6 == 0
        nid = rnd.choice([0, 1, 12345, 0xFFFFFFFF, 1 << 33, rnd.getrandbits(64)])
        p0 = bytes(rnd.getrandbits(8) for _ in range(32)) if case % 3 != 2 else None
        p1 = bytes(rnd.getrandbits(8) for _ in range(32)) if case % 4 != 3 else None
        seed = bytes(rnd.getrandbits(8) for _ in range(32))
        desc = d10.raw if case % 2 == 0 else bytes(rnd.getrandbits(8) for _ in range(32))
        o = C.create_string_buffer(32)
        lib.pose_label_node(o, int(is_src), nid, p0, p1, seed, desc, 0)
        ln.append({"is_source": int(is_src), "id": str(nid),
                   "p0": p0.hex() if p0 else None, "p1": p1.hex() if p1 else None,
                   "seed": seed.hex(), "desc": desc.hex(), "out": o.raw.hex()})
    g["label_node"] = ln

    topo = []
    for m in list(range(1, 33)) + [100, 1000] + [1 << k for k in range(6, 18)]:
        npred, p0, p1, rank = topology(lib, libc, m)
        rec = {"m": m, "count": int(len(npred)),
               "scaffold_count": int(lib.pose_graph_scaffold_node_count(m)),
               "digest": topo_digest(npred, p0, p1, rank),
               "n_one_pred": int((npred == 1).sum()), "n_two_pred": int((npred == 2).sum()),
               "n_source": int((npred == 0).sum())}
        if m <= (1 << 14):
            rec.update(level_stats(npred, p0, p1))
        topo.append(rec)
        print("topology", m, rec["count"], file=sys.stderr)
    g["topology"] = topo

    pools = []
    for k in [4, 8, 10, 12, 14]:
        cb = 1 << k
        cnt = lib.pose_graph_scaffold_node_count(cb)
        slot = (C.c_uint32 * cnt)()
        ps = C.c_uint32()
        assert lib.pose_graph_slot_map(cb, slot, C.byref(ps)) == 0
        pools.append({"cb": cb, "pool": ps.value})
    g["slot_pool"] = pools

    # Whole super-chunk digests (rank order, chunk_index from region offsets).
    regs = []
    cfgs = [(1 << k, 3, 0, 0) for k in range(0, 11)] + [
        (1 << 10, 4, 0, 1), (1 << 11, 2, 1 << 11, 2), (1 << 12, 2, 3 << 12, 3),
        (1 << 12, 1, 0, 4), (16, 8, 16, 5), (1 << 13, 1, 7 << 13, 6),
        (1 << 14, 1, 0, 7), (1 << 15, 1, 5 << 15, 8), (1 << 17, 1, 0, 0)]
    for cb, nsup, off, seedsel in cfgs:
        seed = b"\x00" * 32 if seedsel == 0 else key_from_seed(seedsel - 1) if seedsel <= 4 \
            else pattern(10 * seedsel)
        out = region(lib, seed, cb, nsup, off)
        sb = cb * 32
        regs.append({"cb": cb, "n_super": nsup, "block_offset": off, "seed": seed.hex(),
                     "super_sha256": [hashlib.sha256(out[i * sb:(i + 1) * sb]).hexdigest()
                                      for i in range(nsup)],
                     "rank0": [out[i * sb:i * sb + 32].hex() for i in range(nsup)],
                     "last": [out[(i + 1) * sb - 32:(i + 1) * sb].hex() for i in range(nsup)]})
        print("region", cb, nsup, file=sys.stderr)
    g["regions"] = regs

    ch = []
    for m in [1, 2, 3, 5, 7, 12, 100, 1000, 4097]:
        seed = bytes(range(1, 33))
        for idx in sorted({0, m // 2, m - 1}):
            o = C.create_string_buffer(32)
            assert lib.pose_graph_challenge(seed, 32, m, idx, o, 0) == 0
            ch.append({"m": m, "idx": idx, "seed": seed.hex(), "out": o.raw.hex()})
    g["challenges"] = ch

    g["splitmix_keys"] = [{"prng_seed": s, "key": key_from_seed(s).hex()} for s in range(5)]

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(g, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
