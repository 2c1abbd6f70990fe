"""ctypes access to the CPU checkers (TEST INFRASTRUCTURE ONLY).

``Oracle`` wraps oracle/liboracle.so, this repository's C restatement of the
reference labeling path.  ``RefLib`` wraps oracle/_ref/libposeref_t*.so, the
reference's own common/ CPU library compiled from source, when present.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.
"""
import ctypes as C
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")


def _buf(b):
    return None if b is None else C.c_char_p(bytes(b))


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        L.oracle_blake3_hash.argtypes = [C.c_void_p, C.c_uint32, C.c_char_p, C.c_size_t, C.c_char_p]
        L.oracle_blake3_context_key.argtypes = [C.c_char_p, C.c_void_p]
        L.oracle_blake3_derive_key.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t, C.c_char_p]
        L.oracle_label_node_keys.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_graph_descriptor.argtypes = [C.c_char_p, C.c_uint64, C.c_uint64]
        L.oracle_chunk_seed.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t, C.c_uint64]
        L.oracle_label_node.argtypes = [C.c_char_p, C.c_int, C.c_uint64, C.c_char_p,
                                        C.c_char_p, C.c_char_p, C.c_char_p]
        L.oracle_graph_parameter_n.argtypes = [C.c_uint64]
        L.oracle_graph_parameter_n.restype = C.c_int
        L.oracle_scaffold_node_count.argtypes = [C.c_uint64]
        L.oracle_scaffold_node_count.restype = C.c_uint64
        L.oracle_graph_edges.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_graph_edges.restype = C.c_int
        L.oracle_label_chunk_inplace.argtypes = [C.c_char_p, C.c_uint64, C.c_void_p]
        L.oracle_label_chunk_inplace.restype = C.c_int
        L.oracle_label_chunk_edges.argtypes = [C.c_char_p, C.c_uint64, C.c_void_p]
        L.oracle_label_chunk_edges.restype = C.c_int
        L.oracle_label_region.argtypes = [C.c_char_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                          C.c_uint64, C.c_uint32, C.c_int]
        L.oracle_label_region.restype = C.c_int
        L.oracle_label_superchunks.argtypes = [C.c_char_p, C.c_size_t, C.c_uint64, C.c_uint32,
                                               C.c_void_p, C.c_size_t, C.c_void_p, C.c_int]
        L.oracle_label_superchunks.restype = C.c_int
        L.oracle_splitmix64_next.argtypes = [C.POINTER(C.c_uint64)]
        L.oracle_splitmix64_next.restype = C.c_uint64
        self.L = L

    # BLAKE3 modes
    def blake3(self, data, key=None, flags=0):
        out = C.create_string_buffer(32)
        k = None
        if key is not None:
            k = (C.c_uint32 * 8)(*[int.from_bytes(key[4 * i:4 * i + 4], "little") for i in range(8)])
        self.L.oracle_blake3_hash(k, flags, bytes(data), len(data), out)
        return out.raw

    def keyed(self, key32, data):
        return self.blake3(data, key32, 16)

    def derive_key(self, context, material):
        out = C.create_string_buffer(32)
        self.L.oracle_blake3_derive_key(context.encode(), bytes(material), len(material), out)
        return out.raw

    def keys(self):
        a, b = (C.c_uint32 * 8)(), (C.c_uint32 * 8)()
        self.L.oracle_label_node_keys(a, b)
        return list(a), list(b)

    def descriptor(self, m, n=None):
        if n is None:
            n = self.parameter_n(m)
        out = C.create_string_buffer(32)
        self.L.oracle_graph_descriptor(out, m, n)
        return out.raw

    def chunk_seed(self, seed, idx):
        out = C.create_string_buffer(32)
        self.L.oracle_chunk_seed(out, bytes(seed), len(seed), idx)
        return out.raw

    def label_node(self, is_source, nid, p0, p1, seed, desc):
        out = C.create_string_buffer(32)
        self.L.oracle_label_node(out, int(is_source), nid, _buf(p0), _buf(p1), bytes(seed), bytes(desc))
        return out.raw

    def parameter_n(self, m):
        return self.L.oracle_graph_parameter_n(m)

    def scaffold_count(self, m):
        return self.L.oracle_scaffold_node_count(m)

    def edges(self, m):
        import numpy as np
        n = self.scaffold_count(m)
        npred = np.zeros(n, np.uint8)
        p0 = np.zeros(n, np.uint32)
        p1 = np.zeros(n, np.uint32)
        rank = np.zeros(n, np.int32)
        rc = self.L.oracle_graph_edges(m, npred.ctypes.data, p0.ctypes.data, p1.ctypes.data,
                                       rank.ctypes.data)
        assert rc == 0
        return npred, p0, p1, rank

    def label_chunk(self, cseed, m, formulation="inplace"):
        out = C.create_string_buffer(m * 32)
        f = self.L.oracle_label_chunk_inplace if formulation == "inplace" else self.L.oracle_label_chunk_edges
        assert f(bytes(cseed), m, out) == 0
        return out.raw

    def label_region(self, seed, cb, n_super, block_offset=0, threads=None):
        threads = threads or os.cpu_count() or 1
        out = C.create_string_buffer(n_super * cb * 32)
        rc = self.L.oracle_label_region(bytes(seed), len(seed), out, n_super * cb * 32,
                                        block_offset, cb, threads)
        assert rc == 0
        return out.raw

    def label_superchunks(self, seed, cb, js, block_offset=0, threads=None):
        threads = threads or os.cpu_count() or 1
        arr = (C.c_uint64 * len(js))(*js)
        out = C.create_string_buffer(len(js) * cb * 32)
        rc = self.L.oracle_label_superchunks(bytes(seed), len(seed), block_offset, cb, arr,
                                             len(js), out, threads)
        assert rc == 0
        return out.raw


class RefLib:
    """The reference common/ library built from source by `make -C oracle ref`."""

    def __init__(self, threads=8):
        path = os.path.join(REF_DIR, f"libposeref_t{threads}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.threads = threads
        L = C.CDLL(path)
        L.pose_graph_scratch_bytes_inplace.argtypes = [C.c_uint32]
        L.pose_graph_scratch_bytes_inplace.restype = C.c_size_t
        L.pose_graph_pool_create.restype = C.c_void_p
        L.pose_graph_pool_destroy.argtypes = [C.c_void_p]
        L.pose_graph_label_region_pooled_inplace.argtypes = [
            C.c_char_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_uint64,
            C.c_void_p, C.c_void_p, C.c_int, C.c_uint32]
        L.pose_graph_label_region_pooled_inplace.restype = C.c_int
        self.L = L
        self._pool = None
        self._scratch = {}

    def _scr(self, cb):
        if cb not in self._scratch:
            sb = self.L.pose_graph_scratch_bytes_inplace(cb)
            bufs = [C.create_string_buffer(sb) for _ in range(self.threads)]
            arr = (C.c_void_p * self.threads)(*[C.cast(b, C.c_void_p) for b in bufs])
            self._scratch[cb] = (bufs, arr)
        return self._scratch[cb][1]

    def label_region(self, seed, cb, n_super, block_offset=0):
        if self._pool is None:
            self._pool = self.L.pose_graph_pool_create()
        out = C.create_string_buffer(n_super * cb * 32)
        rc = self.L.pose_graph_label_region_pooled_inplace(
            bytes(seed), len(seed), out, n_super * cb * 32, block_offset,
            self._scr(cb), self._pool, 0, cb)
        assert rc == 0
        return out.raw

    def challenge(self, chunk_seed, m, idx):
        """The reference's pose_graph_challenge (common/src/graph.c:1019-1042):
        labels the whole graph on one core and returns rank idx."""
        L = self.L
        if not getattr(L, "_challenge_bound", False):
            L.pose_graph_challenge.argtypes = [C.c_char_p, C.c_size_t, C.c_uint64, C.c_uint64,
                                               C.c_char_p, C.c_int]
            L.pose_graph_challenge.restype = C.c_int
            L._challenge_bound = True
        out = C.create_string_buffer(32)
        assert L.pose_graph_challenge(bytes(chunk_seed), 32, m, idx, out, 0) == 0
        return out.raw

    def topology_build_seconds(self, cb):
        """Host time of the reference's seed-independent topology for one
        chunk_blocks: pose_graph_edges (common/src/graph.c:1044-1100) and
        pose_graph_slot_map (common/src/graph.c:1102-1211), the two calls its
        GPU topology build makes (common/src/cuda/label_gpu_common.cu:59-254).
        Returns (edges seconds, slot map seconds, node count)."""
        L = self.L
        if not getattr(L, "_topo_bound", False):
            L.pose_graph_edges.argtypes = [C.c_uint64, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
            L.pose_graph_edges.restype = C.c_int
            L.pose_graph_slot_map.argtypes = [C.c_uint32, C.c_void_p, C.POINTER(C.c_uint32)]
            L.pose_graph_slot_map.restype = C.c_int
            L._topo_bound = True
        libc = C.CDLL(None)
        libc.free.argtypes = [C.c_void_p]
        nodes, count = C.c_void_p(), C.c_uint64()
        t0 = time.perf_counter()
        assert L.pose_graph_edges(cb, C.byref(nodes), C.byref(count)) == 0
        t1 = time.perf_counter()
        libc.free(nodes)
        slots = (C.c_uint32 * count.value)()
        pool = C.c_uint32()
        t2 = time.perf_counter()
        assert L.pose_graph_slot_map(cb, slots, C.byref(pool)) == 0
        t3 = time.perf_counter()
        return t1 - t0, t3 - t2, count.value

    def close(self):
        if self._pool is not None:
            self.L.pose_graph_pool_destroy(self._pool)
            self._pool = None


def ref_available(threads=8):
    return os.path.exists(os.path.join(REF_DIR, f"libposeref_t{threads}.so"))


def splitmix_key(prng_seed):
    """32-byte session key: the first four SplitMix64 outputs, little-endian."""
    st = prng_seed & 0xFFFFFFFFFFFFFFFF
    out = b""
    for _ in range(4):
        st = (st + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        out += (z ^ (z >> 31)).to_bytes(8, "little")
    return out
