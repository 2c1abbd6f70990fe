"""Python view of the MI355X labeler's C ABI (include/gpu_label.h).

Mirrors the reference's GPU labeling interface (common/include/gpu_label.h):
``label_hbm`` is pose_graph_label_hbm, ``Topology.label`` is
pose_graph_hbm_topo_label, ``label_node_single`` is pose_gpu_label_node_single;
``label_batch`` is the additive one-launch batch of many sessions.  Device
buffers are passed as raw device pointers (for example ``tensor.data_ptr()`` of
a torch tensor on the HIP device).

The library must be built (``make -C pose-memory-wiping_amd/csrc``); importing
this module without it raises, there is no fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpose_gpu.so")

POSE_HASH_BLAKE3 = 0
SCHED_LEVELS = 0
SCHED_LIST = 1


class TopoInfo(C.Structure):
    _fields_ = [("chunk_blocks", C.c_uint32), ("lanes", C.c_uint32), ("mode", C.c_int32),
                ("levels", C.c_uint32), ("steps", C.c_uint32), ("pool_slots", C.c_uint32),
                ("nodes", C.c_uint64), ("device_bytes", C.c_uint64),
                ("build_seconds", C.c_double)]


class Session(C.Structure):
    _fields_ = [("topo", C.c_void_p), ("session_seed", C.c_char_p), ("seed_len", C.c_size_t),
                ("region_block_offset", C.c_uint64), ("n_super", C.c_uint64),
                ("out_dev", C.c_void_p)]


class Launch(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("waves_per_simd", C.c_uint32)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `make -C pose-memory-wiping_amd/csrc`")
    L = C.CDLL(LIB_PATH)
    L.pose_graph_label_hbm.argtypes = [C.c_char_p, C.c_size_t, C.c_uint32, C.c_void_p,
                                       C.c_size_t, C.c_uint64, C.c_int]
    L.pose_graph_label_hbm.restype = C.c_int
    L.pose_graph_hbm_topo_build.argtypes = [C.c_uint32, C.c_int]
    L.pose_graph_hbm_topo_build.restype = C.c_void_p
    L.pose_graph_hbm_topo_label.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t, C.c_void_p,
                                            C.c_size_t, C.c_uint64]
    L.pose_graph_hbm_topo_label.restype = C.c_int
    L.pose_graph_hbm_topo_free.argtypes = [C.c_void_p]
    L.pose_graph_hbm_topo_free.restype = None
    L.pose_gpu_label_node_single.argtypes = [C.c_char_p, C.c_int, C.c_uint64, C.c_char_p,
                                             C.c_char_p, C.c_char_p, C.c_char_p,
                                             C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.pose_gpu_label_node_single.restype = C.c_int
    L.pose_gpu_topo_build_ex.argtypes = [C.c_uint32, C.c_uint32, C.c_int]
    L.pose_gpu_topo_build_ex.restype = C.c_void_p
    L.pose_gpu_topo_build_host.argtypes = [C.c_uint32, C.c_uint32, C.c_int]
    L.pose_gpu_topo_build_host.restype = C.c_void_p
    L.pose_gpu_topo_free_host.argtypes = [C.c_void_p]
    L.pose_gpu_topo_free_host.restype = None
    L.pose_gpu_topo_info.argtypes = [C.c_void_p, C.POINTER(TopoInfo)]
    L.pose_gpu_topo_info.restype = C.c_int
    L.pose_gpu_topo_export.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.pose_gpu_topo_export.restype = C.c_int
    L.pose_gpu_label_batch.argtypes = [C.POINTER(Session), C.c_size_t, C.POINTER(Launch),
                                       C.POINTER(C.c_double)]
    L.pose_gpu_label_batch.restype = C.c_int
    return L


lib = _load()


class Topology:
    """Resident scaffold schedule for one chunk_blocks (pose_graph_hbm_topo_*)."""

    def __init__(self, chunk_blocks, lanes=64, mode=SCHED_LIST, host_only=False):
        self.host_only = host_only
        if host_only:
            h = lib.pose_gpu_topo_build_host(chunk_blocks, lanes, mode)
        elif lanes == 64 and mode == SCHED_LIST:
            h = lib.pose_graph_hbm_topo_build(chunk_blocks, POSE_HASH_BLAKE3)
        else:
            h = lib.pose_gpu_topo_build_ex(chunk_blocks, lanes, mode)
        if not h:
            raise RuntimeError(f"topology build failed for chunk_blocks={chunk_blocks}")
        self.h = h
        self.info = TopoInfo()
        lib.pose_gpu_topo_info(h, C.byref(self.info))
        self.cb = self.info.chunk_blocks

    @property
    def super_bytes(self):
        return self.cb * 32

    def label(self, seed, dev_ptr, length, region_block_offset=0):
        return lib.pose_graph_hbm_topo_label(self.h, bytes(seed), len(seed), C.c_void_p(dev_ptr),
                                             length, region_block_offset)

    def export(self):
        import numpy as np
        i = self.info
        steps = np.zeros(i.steps + 1, np.uint32)
        rec = np.zeros(4 * (i.steps + 1) * i.lanes, np.uint32)
        assert lib.pose_gpu_topo_export(self.h, steps.ctypes.data, rec.ctypes.data) == 0
        return steps, rec.reshape(i.steps + 1, i.lanes, 4)

    def close(self):
        if self.h:
            if self.host_only:
                lib.pose_gpu_topo_free_host(self.h)
            else:
                lib.pose_graph_hbm_topo_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def label_hbm(seed, chunk_blocks, dev_ptr, length, region_block_offset=0, algo=POSE_HASH_BLAKE3):
    return lib.pose_graph_label_hbm(bytes(seed), len(seed), chunk_blocks, C.c_void_p(dev_ptr),
                                    length, region_block_offset, algo)


def label_node_single(is_source, node_index, pred0, pred1, seed, desc, key_src, key_int):
    out = C.create_string_buffer(32)
    ks = (C.c_uint32 * 8)(*key_src)
    ki = (C.c_uint32 * 8)(*key_int)
    rc = lib.pose_gpu_label_node_single(out, int(is_source), node_index,
                                        None if pred0 is None else bytes(pred0),
                                        None if pred1 is None else bytes(pred1),
                                        bytes(seed), bytes(desc), ks, ki)
    if rc != 0:
        raise RuntimeError("pose_gpu_label_node_single failed")
    return out.raw


def label_batch(sessions, slots=0, waves_per_simd=0):
    """sessions: iterable of (Topology, seed, region_block_offset, n_super, dev_ptr).
    Returns the kernel time in milliseconds."""
    sessions = list(sessions)
    arr = (Session * max(1, len(sessions)))()
    keep = []
    for k, (t, seed, rbo, n, ptr) in enumerate(sessions):
        sb = bytes(seed)
        keep.append(sb)
        arr[k].topo = t.h
        arr[k].session_seed = sb
        arr[k].seed_len = len(sb)
        arr[k].region_block_offset = rbo
        arr[k].n_super = n
        arr[k].out_dev = ptr
    ms = C.c_double()
    ln = Launch(slots, waves_per_simd)
    rc = lib.pose_gpu_label_batch(arr, len(sessions), C.byref(ln), C.byref(ms))
    if rc != 0:
        raise RuntimeError("pose_gpu_label_batch failed")
    return ms.value
