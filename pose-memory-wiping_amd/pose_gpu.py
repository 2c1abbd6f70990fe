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
LANES_AUTO = 0xFFFFFFFF   # lanes_per_node: four-lane latency form for small batches
SPLIT_AUTO, SPLIT_OFF, SPLIT_ON = 0, 1, 2   # workgroups per super-chunk: auto, one, one per copy


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
    _fields_ = [("slots", C.c_uint32), ("waves_per_simd", C.c_uint32), ("lanes_per_node", C.c_uint32),
                ("split", C.c_uint32)]


class TopoPart(C.Structure):
    _fields_ = [("first_node", C.c_uint32), ("end_node", C.c_uint32), ("first_rank", C.c_uint32),
                ("end_rank", C.c_uint32), ("steps", C.c_uint32), ("pool_slots", C.c_uint32)]


class PlanInfo(C.Structure):
    _fields_ = [("super_chunks", C.c_uint32), ("slots", C.c_uint32), ("lanes", C.c_uint32),
                ("lanes_per_node", C.c_uint32), ("uses_graph", C.c_int32), ("scratch_bytes", C.c_uint64), ("nodes", C.c_uint64),
                ("output_bytes", C.c_uint64), ("groups", C.c_uint32)]


class PlanGroup(C.Structure):
    _fields_ = [("super_chunks", C.c_uint32), ("slots", C.c_uint32), ("lanes", C.c_uint32),
                ("lanes_per_node", C.c_uint32), ("max_steps", C.c_uint64), ("scratch_bytes", C.c_uint64),
                ("jobs", C.c_uint32), ("split", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("pinned_bytes", C.c_uint64),
                ("max_inflight", C.c_uint32), ("max_super_per_launch", C.c_uint32),
                ("waves_per_simd", C.c_uint32), ("sink_fd", C.c_int32),
                ("preload_mask", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("session_seed", C.c_char_p), ("seed_len", C.c_size_t),
                ("chunk_blocks", C.c_uint32), ("region_block_offset", C.c_uint64),
                ("n_super", C.c_uint64), ("sink_offset", C.c_uint64), ("host_out", C.c_void_p)]


class ServerStats(C.Structure):
    _fields_ = [("requests_done", C.c_uint64), ("super_chunks_done", C.c_uint64),
                ("bytes_done", C.c_uint64), ("elapsed_s", C.c_double), ("mib_per_s", C.c_double),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double),
                ("mean_ms", C.c_double), ("peak_inflight", C.c_uint32),
                ("peak_hbm_bytes", C.c_uint64), ("launches", C.c_uint64),
                ("kernel_ms_total", C.c_double), ("copy_ms_total", C.c_double),
                ("sink_ms_total", C.c_double), ("memory_waits", C.c_uint64),
                ("memory_cut_launches", C.c_uint64)]


class Challenge(C.Structure):
    _fields_ = [("chunk_seed", C.c_uint8 * 32), ("chunk_blocks", C.c_uint32), ("rank", C.c_uint32)]


class Tuned(C.Structure):
    _fields_ = [("chunk_blocks", C.c_uint32), ("lanes", C.c_uint32), ("mode", C.c_int32),
                ("lanes_per_node", C.c_uint32), ("waves_per_simd", C.c_uint32), ("slots", C.c_uint32), ("steps", C.c_uint32),
                ("trial_super", C.c_uint32), ("trial_kernel_ms", C.c_double),
                ("batch_mib_per_s", C.c_double), ("split", C.c_uint32)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `make -C pose-memory-wiping_amd/csrc`")
    # One HIP runtime per process.  torch carries its own copy of the HIP
    # runtime (torch/lib), which it loads by path; a process that loads the
    # system runtime first (through this library) and torch's later holds two,
    # and the library's device calls then find no device.  With torch loaded
    # first, the library's libamdhip64.so.7 resolves to torch's copy.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
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
    L.pose_gpu_topo_export_near.argtypes = [C.c_void_p, C.c_void_p]
    L.pose_gpu_topo_export_near.restype = C.c_int
    L.pose_gpu_topo_part.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(TopoPart), C.c_void_p, C.c_void_p,
                                     C.c_void_p]
    L.pose_gpu_topo_part.restype = C.c_int
    L.pose_gpu_label_batch.argtypes = [C.POINTER(Session), C.c_size_t, C.POINTER(Launch),
                                       C.POINTER(C.c_double)]
    L.pose_gpu_label_batch.restype = C.c_int
    L.pose_gpu_plan_create.argtypes = [C.POINTER(Session), C.c_size_t, C.POINTER(Launch), C.c_int]
    L.pose_gpu_plan_create.restype = C.c_void_p
    L.pose_gpu_plan_create_groups.argtypes = [C.POINTER(Session), C.c_size_t, C.POINTER(Launch), C.c_int]
    L.pose_gpu_plan_create_groups.restype = C.c_void_p
    L.pose_gpu_plan_group.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(PlanGroup)]
    L.pose_gpu_plan_group.restype = C.c_int
    L.pose_gpu_plan_run.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.pose_gpu_plan_run.restype = C.c_int
    L.pose_gpu_plan_info.argtypes = [C.c_void_p, C.POINTER(PlanInfo)]
    L.pose_gpu_plan_info.restype = C.c_int
    L.pose_gpu_plan_destroy.argtypes = [C.c_void_p]
    L.pose_gpu_plan_destroy.restype = None
    L.pose_gpu_server_create.argtypes = [C.POINTER(ServerConfig)]
    L.pose_gpu_server_create.restype = C.c_void_p
    L.pose_gpu_server_submit.argtypes = [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]
    L.pose_gpu_server_submit.restype = C.c_int
    L.pose_gpu_server_wait.argtypes = [C.c_void_p, C.c_uint64]
    L.pose_gpu_server_wait.restype = C.c_int
    L.pose_gpu_server_drain.argtypes = [C.c_void_p]
    L.pose_gpu_server_drain.restype = C.c_int
    L.pose_gpu_server_stats.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
    L.pose_gpu_server_stats.restype = C.c_int
    L.pose_gpu_server_destroy.argtypes = [C.c_void_p]
    L.pose_gpu_server_destroy.restype = None
    L.pose_gpu_challenge_batch.argtypes = [C.POINTER(Challenge), C.c_size_t, C.c_char_p,
                                           C.POINTER(C.c_double)]
    L.pose_gpu_challenge_batch.restype = C.c_int
    L.pose_gpu_challenge_release.argtypes = []
    L.pose_gpu_challenge_release.restype = None
    L.pose_gpu_chunk_seed.argtypes = [C.c_char_p, C.c_size_t, C.c_uint64, C.c_char_p]
    L.pose_gpu_chunk_seed.restype = C.c_int
    L.pose_gpu_autotune.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32, C.POINTER(Tuned),
                                    C.POINTER(Tuned), C.c_size_t, C.POINTER(C.c_size_t)]
    L.pose_gpu_autotune.restype = C.c_int
    return L


lib = _load()

TUNE_LEVELS = 1
TUNE_FORCE = 2


def chunk_seed(session_seed, chunk_index):
    """32-byte chunk seed of super-chunk `chunk_index` (pose_gpu_chunk_seed)."""
    out = C.create_string_buffer(32)
    sb = bytes(session_seed)
    if lib.pose_gpu_chunk_seed(sb, len(sb), chunk_index, out) != 0:
        raise ValueError("pose_gpu_chunk_seed failed")
    return out.raw


def challenge_batch(challenges, with_time=False):
    """challenges: iterable of (chunk_seed 32 B, chunk_blocks, rank).  Returns the
    32-byte labels in order (pose_gpu_challenge_batch), and the kernel time in
    milliseconds if with_time."""
    challenges = list(challenges)
    arr = (Challenge * max(1, len(challenges)))()
    for k, (seed, cb, rank) in enumerate(challenges):
        seed = bytes(seed)
        if len(seed) != 32:
            raise ValueError("chunk seeds are 32 bytes")
        C.memmove(arr[k].chunk_seed, seed, 32)
        arr[k].chunk_blocks = cb
        arr[k].rank = rank
    out = C.create_string_buffer(32 * max(1, len(challenges)))
    ms = C.c_double()
    if lib.pose_gpu_challenge_batch(arr, len(challenges), out, C.byref(ms)) != 0:
        raise RuntimeError("pose_gpu_challenge_batch failed")
    labels = [out.raw[32 * k:32 * (k + 1)] for k in range(len(challenges))]
    return (labels, ms.value) if with_time else labels


def autotune(chunk_blocks, n_super, levels=False, force=False):
    """Launch shape for n_super super-chunks of chunk_blocks (pose_gpu_autotune).
    Returns (best, trials) as dicts."""
    best = Tuned()
    trials = (Tuned * 32)()
    nt = C.c_size_t()
    flags = (TUNE_LEVELS if levels else 0) | (TUNE_FORCE if force else 0)
    if lib.pose_gpu_autotune(chunk_blocks, n_super, flags, C.byref(best), trials, 32, C.byref(nt)) != 0:
        raise RuntimeError("pose_gpu_autotune failed")
    as_dict = lambda t: {f: getattr(t, f) for f, _ in Tuned._fields_}
    return as_dict(best), [as_dict(trials[i]) for i in range(min(nt.value, 32))]


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

    def device_bytes(self):
        """Topology bytes resident on the device now (the copy schedules are
        uploaded by the first launch that labels a copy on its own)."""
        i = TopoInfo()
        lib.pose_gpu_topo_info(self.h, C.byref(i))
        return i.device_bytes

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

    def export_near(self):
        """Near table, (steps+1, lanes) words: per predecessor (bits 0-15, 16-31)
        the lane of the previous step that wrote it, or 0xFFFF."""
        import numpy as np
        i = self.info
        near = np.zeros((i.steps + 1) * i.lanes, np.uint32)
        assert lib.pose_gpu_topo_export_near(self.h, near.ctypes.data) == 0
        return near.reshape(i.steps + 1, i.lanes)

    def part(self, c):
        """Schedule of copy c (0 or 1) of the scaffold (pose_gpu_topo_part):
        (info dict, step_start, rec (steps+1, lanes, 4), near (steps+1, lanes)),
        or None for a topology without copy schedules."""
        import numpy as np
        info = TopoPart()
        if lib.pose_gpu_topo_part(self.h, c, C.byref(info), None, None, None) != 0:
            return None
        n = info.steps + 1
        lanes = self.info.lanes
        steps = np.zeros(n, np.uint32)
        rec = np.zeros(4 * n * lanes, np.uint32)
        near = np.zeros(n * lanes, np.uint32)
        assert lib.pose_gpu_topo_part(self.h, c, None, steps.ctypes.data, rec.ctypes.data, near.ctypes.data) == 0
        return ({f: getattr(info, f) for f, _ in TopoPart._fields_}, steps, rec.reshape(n, lanes, 4),
                near.reshape(n, lanes))

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


def _sessions(sessions):
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
    return arr, len(sessions), keep


class Plan:
    """A prepared batch (pose_gpu_plan_*): one persistent launch per launch
    group, optionally captured in one hipGraph and replayed by run().

    ``launches`` (optional) gives each session its own shape as
    (slots, waves_per_simd, lanes_per_node[, split]) tuples, e.g. the
    autotuner's choice for its chunk_blocks (pose_gpu_plan_create_groups);
    otherwise every session uses the shape given by the keyword arguments
    (pose_gpu_plan_create)."""

    def __init__(self, sessions, slots=0, waves_per_simd=0, graph=True, lanes_per_node=LANES_AUTO,
                 launches=None, split=SPLIT_AUTO):
        arr, n, self._keep = _sessions(sessions)
        if launches is not None:
            launches = list(launches)
            if len(launches) != n:
                raise ValueError("one launch shape per session")
            la = (Launch * max(1, n))(*[Launch(*l) for l in launches])
            self.h = lib.pose_gpu_plan_create_groups(arr, n, la, int(bool(graph)))
        else:
            ln = Launch(slots, waves_per_simd, lanes_per_node, split)
            self.h = lib.pose_gpu_plan_create(arr, n, C.byref(ln), int(bool(graph)))
        if not self.h:
            raise RuntimeError("pose_gpu_plan_create failed")
        self.info = PlanInfo()
        lib.pose_gpu_plan_info(self.h, C.byref(self.info))

    def groups(self):
        out = []
        for g in range(self.info.groups):
            pg_ = PlanGroup()
            assert lib.pose_gpu_plan_group(self.h, g, C.byref(pg_)) == 0
            out.append({f: getattr(pg_, f) for f, _ in PlanGroup._fields_})
        return out

    def run(self):
        ms = C.c_double()
        if lib.pose_gpu_plan_run(self.h, C.byref(ms)) != 0:
            raise RuntimeError("pose_gpu_plan_run failed")
        return ms.value

    def close(self):
        if self.h:
            lib.pose_gpu_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching label server (pose_gpu_server_*)."""

    def __init__(self, hbm_budget_bytes=0, pinned_bytes=0, max_inflight=0, max_super_per_launch=0,
                 waves_per_simd=0, sink_fd=-1, preload=()):
        mask = 0
        for cb in preload:
            mask |= 1 << (int(cb).bit_length() - 1)
        cfg = ServerConfig(hbm_budget_bytes, pinned_bytes, max_inflight, max_super_per_launch,
                           waves_per_simd, sink_fd, mask)
        self.h = lib.pose_gpu_server_create(C.byref(cfg))
        if not self.h:
            raise RuntimeError("pose_gpu_server_create failed")
        self._keep = {}

    def submit(self, seed, chunk_blocks, region_block_offset, n_super, sink_offset=0, host_out=None):
        """host_out: optional writable buffer (bytearray / numpy array) for a copy of the labels."""
        sb = bytes(seed)
        out_ptr = None
        if host_out is not None:
            out_ptr = C.addressof((C.c_char * len(host_out)).from_buffer(host_out))
        req = Request(sb, len(sb), chunk_blocks, region_block_offset, n_super, sink_offset, out_ptr)
        rid = C.c_uint64()
        if lib.pose_gpu_server_submit(self.h, C.byref(req), C.byref(rid)) != 0:
            raise RuntimeError("pose_gpu_server_submit failed")
        self._keep[rid.value] = (sb, host_out)
        return rid.value

    def wait(self, rid):
        rc = lib.pose_gpu_server_wait(self.h, rid)
        self._keep.pop(rid, None)
        return rc

    def drain(self):
        return lib.pose_gpu_server_drain(self.h)

    def stats(self):
        st = ServerStats()
        lib.pose_gpu_server_stats(self.h, C.byref(st))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib.pose_gpu_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def label_batch(sessions, slots=0, waves_per_simd=0, lanes_per_node=LANES_AUTO, split=SPLIT_AUTO):
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
    ln = Launch(slots, waves_per_simd, lanes_per_node, split)
    rc = lib.pose_gpu_label_batch(arr, len(sessions), C.byref(ln), C.byref(ms))
    if rc != 0:
        raise RuntimeError("pose_gpu_label_batch failed")
    return ms.value
