/*
 * gpu_label.h — MI355X (gfx950) labeler for PoSE-DB Protocol 3, C ABI.
 *
 * Drop-in for the reference's GPU labeling library (common/include/gpu_label.h
 * and its CUDA implementation common/src/cuda/).  The five entry points of the
 * reference keep their names, signatures, ownership rules, return codes and
 * stdout/stderr behaviour; the rest are additive (batching, serving,
 * introspection).  Labels are byte-identical to the reference CPU in-place
 * labeler (common/src/graph.c:1604-1689) for the same session seed and
 * absolute position.
 *
 * Link: libpose_gpu.so (self-contained: it carries its own host BLAKE3 and
 * scaffold emitter, so it does not need the reference's libpose_common).
 * Device memory passed in must be allocated on the device current at the call.
 */
#ifndef POSE_MI355X_GPU_LABEL_H
#define POSE_MI355X_GPU_LABEL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same enum and values as the reference (common/include/hash.h:31-35).  When the
 * reference's hash.h is included first its definition is used. */
#ifndef POSE_HASH_H
typedef enum {
    POSE_HASH_BLAKE3  = 0,
    POSE_HASH_SHA256  = 1,
    POSE_HASH_AES_PRF = 2,
} pose_hash_algo_t;
#endif

/* ── reference entry points ─────────────────────────────────────────────── */

/* Replaces pose_graph_label_hbm (common/include/gpu_label.h:40-44;
 * common/src/cuda/label_gpu_common.cu:338-365): build -> label -> free.
 * chunk_blocks 0 means 4096.  Returns 0, or -1 (algo not BLAKE3, bad
 * chunk_blocks, device error).  A buffer smaller than one super-chunk labels
 * nothing and returns 0. */
int pose_graph_label_hbm(const uint8_t *session_seed, size_t seed_len,
                         uint32_t chunk_blocks,
                         void *hbm_dev_ptr, size_t hbm_len,
                         uint64_t region_block_offset,
                         pose_hash_algo_t algo);

typedef struct pose_hbm_topo pose_hbm_topo_t;

/* Replaces pose_graph_hbm_topo_build (common/include/gpu_label.h:74-75;
 * label_gpu_common.cu:72-254).  Builds the seed-independent scaffold schedule
 * for chunk_blocks on the host and uploads it to the CURRENT device.  Returns
 * NULL on error (message on stderr). */
pose_hbm_topo_t *pose_graph_hbm_topo_build(uint32_t chunk_blocks,
                                           pose_hash_algo_t algo);

/* Replaces pose_graph_hbm_topo_label (common/include/gpu_label.h:77-80;
 * label_gpu_common.cu:256-319).  Labels floor(hbm_len / (cb*32)) super-chunks of
 * hbm_dev_ptr in place; super-chunk j uses chunk index
 * (region_block_offset + j*cb) / cb.  Synchronous.  Prints the
 * "GPU wipe: X MiB in Y s = Z MiB/s" line unless POSE_GPU_QUIET is set.
 * Call from one thread per topology. */
int pose_graph_hbm_topo_label(const pose_hbm_topo_t *topo,
                              const uint8_t *session_seed, size_t seed_len,
                              void *hbm_dev_ptr, size_t hbm_len,
                              uint64_t region_block_offset);

/* Replaces pose_graph_hbm_topo_free (common/include/gpu_label.h:82;
 * label_gpu_common.cu:321-331). */
void pose_graph_hbm_topo_free(pose_hbm_topo_t *topo);

/* Replaces pose_gpu_label_node_single (common/include/gpu_label.h:92-97;
 * common/src/cuda/blake3_node.cu:33-73): one node label on the GPU for explicit
 * inputs.  pred0/pred1 may be NULL (zero blocks); a source node ignores them.
 * Returns 0 or -1. */
int pose_gpu_label_node_single(uint8_t out[32], int is_source,
                               uint64_t node_index,
                               const uint8_t *pred0, const uint8_t *pred1,
                               const uint8_t *seed, const uint8_t *descriptor,
                               const uint32_t key_src[8],
                               const uint32_t key_int[8]);

/* ── additive: topology options and introspection ───────────────────────── */

/* Schedule options for a topology: lanes = nodes per step (64 * waves per
 * workgroup: 64, 128, 256 or 512); mode 0 = dependency levels split into steps
 * (the reference's level order), 1 = list schedule by longest remaining path.
 * No counterpart in the reference (its strategy is a build-time choice,
 * common/CMakeLists.txt:152-171). */
pose_hbm_topo_t *pose_gpu_topo_build_ex(uint32_t chunk_blocks, uint32_t lanes,
                                        int mode);

typedef struct {
    uint32_t chunk_blocks;
    uint32_t lanes;
    int32_t  mode;
    uint32_t levels;          /* ASAP dependency levels */
    uint32_t steps;           /* scheduled steps */
    uint32_t pool_slots;      /* scratch labels per super-chunk in flight */
    uint64_t nodes;           /* scaffold nodes per super-chunk */
    uint64_t device_bytes;    /* resident topology bytes on the device */
    double   build_seconds;   /* host schedule build */
} pose_gpu_topo_info_t;

int pose_gpu_topo_info(const pose_hbm_topo_t *topo, pose_gpu_topo_info_t *info);

/* Copy the schedule (for CPU-side validation): step_start receives steps+1
 * offsets into the unpadded node order; rec receives 4*(steps+1)*lanes words,
 * one 16-byte record per lane per step: {dst, pred0, pred1 locations,
 * id<<2|np} with np == 3 for an idle lane (the last row is all idle); a
 * location with bit 31 set is an output rank, otherwise a scratch slot (slot 0
 * reads as zeros, slot 1 is discarded).  Either pointer may be NULL. */
int pose_gpu_topo_export(const pose_hbm_topo_t *topo, uint32_t *step_start,
                         uint32_t *rec);

/* Host-only schedule build (no device): same records as topo_build_ex, for
 * tests that run without a GPU.  Free with pose_gpu_topo_free_host. */
pose_hbm_topo_t *pose_gpu_topo_build_host(uint32_t chunk_blocks, uint32_t lanes,
                                          int mode);
void pose_gpu_topo_free_host(pose_hbm_topo_t *topo);

/* ── additive: batched labeling of many sessions in one launch ─────────── */

/* One session of a batch: n_super consecutive super-chunks of the session
 * starting at region_block_offset, written to out_dev (n_super * cb * 32 B,
 * device memory).  The topology fixes chunk_blocks. */
typedef struct {
    const pose_hbm_topo_t *topo;
    const uint8_t *session_seed;
    size_t         seed_len;
    uint64_t       region_block_offset;
    uint64_t       n_super;
    void          *out_dev;
} pose_gpu_session_t;

/* Launch parameters chosen by the autotuner; zero fields mean "auto". */
typedef struct {
    uint32_t slots;           /* resident workgroups (0 = from occupancy) */
    uint32_t waves_per_simd;  /* target occupancy used when slots == 0 */
} pose_gpu_launch_t;

/* Label every super-chunk of every session in ONE persistent launch
 * (longest-first work queue).  Synchronous; scratch is zeroed before return.
 * stats_ms (optional) receives the kernel time in milliseconds.
 * No counterpart in the reference; additive to its ABI. */
int pose_gpu_label_batch(const pose_gpu_session_t *sessions, size_t n_sessions,
                         const pose_gpu_launch_t *launch, double *kernel_ms);

#ifdef __cplusplus
}
#endif

#endif /* POSE_MI355X_GPU_LABEL_H */
