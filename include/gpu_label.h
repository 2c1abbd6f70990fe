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

/* pose_hash_algo_t comes from the reference's hash.h, which this header
 * includes itself when it is on the include path, as the reference's
 * gpu_label.h does (common/include/gpu_label.h:7).  A caller may then include
 * graph.h and hash.h before or after this header, in any order
 * (tools/disk-wipe-bench/disk_wipe_bench.cu:55-57 includes gpu_label.h first).
 * Without hash.h the same enum, with the same values
 * (common/include/hash.h:31-35), is declared here. */
#if defined(__has_include)
#if __has_include("hash.h")
#include "hash.h"
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif

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

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the labeling of pose_graph_hbm_topo_label, with the same
 * arguments, labels, return codes and stdout lines, but no device allocation
 * or free.  Its scratch (work counter, job records, slot regions) is one fixed
 * device array of POSE_GPU_NOALLOC_SCRATCH_BYTES per GPU, held by the library
 * until the process ends; the slot regions it holds bound how many
 * super-chunks label at once (at chunk_blocks 2^17: 18 whole super-chunks,
 * or 42 copy jobs when super-chunks are split into their two copies).
 * Every call on a GPU shares that GPU's array, so calls that overlap on one
 * GPU race: serialise them.  Returns -1 if one slot region of the topology's
 * chunk_blocks does not fit the array, or if the current device is not the
 * topology's. */
#define POSE_GPU_NOALLOC_SCRATCH_BYTES (256ull << 20)
int pose_graph_hbm_topo_label_noalloc(const pose_hbm_topo_t *topo,
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

/* Copy the near table: (steps+1)*lanes words, one per record position.  Bits
 * 0-15 (pred0) and 16-31 (pred1) hold the lane of the step just before that
 * wrote the predecessor, or 0xFFFF when it was written earlier or is absent.
 * The four-lanes-per-node kernel takes such predecessors from its on-chip copy
 * of the previous step.  No counterpart in the reference (schedule detail). */
int pose_gpu_topo_export_near(const pose_hbm_topo_t *topo, uint32_t *near);

/* The scaffold of a super-chunk is two copies of one standalone graph with no
 * edge between them: copy 0 holds node ids [0, nodes/2) and output ranks
 * [0, 2^n), copy 1 the rest.  Topologies of 64 lanes also carry a schedule of
 * each copy on its own, so that two workgroups can label one super-chunk (see
 * split in pose_gpu_launch_t): one copy's critical path is the whole
 * scaffold's, with half the nodes.  pose_gpu_topo_part describes copy `part`
 * (0 or 1) and, like pose_gpu_topo_export / _export_near, copies its
 * schedule: step_start receives steps+1 offsets, rec 4*(steps+1)*lanes words
 * (node ids are scaffold ids, outputs carry their absolute rank, scratch slots
 * are numbered within the copy's own pool), near (steps+1)*lanes words.  Any
 * pointer may be NULL.  Returns -1 for a topology without copy schedules.  No
 * counterpart in the reference, which labels the scaffold as one graph. */
typedef struct {
    uint32_t first_node, end_node;   /* node ids [first_node, end_node) */
    uint32_t first_rank, end_rank;   /* output ranks [first_rank, end_rank) */
    uint32_t steps;
    uint32_t pool_slots;
} pose_gpu_topo_part_t;

int pose_gpu_topo_part(const pose_hbm_topo_t *topo, uint32_t part, pose_gpu_topo_part_t *info,
                       uint32_t *step_start, uint32_t *rec, uint32_t *near);

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

/* Launch parameters chosen by the autotuner; zero fields mean "auto".
 * lanes_per_node: 1 = one lane per node hash (throughput form); 4 = a quad of
 * lanes per node, state columns exchanged with DPP and predecessor words
 * staged in LDS (latency form, for batches too small to fill the device;
 * topologies of at most 256 lanes). */
#define POSE_GPU_LANES_AUTO 0xFFFFFFFFu   /* lanes_per_node: pick by batch size */
/* split: how many workgroups label one super-chunk.  AUTO: two (one per copy
 * of the scaffold, see pose_gpu_topo_part) when the four-lane form is used,
 * every topology has copy schedules, the plan holds at most one super-chunk
 * per two CUs and slots is 0 or at least the super-chunk count (each copy then
 * gets a workgroup of its own); else one.  OFF: one.  ON: two, in either form
 * (slots as given). */
#define POSE_GPU_SPLIT_AUTO 0u
#define POSE_GPU_SPLIT_OFF  1u
#define POSE_GPU_SPLIT_ON   2u
typedef struct {
    uint32_t slots;           /* resident workgroups (0 = from occupancy) */
    uint32_t waves_per_simd;  /* target occupancy used when slots == 0 */
    uint32_t lanes_per_node;  /* 1 (0 also means 1), 4, or POSE_GPU_LANES_AUTO: 4 when
                                 the batch has at most 3 super-chunks per CU of its
                                 deepest chunk_blocks and the topologies have 64 lanes */
    uint32_t split;           /* POSE_GPU_SPLIT_AUTO (0), _OFF or _ON */
} pose_gpu_launch_t;

/* Label every super-chunk of every session in one persistent launch per
 * topology width (longest-first work queue; see the plan below for how
 * sessions of different widths are grouped).  Synchronous; scratch is zeroed
 * before return.
 * launch NULL means automatic shape (lanes_per_node POSE_GPU_LANES_AUTO).
 * kernel_ms (optional) receives the kernel time in milliseconds.
 * No counterpart in the reference; additive to its ABI. */
int pose_gpu_label_batch(const pose_gpu_session_t *sessions, size_t n_sessions,
                         const pose_gpu_launch_t *launch, double *kernel_ms);

/* A prepared batch (same sessions as pose_gpu_label_batch): chunk seeds, the
 * job table, per-slot scratch and the work counters are set up once; each run
 * re-labels every super-chunk.  Sessions are split into launch groups by kernel
 * shape (topology lanes, lanes per node, occupancy request): each group is one
 * persistent launch, and the groups of a run execute concurrently (forked
 * streams, joined before the run returns).  With use_graph != 0 the whole
 * sequence is captured once into one hipGraph (the groups as parallel
 * branches) and each run replays it.  Runs are synchronous; destroy zeroes the
 * scratch before freeing it.  Output buffers must stay valid for the plan's
 * lifetime.  Scratch: each group's slots halve until all scratch fits free HBM
 * (less 256 MiB) and, when the environment variable POSE_GPU_SCRATCH_LIMIT is
 * set, that many bytes; one slot per group is tried even if it does not fit,
 * as the reference's pool sizing does (common/src/cuda/label_gpu_warp.cu:163-175).
 * No counterpart in the reference. */
typedef struct pose_gpu_plan pose_gpu_plan_t;

typedef struct {
    uint32_t super_chunks;
    uint32_t slots;           /* resident workgroups, summed over the groups */
    uint32_t lanes;           /* of group 0 (the one with the deepest graph) */
    uint32_t lanes_per_node;  /* of group 0 */
    int32_t  uses_graph;
    uint64_t scratch_bytes;
    uint64_t nodes;           /* scaffold nodes labeled per run */
    uint64_t output_bytes;    /* persisted label bytes per run */
    uint32_t groups;          /* persistent launches per run */
} pose_gpu_plan_info_t;

typedef struct {
    uint32_t super_chunks;
    uint32_t slots;
    uint32_t lanes;
    uint32_t lanes_per_node;
    uint64_t max_steps;       /* schedule steps of the group's deepest job */
    uint64_t scratch_bytes;
    uint32_t jobs;            /* workgroup jobs: super_chunks, or twice that when split */
    uint32_t split;           /* 1 if each super-chunk's copies are labeled separately */
} pose_gpu_plan_group_t;

/* One launch shape for every session (NULL: automatic). */
pose_gpu_plan_t *pose_gpu_plan_create(const pose_gpu_session_t *sessions, size_t n_sessions,
                                      const pose_gpu_launch_t *launch, int use_graph);
/* launches[i] is session i's launch shape, e.g. the autotuner's choice for its
 * chunk_blocks; sessions of equal shape share one persistent launch. */
pose_gpu_plan_t *pose_gpu_plan_create_groups(const pose_gpu_session_t *sessions, size_t n_sessions,
                                             const pose_gpu_launch_t *launches, int use_graph);
int  pose_gpu_plan_run(pose_gpu_plan_t *plan, double *kernel_ms);
int  pose_gpu_plan_info(const pose_gpu_plan_t *plan, pose_gpu_plan_info_t *info);
/* Shape of launch group g (0 <= g < info.groups). */
int  pose_gpu_plan_group(const pose_gpu_plan_t *plan, uint32_t g, pose_gpu_plan_group_t *out);
void pose_gpu_plan_destroy(pose_gpu_plan_t *plan);

/* ── additive: challenge answers (verifier side) ──────────────────────── */

/* One challenge: the label at output rank `rank` of the super-chunk graph of
 * `chunk_blocks` blocks keyed by `chunk_seed` — the value the reference's CPU
 * pose_graph_challenge(chunk_seed, 32, chunk_blocks, rank, out,
 * POSE_HASH_BLAKE3) returns (common/src/graph.c:1019-1042).  Absolute block b
 * of a session written with this chunk_blocks is challenge
 * {pose_gpu_chunk_seed(session, b / chunk_blocks), chunk_blocks, b % chunk_blocks}. */
typedef struct {
    uint8_t  chunk_seed[32];
    uint32_t chunk_blocks;    /* 1 .. 2^20 */
    uint32_t rank;            /* 0 .. chunk_blocks-1 */
} pose_gpu_challenge_t;

/* Answers n challenges into out (n * 32 bytes, host memory).  Challenges on
 * the same (chunk_seed, chunk_blocks) share one labeling; distinct super-chunks
 * are labeled in batched launches (at most 8192 super-chunks / 16 GiB of
 * labels each).  Topologies are built on first use per device and chunk_blocks
 * and kept until pose_gpu_challenge_release.  Synchronous; label buffers are
 * zeroed before they are freed.  kernel_ms (optional) receives the labeling
 * kernel time.  Returns 0, or -1 on a bad challenge or a device error.  No
 * counterpart in the reference's GPU library (its CPU pose_graph_challenge
 * answers one at a time). */
int pose_gpu_challenge_batch(const pose_gpu_challenge_t *challenges, size_t n, uint8_t *out,
                             double *kernel_ms);

/* Frees the topologies pose_gpu_challenge_batch keeps for the current device
 * (later calls build them again).  A topology is seed-independent graph
 * structure, so it holds nothing to scrub; the labels of every call were
 * zeroed when that call returned.  Call when no pose_gpu_challenge_batch call
 * is running on this device.  No counterpart in the reference. */
void pose_gpu_challenge_release(void);

/* Host helper: the 32-byte chunk seed derive_key("pose-db/chunk-seed",
 * session_seed || be64(chunk_index)), as the reference's pose_chunk_seed
 * (common/src/hash.c:218-233).  Returns 0, or -1 on NULL arguments. */
int pose_gpu_chunk_seed(const uint8_t *session_seed, size_t seed_len, uint64_t chunk_index,
                        uint8_t out[32]);

/* ── additive: launch autotuner ───────────────────────────────────────── */

/* Chooses the launch shape for labeling n_super super-chunks of chunk_blocks:
 * nodes per step (64/128/256), lanes per node (1 or 4), split (four-lane form
 * on 64-lane topologies: whole super-chunks or one workgroup per scaffold
 * copy), step schedule (list; levels too with POSE_GPU_TUNE_LEVELS) and
 * occupancy (1/2/4/8 waves per SIMD), by timing one trial round of each on the
 * current device; every trial runs once untimed first.  The winner maximises
 * n_super*cb*32 / (rounds x round time).  Cached per (device, chunk_blocks,
 * batch rounded up to a power of two) unless POSE_GPU_TUNE_FORCE.  trials
 * (optional) receives every measured shape.  No counterpart in the reference,
 * whose strategy is fixed at build time (common/CMakeLists.txt:152-171). */
#define POSE_GPU_TUNE_LEVELS 1u
#define POSE_GPU_TUNE_FORCE  2u

typedef struct {
    uint32_t chunk_blocks;
    uint32_t lanes;
    int32_t  mode;
    uint32_t lanes_per_node;
    uint32_t waves_per_simd;
    uint32_t slots;
    uint32_t steps;
    uint32_t trial_super;
    double   trial_kernel_ms;
    double   batch_mib_per_s;
    uint32_t split;           /* POSE_GPU_SPLIT_OFF or _ON, for pose_gpu_launch_t.split */
} pose_gpu_tuned_t;

int pose_gpu_autotune(uint32_t chunk_blocks, uint64_t n_super, uint32_t flags,
                      pose_gpu_tuned_t *best, pose_gpu_tuned_t *trials, size_t max_trials,
                      size_t *n_trials);

/* ── additive: continuous-batching label server ────────────────────────── */

/* Label requests are queued and packed into persistent launches on up to four
 * streams; a request that arrives while others run is admitted as soon as a
 * stream, device arena space (scratch + output staging) and pinned host space
 * are free, and runs concurrently with them.  Finished super-chunks are copied
 * through pinned host memory to the sink: sink_fd < 0 is a dry run (labels
 * are dropped after the copy), otherwise each super-chunk is pwrite()n at
 * sink_offset + j*cb*32 (the fd may be O_DIRECT: buffers and sizes are 4 KiB
 * aligned).  Topologies are built once per chunk_blocks and stay resident.
 * Latency of a super-chunk = submit -> handed to the sink.  No counterpart in
 * the reference; it replaces the serial label -> D2H -> pwrite loop of its
 * producer (tools/disk-wipe-bench/disk_wipe_bench.cu:646-697). */
typedef struct pose_gpu_server pose_gpu_server_t;

typedef struct {
    uint64_t hbm_budget_bytes;     /* device arena (0 = a third of free HBM) */
    uint64_t pinned_bytes;         /* pinned host staging (0 = 1 GiB) */
    uint32_t max_inflight;         /* concurrent launches, 1..4 (0 = 4) */
    uint32_t max_super_per_launch; /* 0 = four per CU */
    uint32_t waves_per_simd;       /* per-launch occupancy target (0 = 2) */
    int32_t  sink_fd;              /* < 0: dry-run sink */
    uint32_t preload_mask;         /* bit k set: build the 2^k topology up front */
} pose_gpu_server_config_t;

typedef struct {
    const uint8_t *session_seed;
    size_t         seed_len;
    uint32_t       chunk_blocks;        /* power of two, 0 = 4096 */
    uint64_t       region_block_offset; /* absolute block of super-chunk 0 */
    uint64_t       n_super;
    uint64_t       sink_offset;         /* sink byte offset of super-chunk 0 */
    uint8_t       *host_out;            /* optional host copy (n_super*cb*32 B) */
} pose_gpu_request_t;

typedef struct {
    uint64_t requests_done;
    uint64_t super_chunks_done;
    uint64_t bytes_done;
    double   elapsed_s;           /* first submit -> last super-chunk in the sink */
    double   mib_per_s;
    double   p50_ms, p99_ms, max_ms, mean_ms;   /* per super-chunk latency */
    uint32_t peak_inflight;
    uint64_t peak_hbm_bytes;
    uint64_t launches;
    double   kernel_ms_total;
    double   copy_ms_total;       /* device -> pinned copies (GPU time) */
    double   sink_ms_total;       /* writer thread: host copies and sink writes */
    /* Admission by footprint: requests that found a free stream but no arena or
     * pinned space for even one super-chunk (counted once per request), and
     * launches admitted with fewer super-chunks than were pending (up to
     * max_super_per_launch) because the arena or the pinned ring was full. */
    uint64_t memory_waits;
    uint64_t memory_cut_launches;
} pose_gpu_server_stats_t;

pose_gpu_server_t *pose_gpu_server_create(const pose_gpu_server_config_t *cfg);
int  pose_gpu_server_submit(pose_gpu_server_t *srv, const pose_gpu_request_t *req, uint64_t *id);
/* wait (one request) and drain (all) return 0 when complete, -1 when failed;
 * either way only once no launch can still write the request's host_out or
 * sink range, so the caller may release them on return. */
int  pose_gpu_server_wait(pose_gpu_server_t *srv, uint64_t id);
int  pose_gpu_server_drain(pose_gpu_server_t *srv);
int  pose_gpu_server_stats(pose_gpu_server_t *srv, pose_gpu_server_stats_t *stats);
void pose_gpu_server_destroy(pose_gpu_server_t *srv);   /* drains, zeroes the arena, frees */

#ifdef __cplusplus
}
#endif

#endif /* POSE_MI355X_GPU_LABEL_H */
