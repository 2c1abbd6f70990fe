// labeler_internal.h — shared between the labeler (labeler.hip) and the
// serving engine (server.hip): device tables, the topology handle and the
// launch helpers.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <atomic>
#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/gpu_label.h"
#include "topology.h"

namespace pose {
namespace internal {

// Device form of the step records: every location is a byte offset in the
// labeling workgroup's slot region, [scratch pool | one super-chunk of
// labels]: scratch slot s at 32*s, output rank r at out_off + 32*r.  The
// kernels copy the finished labels from out_off to the caller's buffer.
struct DevTopo {                 // one entry per topology in a launch
    const uint4* rec;            // (n_steps + 1) * lanes records, device locations
    const uint4* cv0;            // (n_steps + 1) * lanes * 2
    uint32_t n_steps;
    uint32_t pool;
    uint32_t desc[8];
    const uint32_t* near;        // (n_steps + 1) * lanes near-table entries
    uint32_t out_off;            // byte offset of the labels in the slot region
    uint32_t out16;              // super-chunk label bytes / 16
    uint64_t pad;
};
static_assert(sizeof(DevTopo) == 80, "DevTopo layout");

// Slot region layout of a topology: labels start after the scratch pool.
inline uint32_t slot_out_off(uint32_t pool) { return (uint32_t)(((uint64_t)pool * 32 + 255) / 256 * 256); }

struct DevJob {                  // one entry per super-chunk (or per copy of one)
    uint8_t* persist;            // output of this job: rank r at persist + 32*(r - first rank)
    uint32_t topo;               // index into the DevTopo table
    uint32_t part;               // host side, before plan_build: kWhole or kCopy0 + p
    uint32_t seed[8];            // chunk seed words
};
static_assert(sizeof(DevJob) == 48, "DevJob layout");

}  // namespace internal
}  // namespace pose

struct pose_hbm_topo {
    pose::HostTopology host;     // host schedule; rec and near freed after upload unless host-only
    bool on_device = false;
    int device = -1;
    uint4* d_rec = nullptr;
    uint4* d_cv0 = nullptr;
    uint32_t* d_near = nullptr;
    // Schedules of the two copies (host.part[0], host.part[1]), rows of copy 0
    // then copy 1, when host.has_parts.  They are uploaded by the first launch
    // that labels a copy on its own (pose::internal::ensure_parts); until then
    // their records stay on the host and the device holds only the whole
    // scaffold's schedule.
    mutable std::mutex parts_mu;
    mutable std::atomic<bool> parts_on_device{false};
    mutable uint4* d_prec = nullptr;
    mutable uint4* d_pcv0 = nullptr;
    mutable uint32_t* d_pnear = nullptr;
    mutable std::atomic<uint64_t> device_bytes{0};
};


namespace pose {
namespace internal {

int lanes_to_waves(uint32_t lanes);
// waves = topology lanes / 64; quad selects the four-lanes-per-node kernel.
int max_blocks_per_cu(int waves, bool quad = false);

// Automatic form choice: a launch whose deepest chunk_blocks has at most this
// many super-chunks per CU takes the four-lane form.  Measured at 2^17 (64
// lanes): 512 super-chunks run at 2,054 MiB/s with four lanes against 1,271
// with one, 768 at 2,175 against 1,875, 1,024 at 2,229 against 2,452.
constexpr uint32_t kQuadMaxPerCu = 3;
// pipe_ok: the launch may take the software-pipelined one-lane kernel when its
// occupancy is low.  That kernel needs more VGPRs per wave (at most 5 waves per
// SIMD), so callers that run several launches concurrently pass false.
hipError_t launch_label(int waves, uint32_t grid, const DevTopo* topos, const DevJob* jobs,
                        uint32_t n_jobs, uint32_t* counter, uint8_t* scratch, uint64_t slot_bytes,
                        hipStream_t st, bool quad = false, bool pipe_ok = true);
// Zero a device range with a kernel (no memset nodes, so this is graph-safe).
hipError_t zero_device(void* p, uint64_t bytes, hipStream_t st);
// Device table entry of the whole scaffold (part kWhole) or of copy p
// (part kCopy0 + p).
constexpr uint32_t kWhole = 0, kCopy0 = 1;
void fill_dev_topo(const pose_hbm_topo* t, DevTopo* D, uint32_t part = kWhole);
// Uploads the copy schedules of a 64-lane topology if they are not on the
// device yet (thread-safe; allocates device memory).  Returns hipSuccess, or
// the error of the upload.
hipError_t ensure_parts(const pose_hbm_topo* t);
// Bytes of one workgroup's slot region for this topology (or one copy of it).
uint64_t slot_region_bytes(const pose_hbm_topo* t, uint32_t part = kWhole);
// Schedule steps of the whole scaffold or of one copy.
uint32_t schedule_steps(const pose_hbm_topo* t, uint32_t part = kWhole);
// Output ranks [first, end) written by a job of this part.
uint32_t part_first_rank(const pose_hbm_topo* t, uint32_t part);
uint32_t part_end_rank(const pose_hbm_topo* t, uint32_t part);

// Requested shape of a launch group; jobs with equal requests share one.
struct LaunchReq {
    int waves;                   // topology lanes / 64
    uint32_t lpn;                // 1, 4 or POSE_GPU_LANES_AUTO
    uint32_t slots, wps;         // as in pose_gpu_launch_t
    uint32_t split = POSE_GPU_SPLIT_AUTO;   // as in pose_gpu_launch_t
    bool operator==(const LaunchReq& o) const {
        return waves == o.waves && lpn == o.lpn && slots == o.slots && wps == o.wps && split == o.split;
    }
};

// Jobs of one requested shape.  A job with part kWhole labels a whole
// super-chunk and may be split into its two copies by plan_build; a job with
// part kCopy0 + p labels copy p only.
struct PendingGroup {
    LaunchReq req;
    std::vector<DevJob> jobs;
};

// Builds a plan from prepared jobs.  tlist holds the topologies the jobs'
// `topo` fields index.  idle_waves / idle_quad shape the idle group of a plan
// without jobs.  Returns nullptr (after printing why) on failure.
pose_gpu_plan* plan_build(std::vector<PendingGroup>& pend, const std::vector<const pose_hbm_topo*>& tlist,
                          uint64_t nodes, uint64_t out_bytes, int use_graph, int idle_waves, bool idle_quad);

}  // namespace internal
}  // namespace pose
