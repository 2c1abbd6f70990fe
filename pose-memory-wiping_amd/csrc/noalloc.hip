// noalloc.hip — pose_graph_hbm_topo_label_noalloc: the labeling of
// pose_graph_hbm_topo_label without any device allocation.
//
// Its scratch — the work counter, the device topology table, a window of job
// records and the workgroups' slot regions — lives in one file-scope device
// array of kScratchBytes per GPU, which the library holds until the process
// ends.  The array's address on each device comes from hipGetSymbolAddress
// once per device (each GPU has its own copy); host code uses that address,
// never the array's name.  Super-chunks are labeled in windows of at most
// kWindowJobs workgroup jobs, with as many resident slot regions as the array
// holds.  Every call on a GPU shares that GPU's array: calls that overlap on
// one GPU race, so callers serialise them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "../../include/gpu_label.h"
#include "b3_host.h"
#include "labeler_internal.h"

using namespace pose;
using namespace pose::internal;

namespace {

constexpr uint64_t kScratchBytes = POSE_GPU_NOALLOC_SCRATCH_BYTES;
constexpr uint32_t kWindowJobs = 1024;        // job records per launch
constexpr uint64_t kCounterOff = 0;           // work counter (256 B)
constexpr uint64_t kTopoOff = 256;            // whole scaffold + two copies
constexpr uint64_t kJobsOff = 1024;
constexpr uint64_t kSlotsOff = 64 * 1024;     // slot regions, 256-B aligned
constexpr int kMaxDevices = 64;
static_assert(kTopoOff + 3 * sizeof(DevTopo) <= kJobsOff, "topology table");
static_assert(kJobsOff + kWindowJobs * sizeof(DevJob) <= kSlotsOff, "job window");

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

std::mutex g_mu;
uint8_t* g_scratch[kMaxDevices] = {nullptr};

// Device address of `scratch` on device `dev`, looked up once per device with
// that device current.
uint8_t* scratch_on(int dev) {
    if (dev < 0 || dev >= kMaxDevices) {
        std::fprintf(stderr, "pose_graph_hbm_topo_label_noalloc: device %d is beyond the %d this library serves\n",
                     dev, kMaxDevices);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            std::fprintf(stderr, "pose_graph_hbm_topo_label_noalloc: hipGetSymbolAddress failed: %s\n",
                         hipGetErrorString(e));
            return nullptr;
        }
        g_scratch[dev] = static_cast<uint8_t*>(p);
    }
    return g_scratch[dev];
}

double now_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool quiet() {
    const char* e = std::getenv("POSE_GPU_QUIET");
    return e && e[0] && e[0] != '0';
}

}  // namespace

extern "C" int pose_graph_hbm_topo_label_noalloc(const pose_hbm_topo_t* t, const uint8_t* session_seed,
                                                 size_t seed_len, void* hbm_dev_ptr, size_t hbm_len,
                                                 uint64_t region_block_offset) {
    static const char* W = "pose_graph_hbm_topo_label_noalloc";
    if (!t || !t->on_device || (!session_seed && seed_len)) return -1;
    const uint32_t cb = t->host.cb;
    const uint64_t sb = (uint64_t)cb * 32;
    const uint64_t n_super = hbm_len / sb;
    if (n_super == 0) return 0;
    if (!hbm_dev_ptr) return -1;
    int dev = -1, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
        std::fprintf(stderr, "%s: no HIP device\n", W);
        return -1;
    }
    if (dev != t->device) {
        std::fprintf(stderr, "%s: the topology was built on device %d, the current device is %d\n", W,
                     t->device, dev);
        return -1;
    }
    uint8_t* base = scratch_on(dev);
    if (!base) return -1;

    // Form as pose_graph_hbm_topo_label chooses it (automatic plan shape):
    // four lanes per node for at most kQuadMaxPerCu super-chunks per CU on a
    // 64-lane topology, each super-chunk as two copy jobs when that leaves
    // every copy a workgroup of its own; slots bounded by the array.  Copies
    // only when their schedules are already on the device: uploading them
    // would allocate (the labels are the same either way).
    const int waves = lanes_to_waves(t->host.lanes);
    const bool quad = waves == 1 && n_super <= (uint64_t)kQuadMaxPerCu * (uint64_t)cus;
    const bool split = quad && t->host.has_parts && t->parts_on_device.load(std::memory_order_acquire) &&
                       2 * n_super <= (uint64_t)cus;
    const uint32_t parts[2] = {split ? kCopy0 : kWhole, kCopy0 + 1};
    const uint32_t jobs_per_super = split ? 2 : 1;
    uint64_t slot_bytes = slot_region_bytes(t, parts[0]);
    if (split) slot_bytes = std::max(slot_bytes, slot_region_bytes(t, parts[1]));
    slot_bytes = (slot_bytes + 255) / 256 * 256;
    const uint64_t fit = (kScratchBytes - kSlotsOff) / slot_bytes;
    if (fit == 0) {
        std::fprintf(stderr, "%s: one slot region of chunk_blocks %u needs %llu bytes, more than the %llu-byte "
                     "scratch array\n", W, cb, (unsigned long long)slot_bytes,
                     (unsigned long long)(kScratchBytes - kSlotsOff));
        return -1;
    }
    const int per_cu = std::max(1, max_blocks_per_cu(waves, quad));
    const uint32_t wg_waves = (uint32_t)waves * (quad ? 4u : 1u);
    const uint32_t per_cu_target = std::max<uint32_t>(1, 8 / wg_waves);   // two waves per SIMD
    uint64_t want = (uint64_t)cus * std::min<uint32_t>((uint32_t)per_cu, per_cu_target);
    if (split) want = std::max<uint64_t>(want, n_super * jobs_per_super);
    const uint64_t slot_cap = std::min<uint64_t>(want, fit);

    DevTopo topos[3];
    fill_dev_topo(t, &topos[0]);
    for (uint32_t c = 0; c < 2; ++c) fill_dev_topo(t, &topos[1 + c], kCopy0 + c);   // zeros unless resident
    DevJob jobs[kWindowJobs];
    ChunkSeeder seeder(session_seed, seed_len);
    const double t0 = now_seconds();
    if (!quiet()) {
        std::printf("  GPU label: strategy=persistent-step(lanes=%u, no-alloc)  %llu super-chunks  scaffold=%llu nodes\n",
                    t->host.lanes, (unsigned long long)n_super, (unsigned long long)t->host.count);
        std::fflush(stdout);
    }
    uint32_t* counter = reinterpret_cast<uint32_t*>(base + kCounterOff);
    DevTopo* d_topos = reinterpret_cast<DevTopo*>(base + kTopoOff);
    DevJob* d_jobs = reinterpret_cast<DevJob*>(base + kJobsOff);
    uint8_t* d_slots = base + kSlotsOff;
    uint64_t used_slots = 0;
    hipError_t e = hipMemcpy(d_topos, topos, sizeof topos, hipMemcpyHostToDevice);
    const uint64_t per_window = kWindowJobs / jobs_per_super;
    for (uint64_t j0 = 0; e == hipSuccess && j0 < n_super; j0 += per_window) {
        const uint64_t n = std::min<uint64_t>(per_window, n_super - j0);
        uint32_t nj = 0;
        // Deepest jobs first: with copies, copy 0 (the longer schedule) of
        // every super-chunk, then copy 1.
        for (uint32_t c = 0; c < jobs_per_super; ++c) {
            for (uint64_t k = 0; k < n; ++k) {
                DevJob& J = jobs[nj++];
                std::memset(&J, 0, sizeof J);
                J.persist = static_cast<uint8_t*>(hbm_dev_ptr) + (j0 + k) * sb +
                            (uint64_t)part_first_rank(t, parts[c]) * 32;
                J.topo = parts[c] == kWhole ? 0 : 1 + (parts[c] - kCopy0);
                seeder.seed_words((region_block_offset + (j0 + k) * (uint64_t)cb) / cb, J.seed);
            }
        }
        if (split && schedule_steps(t, parts[1]) > schedule_steps(t, parts[0])) {
            // Keep the deeper copy's jobs first.
            std::rotate(jobs, jobs + n, jobs + nj);
        }
        const uint32_t slots = (uint32_t)std::min<uint64_t>(slot_cap, nj);
        used_slots = std::max<uint64_t>(used_slots, slots);
        // hipMemcpy is ordered after the previous window's launch on the null
        // stream, so the job records it overwrites are no longer read.
        if ((e = hipMemcpy(d_jobs, jobs, nj * sizeof(DevJob), hipMemcpyHostToDevice)) != hipSuccess) break;
        if ((e = zero_device(counter, 256, nullptr)) != hipSuccess) break;
        e = launch_label(waves, slots, d_topos, d_jobs, nj, counter, d_slots, slot_bytes, nullptr, quad, true);
    }
    // Scrub the transient scaffold labels (the reference zeroes its scratch
    // before freeing it, common/src/cuda/label_gpu_warp.cu:236-238).
    if (e == hipSuccess) e = zero_device(d_slots, used_slots * slot_bytes, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", W, hipGetErrorString(e));
        return -1;
    }
    const double el = now_seconds() - t0;
    if (!quiet() && el > 0) {
        const double mib = (double)(n_super * sb) / (double)(1u << 20);
        std::printf("  GPU wipe: %.0f MiB in %.3f s = %.1f MiB/s\n", mib, el, mib / el);
        std::fflush(stdout);
    }
    return 0;
}
