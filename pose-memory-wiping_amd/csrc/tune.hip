// tune.hip — autotuner for the labeling launch (pose_gpu_autotune in
// include/gpu_label.h).
//
// For one chunk_blocks and a batch size it times short trial launches over the
// launch-shape space and keeps the fastest for that batch:
//   nodes per step       64, 128, 256
//   lanes per node       1 (throughput form) or 4 (latency form)
//   split                four-lane form on 64-lane topologies: whole
//                        super-chunks or one workgroup per scaffold copy
//   step schedule        list (and, on request, dependency levels)
//   occupancy            1, 2, 4 or 8 resident waves per SIMD (-> workgroup slots)
// A trial labels one round of super-chunks (one per slot, or the whole batch if
// it is smaller), so its kernel time is the round time of that shape; the
// batch time is rounds x round time.  Each trial's plan runs once untimed
// before the timed run, so no trial pays a first-launch cost the batch would
// not.  Results are cached per (device, chunk_blocks, batch) for the life of
// the process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/gpu_label.h"
#include "labeler_internal.h"

namespace {

std::mutex g_mu;
// (device, chunk_blocks, batch bucket) -> choice
std::map<std::tuple<int, uint32_t, uint64_t>, pose_gpu_tuned_t> g_cache;

uint64_t batch_bucket(uint64_t n) {   // cache key: next power of two
    uint64_t b = 1;
    while (b < n) b <<= 1;
    return b;
}

}  // namespace

extern "C" int pose_gpu_autotune(uint32_t chunk_blocks, uint64_t n_super, uint32_t flags,
                                 pose_gpu_tuned_t* best, pose_gpu_tuned_t* trials, size_t max_trials,
                                 size_t* n_trials) {
    static const char* W = "pose_gpu_autotune";
    if (!best || chunk_blocks == 0 || (chunk_blocks & (chunk_blocks - 1)) || chunk_blocks > (1u << 20) ||
        n_super == 0) {
        std::fprintf(stderr, "%s: bad arguments\n", W);
        return -1;
    }
    if (n_trials) *n_trials = 0;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -1;
    const auto key = std::make_tuple(dev, chunk_blocks, batch_bucket(n_super));
    if (!(flags & POSE_GPU_TUNE_FORCE)) {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_cache.find(key);
        if (it != g_cache.end()) { *best = it->second; return 0; }
    }
    const uint64_t sb = (uint64_t)chunk_blocks * 32;
    std::vector<std::pair<uint32_t, int>> shapes = {{64, 1}, {128, 1}, {256, 1}};
    if (flags & POSE_GPU_TUNE_LEVELS) shapes.push_back({64, 0});
    const uint32_t lpn_opts[2] = {1, 4};
    const uint32_t wps_opts[4] = {1, 2, 4, 8};

    // One output buffer big enough for the largest trial round.
    uint64_t max_round = std::min<uint64_t>(n_super, (uint64_t)cus * 4 * 8);
    uint8_t* d_out = nullptr;
    if (hipMalloc(&d_out, max_round * sb) != hipSuccess) {
        std::fprintf(stderr, "%s: trial buffer of %llu bytes does not fit\n", W,
                     (unsigned long long)(max_round * sb));
        return -1;
    }
    const uint8_t seed[32] = {0x70, 0x6f, 0x73, 0x65, 0x2d, 0x74, 0x75, 0x6e, 0x65};
    pose_gpu_tuned_t top{};
    double top_rate = -1;
    size_t nt = 0;
    for (const auto& sh : shapes) {
        pose_hbm_topo_t* topo = pose_gpu_topo_build_ex(chunk_blocks, sh.first, sh.second);
        if (!topo) continue;
        pose_gpu_topo_info_t ti{};
        pose_gpu_topo_info(topo, &ti);
        const uint32_t waves = sh.first / 64;
        for (uint32_t lpn : lpn_opts) {
        const int per_cu = std::max(1, pose::internal::max_blocks_per_cu((int)waves, lpn == 4));
        const uint32_t wg_waves = waves * lpn;
        // Split into copy jobs: four-lane form on topologies with copy
        // schedules only (one copy's critical path is the whole scaffold's).
        const bool can_split = lpn == 4 && pose_gpu_topo_part(topo, 0, nullptr, nullptr, nullptr, nullptr) == 0;
        for (uint32_t split : {(uint32_t)POSE_GPU_SPLIT_OFF, (uint32_t)POSE_GPU_SPLIT_ON}) {
        if (split == POSE_GPU_SPLIT_ON && !can_split) continue;
        const uint32_t jobs_per_super = split == POSE_GPU_SPLIT_ON ? 2 : 1;
        uint32_t last_slots = 0;
        for (uint32_t wps : wps_opts) {
            const uint32_t per_cu_target = std::max<uint32_t>(1, 4 * wps / wg_waves);
            const uint32_t slots_full = (uint32_t)cus * std::min<uint32_t>((uint32_t)per_cu, per_cu_target);
            // Workgroup slots (copy jobs when split); a round is one job per slot.
            const uint32_t slots = (uint32_t)std::min<uint64_t>(slots_full, n_super * jobs_per_super);
            if (slots == last_slots) continue;      // batch smaller than this occupancy
            last_slots = slots;
            const uint64_t round = std::min<uint64_t>(std::max<uint64_t>(1, slots / jobs_per_super), max_round);
            pose_gpu_session_t S{topo, seed, sizeof seed, 0, round, d_out};
            pose_gpu_launch_t L{slots, 0, lpn, split};
            double ms = 0;
            pose_gpu_plan_t* plan = pose_gpu_plan_create(&S, 1, &L, 0);
            if (!plan) continue;
            const bool ok = pose_gpu_plan_run(plan, nullptr) == 0 && pose_gpu_plan_run(plan, &ms) == 0;
            pose_gpu_plan_destroy(plan);
            if (!ok || ms <= 0) continue;
            const double rounds = std::ceil((double)n_super * jobs_per_super / (double)slots);
            const double batch_s = rounds * ms / 1e3;
            pose_gpu_tuned_t r{};
            r.chunk_blocks = chunk_blocks;
            r.lanes = sh.first;
            r.mode = sh.second;
            r.lanes_per_node = lpn;
            r.waves_per_simd = wps;
            r.slots = slots;
            r.steps = ti.steps;
            r.trial_super = (uint32_t)round;
            r.trial_kernel_ms = ms;
            r.batch_mib_per_s = (double)n_super * (double)sb / batch_s / 1048576.0;
            r.split = split;
            if (trials && nt < max_trials) trials[nt] = r;
            ++nt;
            if (r.batch_mib_per_s > top_rate) { top_rate = r.batch_mib_per_s; top = r; }
        }
        }
        }
        pose_graph_hbm_topo_free(topo);
    }
    (void)hipFree(d_out);
    if (n_trials) *n_trials = nt;
    if (top_rate < 0) return -1;
    *best = top;
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = top;
    return 0;
}
