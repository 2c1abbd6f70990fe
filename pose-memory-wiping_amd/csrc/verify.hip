// verify.hip — challenge answers on the GPU (pose_gpu_challenge_batch and
// pose_gpu_chunk_seed in include/gpu_label.h).
//
// A challenge asks for the label at output rank r of one super-chunk graph,
// identified by its 32-byte chunk seed and chunk_blocks: the value the
// reference's CPU pose_graph_challenge returns (common/src/graph.c:1019-1042),
// which labels the whole graph and keeps rank r.  Here a batch of challenges
// is answered with one batched labeling launch per window of distinct
// super-chunks (challenges on the same super-chunk share its labeling), then a
// gather kernel copies the challenged 32-byte labels out.  A super-chunk's
// scaffold is two copies with no edge between them, and each output rank
// belongs to one copy; a super-chunk whose challenges all fall in one copy
// has only that copy labeled.  Topologies are built on first use per device
// and chunk_blocks and kept until pose_gpu_challenge_release (or the process
// ends).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/gpu_label.h"
#include "b3_host.h"
#include "labeler_internal.h"

using namespace pose;
using namespace pose::internal;

namespace {

constexpr uint32_t kMaxChunkBlocks = 1u << 20;
constexpr uint64_t kWindowBytes = 16ull << 30;   // label bytes per labeling launch
constexpr uint32_t kWindowJobs = 8192;            // super-chunks per labeling launch

std::mutex g_mu;
std::map<std::pair<int, uint32_t>, pose_hbm_topo*> g_topos;

pose_hbm_topo* cached_topology(int dev, uint32_t cb) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_topos.find({dev, cb});
    if (it != g_topos.end()) return it->second;
    pose_hbm_topo* t = pose_gpu_topo_build_ex(cb, 64, kSchedList);
    if (t) g_topos[{dev, cb}] = t;
    return t;
}

// out[i] = the 32 bytes at base + off[i]; two lanes per label, 16 bytes each.
__global__ void k_gather(const uint8_t* __restrict__ base, const uint64_t* __restrict__ off,
                         uint4* __restrict__ out, uint32_t n) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * n) return;
    const uint32_t i = t >> 1, h = t & 1;
    out[2 * i + h] = reinterpret_cast<const uint4*>(base + off[i])[h];
}

}  // namespace

extern "C" int pose_gpu_chunk_seed(const uint8_t* session_seed, size_t seed_len, uint64_t chunk_index,
                                   uint8_t out[32]) {
    if (!out || (!session_seed && seed_len)) return -1;
    uint32_t w[8];
    ChunkSeeder seeder(session_seed, seed_len);
    seeder.seed_words(chunk_index, w);
    for (int i = 0; i < 8; ++i) {
        out[4 * i] = (uint8_t)w[i];
        out[4 * i + 1] = (uint8_t)(w[i] >> 8);
        out[4 * i + 2] = (uint8_t)(w[i] >> 16);
        out[4 * i + 3] = (uint8_t)(w[i] >> 24);
    }
    return 0;
}

namespace {

// Device buffers of one call.  The labels of challenged super-chunks are
// scrubbed before release, on every return path.
struct ChallengeBuffers {
    uint8_t* out = nullptr;
    uint64_t* off = nullptr;
    uint4* res = nullptr;
    uint64_t cap_bytes = 0;
    uint32_t cap_chal = 0;
    ~ChallengeBuffers() {
        if (out) {
            if (zero_device(out, cap_bytes, nullptr) == hipSuccess) (void)hipStreamSynchronize(nullptr);
            (void)hipFree(out);
        }
        if (off) (void)hipFree(off);
        if (res) (void)hipFree(res);
    }
};

int challenge_batch(const pose_gpu_challenge_t* ch, size_t n, uint8_t* out, double* kernel_ms) {
    static const char* W = "pose_gpu_challenge_batch";
    if (kernel_ms) *kernel_ms = 0;
    if (n == 0) return 0;
    if (!ch || !out) {
        std::fprintf(stderr, "%s: NULL challenges or output\n", W);
        return -1;
    }
    for (size_t i = 0; i < n; ++i) {
        if (ch[i].chunk_blocks == 0 || ch[i].chunk_blocks > kMaxChunkBlocks || ch[i].rank >= ch[i].chunk_blocks) {
            std::fprintf(stderr, "%s: challenge %zu: chunk_blocks %u / rank %u out of range\n", W, i,
                         ch[i].chunk_blocks, ch[i].rank);
            return -1;
        }
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        std::fprintf(stderr, "%s: no HIP device\n", W);
        return -1;
    }

    // Distinct super-chunks, deepest first so each window groups similar depths.
    using Key = std::pair<uint32_t, std::array<uint8_t, 32>>;
    std::map<Key, uint32_t> index;
    std::vector<Key> sc;
    std::vector<uint32_t> sc_of(n);
    for (size_t i = 0; i < n; ++i) {
        Key k{ch[i].chunk_blocks, {}};
        std::memcpy(k.second.data(), ch[i].chunk_seed, 32);
        auto it = index.find(k);
        if (it == index.end()) {
            it = index.emplace(k, (uint32_t)sc.size()).first;
            sc.push_back(k);
        }
        sc_of[i] = it->second;
    }
    std::vector<uint32_t> order(sc.size());
    for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return sc[a].first > sc[b].first; });
    std::vector<std::vector<uint32_t>> chal_of(sc.size());
    for (size_t i = 0; i < n; ++i) chal_of[sc_of[i]].push_back((uint32_t)i);

    ChallengeBuffers B;
    uint8_t*& d_out = B.out;
    uint64_t*& d_off = B.off;
    uint4*& d_res = B.res;
    uint64_t& cap_bytes = B.cap_bytes;
    uint32_t& cap_chal = B.cap_chal;
    int rc = -1;
    double total_ms = 0;
    size_t w0 = 0;
    while (w0 < order.size()) {
        // One window: distinct super-chunks up to the byte and job limits.
        size_t w1 = w0;
        uint64_t bytes = 0;
        uint32_t n_chal = 0;
        while (w1 < order.size() && w1 - w0 < kWindowJobs) {
            const uint64_t b = (uint64_t)sc[order[w1]].first * 32;
            if (w1 > w0 && bytes + b > kWindowBytes) break;
            bytes += b;
            n_chal += (uint32_t)chal_of[order[w1]].size();
            ++w1;
        }
        if (bytes > cap_bytes) {
            if (d_out) {   // scrubbed before release, like the final buffer
                if (zero_device(d_out, cap_bytes, nullptr) == hipSuccess) (void)hipStreamSynchronize(nullptr);
                (void)hipFree(d_out);
                d_out = nullptr;
                cap_bytes = 0;
            }
            if (hipMalloc(&d_out, bytes) != hipSuccess) {
                std::fprintf(stderr, "%s: label buffer of %llu bytes does not fit\n", W, (unsigned long long)bytes);
                d_out = nullptr;
                goto done;
            }
            cap_bytes = bytes;
        }
        if (n_chal > cap_chal) {
            if (d_off) { (void)hipFree(d_off); d_off = nullptr; }
            if (d_res) { (void)hipFree(d_res); d_res = nullptr; }
            if (hipMalloc(&d_off, (size_t)n_chal * 8) != hipSuccess ||
                hipMalloc(&d_res, (size_t)n_chal * 32) != hipSuccess) {
                std::fprintf(stderr, "%s: gather buffers do not fit\n", W);
                goto done;
            }
            cap_chal = n_chal;
        }
        {
            // Jobs: one per distinct super-chunk, the chunk seed as its seed words.
            std::vector<const pose_hbm_topo*> tlist;
            std::vector<PendingGroup> pend(1);
            pend[0].req = LaunchReq{1, POSE_GPU_LANES_AUTO, 0, 0};
            std::vector<uint64_t> h_off;
            std::vector<uint32_t> h_idx;
            uint64_t off = 0, nodes = 0;
            for (size_t k = w0; k < w1; ++k) {
                const Key& K = sc[order[k]];
                pose_hbm_topo* t = cached_topology(dev, K.first);
                if (!t) goto done;
                uint32_t ti;
                auto it = std::find(tlist.begin(), tlist.end(), t);
                if (it == tlist.end()) { ti = (uint32_t)tlist.size(); tlist.push_back(t); }
                else ti = (uint32_t)(it - tlist.begin());
                DevJob J{};
                J.persist = d_out + off;
                J.topo = ti;
                J.part = kWhole;
                // A challenge reads one label.  When every challenged rank of
                // this super-chunk lies in one copy of its scaffold (the copies
                // share no edge), only that copy is labeled: half the nodes.
                if (t->host.has_parts) {
                    bool need[2] = {false, false};
                    for (uint32_t c : chal_of[order[k]]) {
                        const Schedule& P0 = t->host.part[0];
                        need[ch[c].rank >= P0.rank_lo && ch[c].rank < P0.rank_hi ? 0 : 1] = true;
                    }
                    if (need[0] != need[1]) {
                        J.part = kCopy0 + (need[0] ? 0u : 1u);
                        J.persist += (uint64_t)part_first_rank(t, J.part) * 32;
                    }
                }
                for (int q = 0; q < 8; ++q) J.seed[q] = b3::load_le(K.second.data() + 4 * q);
                pend[0].jobs.push_back(J);
                nodes += J.part == kWhole ? t->host.count
                                          : t->host.part[J.part - kCopy0].id_hi - t->host.part[J.part - kCopy0].id_lo;
                for (uint32_t c : chal_of[order[k]]) {
                    h_off.push_back(off + (uint64_t)ch[c].rank * 32);
                    h_idx.push_back(c);
                }
                off += (uint64_t)K.first * 32;
            }
            pose_gpu_plan_t* p = plan_build(pend, tlist, nodes, off, 0, 1, false);
            if (!p) goto done;
            double ms = 0;
            const int prc = pose_gpu_plan_run(p, &ms);
            pose_gpu_plan_destroy(p);
            if (prc != 0) goto done;
            total_ms += ms;
            const uint32_t nc = (uint32_t)h_off.size();
            std::vector<uint8_t> res((size_t)nc * 32);
            if (hipMemcpy(d_off, h_off.data(), (size_t)nc * 8, hipMemcpyHostToDevice) != hipSuccess) goto done;
            k_gather<<<(2 * nc + 255) / 256, 256>>>(d_out, d_off, d_res, nc);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpy(res.data(), d_res, (size_t)nc * 32, hipMemcpyDeviceToHost) != hipSuccess) {
                std::fprintf(stderr, "%s: gather failed\n", W);
                goto done;
            }
            for (uint32_t q = 0; q < nc; ++q) std::memcpy(out + (size_t)h_idx[q] * 32, res.data() + (size_t)q * 32, 32);
        }
        w0 = w1;
    }
    rc = 0;
    if (kernel_ms) *kernel_ms = total_ms;
done:
    return rc;
}

}  // namespace

extern "C" void pose_gpu_challenge_release(void) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return;
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto it = g_topos.begin(); it != g_topos.end();) {
        if (it->first.first == dev) {
            pose_graph_hbm_topo_free(it->second);
            it = g_topos.erase(it);
        } else {
            ++it;
        }
    }
}

extern "C" int pose_gpu_challenge_batch(const pose_gpu_challenge_t* ch, size_t n, uint8_t* out,
                                        double* kernel_ms) {
    try {
        return challenge_batch(ch, n, out, kernel_ms);
    } catch (const std::exception& e) {   // host allocation failure: nothing escapes the C ABI
        std::fprintf(stderr, "pose_gpu_challenge_batch: %s\n", e.what());
        return -1;
    }
}
