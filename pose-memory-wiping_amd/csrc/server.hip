// server.hip — continuous-batching label server (pose_gpu_server_* in
// include/gpu_label.h).
//
// Requests (session seed, chunk_blocks, start block, super-chunk count) are
// queued; a scheduler thread packs pending super-chunks into persistent
// launches on up to four streams.  Admission is by HBM footprint: each launch
// needs its scratch (slots x pool x 32 B) and its output staging (super-chunks x
// cb x 32 B) from a fixed device arena, plus a pinned host region of the same
// output size.  A launch only takes super-chunks of one chunk_blocks, so a
// short graph is never held back by a deep one in the same launch.  Requests
// that arrive while launches are running are admitted as soon as a stream,
// arena and pinned space are free, and run concurrently with them on the
// device.  Each launch's output is copied device->pinned on its own stream; a
// writer thread hands finished super-chunks to the sink (dry run, or pwrite to
// a file descriptor, O_DIRECT-aligned) and records per-super-chunk latency from
// submit to sink.  Topologies are built once per chunk_blocks and stay
// resident; the arena is zeroed before it is freed.
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/gpu_label.h"
#include "b3_host.h"
#include "labeler_internal.h"

using namespace pose;
using namespace pose::internal;

namespace {

constexpr int kMaxStreams = 4;
constexpr uint32_t kServerQuadMaxPerCu = 2;
constexpr uint64_t kAlign = 4096;   // arena and pinned-ring granularity (O_DIRECT safe)

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

uint64_t round_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

// First-fit range allocator over [0, size) with coalescing frees.
class RangeAlloc {
public:
    explicit RangeAlloc(uint64_t size = 0) { reset(size); }
    void reset(uint64_t size) {
        free_.clear();
        if (size) free_[0] = size;
        size_ = size;
        used_ = 0;
    }
    bool alloc(uint64_t n, uint64_t* off) {
        n = round_up(std::max<uint64_t>(n, 1), kAlign);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const uint64_t rest = it->second - n, at = it->first + n;
                free_.erase(it);
                if (rest) free_[at] = rest;
                used_ += n;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        n = round_up(std::max<uint64_t>(n, 1), kAlign);
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t used() const { return used_; }
    uint64_t size() const { return size_; }

private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t size_ = 0, used_ = 0;
};

struct Request {
    uint64_t id = 0;
    std::vector<uint8_t> seed;
    uint32_t cb = 0;
    uint64_t rbo = 0, n_super = 0, sink_offset = 0;
    uint8_t* host_out = nullptr;
    double t_submit = 0;
    uint64_t next = 0;        // next super-chunk to admit
    uint64_t done = 0;        // super-chunks in the sink
    // Super-chunks admitted into launches and not yet written or dropped.  A
    // request leaves `live` (and wait returns) only when this is zero, so no
    // launch writes into host_out after its caller has been told the outcome.
    uint64_t outstanding = 0;
    bool failed = false;
    bool mem_waited = false;  // found a free stream but no arena / pinned space
    uint32_t topo_index = 0;
};

struct Piece {
    std::shared_ptr<Request> req;
    uint64_t first = 0, count = 0;
    uint64_t out_off = 0;     // offset of this piece's output in the launch's region
};

struct Launch {
    int stream = -1;
    std::vector<Piece> pieces;
    uint64_t arena_off = 0, arena_bytes = 0, scratch_bytes = 0, slot_bytes = 0;
    uint64_t pin_off = 0, out_bytes = 0;
    uint32_t n_jobs = 0, slots = 0;   // super-chunks, resident workgroups
    bool quad = false;
    bool split = false;               // one workgroup job per copy of each scaffold
    double t_launch = 0;
    hipEvent_t ev_k0 = nullptr, ev_k1 = nullptr, ev_done = nullptr;
};

struct StreamSlot {
    hipStream_t stream = nullptr;
    DevJob* h_jobs = nullptr;     // pinned job table
    DevJob* d_jobs = nullptr;
    uint32_t* d_counter = nullptr;
    hipEvent_t ev_k0 = nullptr, ev_k1 = nullptr, ev_done = nullptr;
};

}  // namespace

struct pose_gpu_server {
    pose_gpu_server_config_t cfg{};
    int device = 0;
    int cus = 0;
    uint32_t max_jobs = 0;       // per launch
    uint32_t slot_cap = 0;       // resident workgroups per launch

    uint8_t* d_arena = nullptr;
    uint8_t* h_pinned = nullptr;
    DevTopo* d_topos = nullptr;  // [log2(cb)]: whole scaffold; [kTopoSlots + 2 log2(cb) + c]: copy c
    pose_hbm_topo* topos[21] = {nullptr};
    StreamSlot slots[kMaxStreams];
    int n_streams = 0;

    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_writer;
    std::deque<std::shared_ptr<Request>> pending;
    std::map<uint64_t, std::shared_ptr<Request>> live;   // submitted, not yet settled
    std::vector<uint64_t> failed_ids;
    std::mutex build_mu;         // serialises topology builds (done outside mu)
    std::vector<std::unique_ptr<Launch>> inflight;
    std::deque<std::unique_ptr<Launch>> to_write;
    RangeAlloc arena, pinned;
    uint64_t next_id = 1;
    bool stop = false;
    bool error = false;
    // Host-side fault injection for tests: the fail_retire_at-th launch to
    // retire reports a failure (POSE_GPU_SERVER_FAIL_RETIRE; 0 = off).
    uint64_t fail_retire_at = 0, retired = 0;

    // statistics
    std::vector<double> lat_ms;
    uint64_t req_done = 0, sc_done = 0, bytes_done = 0, launches = 0;
    double t_first = 0, t_last = 0, kernel_ms = 0, copy_ms = 0, sink_ms = 0;
    uint32_t peak_inflight = 0;
    uint64_t peak_hbm = 0;
    uint64_t memory_waits = 0;   // requests that found no room for one super-chunk
    uint64_t memory_cut = 0;     // launches made smaller by arena / pinned space

    // Sink writes: straight from pinned staging, or through an aligned bounce
    // buffer when the sink refuses the pinned pages (some O_DIRECT paths cannot
    // pin driver-mapped host memory).
    uint8_t* bounce = nullptr;
    bool use_bounce = false;

    std::thread sched, writer;
};

namespace {

int topo_index_of(uint32_t cb) { return 31 - __builtin_clz(cb); }

constexpr int kTopoSlots = 21;   // chunk_blocks 2^0 .. 2^20

// Build (once) the 2^k topology and publish its device table entry.  Called
// without s->mu held: a deep topology takes seconds to schedule on the host.
int ensure_topology(pose_gpu_server* s, uint32_t cb) {
    const int k = topo_index_of(cb);
    std::lock_guard<std::mutex> bl(s->build_mu);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->topos[k]) return 0;
    }
    // The build runs on the server's device; the caller's current device is
    // restored on every return path.
    struct DeviceGuard {
        int prev = -1;
        explicit DeviceGuard(int dev) {
            if (hipGetDevice(&prev) != hipSuccess) prev = -1;
            if (prev != dev) (void)hipSetDevice(dev);
        }
        ~DeviceGuard() {
            int cur = -1;
            if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
        }
    } guard(s->device);
    pose_hbm_topo* t = pose_gpu_topo_build_ex(cb, 64, kSchedList);
    if (!t) return -1;
    // Small launches label each copy of a scaffold on its own: the server
    // keeps the copy schedules resident from the start.
    if (ensure_parts(t) != hipSuccess) {
        pose_graph_hbm_topo_free(t);
        return -1;
    }
    DevTopo D[3];
    fill_dev_topo(t, &D[0]);
    for (uint32_t c = 0; c < 2; ++c) {
        if (t->host.has_parts) fill_dev_topo(t, &D[1 + c], kCopy0 + c);
        else std::memset(&D[1 + c], 0, sizeof D[0]);
    }
    // Entries of 2^k are read only by launches of 2^k super-chunks, none of
    // which can exist before they are published.
    if (hipMemcpy(s->d_topos + k, &D[0], sizeof D[0], hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_topos + kTopoSlots + 2 * k, &D[1], 2 * sizeof D[0], hipMemcpyHostToDevice) != hipSuccess) {
        pose_graph_hbm_topo_free(t);
        return -1;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->topos[k] = t;
    return 0;
}

bool work_pending(const pose_gpu_server* s) {
    for (const auto& r : s->pending)
        if (!r->failed && r->next < r->n_super) return true;
    return false;
}

// A request is settled when none of its super-chunks is in a launch any more
// and it is complete or failed; it then leaves `live`.  Called with the lock.
void settle(pose_gpu_server* s, Request& r) {
    if (r.outstanding || !(r.failed || r.done == r.n_super)) return;
    if (s->live.erase(r.id)) {
        if (!r.failed) s->req_done++;
        s->cv_done.notify_all();
    }
}

void mark_failed(pose_gpu_server* s, Request& r) {
    if (r.failed) return;
    r.failed = true;
    s->failed_ids.push_back(r.id);
}

// Marks every live request failed and stops admission.  Reports only the
// first error: later calls (other launches of a failed device) are silent.
// Launches already running still finish (or fail) before their requests
// settle.
void fail_all(pose_gpu_server* s, const char* why) {
    if (s->error) return;
    std::fprintf(stderr, "pose_gpu_server: %s\n", why);
    s->error = true;
    s->pending.clear();
    std::vector<std::shared_ptr<Request>> reqs;
    for (auto& kv : s->live) reqs.push_back(kv.second);
    for (auto& r : reqs) {
        mark_failed(s, *r);
        settle(s, *r);
    }
    s->cv_done.notify_all();
}

// Pick super-chunks for one launch on stream `si`: the chunk_blocks of the
// oldest pending request, then pending super-chunks of that size in arrival
// order, while the arena and pinned ring can hold them.  Returns nullptr when
// nothing fits now.  Called with the lock held.
std::unique_ptr<Launch> admit(pose_gpu_server* s, int si) {
    while (!s->pending.empty() &&
           (s->pending.front()->failed || s->pending.front()->next >= s->pending.front()->n_super))
        s->pending.pop_front();
    if (s->pending.empty()) return nullptr;
    const uint32_t cb = s->pending.front()->cb;
    const pose_hbm_topo* t = s->topos[topo_index_of(cb)];
    const uint64_t sb = (uint64_t)cb * 32;
    const uint64_t whole_bytes = slot_region_bytes(t);
    const uint64_t copy_bytes = t->host.has_parts
        ? std::max(slot_region_bytes(t, kCopy0), slot_region_bytes(t, kCopy0 + 1)) : 0;

    // How many super-chunks: bounded by the per-launch cap and by memory.
    uint64_t avail = 0;
    for (auto& r : s->pending)
        if (r->cb == cb && !r->failed) avail += r->n_super - r->next;
    uint64_t n = std::min<uint64_t>(avail, s->max_jobs);
    const uint64_t n_wanted = n;
    // CUs held by launches in flight: a four-lane workgroup fills a CU (one
    // wave per SIMD), a one-lane workgroup one SIMD.
    uint64_t busy_cus = 0;
    for (const auto& L : s->inflight) busy_cus += L->quad ? L->slots : (L->slots + 3) / 4;
    auto launch = std::make_unique<Launch>();
    for (; n >= 1 && !launch->n_jobs; n = n > 1 ? n / 2 : 0) {
        // Small launches use the four-lane latency form: they cannot fill the
        // device one lane per node.  The bound is two super-chunks per CU,
        // below the single-launch crossover (kQuadMaxPerCu): launches here
        // overlap, and four-lane workgroups (147 VGPRs per wave) leave less
        // room for the others.  Each copy of a scaffold gets a workgroup of
        // its own (split) when the launch has at most one super-chunk per two
        // CUs and, with the launches in flight, at most two four-lane
        // workgroups per CU result, if the arena holds the two copies'
        // scratch; else the launch is unsplit.  (A split super-chunk takes
        // 38 % more CU time at 2^17, so splitting onto CUs that one-lane
        // launches keep busy only slows those.)
        const bool quad = n <= kServerQuadMaxPerCu * (uint64_t)s->cus;
        const bool can_split = quad && t->host.has_parts && 2 * n <= (uint64_t)s->cus &&
                               2 * n + busy_cus <= 2 * (uint64_t)s->cus;
        for (int attempt = can_split ? 0 : 1; attempt < 2; ++attempt) {
            const bool split = attempt == 0;
            const uint64_t slot_bytes = split ? copy_bytes : whole_bytes;
            const uint32_t slots = split ? (uint32_t)(2 * n) : (uint32_t)std::min<uint64_t>(n, s->slot_cap);
            const uint64_t scratch = round_up((uint64_t)slots * slot_bytes, kAlign);
            const uint64_t out = n * sb;
            uint64_t a_off, p_off;
            if (!s->arena.alloc(scratch + out, &a_off)) continue;
            if (!s->pinned.alloc(out, &p_off)) { s->arena.release(a_off, scratch + out); continue; }
            launch->arena_off = a_off;
            launch->arena_bytes = scratch + out;
            launch->scratch_bytes = scratch;
            launch->slot_bytes = slot_bytes;
            launch->pin_off = p_off;
            launch->out_bytes = out;
            launch->slots = slots;
            launch->n_jobs = (uint32_t)n;
            launch->quad = quad;
            launch->split = split;
            break;
        }
    }
    if (!launch->n_jobs) {
        // A stream is free but the oldest request's next super-chunk does not
        // fit: it waits for running launches to release memory.
        Request& f = *s->pending.front();
        if (!f.mem_waited) {
            f.mem_waited = true;
            s->memory_waits++;
        }
        return nullptr;
    }
    if (launch->n_jobs < n_wanted) s->memory_cut++;
    uint64_t left = launch->n_jobs, off = 0;
    for (auto& r : s->pending) {
        if (!left) break;
        if (r->cb != cb || r->failed || r->next >= r->n_super) continue;
        const uint64_t take = std::min(left, r->n_super - r->next);
        launch->pieces.push_back(Piece{r, r->next, take, off});
        r->next += take;
        r->outstanding += take;
        off += take * sb;
        left -= take;
    }
    launch->stream = si;
    s->peak_hbm = std::max(s->peak_hbm, s->arena.used());
    return launch;
}

// Enqueue job table upload, counter reset, labeling, and device->pinned copy.
bool enqueue(pose_gpu_server* s, Launch* L) {
    StreamSlot& S = s->slots[L->stream];
    const uint32_t cb = L->pieces.front().req->cb;
    const uint64_t sb = (uint64_t)cb * 32;
    uint8_t* out_dev = s->d_arena + L->arena_off + L->scratch_bytes;
    const pose_hbm_topo* t = s->topos[topo_index_of(cb)];
    uint32_t j = 0;
    for (const Piece& P : L->pieces) {
        ChunkSeeder seeder(P.req->seed.data(), P.req->seed.size());
        for (uint64_t k = 0; k < P.count; ++k) {
            DevJob J;
            std::memset(&J, 0, sizeof J);
            J.persist = out_dev + P.out_off + k * sb;
            J.topo = P.req->topo_index;
            seeder.seed_words((P.req->rbo + (P.first + k) * (uint64_t)cb) / cb, J.seed);
            if (!L->split) { S.h_jobs[j++] = J; continue; }
            // Copy c of the scaffold writes ranks [first, end) of the super-chunk.
            for (uint32_t c = 0; c < 2; ++c) {
                DevJob H = J;
                H.topo = kTopoSlots + 2 * P.req->topo_index + c;
                H.persist = J.persist + (uint64_t)part_first_rank(t, kCopy0 + c) * 32;
                S.h_jobs[j++] = H;
            }
        }
    }
    const uint32_t wg_jobs = j;
    hipStream_t st = S.stream;
    L->ev_k0 = S.ev_k0; L->ev_k1 = S.ev_k1; L->ev_done = S.ev_done;
    L->t_launch = now_s();
    if (hipMemcpyAsync(S.d_jobs, S.h_jobs, (size_t)wg_jobs * sizeof(DevJob), hipMemcpyHostToDevice, st) != hipSuccess) return false;
    if (zero_device(S.d_counter, 256, st) != hipSuccess) return false;
    if (hipEventRecord(L->ev_k0, st) != hipSuccess) return false;
    // Launches overlap on up to n_streams streams. Under the default cap of
    // four super-chunks per CU per launch each holds at most one runner per
    // SIMD (four-runner workgroups), so four of them
    // together stay within the five waves per SIMD that the pipelined
    // one-lane kernel's 87 VGPRs allow: it may run beside the others. Measured at
    // 64 requests/s (2^12..2^17, 1..128 super-chunks per request): 2,502-2,507
    // against 2,221 MiB/s with k_label, p99 8.1 against 9.9 s.
    if (launch_label(1, L->slots, s->d_topos, S.d_jobs, wg_jobs, S.d_counter,
                     s->d_arena + L->arena_off, L->slot_bytes, st, L->quad, true) != hipSuccess)
        return false;
    if (hipEventRecord(L->ev_k1, st) != hipSuccess) return false;
    if (hipMemcpyAsync(s->h_pinned + L->pin_off, out_dev, L->out_bytes, hipMemcpyDeviceToHost, st) != hipSuccess) return false;
    if (hipEventRecord(L->ev_done, st) != hipSuccess) return false;
    return true;
}

void scheduler_loop(pose_gpu_server* s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        // Retire finished launches.
        for (size_t i = 0; i < s->inflight.size();) {
            Launch* L = s->inflight[i].get();
            hipError_t q = hipEventQuery(L->ev_done);
            if (q == hipErrorNotReady) { ++i; continue; }
            if (q == hipSuccess && s->fail_retire_at && ++s->retired == s->fail_retire_at)
                q = hipErrorLaunchFailure;   // injected (POSE_GPU_SERVER_FAIL_RETIRE)
            if (q != hipSuccess) {
                // The launch leaves the in-flight set, so the loops can end; its
                // arena and pinned ranges stay allocated (the device may still
                // be using them) until the server is destroyed.
                fail_all(s, hipGetErrorString(q));
                for (const Piece& P : L->pieces) {
                    P.req->outstanding -= P.count;
                    settle(s, *P.req);
                }
                s->inflight.erase(s->inflight.begin() + (long)i);
                continue;
            }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, L->ev_k0, L->ev_k1);
            s->kernel_ms += ms;
            if (hipEventElapsedTime(&ms, L->ev_k1, L->ev_done) == hipSuccess) s->copy_ms += ms;
            s->arena.release(L->arena_off, L->arena_bytes);
            s->to_write.push_back(std::move(s->inflight[i]));
            s->inflight.erase(s->inflight.begin() + (long)i);
            s->cv_writer.notify_one();
        }
        // Admit new work onto free streams.
        bool admitted = false;
        if (!s->error) {
            for (int si = 0; si < s->n_streams; ++si) {
                bool slot_free = true;
                for (auto& L : s->inflight) if (L->stream == si) slot_free = false;
                if (!slot_free) continue;
                auto L = admit(s, si);
                if (!L) break;
                if (!enqueue(s, L.get())) { fail_all(s, "launch enqueue failed"); break; }
                s->launches++;
                s->inflight.push_back(std::move(L));
                s->peak_inflight = std::max<uint32_t>(s->peak_inflight, (uint32_t)s->inflight.size());
                admitted = true;
            }
        }
        if (s->stop && (s->error || (s->inflight.empty() && !work_pending(s)))) break;
        if (!admitted)
            s->cv_sched.wait_for(lk, std::chrono::microseconds(s->inflight.empty() ? 20000 : 200));
    }
}

bool sink_write(pose_gpu_server* s, const uint8_t* src, uint64_t len, uint64_t off);

// Copy into a caller's host buffer.  One thread copies pinned staging at
// about 27 GB/s on the GPU box's host (256 MiB of a 64-super-chunk 2^17
// request in 9-10 ms), so large copies are split over up to 8 threads in
// contiguous parts; small ones stay on the writer thread.
void copy_host(uint8_t* dst, const uint8_t* src, uint64_t len) {
    constexpr uint64_t kPart = 32ull << 20;
    const unsigned hw = std::thread::hardware_concurrency();
    const uint64_t nt = std::min<uint64_t>({len / kPart, 8u, hw ? hw : 1u});
    if (nt < 2) {
        std::memcpy(dst, src, len);
        return;
    }
    std::vector<std::thread> th;
    for (uint64_t t = 1; t < nt; ++t)
        th.emplace_back([=]() {
            const uint64_t lo = len * t / nt, hi = len * (t + 1) / nt;
            std::memcpy(dst + lo, src + lo, hi - lo);
        });
    std::memcpy(dst, src, len / nt);
    for (std::thread& x : th) x.join();
}

void writer_loop(pose_gpu_server* s) {
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        if (s->to_write.empty()) {
            if (s->stop && (s->error || (s->inflight.empty() && !work_pending(s)))) break;
            s->cv_writer.wait_for(lk, std::chrono::milliseconds(20));
            continue;
        }
        std::unique_ptr<Launch> L = std::move(s->to_write.front());
        s->to_write.pop_front();
        // Pieces of requests that have failed meanwhile are dropped, not
        // written (their requests are still live, so host_out is valid).
        std::vector<char> skip;
        for (const Piece& P : L->pieces) skip.push_back(P.req->failed ? 1 : 0);
        lk.unlock();
        const uint32_t cb = L->pieces.front().req->cb;
        const uint64_t sb = (uint64_t)cb * 32;
        std::vector<char> ok(L->pieces.size(), 1);
        std::vector<double> lat;
        const double t_sink0 = now_s();
        for (size_t i = 0; i < L->pieces.size(); ++i) {
            const Piece& P = L->pieces[i];
            if (skip[i]) continue;
            const uint8_t* src = s->h_pinned + L->pin_off + P.out_off;
            if (P.req->host_out) copy_host(P.req->host_out + P.first * sb, src, P.count * sb);
            if (s->cfg.sink_fd >= 0 && !sink_write(s, src, P.count * sb, P.req->sink_offset + P.first * sb))
                ok[i] = 0;
            const double t = now_s();
            if (ok[i])
                for (uint64_t k = 0; k < P.count; ++k) lat.push_back((t - P.req->t_submit) * 1e3);
        }
        const double t_sink1 = now_s();
        lk.lock();
        s->sink_ms += (t_sink1 - t_sink0) * 1e3;
        s->pinned.release(L->pin_off, L->out_bytes);
        const double t = now_s();
        for (size_t i = 0; i < L->pieces.size(); ++i) {
            const Piece& P = L->pieces[i];
            P.req->outstanding -= P.count;
            if (!skip[i] && !ok[i]) mark_failed(s, *P.req);
            if (!skip[i] && ok[i]) {
                P.req->done += P.count;
                s->sc_done += P.count;
                s->bytes_done += P.count * sb;
            }
            settle(s, *P.req);
        }
        s->lat_ms.insert(s->lat_ms.end(), lat.begin(), lat.end());
        s->t_last = t;
        s->cv_done.notify_all();
        s->cv_sched.notify_one();
    }
}

constexpr uint64_t kBounceBytes = 64ull << 20;

bool pwrite_all(int fd, const uint8_t* src, uint64_t len, uint64_t off) {
    uint64_t done = 0;
    while (done < len) {
        ssize_t w = pwrite(fd, src + done, len - done, (off_t)(off + done));
        if (w < 0 && errno == EINTR) continue;
        if (w <= 0) { if (w == 0) errno = EIO; return false; }
        done += (uint64_t)w;
    }
    return true;
}

// Write one finished range to the sink fd (writer thread only).
bool sink_write(pose_gpu_server* s, const uint8_t* src, uint64_t len, uint64_t off) {
    if (!s->use_bounce) {
        if (pwrite_all(s->cfg.sink_fd, src, len, off)) return true;
        if (errno != EFAULT && errno != EINVAL) {
            std::fprintf(stderr, "pose_gpu_server: sink write at %llu failed: %s\n",
                         (unsigned long long)off, std::strerror(errno));
            return false;
        }
        std::fprintf(stderr, "pose_gpu_server: sink refused pinned staging (%s); "
                     "writing through an aligned bounce buffer\n", std::strerror(errno));
        if (posix_memalign(reinterpret_cast<void**>(&s->bounce), kAlign, kBounceBytes) != 0) {
            s->bounce = nullptr;
            return false;
        }
        s->use_bounce = true;
    }
    for (uint64_t done = 0; done < len; done += kBounceBytes) {
        const uint64_t n = std::min(kBounceBytes, len - done);
        std::memcpy(s->bounce, src + done, n);
        if (!pwrite_all(s->cfg.sink_fd, s->bounce, n, off + done)) {
            std::fprintf(stderr, "pose_gpu_server: sink write at %llu failed: %s\n",
                         (unsigned long long)(off + done), std::strerror(errno));
            return false;
        }
    }
    return true;
}

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const double idx = p * (double)(v.size() - 1);
    const size_t lo = (size_t)idx;
    const size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (idx - (double)lo);
}

// The CPU's page-table entries of pinned memory are filled on first access:
// a writer's first read of a freshly allocated staging page costs a minor
// fault (measured: 256 MiB read at 2.7 GB/s the first time, 28-30 GB/s after).
// Touch every page once at creation, from several threads, so that sink
// writes and host copies run at memory speed from the first launch on.
void prefault_host(uint8_t* p, uint64_t bytes) {
    constexpr uint64_t kPage = 4096;
    const uint64_t pages = (bytes + kPage - 1) / kPage;
    const unsigned hw = std::thread::hardware_concurrency();
    const uint64_t nt = std::max<uint64_t>(1, std::min<uint64_t>({hw ? hw : 1u, 16u, pages / 4096 + 1}));
    std::vector<std::thread> th;
    for (uint64_t t = 0; t < nt; ++t)
        th.emplace_back([=]() {   // a contiguous range per thread: page-table locks are per range
            const uint64_t lo = pages * t / nt, hi = pages * (t + 1) / nt;
            for (uint64_t i = lo; i < hi; ++i) reinterpret_cast<volatile uint8_t*>(p)[i * kPage] = 0;
        });
    for (std::thread& x : th) x.join();
}

void server_release(pose_gpu_server* s) {
    for (int i = 0; i < s->n_streams; ++i) {
        StreamSlot& S = s->slots[i];
        if (S.h_jobs) (void)hipHostFree(S.h_jobs);
        if (S.d_jobs) (void)hipFree(S.d_jobs);
        if (S.d_counter) (void)hipFree(S.d_counter);
        if (S.ev_k0) (void)hipEventDestroy(S.ev_k0);
        if (S.ev_k1) (void)hipEventDestroy(S.ev_k1);
        if (S.ev_done) (void)hipEventDestroy(S.ev_done);
        if (S.stream) (void)hipStreamDestroy(S.stream);
    }
    if (s->d_arena) {
        if (zero_device(s->d_arena, s->arena.size(), 0) == hipSuccess) (void)hipDeviceSynchronize();
        (void)hipFree(s->d_arena);
    }
    if (s->h_pinned) (void)hipHostFree(s->h_pinned);
    std::free(s->bounce);
    if (s->d_topos) (void)hipFree(s->d_topos);
    for (auto* t : s->topos) if (t) pose_graph_hbm_topo_free(t);
    delete s;
}

}  // namespace

extern "C" pose_gpu_server_t* pose_gpu_server_create(const pose_gpu_server_config_t* cfg_in) {
    static const char* W = "pose_gpu_server_create";
    pose_gpu_server* s = new (std::nothrow) pose_gpu_server();
    if (!s) return nullptr;
    if (cfg_in) s->cfg = *cfg_in;
    else s->cfg.sink_fd = -1;
    pose_gpu_server_config_t& c = s->cfg;
    size_t free_b = 0, total_b = 0;
    if (hipGetDevice(&s->device) != hipSuccess ||
        hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess ||
        hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        std::fprintf(stderr, "%s: no HIP device\n", W);
        delete s;
        return nullptr;
    }
    if (const char* e = std::getenv("POSE_GPU_SERVER_FAIL_RETIRE")) s->fail_retire_at = std::strtoull(e, nullptr, 0);
    if (!c.max_inflight || c.max_inflight > (uint32_t)kMaxStreams) c.max_inflight = kMaxStreams;
    if (!c.waves_per_simd) c.waves_per_simd = 2;
    // A third of free HBM: room for several full launches, and small enough
    // that a server created right after another process's does not wait for
    // the driver to reclaim that one's arena (with half of HBM the allocation
    // then took about 4 s instead of 0.1 s).
    if (!c.hbm_budget_bytes) c.hbm_budget_bytes = free_b / 3;
    if (!c.pinned_bytes) c.pinned_bytes = 1ull << 30;
    c.hbm_budget_bytes = round_up(c.hbm_budget_bytes, kAlign);
    c.pinned_bytes = round_up(c.pinned_bytes, kAlign);
    const int per_cu = std::max(1, max_blocks_per_cu(1));
    s->slot_cap = (uint32_t)s->cus * std::min<uint32_t>((uint32_t)per_cu, 4 * c.waves_per_simd);
    // Default cap: four super-chunks per CU per launch.  Smaller launches
    // release their output sooner and leave room for launches of other
    // chunk_blocks beside them; with Poisson arrivals at 64 req/s this served
    // 2,777-2,872 MiB/s at p99 8.8-9.3 s, against 2,632 MiB/s at p99 11.1 s
    // with two super-chunks per resident workgroup (about 16 per CU).
    s->max_jobs = c.max_super_per_launch ? c.max_super_per_launch : 4 * (uint32_t)s->cus;
    s->n_streams = (int)c.max_inflight;
    s->arena.reset(c.hbm_budget_bytes);
    s->pinned.reset(c.pinned_bytes);
    if (hipMalloc(&s->d_arena, c.hbm_budget_bytes) != hipSuccess) {
        std::fprintf(stderr, "%s: arena of %llu bytes does not fit\n", W, (unsigned long long)c.hbm_budget_bytes);
        server_release(s);
        return nullptr;
    }
    if (hipHostMalloc(&s->h_pinned, c.pinned_bytes, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&s->d_topos, 3 * kTopoSlots * sizeof(DevTopo)) != hipSuccess) {
        std::fprintf(stderr, "%s: pinned/table allocation failed\n", W);
        server_release(s);
        return nullptr;
    }
    prefault_host(s->h_pinned, c.pinned_bytes);
    for (int i = 0; i < s->n_streams; ++i) {
        StreamSlot& S = s->slots[i];
        if (hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking) != hipSuccess ||
            hipHostMalloc(&S.h_jobs, 2 * (size_t)s->max_jobs * sizeof(DevJob), hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&S.d_jobs, 2 * (size_t)s->max_jobs * sizeof(DevJob)) != hipSuccess ||
            hipMalloc(&S.d_counter, 256) != hipSuccess ||
            hipEventCreate(&S.ev_k0) != hipSuccess || hipEventCreate(&S.ev_k1) != hipSuccess ||
            hipEventCreate(&S.ev_done) != hipSuccess) {
            std::fprintf(stderr, "%s: stream setup failed\n", W);
            server_release(s);
            return nullptr;
        }
    }
    for (int k = 0; k <= 20; ++k) {
        if ((c.preload_mask >> k) & 1u) {
            if (ensure_topology(s, 1u << k) != 0) {
                std::fprintf(stderr, "%s: topology 2^%d failed\n", W, k);
                server_release(s);
                return nullptr;
            }
        }
    }
    s->sched = std::thread(scheduler_loop, s);
    s->writer = std::thread(writer_loop, s);
    return s;
}

static int server_submit(pose_gpu_server_t* s, const pose_gpu_request_t* req, uint64_t* id) {
    if (!s || !req) return -1;
    const uint32_t cb = req->chunk_blocks ? req->chunk_blocks : 4096;
    if ((cb & (cb - 1)) != 0 || cb > (1u << 20)) {
        std::fprintf(stderr, "pose_gpu_server_submit: chunk_blocks must be a power of two <= 2^20\n");
        return -1;
    }
    if (!req->session_seed && req->seed_len) return -1;
    if (s->cfg.sink_fd >= 0 && ((uint64_t)cb * 32) % 512 != 0) {
        std::fprintf(stderr, "pose_gpu_server_submit: super-chunk not sector aligned for the sink\n");
        return -1;
    }
    auto r = std::make_shared<Request>();
    r->seed.assign(req->session_seed, req->session_seed + req->seed_len);
    r->cb = cb;
    r->rbo = req->region_block_offset;
    r->n_super = req->n_super;
    r->sink_offset = req->sink_offset;
    r->host_out = req->host_out;
    r->topo_index = (uint32_t)topo_index_of(cb);
    // Topologies are built on first use, before the request is queued and
    // outside the server lock, so running launches are never held back.
    if (ensure_topology(s, cb) != 0) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->error) return -1;
    {
        // One super-chunk must fit on its own, or the request could never run.
        const uint64_t one = round_up(slot_region_bytes(s->topos[r->topo_index]), kAlign) +
                             (uint64_t)cb * 32;
        if (one > s->arena.size() || (uint64_t)cb * 32 > s->pinned.size()) {
            std::fprintf(stderr, "pose_gpu_server_submit: one 2^%u super-chunk needs %llu device bytes, "
                         "more than the server budget\n", r->topo_index, (unsigned long long)one);
            return -1;
        }
    }
    r->id = s->next_id++;
    r->t_submit = now_s();
    if (s->t_first == 0) s->t_first = r->t_submit;
    if (id) *id = r->id;
    if (r->n_super == 0) return 0;
    s->live[r->id] = r;
    s->pending.push_back(r);
    s->cv_sched.notify_one();
    return 0;
}

extern "C" int pose_gpu_server_submit(pose_gpu_server_t* s, const pose_gpu_request_t* req, uint64_t* id) {
    try {
        return server_submit(s, req, id);
    } catch (const std::exception& e) {   // host allocation failure: nothing escapes the C ABI
        std::fprintf(stderr, "pose_gpu_server_submit: %s\n", e.what());
        return -1;
    }
}

// Both return once the request (every request) has settled: complete, or
// failed with none of its super-chunks still in a launch.
extern "C" int pose_gpu_server_wait(pose_gpu_server_t* s, uint64_t id) {
    if (!s) return -1;
    std::unique_lock<std::mutex> lk(s->mu);
    while (s->live.find(id) != s->live.end()) s->cv_done.wait_for(lk, std::chrono::milliseconds(50));
    if (id == 0 || id >= s->next_id) return -1;
    return std::find(s->failed_ids.begin(), s->failed_ids.end(), id) != s->failed_ids.end() ? -1 : 0;
}

extern "C" int pose_gpu_server_drain(pose_gpu_server_t* s) {
    if (!s) return -1;
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->live.empty()) s->cv_done.wait_for(lk, std::chrono::milliseconds(50));
    return (s->error || !s->failed_ids.empty()) ? -1 : 0;
}

extern "C" int pose_gpu_server_stats(pose_gpu_server_t* s, pose_gpu_server_stats_t* st) {
    if (!s || !st) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(st, 0, sizeof *st);
    st->requests_done = s->req_done;
    st->super_chunks_done = s->sc_done;
    st->bytes_done = s->bytes_done;
    st->elapsed_s = s->t_last > s->t_first ? s->t_last - s->t_first : 0;
    st->mib_per_s = st->elapsed_s > 0 ? (double)s->bytes_done / st->elapsed_s / 1048576.0 : 0;
    st->p50_ms = percentile(s->lat_ms, 0.50);
    st->p99_ms = percentile(s->lat_ms, 0.99);
    st->max_ms = s->lat_ms.empty() ? 0 : *std::max_element(s->lat_ms.begin(), s->lat_ms.end());
    double sum = 0;
    for (double v : s->lat_ms) sum += v;
    st->mean_ms = s->lat_ms.empty() ? 0 : sum / (double)s->lat_ms.size();
    st->peak_inflight = s->peak_inflight;
    st->peak_hbm_bytes = s->peak_hbm;
    st->memory_waits = s->memory_waits;
    st->memory_cut_launches = s->memory_cut;
    st->launches = s->launches;
    st->kernel_ms_total = s->kernel_ms;
    st->copy_ms_total = s->copy_ms;
    st->sink_ms_total = s->sink_ms;
    return 0;
}

extern "C" void pose_gpu_server_destroy(pose_gpu_server_t* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_sched.notify_all();
    s->cv_writer.notify_all();
    if (s->sched.joinable()) s->sched.join();
    if (s->writer.joinable()) s->writer.join();
    server_release(s);
}
