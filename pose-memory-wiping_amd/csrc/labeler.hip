// labeler.hip — MI355X labeling kernels and the C ABI of include/gpu_label.h.
//
// Data layout in HBM (per topology, seed-independent, uploaded once):
//   rec[s*lanes+t]  uint4    {dst, pred0, pred1, id<<2 | np} of lane t in step s
//                            (np == 3: idle lane); locations are byte offsets
//                            into the workgroup's slot region
//   cv0[s*lanes+t]  2*uint4  chaining value after block 0 (depends only on id, np),
//                            words in the order 0 4 1 5 | 2 6 3 7, so that words
//                            j and 4+j (a four-lane quad's lane j) are adjacent
//   near[s*lanes+t] uint32   lane of step s-1 that wrote pred0 (bits 0-15) and
//                            pred1 (16-31), 0xFFFF where it was written earlier
// Per super-chunk in flight: a slot region holding a scratch pool of `pool`
// 32-byte slots, then the super-chunk's labels by rank, which later nodes read
// back from there; the labels are copied to the caller's buffer at the end.
//
// Kernel: one workgroup of 64*W lanes labels one super-chunk at a time, one
// step per iteration (a step holds at most 64*W independent nodes), with a
// workgroup barrier between steps.  Workgroups are persistent and pull
// super-chunks from a device work counter, longest topology first, so a single
// launch covers a batch of sessions of different chunk_blocks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gpu_label.h"
#include "b3_host.h"
#include "blake3_dev.h"
#include "label_quad.h"
#include "labeler_internal.h"
#include "topology.h"

using namespace pose;
using namespace pose::internal;

static constexpr uint32_t kLanesPerNodeAuto = POSE_GPU_LANES_AUTO;
// Independent one-wave job runners per workgroup for 64-lane rows.
static constexpr uint32_t kRunnersPerWorkgroup = 4;

// ── kernels ───────────────────────────────────────────────────────────────

__global__ void k_node_single(uint32_t* out, uint64_t node_index, uint32_t is_source,
                              const uint32_t* __restrict__ in /* p0[8] p1[8] seed[8] desc[8] ksrc[8] kint[8] */) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t cv[8], m[16];
    for (int i = 0; i < 8; ++i) cv[i] = is_source ? in[32 + i] : in[40 + i];
    dev::block0_words(node_index, m);
    dev::compress(cv, m, dev::kFlagKeyed | dev::kFlagChunkStart);
    for (int i = 0; i < 16; ++i) m[i] = is_source ? 0u : in[i];
    dev::compress(cv, m, dev::kFlagKeyed);
    for (int i = 0; i < 16; ++i) m[i] = in[16 + i];
    dev::compress(cv, m, dev::kFlagKeyed | dev::kFlagChunkEnd | dev::kFlagRoot);
    for (int i = 0; i < 8; ++i) out[i] = cv[i];
}

// Block-0 chaining value of every scheduled node: keyed compression of
// be64(id + id_add) with the source or internal key.  Seed-independent, so it
// is part of the topology and the labeling kernel runs two compressions per
// node.  id_add > 0 gives the CVs of a copy whose ids are these records' ids
// shifted (copy 1 of mirrored copies, from copy 0's records).
__global__ void __launch_bounds__(256)
k_cv0(const uint4* __restrict__ rec, uint4* __restrict__ cv0, uint64_t count,
      const uint32_t* __restrict__ keys /* ksrc[8] kint[8] */, uint32_t id_add) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const uint32_t w = rec[i].w;
    const uint32_t id = (w >> 2) + id_add, np = w & 3;
    if (np == 3) { cv0[2 * i] = make_uint4(0, 0, 0, 0); cv0[2 * i + 1] = cv0[2 * i]; return; }
    uint32_t cv[8], m[16];
    const uint32_t* k = np == 0 ? keys : keys + 8;
#pragma unroll
    for (int j = 0; j < 8; ++j) cv[j] = k[j];
    dev::block0_words(id, m);
    dev::compress(cv, m, dev::kFlagKeyed | dev::kFlagChunkStart);
    cv0[2 * i] = make_uint4(cv[0], cv[4], cv[1], cv[5]);
    cv0[2 * i + 1] = make_uint4(cv[2], cv[6], cv[3], cv[7]);
}

// Host locations of a schedule's records (slot index, or kPersistBit | rank)
// -> byte offsets in a slot region: slot k at 32k, output rank r at
// out_off + 32 (r - rank_lo).  The node word (w) is kept.
__global__ void __launch_bounds__(256)
k_rec_offsets(uint4* __restrict__ rec, uint64_t count, uint32_t out_off, uint32_t rank_lo) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    uint4 r = rec[i];
    auto off = [&](uint32_t loc) {
        return (loc & kPersistBit) ? out_off + ((loc & ~kPersistBit) - rank_lo) * 32 : loc * 32;
    };
    r.x = off(r.x);
    r.y = off(r.y);
    r.z = off(r.z);
    rec[i] = r;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

using dev::gcu4;
using dev::gu4;
using dev::u32x4;

template <int W>
__device__ __forceinline__ void step_barrier() {
    if constexpr (W == 1) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        __syncthreads();
    }
}

// Job runners.  G == 1: the workgroup's 64*W lanes label one job at a time
// together.  G > 1 (only with W == 1): each of the workgroup's G waves is an
// independent job runner with its own slot region, synchronised only within
// itself.  One-wave workgroups let the dispatcher place two of them on one
// SIMD while another SIMD of the CU idles (measured on the first large launch
// after a topology build: 107 of 1,024 SIMDs doubled, their second wave
// finishing at 2.72 s against 1.43 s); a workgroup of four waves puts one on
// each SIMD of its CU.
template <int W, int G>
__device__ __forceinline__ void job_barrier() {
    if constexpr (G == 1) __syncthreads();
    else step_barrier<1>();
}

// Claims the next job: lane 0 takes it from the work counter and the job
// runner's lanes all see it.
template <int W, int G>
__device__ __forceinline__ uint32_t claim_job(uint32_t* job_counter, uint32_t* s_job, uint32_t lane) {
    if constexpr (G == 1) {
        if (lane == 0) *s_job = atomicAdd(job_counter, 1u);
        __syncthreads();
        const uint32_t j = __builtin_amdgcn_readfirstlane(*s_job);
        __syncthreads();
        return j;
    } else {
        uint32_t j = 0;
        if (lane == 0) j = atomicAdd(job_counter, 1u);
        return __builtin_amdgcn_readlane(j, 0);
    }
}

// Persistent labeler.  Workgroup b owns the slot region [b*slot_bytes,
// (b+1)*slot_bytes): scratch slots, then the labels of the super-chunk in hand
// (device locations are byte offsets in this region, so every label access is
// the region base in SGPRs plus a 32-bit lane offset, no address arithmetic).
// Slot 0 is kept zero (absent predecessors read it), slot 1 absorbs idle
// lanes' results.  Per step every lane runs the same straight-line code: load
// two predecessor labels, fetch the next row's record and block-0 CV, run the
// block-1 and block-2 compressions, store the label, barrier.  After the last
// step the workgroup copies the labels to the caller's buffer.
template <int W, int G>
__global__ void __launch_bounds__(64 * W * G)
k_label(const DevTopo* __restrict__ topos, const DevJob* __restrict__ jobs, uint32_t n_jobs,
        uint32_t* __restrict__ job_counter, uint8_t* __restrict__ scratch, uint64_t slot_bytes,
        uint32_t n_slots) {
    static_assert(G == 1 || W == 1, "independent waves only for 64-lane rows");
    constexpr uint32_t P = 64 * W;
    __shared__ uint32_t s_job;
    const uint32_t tid = threadIdx.x % P;
    const uint32_t slot = blockIdx.x * G + (G > 1 ? __builtin_amdgcn_readfirstlane(threadIdx.x / 64) : 0);
    if (G > 1 && slot >= n_slots) return;
    GAS uint8_t* const sbase = (GAS uint8_t*)(scratch + (uint64_t)slot * slot_bytes);
    if (tid < 2) ((gu4)sbase)[tid] = u32x4{0, 0, 0, 0};

    for (;;) {
        const uint32_t j = claim_job<W, G>(job_counter, &s_job, tid);
        if (j >= n_jobs) break;

        const DevJob& J = jobs[j];
        const DevTopo& T = topos[J.topo];
        // Row pointers are uniform (SGPRs); lanes index them by tid.
        const GAS u32x4* rec = (const GAS u32x4*)T.rec;
        const GAS u32x4* cv0 = (const GAS u32x4*)T.cv0;
        const uint32_t n_steps = T.n_steps;
        uint32_t tail[16];
#pragma unroll
        for (int k = 0; k < 8; ++k) { tail[k] = J.seed[k]; tail[8 + k] = T.desc[k]; }

        // One step: labels the row whose record/CV are in (r, c0, c1) and
        // fetches the next row into (rn, cn0, cn1).  Two steps per iteration
        // with the register sets swapped, so no loop-carried moves remain.
        auto step = [&](const u32x4& r, const u32x4& c0, const u32x4& c1, u32x4& rn, u32x4& cn0, u32x4& cn1) {
            gcu4 qa = (gcu4)(sbase + r.y);
            gcu4 qb = (gcu4)(sbase + r.z);
            const u32x4 a0 = qa[0], a1 = qa[1], b0 = qb[0], b1 = qb[1];
            // Keep the predecessor loads ahead of the next-row fetch: the wait
            // for the predecessors then leaves the (HBM-latency) fetch in flight.
            __builtin_amdgcn_sched_barrier(0);
            rec += P;
            cv0 += 2 * P;
            rn = rec[tid];
            cn0 = cv0[2 * tid];
            cn1 = cv0[2 * tid + 1];
            __builtin_amdgcn_sched_barrier(0);
            uint32_t m[16] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w,
                              b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
            uint32_t cv[8] = {c0.x, c0.z, c1.x, c1.z, c0.y, c0.w, c1.y, c1.w};   // stored 0 4 1 5 | 2 6 3 7
            dev::compress(cv, m, dev::kFlagKeyed);
            dev::compress(cv, tail, dev::kFlagKeyed | dev::kFlagChunkEnd | dev::kFlagRoot);
            // The next row's registers are consumed only after the hash, so the
            // fetch is never waited for inside it.
            __builtin_amdgcn_sched_barrier(0);
            gu4 d = (gu4)(sbase + r.x);
            d[0] = u32x4{cv[0], cv[1], cv[2], cv[3]};
            d[1] = u32x4{cv[4], cv[5], cv[6], cv[7]};
            step_barrier<W>();
        };
        u32x4 ra = rec[tid], c0a = cv0[2 * tid], c1a = cv0[2 * tid + 1];
        u32x4 rb, c0b, c1b;
        uint32_t s = 0;
        for (; s + 2 <= n_steps; s += 2) {
            step(ra, c0a, c1a, rb, c0b, c1b);
            step(rb, c0b, c1b, ra, c0a, c1a);
        }
        if (s < n_steps) step(ra, c0a, c1a, rb, c0b, c1b);
        job_barrier<W, G>();
        // The finished labels, from the slot region to the caller's buffer.
        const gcu4 src = (gcu4)(sbase + T.out_off);
        const gu4 dst = (gu4)J.persist;
        const uint32_t n16 = T.out16;
        for (uint32_t i = tid; i < n16; i += P) dst[i] = src[i];
        job_barrier<W, G>();
    }
}

// One lane per node, software-pipelined, for launches with few resident waves
// per SIMD, where k_label's per-step wait for its predecessor loads is not
// hidden by other waves.  Each step issues the loads of the next step's
// predecessors written two or more steps back (a predecessor of the step just
// before loads the zero slot instead), so they have a whole step to arrive.
// Predecessors written in the step just before come from an LDS copy of that
// step's labels (two step buffers, alternating), at the lane the near table
// names.  Early loads are race-free for the reason given for the latency form:
// a slot read by step c is rewritten no earlier than step c+1.
template <int W, int G>
__global__ void __launch_bounds__(64 * W * G)
k_label_pipe(const DevTopo* __restrict__ topos, const DevJob* __restrict__ jobs, uint32_t n_jobs,
             uint32_t* __restrict__ job_counter, uint8_t* __restrict__ scratch, uint64_t slot_bytes,
             uint32_t n_slots) {
    static_assert(G == 1 || W == 1, "independent waves only for 64-lane rows");
    constexpr uint32_t P = 64 * W;
    __shared__ uint32_t s_job;
    __shared__ u32x4 s_prev[G * 2 * 2 * P];      // [job runner][step parity][lane][half]
    const uint32_t tid = threadIdx.x % P;
    const uint32_t runner = G > 1 ? __builtin_amdgcn_readfirstlane(threadIdx.x / 64) : 0;
    const uint32_t slot = blockIdx.x * G + runner;
    if (G > 1 && slot >= n_slots) return;
    GAS uint8_t* const sbase = (GAS uint8_t*)(scratch + (uint64_t)slot * slot_bytes);
    if (tid < 2) ((gu4)sbase)[tid] = u32x4{0, 0, 0, 0};
    dev::lds_u32x4* const prev = (dev::lds_u32x4*)(u32x4*)s_prev + runner * 4 * P;

    for (;;) {
        const uint32_t j = claim_job<W, G>(job_counter, &s_job, tid);
        if (j >= n_jobs) break;

        const DevJob& J = jobs[j];
        const DevTopo& T = topos[J.topo];
        const gcu4 rec = (gcu4)T.rec;
        const gcu4 cv0 = (gcu4)T.cv0;
        const GAS uint32_t* const nearp = (const GAS uint32_t*)T.near;
        const uint32_t n_steps = T.n_steps;   // rows 0..n_steps exist (the last is idle)
        // Pointers to row s+2 of the records and near entries and to row s+1
        // of the block-0 CVs, advanced one row per step instead of computed
        // from the step index.  Every uploaded schedule is followed by one
        // more row of records and near entries, so the last step's look-ahead
        // needs no bound.
        gcu4 rec2 = rec + 2 * P;
        const GAS uint32_t* near2 = nearp + 2 * P;
        gcu4 cv1 = cv0 + 2 * P;
        uint32_t tail[16];
#pragma unroll
        for (int k = 0; k < 8; ++k) { tail[k] = J.seed[k]; tail[8 + k] = T.desc[k]; }

        // One step.  In: this row's record r, near entry hn, far predecessor
        // words (fa, fb) and block-0 CV (c0, c1); the next row's record rn and
        // near entry hnn.  Out: the same one row later.
        // r and hn are taken by value: the second step of an iteration
        // writes the registers they came from.
        auto step = [&](uint32_t s, const u32x4 r, const uint32_t hn, const u32x4& fa0, const u32x4& fa1,
                        const u32x4& fb0, const u32x4& fb1, const u32x4& c0, const u32x4& c1,
                        const u32x4& rn, uint32_t hnn,
                        u32x4& fa0o, u32x4& fa1o, u32x4& fb0o, u32x4& fb1o, u32x4& c0o, u32x4& c1o,
                        u32x4& rno, uint32_t& hnno) {
            static_assert(kZeroSlot == 0, "the zero slot is at offset 0");
            static_assert(kFarPred == 0xFFFFu, "near entries below 2^15 are lanes, kFarPred has bit 15 set");
            // Bit 15 of each half, sign-extended: all ones for a far
            // predecessor, zero for a near one (whose lane loads the zero slot).
            const uint32_t pa = rn.y & (uint32_t)__builtin_amdgcn_sbfe((int)hnn, 15, 1);
            const uint32_t pb = rn.z & (uint32_t)((int)hnn >> 31);
            fa0o = ((gcu4)(sbase + pa))[0];
            fa1o = ((gcu4)(sbase + pa))[1];
            fb0o = ((gcu4)(sbase + pb))[0];
            fb1o = ((gcu4)(sbase + pb))[1];
            __builtin_amdgcn_sched_barrier(0);
            rno = rec2[tid];
            hnno = near2[tid];
            c0o = cv1[2 * tid];
            c1o = cv1[2 * tid + 1];
            rec2 += P;
            near2 += P;
            cv1 += 2 * P;
            __builtin_amdgcn_sched_barrier(0);
            // Predecessors written in the step before, from its LDS copy.
            const dev::lds_u32x4* const pv = prev + ((s + 1) & 1) * 2 * P;
            const uint32_t ha = hn & 0xFFFFu, hb = hn >> 16;
            const bool na = ha != kFarPred, nb = hb != kFarPred;
            // Near lanes read under an execution mask into the registers the
            // far words arrived in: no per-word selects.
            u32x4 a0 = fa0, a1 = fa1, b0 = fb0, b1 = fb1;
            if (na) { a0 = pv[2 * ha]; a1 = pv[2 * ha + 1]; }
            if (nb) { b0 = pv[2 * hb]; b1 = pv[2 * hb + 1]; }
            uint32_t m[16] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w,
                              b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
            uint32_t cv[8] = {c0.x, c0.z, c1.x, c1.z, c0.y, c0.w, c1.y, c1.w};   // stored 0 4 1 5 | 2 6 3 7
            dev::compress(cv, m, dev::kFlagKeyed);
            dev::compress(cv, tail, dev::kFlagKeyed | dev::kFlagChunkEnd | dev::kFlagRoot);
            __builtin_amdgcn_sched_barrier(0);
            const u32x4 lo = u32x4{cv[0], cv[1], cv[2], cv[3]}, hi = u32x4{cv[4], cv[5], cv[6], cv[7]};
            gu4 d = (gu4)(sbase + r.x);
            d[0] = lo;
            d[1] = hi;
            dev::lds_u32x4* const po = prev + (s & 1) * 2 * P + 2 * tid;
            po[0] = lo;
            po[1] = hi;
            step_barrier<W>();
        };

        // Row 0 (no near predecessors) with its loads, and row 1's record.
        const u32x4 r0 = rec[tid];
        u32x4 fa0A = ((gcu4)(sbase + r0.y))[0], fa1A = ((gcu4)(sbase + r0.y))[1];
        u32x4 fb0A = ((gcu4)(sbase + r0.z))[0], fb1A = ((gcu4)(sbase + r0.z))[1];
        u32x4 c0A = cv0[2 * tid], c1A = cv0[2 * tid + 1];
        u32x4 rA = r0, rnA = (rec + P)[tid];
        uint32_t hnA = 0xFFFFFFFFu, hnnA = (nearp + P)[tid];
        u32x4 fa0B, fa1B, fb0B, fb1B, c0B, c1B, rnB;
        uint32_t hnnB;
        // Drain these loads here, so that the loop's waits for them do not
        // also wait for label stores.
        __builtin_amdgcn_s_waitcnt(0);
        uint32_t s = 0;
        for (; s + 2 <= n_steps; s += 2) {
            step(s, rA, hnA, fa0A, fa1A, fb0A, fb1A, c0A, c1A, rnA, hnnA,
                 fa0B, fa1B, fb0B, fb1B, c0B, c1B, rnB, hnnB);
            step(s + 1, rnA, hnnA, fa0B, fa1B, fb0B, fb1B, c0B, c1B, rnB, hnnB,
                 fa0A, fa1A, fb0A, fb1A, c0A, c1A, rnA, hnnA);
            rA = rnB;     // row s+2; rnA / hnnA now hold row s+3
            hnA = hnnB;
        }
        if (s < n_steps)
            step(s, rA, hnA, fa0A, fa1A, fb0A, fb1A, c0A, c1A, rnA, hnnA,
                 fa0B, fa1B, fb0B, fb1B, c0B, c1B, rnB, hnnB);
        job_barrier<W, G>();
        // The finished labels, from the slot region to the caller's buffer.
        const gcu4 src = (gcu4)(sbase + T.out_off);
        const gu4 dst = (gu4)J.persist;
        const uint32_t n16 = T.out16;
        for (uint32_t i = tid; i < n16; i += P) dst[i] = src[i];
        job_barrier<W, G>();
    }
}

// Latency form: four lanes per node (the per-super-chunk work is
// quad_label_superchunk in label_quad.h).  Workgroups are persistent and pull
// super-chunks from the work counter like k_label.
template <int W>
__global__ void __launch_bounds__(64 * W)
k_label_quad(const DevTopo* __restrict__ topos, const DevJob* __restrict__ jobs, uint32_t n_jobs,
             uint32_t* __restrict__ job_counter, uint8_t* __restrict__ scratch, uint64_t slot_bytes) {
    __shared__ uint32_t s_job;
    __shared__ dev::QuadShared<W> S;
    const uint32_t tid = threadIdx.x;
    GAS uint8_t* const sbase = (GAS uint8_t*)(scratch + (uint64_t)blockIdx.x * slot_bytes);
    if (tid < 2) ((gu4)sbase)[tid] = u32x4{0, 0, 0, 0};
    dev::QuadLane c;
    dev::quad_lane_init<W>(S, tid, c);
    for (;;) {
        if (tid == 0) s_job = atomicAdd(job_counter, 1u);
        __syncthreads();
        const uint32_t jb = __builtin_amdgcn_readfirstlane(s_job);
        if (jb >= n_jobs) break;
        const DevJob& J = jobs[jb];
        dev::quad_label_superchunk<W>(S, c, topos[J.topo], J.seed, J.persist, sbase, tid);
    }
}

// ── host side ─────────────────────────────────────────────────────────────

namespace {

bool quiet() {
    static int q = -1;
    if (q < 0) {
        const char* e = std::getenv("POSE_GPU_QUIET");
        q = (e && e[0] && e[0] != '0') ? 1 : 0;
    }
    return q == 1;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Upper bound on a plan's scratch from POSE_GPU_SCRATCH_LIMIT (bytes, read at
// each plan creation); unset or 0 means no bound beyond free HBM.
uint64_t scratch_limit() {
    const char* e = std::getenv("POSE_GPU_SCRATCH_LIMIT");
    if (!e || !*e) return UINT64_MAX;
    const unsigned long long v = std::strtoull(e, nullptr, 0);
    return v ? (uint64_t)v : UINT64_MAX;
}

#define HIP_OK(call, what)                                                              \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "%s: %s failed: %s\n", what, #call, hipGetErrorString(e_)); \
            goto fail;                                                                  \
        }                                                                               \
    } while (0)

}  // namespace

int pose::internal::lanes_to_waves(uint32_t lanes) {
    switch (lanes) {
        case 64: return 1;
        case 128: return 2;
        case 256: return 4;
        case 512: return 8;
        default: return 0;
    }
}

// Device location (byte offset in the slot region, outputs from out_off on)
// -> host location (slot index, or kPersistBit | rank).  k_rec_offsets does
// the forward conversion.
static uint32_t host_loc(uint32_t off, uint32_t out_off) {
    return off >= out_off ? kPersistBit | ((off - out_off) / 32) : off / 32;
}

static void topo_release_device(pose_hbm_topo* t) {
    for (void* p : {(void*)t->d_rec, (void*)t->d_cv0, (void*)t->d_near,
                    (void*)t->d_prec, (void*)t->d_pcv0, (void*)t->d_pnear})
        if (p) (void)hipFree(p);
    t->d_rec = nullptr; t->d_cv0 = nullptr; t->d_near = nullptr;
    t->d_prec = nullptr; t->d_pcv0 = nullptr; t->d_pnear = nullptr;
}

// Copy schedules are built for 64-lane topologies, the width the four-lane
// form runs at one wave per SIMD (wider steps do not shorten a copy's
// 393,197-step critical path at 2^17, and would double the device footprint).
static bool wants_parts(uint32_t lanes) { return lanes == 64; }

// Uploads schedules' records, near tables and block-0 CVs; rows of `sc`
// schedules back to back.  The records go up as the host holds them and are
// converted on the device (k_rec_offsets).  The host arrays are read by the
// copies, so the call waits for its stream before it returns.  With `mirror`
// the records serve two copies whose ids differ by mirror_ids: the CVs are
// computed twice over them, the second time with shifted ids, into a CV array
// of twice the positions.
static hipError_t upload_schedules(const Schedule* const* sc, int n_sc, uint32_t lanes, const uint32_t* d_keys_src,
                                   bool mirror, uint32_t mirror_ids, uint4** d_rec, uint4** d_cv0,
                                   uint32_t** d_near, uint64_t* positions, hipStream_t st) {
    uint64_t N = 0;
    for (int i = 0; i < n_sc; ++i) N += (uint64_t)(sc[i]->n_steps + 1) * lanes;
    *positions = N;
    const uint64_t n_cv = mirror ? 2 * N : N;
    hipError_t e;
    // Records and near entries carry one zero row past the last schedule: the
    // four-lane form loads row s+2 during step s, and its last step's load
    // (row n_steps+1, never used) must stay in bounds.
    if ((e = hipMalloc(d_rec, (N + lanes) * 16)) != hipSuccess) return e;
    if ((e = hipMalloc(d_cv0, n_cv * 32)) != hipSuccess) return e;
    if ((e = hipMalloc(d_near, (N + lanes) * 4)) != hipSuccess) return e;
    if ((e = zero_device(*d_rec + N, (uint64_t)lanes * 16, st)) != hipSuccess) return e;
    if ((e = zero_device(*d_near + N, (uint64_t)lanes * 4, st)) != hipSuccess) return e;
    uint64_t at = 0;
    for (int i = 0; i < n_sc; ++i) {
        const Schedule& S = *sc[i];
        const uint64_t P = S.rec.size() / 4;
        if ((e = hipMemcpyAsync(*d_rec + at, S.rec.data(), P * 16, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(*d_near + at, S.near.data(), P * 4, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
        k_rec_offsets<<<(unsigned)((P + 255) / 256), 256, 0, st>>>(*d_rec + at, P, slot_out_off(S.pool), S.rank_lo);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        at += P;
    }
    k_cv0<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(*d_rec, *d_cv0, N, d_keys_src, 0u);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (mirror) {
        k_cv0<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(*d_rec, *d_cv0 + 2 * N, N, d_keys_src, mirror_ids);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipStreamSynchronize(st);
}

extern "C" pose_hbm_topo_t* pose_gpu_topo_build_host(uint32_t chunk_blocks, uint32_t lanes, int mode) {
    if (chunk_blocks == 0) chunk_blocks = 4096;
    std::unique_ptr<pose_hbm_topo> t(new (std::nothrow) pose_hbm_topo());
    if (!t) return nullptr;
    std::string err;
    int rc = -1;
    try {
        rc = build_host_topology(chunk_blocks, lanes, mode, t->host, &err, wants_parts(lanes));
    } catch (const std::exception&) {   // host allocation failure: nothing escapes the C ABI
        err = "out of host memory";
    }
    if (rc != 0) {
        std::fprintf(stderr, "pose_gpu_topo_build_host: %s\n", err.c_str());
        return nullptr;
    }
    return t.release();
}

extern "C" void pose_gpu_topo_free_host(pose_hbm_topo_t* t) {
    if (t && !t->on_device) delete t;
}

extern "C" pose_hbm_topo_t* pose_gpu_topo_build_ex(uint32_t chunk_blocks, uint32_t lanes, int mode) {
    if (chunk_blocks == 0) chunk_blocks = 4096;
    if (chunk_blocks > (1u << 20)) {
        std::fprintf(stderr, "pose_graph_hbm_topo_build: chunk_blocks %u out of range\n", chunk_blocks);
        return nullptr;
    }
    if (lanes_to_waves(lanes) == 0) {
        std::fprintf(stderr, "pose_gpu_topo_build_ex: lanes must be 64, 128, 256 or 512\n");
        return nullptr;
    }
    pose_hbm_topo* t = pose_gpu_topo_build_host(chunk_blocks, lanes, mode);
    if (!t) return nullptr;
    HostTopology& H = t->host;
    hipStream_t st = nullptr;
    uint32_t keys[16];
    uint32_t* d_keys = nullptr;
    uint64_t n_whole = 0;
    Schedule whole;   // view of the whole-scaffold schedule for the upload
    std::memcpy(keys, H.key_src, 32);
    std::memcpy(keys + 8, H.key_int, 32);
    whole.n_steps = H.n_steps;
    whole.pool = H.pool;
    whole.rank_lo = 0;
    whole.rec.swap(H.rec);
    whole.near.swap(H.near);
    // Device-local work only (a private stream, no device-wide synchronisation
    // and no frees), so a build does not wait for labeling already running.
    HIP_OK(hipGetDevice(&t->device), "pose_graph_hbm_topo_build");
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "pose_graph_hbm_topo_build");
    HIP_OK(hipMalloc(&d_keys, 64), "pose_graph_hbm_topo_build");
    HIP_OK(hipMemcpyAsync(d_keys, keys, 64, hipMemcpyHostToDevice, st), "pose_graph_hbm_topo_build");
    {
        const Schedule* one[1] = {&whole};
        HIP_OK(upload_schedules(one, 1, H.lanes, d_keys, false, 0, &t->d_rec, &t->d_cv0, &t->d_near, &n_whole, st),
               "pose_graph_hbm_topo_build");
    }
    (void)hipFree(d_keys);
    d_keys = nullptr;
    (void)hipStreamDestroy(st);
    t->on_device = true;
    // 16-byte record + 32-byte block-0 CV + 4-byte near entry per position.
    t->device_bytes = n_whole * 52;
    // The copy schedules stay on the host until a launch labels a copy on its
    // own (ensure_parts).
    return t;
fail:
    if (d_keys) (void)hipFree(d_keys);
    if (st) (void)hipStreamDestroy(st);
    topo_release_device(t);
    delete t;
    return nullptr;
}

hipError_t pose::internal::ensure_parts(const pose_hbm_topo* t) {
    if (!t->host.has_parts || t->parts_on_device.load(std::memory_order_acquire)) return hipSuccess;
    std::lock_guard<std::mutex> lk(t->parts_mu);
    if (t->parts_on_device.load(std::memory_order_relaxed)) return hipSuccess;
    // Upload on the topology's device; the caller's current device is restored.
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != t->device) (void)hipSetDevice(t->device);
    HostTopology& H = const_cast<HostTopology&>(t->host);
    hipStream_t st = nullptr;
    uint32_t* d_keys = nullptr;
    uint32_t keys[16];
    std::memcpy(keys, H.key_src, 32);
    std::memcpy(keys + 8, H.key_int, 32);
    uint64_t n_parts = 0;
    // Mirrored copies share copy 0's records and near entries (their device
    // locations are equal); only the block-0 CVs are per copy.
    const bool shared = H.part1_shifted;
    const Schedule* two[2] = {&H.part[0], &H.part[1]};
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&d_keys, 64);
    if (e == hipSuccess) e = hipMemcpyAsync(d_keys, keys, 64, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = upload_schedules(two, shared ? 1 : 2, H.lanes, d_keys, shared, H.part1_id_shift, &t->d_prec, &t->d_pcv0,
                             &t->d_pnear, &n_parts, st);
    if (d_keys) (void)hipFree(d_keys);
    if (st) (void)hipStreamDestroy(st);
    if (e == hipSuccess) {
        // Records and near entries (20 B per position) once, block-0 CVs
        // (32 B) per copy.
        t->device_bytes += shared ? n_parts * 20 + 2 * n_parts * 32 : n_parts * 52;
        // The device holds the copy schedules now; keep only their step tables.
        for (Schedule& P : H.part) {
            U32Buf().swap(P.rec);
            U32Buf().swap(P.near);
        }
        t->parts_on_device.store(true, std::memory_order_release);
    } else {
        for (void* p : {(void*)t->d_prec, (void*)t->d_pcv0, (void*)t->d_pnear})
            if (p) (void)hipFree(p);
        t->d_prec = nullptr; t->d_pcv0 = nullptr; t->d_pnear = nullptr;
        std::fprintf(stderr, "pose_gpu: uploading the copy schedules failed: %s\n", hipGetErrorString(e));
    }
    if (prev >= 0 && prev != t->device) (void)hipSetDevice(prev);
    return e;
}

extern "C" pose_hbm_topo_t* pose_graph_hbm_topo_build(uint32_t chunk_blocks, pose_hash_algo_t algo) {
    if (algo != POSE_HASH_BLAKE3) {
        std::fprintf(stderr, "pose_graph_hbm_topo_build: only BLAKE3 supported on GPU\n");
        return nullptr;
    }
    return pose_gpu_topo_build_ex(chunk_blocks, 64, kSchedList);
}

extern "C" void pose_graph_hbm_topo_free(pose_hbm_topo_t* t) {
    if (!t) return;
    topo_release_device(t);
    delete t;
}

extern "C" int pose_gpu_topo_info(const pose_hbm_topo_t* t, pose_gpu_topo_info_t* info) {
    if (!t || !info) return -1;
    const HostTopology& H = t->host;
    info->chunk_blocks = H.cb;
    info->lanes = H.lanes;
    info->mode = H.mode;
    info->levels = H.levels;
    info->steps = H.n_steps;
    info->pool_slots = H.pool;
    info->nodes = H.count;
    info->device_bytes = t->device_bytes.load();
    info->build_seconds = H.build_seconds;
    return 0;
}

extern "C" int pose_gpu_topo_export(const pose_hbm_topo_t* t, uint32_t* step_start, uint32_t* rec) {
    if (!t) return -1;
    const HostTopology& H = t->host;
    if (step_start) std::memcpy(step_start, H.step_start.data(), (size_t)(H.n_steps + 1) * 4);
    const uint64_t positions = (uint64_t)(H.n_steps + 1) * H.lanes;
    if (rec) {
        if (!H.rec.empty()) {
            std::memcpy(rec, H.rec.data(), positions * 16);
        } else if (t->on_device) {
            if (hipMemcpy(rec, t->d_rec, positions * 16, hipMemcpyDeviceToHost) != hipSuccess) return -1;
            const uint32_t out_off = slot_out_off(H.pool);
            for (uint64_t i = 0; i < positions * 4; i += 4)
                for (int k = 0; k < 3; ++k) rec[i + k] = host_loc(rec[i + k], out_off);
        } else {
            return -1;
        }
    }
    return 0;
}

extern "C" int pose_gpu_topo_export_near(const pose_hbm_topo_t* t, uint32_t* near) {
    if (!t || !near) return -1;
    const HostTopology& H = t->host;
    const uint64_t positions = (uint64_t)(H.n_steps + 1) * H.lanes;
    if (!H.near.empty()) {
        std::memcpy(near, H.near.data(), positions * 4);
    } else if (t->on_device) {
        if (hipMemcpy(near, t->d_near, positions * 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    } else {
        return -1;
    }
    return 0;
}

// ── launch ────────────────────────────────────────────────────────────────

namespace {

// Resident job runners per CU of the one-lane form (workgroups of W waves, or
// for W == 1 four one-wave runners per workgroup).
template <int W>
int occupancy_blocks_per_cu() {
    constexpr int G = W == 1 ? kRunnersPerWorkgroup : 1;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_label<W, G>, 64 * W * G, 0) != hipSuccess) return 0;
    return nb * G;
}

template <int W>
int occupancy_blocks_per_cu_quad() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_label_quad<W>, 64 * W, 0) != hipSuccess) return 0;
    return nb;
}

}  // namespace

int pose::internal::max_blocks_per_cu(int waves, bool quad) {
    if (quad) {
        switch (waves) {
            case 1: return occupancy_blocks_per_cu_quad<4>();
            case 2: return occupancy_blocks_per_cu_quad<8>();
            case 4: return occupancy_blocks_per_cu_quad<16>();
        }
        return 0;
    }
    switch (waves) {
        case 1: return occupancy_blocks_per_cu<1>();
        case 2: return occupancy_blocks_per_cu<2>();
        case 4: return occupancy_blocks_per_cu<4>();
        case 8: return occupancy_blocks_per_cu<8>();
    }
    return 0;
}

namespace {

// Launches of the one-lane form with at most this many resident waves per SIMD
// run k_label_pipe.
#ifndef POSE_PIPE_MAX_WPS
#define POSE_PIPE_MAX_WPS 4
#endif
constexpr uint32_t kPipeMaxWavesPerSimd = POSE_PIPE_MAX_WPS;

// Compute units of the current device, cached per device.
int current_cus() {
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 0;
    std::lock_guard<std::mutex> lk(mu);
    if ((size_t)dev >= cache.size()) cache.resize(dev + 1, 0);
    if (!cache[dev]) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess) cache[dev] = cus;
    }
    return cache[dev];
}

}  // namespace

hipError_t pose::internal::launch_label(int waves, uint32_t grid, const DevTopo* topos, const DevJob* jobs,
                        uint32_t n_jobs, uint32_t* counter, uint8_t* scratch, uint64_t slot_bytes,
                        hipStream_t st, bool quad, bool pipe_ok) {
    if (quad) {
        // Topology of 16*Wq nodes per step -> workgroups of Wq wave64s.
        switch (waves) {
            case 1: k_label_quad<4><<<grid, 256, 0, st>>>(topos, jobs, n_jobs, counter, scratch, slot_bytes); break;
            case 2: k_label_quad<8><<<grid, 512, 0, st>>>(topos, jobs, n_jobs, counter, scratch, slot_bytes); break;
            case 4: k_label_quad<16><<<grid, 1024, 0, st>>>(topos, jobs, n_jobs, counter, scratch, slot_bytes); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    // grid counts job runners (slot regions); 64-lane rows run four runners
    // per workgroup.
    constexpr uint32_t G = kRunnersPerWorkgroup;
    const uint32_t wg1 = (grid + G - 1) / G;
    const uint64_t cus = (uint64_t)current_cus();
    if (pipe_ok && (uint64_t)grid * (uint64_t)waves <= cus * 4 * kPipeMaxWavesPerSimd) {
        switch (waves) {
            case 1: k_label_pipe<1, G><<<wg1, 64 * G, 0, st>>>(topos, jobs, n_jobs, counter, scratch, slot_bytes, grid); break;
            case 2: k_label_pipe<2, 1><<<grid, 128, 0, st>>>(topos, jobs, n_jobs, counter, scratch, slot_bytes, grid); break;
            case 4: k_label_pipe<4, 1><<<grid, 256, 0, st>>>(topos, jobs, n_jobs, counter, scratch, slot_bytes, grid); break;
            case 8: k_label_pipe<8, 1><<<grid, 512, 0, st>>>(topos, jobs, n_jobs, counter, scratch, slot_bytes, grid); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (waves) {
        case 1: k_label<1, G><<<wg1, 64 * G, 0, st>>>(topos, jobs, n_jobs, counter, scratch, slot_bytes, grid); break;
        case 2: k_label<2, 1><<<grid, 128, 0, st>>>(topos, jobs, n_jobs, counter, scratch, slot_bytes, grid); break;
        case 4: k_label<4, 1><<<grid, 256, 0, st>>>(topos, jobs, n_jobs, counter, scratch, slot_bytes, grid); break;
        case 8: k_label<8, 1><<<grid, 512, 0, st>>>(topos, jobs, n_jobs, counter, scratch, slot_bytes, grid); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Zero a device range with a kernel (no memset nodes, so this is graph-safe).
hipError_t pose::internal::zero_device(void* p, uint64_t bytes, hipStream_t st) {
    if (!bytes) return hipSuccess;
    uint64_t words = bytes / 4;
    k_fill_u32<<<1024, 256, 0, st>>>(reinterpret_cast<uint32_t*>(p), words, 0u);
    return hipGetLastError();
}

static const Schedule* part_schedule(const pose_hbm_topo* t, uint32_t part) {
    return part == kWhole ? nullptr : &t->host.part[part - kCopy0];
}

uint32_t pose::internal::schedule_steps(const pose_hbm_topo* t, uint32_t part) {
    return part == kWhole ? t->host.n_steps : part_schedule(t, part)->n_steps;
}

uint32_t pose::internal::part_first_rank(const pose_hbm_topo* t, uint32_t part) {
    return part == kWhole ? 0 : part_schedule(t, part)->rank_lo;
}

uint32_t pose::internal::part_end_rank(const pose_hbm_topo* t, uint32_t part) {
    return part == kWhole ? t->host.cb : part_schedule(t, part)->rank_hi;
}

uint64_t pose::internal::slot_region_bytes(const pose_hbm_topo* t, uint32_t part) {
    const uint32_t pool = part == kWhole ? t->host.pool : part_schedule(t, part)->pool;
    const uint32_t outs = part_end_rank(t, part) - part_first_rank(t, part);
    return (uint64_t)slot_out_off(pool) + (uint64_t)outs * 32;
}

void pose::internal::fill_dev_topo(const pose_hbm_topo* t, DevTopo* D, uint32_t part) {
    std::memset(D, 0, sizeof *D);
    const uint32_t lanes = t->host.lanes;
    if (part == kWhole) {
        D->rec = t->d_rec;
        D->cv0 = t->d_cv0;
        D->near = t->d_near;
        D->n_steps = t->host.n_steps;
        D->pool = t->host.pool;
    } else if (t->parts_on_device.load(std::memory_order_acquire)) {
        // Copy 1's rows follow copy 0's (n_steps + 1 rows each); mirrored
        // copies share copy 0's records and near entries, and copy 1's CVs
        // follow copy 0's.
        const bool shared = t->host.part1_shifted;
        const uint64_t row0 = part == kCopy0 ? 0 : (uint64_t)t->host.part[0].n_steps + 1;
        const Schedule& P = *part_schedule(t, part);
        D->rec = t->d_prec + (shared ? 0 : row0 * lanes);
        D->cv0 = t->d_pcv0 + 2 * row0 * lanes;
        D->near = t->d_pnear + (shared ? 0 : row0 * lanes);
        D->n_steps = P.n_steps;
        D->pool = P.pool;
    }
    D->out_off = slot_out_off(D->pool);
    D->out16 = (part_end_rank(t, part) - part_first_rank(t, part)) * 2;
    for (int k = 0; k < 8; ++k) D->desc[k] = b3::load_le(t->host.desc + 4 * k);
}

extern "C" int pose_gpu_topo_part(const pose_hbm_topo_t* t, uint32_t part, pose_gpu_topo_part_t* info,
                                  uint32_t* step_start, uint32_t* rec, uint32_t* near) {
    if (!t || part > 1 || !t->host.has_parts) return -1;
    const Schedule& P = t->host.part[part];
    const uint32_t lanes = t->host.lanes;
    const uint64_t positions = (uint64_t)(P.n_steps + 1) * lanes;
    // Mirrored copies share copy 0's device records and near entries.
    const bool shared = t->host.part1_shifted;
    const uint64_t row0 = part == 0 || shared ? 0 : (uint64_t)t->host.part[0].n_steps + 1;
    if (info) {
        info->first_node = P.id_lo;
        info->end_node = P.id_hi;
        info->first_rank = P.rank_lo;
        info->end_rank = P.rank_hi;
        info->steps = P.n_steps;
        info->pool_slots = P.pool;
    }
    if (step_start) std::memcpy(step_start, P.step_start.data(), (size_t)(P.n_steps + 1) * 4);
    if (rec) {
        if (!P.rec.empty()) {
            std::memcpy(rec, P.rec.data(), positions * 16);
        } else if (t->on_device && t->parts_on_device.load()) {
            if (hipMemcpy(rec, t->d_prec + row0 * lanes, positions * 16, hipMemcpyDeviceToHost) != hipSuccess) return -1;
            // Device locations are offsets in the copy's own slot region, so
            // the same records give either copy's ranks; copy 1's node ids are
            // copy 0's shifted when the records are shared.
            const uint32_t out_off = slot_out_off(P.pool);
            const uint32_t id_add = part == 1 && shared ? t->host.part1_id_shift : 0;
            for (uint64_t i = 0; i < positions * 4; i += 4) {
                for (int k = 0; k < 3; ++k) {
                    const uint32_t off = rec[i + k];
                    rec[i + k] = off >= out_off ? kPersistBit | ((off - out_off) / 32 + P.rank_lo) : off / 32;
                }
                if (rec[i + 3] != kEmptyRecord) rec[i + 3] += id_add << 2;
            }
        } else {
            return -1;
        }
    }
    if (near) {
        if (!P.near.empty()) {
            std::memcpy(near, P.near.data(), positions * 4);
        } else if (t->on_device && t->parts_on_device.load()) {
            if (hipMemcpy(near, t->d_pnear + row0 * lanes, positions * 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        } else {
            return -1;
        }
    }
    return 0;
}

// A prepared batch.  Its sessions are split into launch groups by kernel
// shape (topology width, lanes per node, occupancy request); each group is one
// persistent launch over its own job range, work counter and scratch, and the
// groups run concurrently (side streams forked from the plan stream and joined
// back).  Job table, scratch and counters live on the device for the plan's
// lifetime; running it replays a captured hipGraph (counter reset, then the
// group launches as parallel branches) or enqueues the same work directly.
struct PlanGroup {
    int waves = 1;               // wave64s per node-lane row (topology lanes / 64)
    bool quad = false;           // four lanes per node (latency form)
    uint32_t slots = 0;          // resident workgroups of this launch
    uint32_t job0 = 0;           // first job of the group in the job table
    uint32_t n_jobs = 0;         // workgroup jobs
    uint32_t n_super = 0;        // jobs as submitted (before any split)
    bool split = false;          // super-chunks labeled as two copies
    uint64_t slot_bytes = 0;     // scratch per workgroup (largest pool in the group)
    uint64_t scratch_off = 0;    // byte offset of the group's scratch
    uint64_t depth = 0;          // steps of the group's deepest job
};

struct pose_gpu_plan {
    int device = -1;
    std::vector<PlanGroup> groups;   // deepest first
    uint32_t n_jobs = 0;             // workgroup jobs
    uint32_t n_super = 0;            // jobs as submitted
    uint64_t scratch_bytes = 0;
    uint64_t nodes = 0;
    uint64_t out_bytes = 0;
    DevTopo* d_topos = nullptr;
    DevJob* d_jobs = nullptr;
    uint32_t* d_counter = nullptr;   // one counter per group, 64 bytes apart
    uint8_t* d_scratch = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<hipStream_t> side;   // streams of groups 1..G-1
    std::vector<hipEvent_t> fj;      // fork event, then one join event per side stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

constexpr uint32_t kCounterStride = 16;   // u32 words between group counters

uint64_t counter_bytes(const pose_gpu_plan* p) {
    return std::max<uint64_t>(256, (uint64_t)p->groups.size() * kCounterStride * 4);
}

void plan_release(pose_gpu_plan* p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    for (hipEvent_t e : p->fj) if (e) (void)hipEventDestroy(e);
    for (hipStream_t s : p->side) if (s) (void)hipStreamDestroy(s);
    if (p->d_topos) (void)hipFree(p->d_topos);
    if (p->d_jobs) (void)hipFree(p->d_jobs);
    if (p->d_counter) (void)hipFree(p->d_counter);
    if (p->d_scratch) (void)hipFree(p->d_scratch);
    if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

hipError_t launch_group(const pose_gpu_plan* p, size_t g, hipStream_t st) {
    const PlanGroup& G = p->groups[g];
    // Groups run concurrently: only a plan of one group may take the
    // pipelined kernel, whose waves need more VGPRs.
    return launch_label(G.waves, G.slots, p->d_topos, p->d_jobs + G.job0, G.n_jobs,
                        p->d_counter + g * kCounterStride, p->d_scratch + G.scratch_off, G.slot_bytes,
                        st, G.quad, p->groups.size() == 1);
}

// Counter reset, then every group: group 0 (the deepest) on the plan stream,
// the others on side streams forked after the reset and joined at the end.
hipError_t plan_enqueue(pose_gpu_plan* p, hipStream_t st) {
    hipError_t e = zero_device(p->d_counter, counter_bytes(p), st);
    if (e != hipSuccess) return e;
    const size_t G = p->groups.size();
    if (G > 1 && (e = hipEventRecord(p->fj[0], st)) != hipSuccess) return e;
    if ((e = launch_group(p, 0, st)) != hipSuccess) return e;
    for (size_t g = 1; g < G; ++g) {
        if ((e = hipStreamWaitEvent(p->side[g - 1], p->fj[0], 0)) != hipSuccess) return e;
        if ((e = launch_group(p, g, p->side[g - 1])) != hipSuccess) return e;
        if ((e = hipEventRecord(p->fj[g], p->side[g - 1])) != hipSuccess) return e;
    }
    for (size_t g = 1; g < G; ++g)
        if ((e = hipStreamWaitEvent(st, p->fj[g], 0)) != hipSuccess) return e;
    return hipSuccess;
}

int device_cus(int* dev, const char* who) {
    int cus = 0;
    if (hipGetDevice(dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, *dev) != hipSuccess) {
        std::fprintf(stderr, "%s: no HIP device\n", who);
        return -1;
    }
    return cus > 0 ? cus : 256;
}

// Builds a plan; per_session[i] (if given) is session i's launch shape, else
// `common` applies to all (nullptr: automatic).
pose_gpu_plan* plan_create(const pose_gpu_session_t* sessions, size_t n_sessions,
                           const pose_gpu_launch_t* per_session, const pose_gpu_launch_t* common,
                           int use_graph) {
    static const char* W = "pose_gpu_plan_create";
    if (!sessions && n_sessions) return nullptr;
    int dev = 0;
    if (device_cus(&dev, W) < 0) return nullptr;
    const pose_gpu_launch_t dflt{0, 0, kLanesPerNodeAuto, POSE_GPU_SPLIT_AUTO};
    using Req = LaunchReq;
    using Pending = PendingGroup;
    std::vector<Pending> pend;
    std::vector<const pose_hbm_topo*> tlist;
    uint64_t nodes = 0, out_bytes = 0;
    for (size_t s = 0; s < n_sessions; ++s) {
        const pose_gpu_session_t& S = sessions[s];
        const pose_hbm_topo* t = S.topo;
        if (!t || !t->on_device) { std::fprintf(stderr, "%s: session %zu has no device topology\n", W, s); return nullptr; }
        const pose_gpu_launch_t& L = per_session ? per_session[s] : (common ? *common : dflt);
        if (L.lanes_per_node != 0 && L.lanes_per_node != 1 && L.lanes_per_node != 4 &&
            L.lanes_per_node != kLanesPerNodeAuto) {
            std::fprintf(stderr, "%s: lanes_per_node must be 1 or 4\n", W);
            return nullptr;
        }
        if (L.split > POSE_GPU_SPLIT_ON) {
            std::fprintf(stderr, "%s: split must be POSE_GPU_SPLIT_AUTO, _OFF or _ON\n", W);
            return nullptr;
        }
        if (L.split == POSE_GPU_SPLIT_ON && !t->host.has_parts) {
            std::fprintf(stderr, "%s: split needs a topology with copy schedules (64 lanes)\n", W);
            return nullptr;
        }
        const int w = lanes_to_waves(t->host.lanes);
        if (L.lanes_per_node == 4 && w > 4) {
            std::fprintf(stderr, "%s: four lanes per node needs a topology of at most 256 lanes\n", W);
            return nullptr;
        }
        if (S.n_super == 0) continue;
        if (!S.out_dev) { std::fprintf(stderr, "%s: session %zu has no output buffer\n", W, s); return nullptr; }
        const Req rq{w, L.lanes_per_node == 0 ? 1u : L.lanes_per_node, L.slots, L.waves_per_simd, L.split};
        size_t gi = 0;
        while (gi < pend.size() && !(pend[gi].req == rq)) ++gi;
        if (gi == pend.size()) pend.push_back(Pending{rq, {}});
        Pending& P = pend[gi];
        uint32_t ti;
        auto it = std::find(tlist.begin(), tlist.end(), t);
        if (it == tlist.end()) { ti = (uint32_t)tlist.size(); tlist.push_back(t); }
        else ti = (uint32_t)(it - tlist.begin());
        const uint32_t cb = t->host.cb;
        ChunkSeeder seeder(S.session_seed, S.seed_len);
        for (uint64_t j = 0; j < S.n_super; ++j) {
            DevJob J{};
            J.persist = reinterpret_cast<uint8_t*>(S.out_dev) + j * (uint64_t)cb * 32;
            J.topo = ti;
            J.part = kWhole;
            seeder.seed_words((S.region_block_offset + j * (uint64_t)cb) / cb, J.seed);
            P.jobs.push_back(J);
        }
        nodes += t->host.count * S.n_super;
        out_bytes += (uint64_t)cb * 32 * S.n_super;
    }
    // Shape of the idle group of a plan without work (keeps info() meaningful).
    int idle_waves = 1;
    if (n_sessions) idle_waves = std::max(1, lanes_to_waves(sessions[0].topo->host.lanes));
    const pose_gpu_launch_t& L0 = per_session && n_sessions ? per_session[0] : (common ? *common : dflt);
    return plan_build(pend, tlist, nodes, out_bytes, use_graph, idle_waves, L0.lanes_per_node == 4);
}

}  // namespace

// Fills a fresh plan; false (after printing why) on failure, leaving the
// partly built plan to the caller's plan_release.
static bool plan_fill(pose_gpu_plan* p, int dev, int cus, std::vector<PendingGroup>& pend,
                      const std::vector<const pose_hbm_topo*>& tlist, uint64_t nodes,
                      uint64_t out_bytes, int use_graph, int idle_waves, bool idle_quad) {
    static const char* W = "pose_gpu_plan_create";
    p->device = dev;
    p->nodes = nodes;
    p->out_bytes = out_bytes;
    std::vector<DevJob> jobs;
    uint64_t submitted = 0;
    for (const PendingGroup& P : pend) submitted += P.jobs.size();
    {
        // Resolve each group's form.  A group with few super-chunks per CU of
        // its deepest chunk_blocks cannot fill the device with one lane per
        // node; the four-lane form then finishes sooner (kQuadMaxPerCu).
        struct Resolved {
            PlanGroup g;
            LaunchReq req;
            std::vector<DevJob> jobs;
        };
        auto cost = [&](const DevJob& J) { return (uint64_t)schedule_steps(tlist[J.topo], J.part); };
        std::vector<Resolved> res;
        for (PendingGroup& P : pend) {
            uint64_t deepest = 0, n_deep = 0;
            for (const DevJob& J : P.jobs) deepest = std::max(deepest, cost(J));
            for (const DevJob& J : P.jobs) n_deep += cost(J) == deepest;
            bool quad = P.req.lpn == 4;
            if (P.req.lpn == kLanesPerNodeAuto) quad = P.req.waves <= 1 && n_deep <= (uint64_t)kQuadMaxPerCu * (uint64_t)cus;
            size_t ri = 0;
            while (ri < res.size() && !(res[ri].g.waves == P.req.waves && res[ri].g.quad == quad &&
                                        res[ri].req.slots == P.req.slots && res[ri].req.wps == P.req.wps &&
                                        res[ri].req.split == P.req.split)) ++ri;
            if (ri == res.size()) {
                Resolved R;
                R.g.waves = P.req.waves;
                R.g.quad = quad;
                R.req = P.req;
                res.push_back(std::move(R));
            }
            Resolved& R = res[ri];
            R.jobs.insert(R.jobs.end(), P.jobs.begin(), P.jobs.end());
        }
        for (Resolved& R : res) {
            R.g.n_super = (uint32_t)R.jobs.size();
            // Split: each whole super-chunk becomes one job per copy of its
            // scaffold.  Automatic only for the four-lane form when every copy
            // gets a workgroup, and so a CU, of its own: a copy's critical
            // path is the whole scaffold's, with half the nodes.
            bool parts_ok = true;
            for (const DevJob& J : R.jobs) parts_ok = parts_ok && tlist[J.topo]->host.has_parts;
            bool split = R.req.split == POSE_GPU_SPLIT_ON;
            if (R.req.split == POSE_GPU_SPLIT_AUTO)
                split = R.g.quad && parts_ok && 2 * submitted <= (uint64_t)cus &&
                        (R.req.slots == 0 || R.req.slots >= R.jobs.size());
            if (split && !parts_ok) { std::fprintf(stderr, "%s: split needs copy schedules\n", W); goto fail; }
            if (split) {
                std::vector<DevJob> two;
                two.reserve(2 * R.jobs.size());
                for (const DevJob& J : R.jobs) {
                    if (J.part != kWhole) { two.push_back(J); continue; }
                    for (uint32_t c = 0; c < 2; ++c) {
                        DevJob H = J;
                        H.part = kCopy0 + c;
                        H.persist = J.persist + (uint64_t)part_first_rank(tlist[J.topo], H.part) * 32;
                        two.push_back(H);
                    }
                }
                R.jobs.swap(two);
            }
            R.g.split = split;
            for (const DevJob& J : R.jobs) {
                R.g.depth = std::max(R.g.depth, cost(J));
                R.g.slot_bytes = std::max(R.g.slot_bytes, slot_region_bytes(tlist[J.topo], J.part));
            }
        }
        std::stable_sort(res.begin(), res.end(), [](const Resolved& a, const Resolved& b) { return a.g.depth > b.g.depth; });
        for (Resolved& R : res) {
            // Longest work first: jobs of the deepest schedule start earliest.
            std::vector<uint32_t> idx(R.jobs.size());
            for (uint32_t i = 0; i < idx.size(); ++i) idx[i] = i;
            std::stable_sort(idx.begin(), idx.end(),
                             [&](uint32_t a, uint32_t b) { return cost(R.jobs[a]) > cost(R.jobs[b]); });
            R.g.job0 = (uint32_t)jobs.size();
            R.g.n_jobs = (uint32_t)R.jobs.size();
            for (uint32_t i : idx) jobs.push_back(R.jobs[i]);
            int per_cu = max_blocks_per_cu(R.g.waves, R.g.quad);
            if (per_cu <= 0) per_cu = 1;
            const uint32_t wg_waves = (uint32_t)R.g.waves * (R.g.quad ? 4u : 1u);
            uint32_t want = R.req.slots;
            if (!want) {
                const uint32_t wps = R.req.wps ? R.req.wps : 2;
                const uint32_t per_cu_target = std::max<uint32_t>(1, (4 * wps) / wg_waves);
                want = (uint32_t)cus * std::min<uint32_t>((uint32_t)per_cu, per_cu_target);
            }
            // An automatic split runs every copy at once.
            if (R.g.split && R.req.split == POSE_GPU_SPLIT_AUTO) want = std::max<uint32_t>(want, R.g.n_jobs);
            R.g.slots = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, R.g.n_jobs));
            p->groups.push_back(R.g);
            p->n_super += R.g.n_super;
            if (jobs.size() > 0xFFFFFFF0ull) { std::fprintf(stderr, "%s: too many super-chunks\n", W); goto fail; }
        }
    }
    // Copy jobs need their topology's copy schedules on the device.
    for (const DevJob& J : jobs) {
        if (J.part == kWhole) continue;
        const hipError_t pe = ensure_parts(tlist[J.topo]);
        if (pe != hipSuccess) { std::fprintf(stderr, "%s: %s\n", W, hipGetErrorString(pe)); goto fail; }
    }
    if (p->groups.empty()) {   // no work: one idle group keeps info() meaningful
        PlanGroup g;
        g.waves = idle_waves;
        g.quad = idle_quad;
        p->groups.push_back(g);
    }
    // Device topology table: entry i is tlist[i]'s whole scaffold, entry
    // T + 2i + c its copy c; the jobs' (topology, part) pairs become indexes.
    for (DevJob& J : jobs) {
        if (J.part != kWhole) J.topo = (uint32_t)tlist.size() + 2 * J.topo + (J.part - kCopy0);
        J.part = 0;
    }
    p->n_jobs = (uint32_t)jobs.size();
    {
        std::vector<DevTopo> dtopos(std::max<size_t>(1, 3 * tlist.size()));
        for (size_t i = 0; i < tlist.size(); ++i) {
            fill_dev_topo(tlist[i], &dtopos[i]);
            if (tlist[i]->host.has_parts)
                for (uint32_t c = 0; c < 2; ++c) fill_dev_topo(tlist[i], &dtopos[tlist.size() + 2 * i + c], kCopy0 + c);
        }
        // Scratch budget: free HBM less a margin, or less if POSE_GPU_SCRATCH_LIMIT
        // (bytes) says so.  Slots halve until the scratch fits; one slot is tried
        // even when it does not, as the reference does
        // (common/src/cuda/label_gpu_warp.cu:163-175).
        size_t free_b = 0, total_b = 0;
        HIP_OK(hipMemGetInfo(&free_b, &total_b), W);
        const uint64_t margin = 256ull << 20;
        uint64_t budget = free_b > margin ? free_b - margin : 0;
        budget = std::min<uint64_t>(budget, scratch_limit());
        auto total_scratch = [&]() {
            uint64_t s = 0;
            for (PlanGroup& g : p->groups) { g.scratch_off = s; s += (uint64_t)g.slots * g.slot_bytes; }
            return s;
        };
        auto halve = [&]() {
            bool any = false;
            for (PlanGroup& g : p->groups) if (g.slots > 1) { g.slots /= 2; any = true; }
            return any;
        };
        while (total_scratch() > budget && halve()) {}
        HIP_OK(hipMalloc(&p->d_topos, dtopos.size() * sizeof(DevTopo)), W);
        HIP_OK(hipMalloc(&p->d_jobs, std::max<size_t>(1, jobs.size()) * sizeof(DevJob)), W);
        HIP_OK(hipMalloc(&p->d_counter, counter_bytes(p)), W);
        for (;;) {
            p->scratch_bytes = total_scratch();
            if (hipMalloc(&p->d_scratch, std::max<uint64_t>(p->scratch_bytes, 256)) == hipSuccess) break;
            (void)hipGetLastError();
            p->d_scratch = nullptr;
            if (!halve()) { std::fprintf(stderr, "%s: scratch allocation failed\n", W); goto fail; }
        }
        HIP_OK(hipMemcpy(p->d_topos, dtopos.data(), dtopos.size() * sizeof(DevTopo), hipMemcpyHostToDevice), W);
        if (!jobs.empty())
            HIP_OK(hipMemcpy(p->d_jobs, jobs.data(), jobs.size() * sizeof(DevJob), hipMemcpyHostToDevice), W);
        // A blocking stream: ordered after work the caller queued on the null
        // stream (as the reference's null-stream launches were).  Side streams
        // only run between the plan stream's fork and join events.
        HIP_OK(hipStreamCreate(&p->stream), W);
        p->own_stream = true;
        HIP_OK(hipEventCreate(&p->ev0), W);
        HIP_OK(hipEventCreate(&p->ev1), W);
        if (p->groups.size() > 1) {
            p->side.assign(p->groups.size() - 1, nullptr);
            p->fj.assign(p->groups.size(), nullptr);
            for (hipStream_t& s : p->side) HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), W);
            for (hipEvent_t& e : p->fj) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming), W);
        }
        if (use_graph && p->n_jobs) {
            HIP_OK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal), W);
            hipError_t ce = plan_enqueue(p, p->stream);
            hipError_t ee = hipStreamEndCapture(p->stream, &p->graph);
            HIP_OK(ce, W);
            HIP_OK(ee, W);
            HIP_OK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0), W);
        }
    }
    return true;
fail:
    return false;
}

pose_gpu_plan* pose::internal::plan_build(std::vector<PendingGroup>& pend,
                                          const std::vector<const pose_hbm_topo*>& tlist, uint64_t nodes,
                                          uint64_t out_bytes, int use_graph, int idle_waves, bool idle_quad) {
    int dev = 0;
    const int cus = device_cus(&dev, "pose_gpu_plan_create");
    if (cus < 0) return nullptr;
    pose_gpu_plan* p = new (std::nothrow) pose_gpu_plan();
    if (!p) return nullptr;
    bool ok = false;
    try {
        ok = plan_fill(p, dev, cus, pend, tlist, nodes, out_bytes, use_graph, idle_waves, idle_quad);
    } catch (const std::exception& e) {   // host allocation failure: nothing escapes the C ABI
        std::fprintf(stderr, "pose_gpu_plan_create: %s\n", e.what());
    }
    if (ok) return p;
    plan_release(p);
    return nullptr;
}

// Host allocation failures end the call with NULL: nothing escapes the C ABI.
static pose_gpu_plan_t* plan_create_guarded(const pose_gpu_session_t* sessions, size_t n_sessions,
                                            const pose_gpu_launch_t* per_session, const pose_gpu_launch_t* common,
                                            int use_graph) {
    try {
        return plan_create(sessions, n_sessions, per_session, common, use_graph);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "pose_gpu_plan_create: %s\n", e.what());
        return nullptr;
    }
}

extern "C" pose_gpu_plan_t* pose_gpu_plan_create(const pose_gpu_session_t* sessions, size_t n_sessions,
                                                 const pose_gpu_launch_t* launch, int use_graph) {
    return plan_create_guarded(sessions, n_sessions, nullptr, launch, use_graph);
}

extern "C" pose_gpu_plan_t* pose_gpu_plan_create_groups(const pose_gpu_session_t* sessions, size_t n_sessions,
                                                        const pose_gpu_launch_t* launches, int use_graph) {
    if (!launches && n_sessions) {
        std::fprintf(stderr, "pose_gpu_plan_create_groups: launches is NULL\n");
        return nullptr;
    }
    return plan_create_guarded(sessions, n_sessions, launches, nullptr, use_graph);
}

extern "C" int pose_gpu_plan_run(pose_gpu_plan_t* p, double* kernel_ms) {
    static const char* W = "pose_gpu_plan_run";
    if (!p) return -1;
    if (kernel_ms) *kernel_ms = 0;
    if (p->n_jobs == 0) return 0;
    float ms = 0;
    HIP_OK(hipEventRecord(p->ev0, p->stream), W);
    if (p->exec) HIP_OK(hipGraphLaunch(p->exec, p->stream), W);
    else HIP_OK(plan_enqueue(p, p->stream), W);
    HIP_OK(hipEventRecord(p->ev1, p->stream), W);
    HIP_OK(hipEventSynchronize(p->ev1), W);
    HIP_OK(hipEventElapsedTime(&ms, p->ev0, p->ev1), W);
    if (kernel_ms) *kernel_ms = ms;
    return 0;
fail:
    return -1;
}

extern "C" int pose_gpu_plan_info(const pose_gpu_plan_t* p, pose_gpu_plan_info_t* info) {
    if (!p || !info) return -1;
    const PlanGroup& g0 = p->groups.front();
    info->super_chunks = p->n_super;
    uint32_t slots = 0;
    for (const PlanGroup& g : p->groups) slots += g.slots;
    info->slots = slots;
    info->lanes = 64u * (uint32_t)g0.waves;
    info->lanes_per_node = g0.quad ? 4u : 1u;
    info->uses_graph = p->exec != nullptr;
    info->scratch_bytes = p->scratch_bytes;
    info->nodes = p->nodes;
    info->output_bytes = p->out_bytes;
    info->groups = (uint32_t)p->groups.size();
    return 0;
}

extern "C" int pose_gpu_plan_group(const pose_gpu_plan_t* p, uint32_t g, pose_gpu_plan_group_t* out) {
    if (!p || !out || g >= p->groups.size()) return -1;
    const PlanGroup& G = p->groups[g];
    out->super_chunks = G.n_super;
    out->jobs = G.n_jobs;
    out->split = G.split ? 1u : 0u;
    out->slots = G.slots;
    out->lanes = 64u * (uint32_t)G.waves;
    out->lanes_per_node = G.quad ? 4u : 1u;
    out->max_steps = G.depth;
    out->scratch_bytes = (uint64_t)G.slots * G.slot_bytes;
    return 0;
}

extern "C" void pose_gpu_plan_destroy(pose_gpu_plan_t* p) {
    if (!p) return;
    // Scrub the transient scaffold before release (the reference zeroes its
    // scratch, common/src/cuda/label_gpu_warp.cu:236-238).
    if (p->d_scratch && p->stream) {
        if (zero_device(p->d_scratch, p->scratch_bytes, p->stream) == hipSuccess)
            (void)hipStreamSynchronize(p->stream);
    }
    plan_release(p);
}

extern "C" int pose_gpu_label_batch(const pose_gpu_session_t* sessions, size_t n_sessions,
                                    const pose_gpu_launch_t* launch, double* kernel_ms) {
    pose_gpu_launch_t dflt{0, 0, kLanesPerNodeAuto, POSE_GPU_SPLIT_AUTO};
    if (!launch) launch = &dflt;
    pose_gpu_plan_t* p = pose_gpu_plan_create(sessions, n_sessions, launch, 0);
    if (!p) return -1;
    const int rc = pose_gpu_plan_run(p, kernel_ms);
    pose_gpu_plan_destroy(p);
    return rc;
}

// Runs a one-group plan and, every 5 s while it runs, prints how many
// super-chunks are finished, as the reference's labeler does between its
// batches (common/src/cuda/label_gpu_warp.cu:181-233).  The count comes from
// the launch's work counter, read on a stream of its own: every job runner
// claims once when it starts and once after each job it finishes, so
// finished jobs = claims - runners.  (The kernels keep a single lane-0 block
// per job iteration: a second one, placed after the job's last barrier, let
// the compiler split the job loop into a lane-0 path and an inner loop around
// the barrier for the other lanes, which then re-labeled a stale job without
// end.)
static int plan_run_with_progress(pose_gpu_plan* p, uint64_t num_super, double t_start) {
    static const char* W = "pose_graph_hbm_topo_label";
    constexpr double kInterval = 5.0;
    if (p->n_jobs == 0) return 0;
    hipStream_t side = nullptr;
    uint32_t* claims = nullptr;
    int rc = -1;
    double t_last = t_start;
    const uint32_t jobs_per_super = p->groups[0].split ? 2u : 1u;
    const uint64_t runners = p->groups[0].slots;
    HIP_OK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), W);
    HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&claims), sizeof(uint32_t), hipHostMallocDefault), W);
    HIP_OK(hipEventRecord(p->ev0, p->stream), W);
    HIP_OK(plan_enqueue(p, p->stream), W);
    HIP_OK(hipEventRecord(p->ev1, p->stream), W);
    for (;;) {
        const hipError_t q = hipEventQuery(p->ev1);
        if (q == hipSuccess) break;
        if (q != hipErrorNotReady) {
            std::fprintf(stderr, "%s: labeling failed: %s\n", W, hipGetErrorString(q));
            goto fail;
        }
        // Poll finely during the first quarter second, so that a short label
        // call returns promptly, then every 20 ms.
        std::this_thread::sleep_for(std::chrono::microseconds(now_s() - t_start < 0.25 ? 200 : 20000));
        const double now = now_s();
        if (now - t_last < kInterval) continue;
        t_last = now;
        HIP_OK(hipMemcpyAsync(claims, p->d_counter, sizeof(uint32_t), hipMemcpyDeviceToHost, side), W);
        HIP_OK(hipStreamSynchronize(side), W);
        const uint64_t jobs_done = *claims > runners ? std::min<uint64_t>(*claims - runners, p->n_jobs) : 0;
        const uint64_t sc = std::min<uint64_t>(jobs_done / jobs_per_super, num_super);
        const double frac = (double)sc / (double)num_super, el = now - t_start;
        std::printf("  GPU label: %llu/%llu super-chunks (%.1f%%)  elapsed %.0fs  eta %.0fs\n",
                    (unsigned long long)sc, (unsigned long long)num_super, frac * 100.0, el,
                    frac > 0 ? el * (1.0 - frac) / frac : 0.0);
        std::fflush(stdout);
    }
    rc = 0;
fail:
    if (claims) (void)hipHostFree(claims);
    if (side) (void)hipStreamDestroy(side);
    if (rc != 0) (void)hipStreamSynchronize(p->stream);
    return rc;
}

extern "C" int pose_graph_hbm_topo_label(const pose_hbm_topo_t* t, const uint8_t* session_seed,
                                         size_t seed_len, void* hbm_dev_ptr, size_t hbm_len,
                                         uint64_t region_block_offset) {
    if (!t || !t->on_device) return -1;
    const uint64_t super_bytes = (uint64_t)t->host.cb * 32;
    const uint64_t num_super = hbm_len / super_bytes;
    if (num_super == 0) return 0;
    if (!quiet()) {
        std::printf("  GPU label: strategy=persistent-step(lanes=%u)  %llu super-chunks  scaffold=%llu nodes\n",
                    t->host.lanes, (unsigned long long)num_super, (unsigned long long)t->host.count);
        std::fflush(stdout);
    }
    pose_gpu_session_t S;
    S.topo = t;
    S.session_seed = session_seed;
    S.seed_len = seed_len;
    S.region_block_offset = region_block_offset;
    S.n_super = num_super;
    S.out_dev = hbm_dev_ptr;
    const double t0 = now_s();
    int rc = -1;
    if (quiet()) {
        rc = pose_gpu_label_batch(&S, 1, nullptr, nullptr);
    } else if (pose_gpu_plan_t* p = pose_gpu_plan_create(&S, 1, nullptr, 0)) {
        rc = plan_run_with_progress(p, num_super, t0);
        pose_gpu_plan_destroy(p);
    }
    const double el = now_s() - t0;
    if (rc == 0 && el > 0 && !quiet()) {
        const double mib = (double)(num_super * super_bytes) / (double)(1u << 20);
        std::printf("  GPU wipe: %.0f MiB in %.3f s = %.1f MiB/s\n", mib, el, mib / el);
        std::fflush(stdout);
    }
    return rc;
}

extern "C" int pose_graph_label_hbm(const uint8_t* session_seed, size_t seed_len, uint32_t chunk_blocks,
                                    void* hbm_dev_ptr, size_t hbm_len, uint64_t region_block_offset,
                                    pose_hash_algo_t algo) {
    if (algo != POSE_HASH_BLAKE3) {
        std::fprintf(stderr, "pose_graph_label_hbm: only BLAKE3 supported on GPU\n");
        return -1;
    }
    if (chunk_blocks == 0) chunk_blocks = 4096;
    if (chunk_blocks > (1u << 20)) {
        std::fprintf(stderr, "pose_graph_label_hbm: chunk_blocks %u out of range\n", chunk_blocks);
        return -1;
    }
    if (hbm_len / ((uint64_t)chunk_blocks * 32) == 0) return 0;
    pose_hbm_topo_t* t = pose_graph_hbm_topo_build(chunk_blocks, algo);
    if (!t) return -1;
    int rc = pose_graph_hbm_topo_label(t, session_seed, seed_len, hbm_dev_ptr, hbm_len, region_block_offset);
    pose_graph_hbm_topo_free(t);
    return rc;
}

extern "C" int pose_gpu_label_node_single(uint8_t out[32], int is_source, uint64_t node_index,
                                          const uint8_t* pred0, const uint8_t* pred1, const uint8_t* seed,
                                          const uint8_t* descriptor, const uint32_t key_src[8],
                                          const uint32_t key_int[8]) {
    static const char* W = "pose_gpu_label_node_single";
    if (!out || !seed || !descriptor || !key_src || !key_int) return -1;
    uint32_t in[48] = {0};
    for (int i = 0; i < 8; ++i) {
        if (pred0) in[i] = b3::load_le(pred0 + 4 * i);
        if (pred1) in[8 + i] = b3::load_le(pred1 + 4 * i);
        in[16 + i] = b3::load_le(seed + 4 * i);
        in[24 + i] = b3::load_le(descriptor + 4 * i);
        in[32 + i] = key_src[i];
        in[40 + i] = key_int[i];
    }
    uint32_t* d = nullptr;
    uint32_t res[8];
    int rc = -1;
    HIP_OK(hipMalloc(&d, 56 * 4), W);
    HIP_OK(hipMemcpy(d + 8, in, sizeof in, hipMemcpyHostToDevice), W);
    k_node_single<<<1, 64>>>(d, node_index, is_source ? 1u : 0u, d + 8);
    HIP_OK(hipGetLastError(), W);
    HIP_OK(hipMemcpy(res, d, 32, hipMemcpyDeviceToHost), W);
    for (int i = 0; i < 8; ++i) b3::store_le(out + 4 * i, res[i]);
    rc = 0;
fail:
    if (d) (void)hipFree(d);
    return rc;
}
