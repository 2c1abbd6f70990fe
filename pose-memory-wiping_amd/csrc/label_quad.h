// label_quad.h — device-side pieces of the labeling kernels (labeler.hip):
// address-space views, and the four-lanes-per-node labeling of one
// super-chunk (the latency form) run by k_label_quad.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "blake3_dev.h"
#include "labeler_internal.h"
#include "topology.h"

namespace pose {
namespace dev {

// Global-address-space views: loads through these compile to global_load_*
// (vmcnt only) instead of flat_load_* (which also holds lgkmcnt).
#define GAS __attribute__((address_space(1)))
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef GAS const u32x4* gcu4;
typedef GAS u32x4* gu4;
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) u32x4 lds_u32x4;

// Latency form: four lanes per node.  A workgroup of 64*W lanes labels 16*W
// nodes per step (the same step records as k_label with lanes = 16*W).  Lane j
// of a quad fetches one 16-byte quarter of its node's two predecessor labels,
// stages it in the quad's 64-byte LDS slot, and every lane reads the 28 message
// words of its column from there; block 2's seed|descriptor words are staged
// once per super-chunk.  Each compression is 7 x (2 G + 6 DPP row rotations)
// per lane, about a third of the single-lane instruction stream, so one step of
// a deep graph takes far less time when the batch is too small to fill the
// device.
//
// No global-memory round trip sits between two steps: every label is also
// written to an LDS copy of its step (two step buffers, alternating), and a
// predecessor written in the step just before is read from there (the near
// table names its lane).  All older predecessors were written at least two
// steps back, so their loads are issued one step early, while the previous
// step is still hashing.  Early loads stay race-free: a slot read by step c is
// rewritten no earlier than step c+1 (ring pages and recycled slots alike), and
// the load issued during step c-1 has returned before step c ends.

template <int W>
struct QuadShared {                      // LDS of one workgroup
    static constexpr uint32_t NODES = 16 * W;
    uint32_t tail[16];                   // block-2 words: chunk seed | descriptor
    uint32_t msg[NODES * 16];            // per quad: 64-byte predecessor staging
    uint32_t prev[2 * NODES * 8];        // labels of the last two steps
};

struct QuadLane {                        // per-lane constants
    uint32_t widx[28];                   // message word (0..15) of each of the 28 per-round words
    lds_u32* mptr[28];                   // its address in this quad's staging slot
    lds_u32* qslot;
    lds_u32x4* prev4;
    lds_u32* prev_out;                   // this lane's words in step buffer 0
    uint32_t q, j, pj, hj;               // quad, lane in quad, predecessor (0/1), half (0/1)
    uint32_t iv_j, d1, d2;               // column j of IV, of the block-1 and block-2 counter words
};

template <int W>
__device__ __forceinline__ void quad_lane_init(QuadShared<W>& S, uint32_t tid, QuadLane& c) {
    c.q = tid >> 2;
    c.j = tid & 3;
    c.qslot = (lds_u32*)(uint32_t*)S.msg + 16 * c.q;
#pragma unroll
    for (int r = 0; r < 7; ++r) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t w = 0;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj)
                if ((int)c.j == jj) w = kSched[r][2 * jj + (k & 1) + (k >> 1) * 8];
            c.widx[4 * r + k] = w;
            c.mptr[4 * r + k] = c.qslot + w;
        }
    }
    const uint32_t j = c.j;
    c.iv_j = j == 0 ? 0x6A09E667u : j == 1 ? 0xBB67AE85u : j == 2 ? 0x3C6EF372u : 0xA54FF53Au;
    c.d1 = j == 2 ? 64u : j == 3 ? (uint32_t)kFlagKeyed : 0u;
    c.d2 = j == 2 ? 64u : j == 3 ? (uint32_t)(kFlagKeyed | kFlagChunkEnd | kFlagRoot) : 0u;
    c.pj = j >> 1;
    c.hj = j & 1;
    c.prev4 = (lds_u32x4*)(uint32_t*)S.prev;
    c.prev_out = (lds_u32*)(uint32_t*)S.prev + c.q * 8 + j;
}

// Labels one super-chunk with all 64*W lanes of the workgroup: topology T,
// chunk-seed words `seed` (readable by lanes 0..7), the workgroup's slot region
// at sbase (slot 0 must hold zeros; device locations are byte offsets in it),
// finished labels copied to pbase (rank r at pbase + 32 r).  Begins and ends
// with a workgroup barrier.
template <int W>
__device__ __forceinline__ void quad_label_superchunk(QuadShared<W>& S, const QuadLane& c,
                                                      const internal::DevTopo& T, const uint32_t* seed,
                                                      uint8_t* pbase, GAS uint8_t* sbase, uint32_t tid) {
    constexpr uint32_t NODES = QuadShared<W>::NODES;
    if (tid < 16) S.tail[tid] = tid < 8 ? seed[tid] : T.desc[tid - 8];
    __syncthreads();
    uint32_t m2[28];
#pragma unroll
    for (int k = 0; k < 28; ++k) m2[k] = S.tail[c.widx[k]];
    const uint32_t q = c.q, j = c.j, pj = c.pj, hj = c.hj;
    const gcu4 rec = (gcu4)T.rec + q;
    const GAS uint32_t* const nearp = (const GAS uint32_t*)T.near + q;
    const GAS uint32_t* const cv0 = (const GAS uint32_t*)T.cv0 + 8 * q + j;
    const uint32_t n_steps = T.n_steps;   // rows 0..n_steps exist (the last is idle)

    // Rows 0 and 1; step 0's predecessors are loaded before the loop.
    const u32x4 r0 = rec[0];
    uint32_t dst = r0.x;
    u32x4 r1 = rec[NODES];
    uint32_t h0 = nearp[0], h1 = nearp[NODES];
    uint32_t ca = cv0[0], cb = cv0[4];
    u32x4 far0 = ((gcu4)(sbase + (pj ? r0.z : r0.y)))[hj];
    // Drain these loads here.  Otherwise the wait for them, merged into the
    // loop body, also waits for the previous step's label stores.
    __builtin_amdgcn_s_waitcnt(0);
    for (uint32_t s = 0; s < n_steps; ++s) {
        // Step s+1's predecessors written before step s: load them now.  A
        // predecessor of the step just before is taken from LDS instead (it
        // does not exist yet); its lane loads the zero slot.
        const uint32_t hn = pj ? h1 >> 16 : h1 & 0xFFFFu;
        const uint32_t pl = hn != kFarPred ? kZeroSlot * 32 : (pj ? r1.z : r1.y);
        const u32x4 far1 = ((gcu4)(sbase + pl))[hj];
        __builtin_amdgcn_sched_barrier(0);
        const uint64_t row2 = (uint64_t)min(s + 2, n_steps) * NODES;
        const u32x4 r2 = rec[row2];
        const uint32_t h2 = nearp[row2];
        const uint64_t row1 = (uint64_t)(s + 1) * NODES * 8;
        const uint32_t can = cv0[row1], cbn = cv0[row1 + 4];
        __builtin_amdgcn_sched_barrier(0);
        // This step's quarter: the previous step's LDS copy, or the load
        // issued one step ago.
        const uint32_t hc = pj ? h0 >> 16 : h0 & 0xFFFFu;
        const bool near = hc != kFarPred;
        const u32x4 lw = c.prev4[((s + 1) & 1) * NODES * 2 + (near ? hc : 0u) * 2 + hj];
        ((lds_u32x4*)c.qslot)[j] = near ? lw : far0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        uint32_t m1[28];
#pragma unroll
        for (int k = 0; k < 28; ++k) m1[k] = *c.mptr[k];
        uint32_t a = ca, b = cb;
        compress_quad(a, b, c.iv_j, c.d1, m1);
        compress_quad(a, b, c.iv_j, c.d2, m2);
        __builtin_amdgcn_sched_barrier(0);
        GAS uint32_t* d = (GAS uint32_t*)(sbase + dst) + j;
        d[0] = a;
        d[4] = b;
        lds_u32* const po = c.prev_out + (s & 1) * NODES * 8;
        po[0] = a;
        po[4] = b;
        __syncthreads();
        dst = r1.x;
        r1 = r2; h0 = h1; h1 = h2; far0 = far1; ca = can; cb = cbn;
    }
    // The finished labels, from the slot region to the caller's buffer.
    const gcu4 src = (gcu4)(sbase + T.out_off);
    const gu4 out = (gu4)pbase;
    const uint32_t n16 = T.out16;
    for (uint32_t i = tid; i < n16; i += 64 * W) out[i] = src[i];
    __syncthreads();
}

}  // namespace dev
}  // namespace pose
