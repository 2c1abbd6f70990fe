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
typedef __attribute__((address_space(3))) uint8_t lds_u8;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Latency form: four lanes per node.  A workgroup of 64*W lanes labels 16*W
// nodes per step (the same step records as k_label with lanes = 16*W).  Lane j
// of a quad fetches one 16-byte quarter of its node's two predecessor labels
// (message words 4j..4j+3), stages it in LDS, and every lane reads the
// message words of its G calls from there (those of round 0 straight from the
// quad's registers by DPP); block 2's seed|descriptor words are staged once
// per super-chunk.  Each compression is 7 x (2 G + 6 DPP row rotations)
// per lane (the last round forms the output words with two of its rotations
// folded into XORs), about a third of the single-lane instruction stream, so
// one step of a deep graph takes far less time when the batch is too small to
// fill the device.
//
// No global-memory round trip sits between two steps: every label is also
// written to an LDS copy of its step (two step buffers, alternating), and a
// predecessor written in the step just before is read from there (the near
// table names its lane).  All older predecessors were written at least two
// steps back, so their loads are issued one step early, while the previous
// step is still hashing, and each lane puts its loaded quarter in an LDS slot
// of its own at the end of that step: a step then reads every lane's quarter
// with one LDS read at a per-lane address, with no execution mask.  Early
// loads stay race-free: a slot read by step c is rewritten no earlier than
// step c+1 (ring pages and recycled slots alike), and the load issued during
// step c-1 has returned before step c ends.

// Workgroups of at least this many waves (two or more per SIMD) fold the
// row moves into their readers.
#ifndef POSE_QUAD_FOLD_MIN_W
#define POSE_QUAD_FOLD_MIN_W 8
#endif
constexpr int kQuadFoldMinW = POSE_QUAD_FOLD_MIN_W;

// Message staging layout.  Every round each lane of a quad reads 4 of its
// node's 16 message words, the same 4 word indices in every quad of the wave.
// With the words of a quad in 16 consecutive LDS words, 4 or 8 quads of a
// 32-lane pass hit the same banks.  Here word w of quad q sits at
//   32 * (8*(qh>>1) + 2*kMsgRank[w] + (qh&1)) + 8*kMsgColour[w] + (q&7),
// qh = q>>3, so its bank (mod 32) is 8*colour + (q&7), and mod 64 the two
// halves of a wave use different banks.  kMsgColour (four colours of four
// words each) makes 22 of the 28 per-step read instructions conflict-free and
// the other 6 two-way, against four-way for all 28 with consecutive words;
// no colouring makes all 28 conflict-free.  kMsgRank is the index of w among
// the words of its colour.
constexpr uint32_t kMsgColour[16] = {1, 0, 3, 1, 2, 3, 0, 2, 2, 3, 3, 1, 1, 0, 0, 2};
constexpr uint32_t kMsgRank[16] = {0, 0, 0, 1, 0, 1, 1, 1, 2, 2, 3, 2, 3, 2, 3, 3};
constexpr uint32_t msg_word_off(uint32_t w) { return 64 * kMsgRank[w] + 8 * kMsgColour[w]; }
__device__ __forceinline__ uint32_t msg_quad_off(uint32_t q) {
    const uint32_t qh = q >> 3;
    return 32 * (8 * (qh >> 1) + (qh & 1)) + (q & 7);
}

template <int W>
struct QuadShared {                      // LDS of one workgroup
    static constexpr uint32_t NODES = 16 * W;
    static_assert(NODES % 16 == 0, "message layout works on pairs of 8-quad groups");
    uint32_t tail[16];                   // block-2 words: chunk seed | descriptor
    uint32_t msg[NODES * 16];            // predecessor staging, 16 words per quad (layout above)
    uint32_t prev[2 * NODES * 8];        // labels of the last two steps
    uint32_t farq[64 * W * 4];           // each lane's quarter loaded for the next step
};

struct QuadLane {                        // per-lane constants
    uint32_t widx[28];                   // message word (0..15) of each of the 28 per-round words
    lds_u32* mptr[28];                   // its staging address
    lds_u32* wptr[4];                    // staging addresses of this lane's words 4j..4j+3
    const lds_u8* prevb;                 // the step buffers (QuadShared::prev)
    lds_u32x4* farq4;                    // this lane's far-quarter slot
    uint32_t farq_off;                   // its byte offset from prevb
    uint32_t hjs;                        // 16 * hj: this lane's quarter in a label
    lds_u32* prev_out;                   // this lane's words in step buffer 0
    uint32_t q, j, pj, hj;               // quad, lane in quad, predecessor (0/1), half (0/1)
    uint32_t iv_j, d1, d2;               // column j of IV, of the block-1 and block-2 counter words
};

template <int W>
__device__ __forceinline__ void quad_lane_init(QuadShared<W>& S, uint32_t tid, QuadLane& c) {
    c.q = tid >> 2;
    c.j = tid & 3;
    lds_u32* const qbase = (lds_u32*)(uint32_t*)S.msg + msg_quad_off(c.q);
#pragma unroll
    for (int r = 0; r < 7; ++r) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            // Column words of G_j, then diagonal words of G_i, i = (j+3)%4.
            uint32_t w = 0, off = 0;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int g = k < 2 ? jj : (jj + 3) & 3;
                const uint32_t wc = kSched[r][2 * g + (k & 1) + (k >> 1) * 8];
                if ((int)c.j == jj) { w = wc; off = msg_word_off(wc); }
            }
            c.widx[4 * r + k] = w;
            c.mptr[4 * r + k] = qbase + off;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        uint32_t off = 0;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
            if ((int)c.j == jj) off = msg_word_off(4 * jj + i);
        c.wptr[i] = qbase + off;
    }
    const uint32_t j = c.j;
    c.iv_j = j == 0 ? 0x6A09E667u : j == 1 ? 0xBB67AE85u : j == 2 ? 0x3C6EF372u : 0xA54FF53Au;
    c.d1 = j == 2 ? 64u : j == 3 ? (uint32_t)kFlagKeyed : 0u;
    c.d2 = j == 2 ? 64u : j == 3 ? (uint32_t)(kFlagKeyed | kFlagChunkEnd | kFlagRoot) : 0u;
    c.pj = j >> 1;
    c.hj = j & 1;
    c.prevb = (const lds_u8*)(uint8_t*)S.prev;
    c.farq4 = (lds_u32x4*)(uint32_t*)S.farq + tid;
    c.farq_off = (uint32_t)((uint32_t*)S.farq - (uint32_t*)S.prev) * 4 + 16 * tid;
    c.hjs = 16 * c.hj;
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
    const uint32_t hsh = 16 * pj;          // this lane's half of a near entry
    const bool p1 = pj != 0;               // this lane loads predecessor 1
    const bool odd = (j & 1) != 0;
    const gcu4 rec = (gcu4)T.rec + q;
    const GAS uint32_t* const nearp = (const GAS uint32_t*)T.near + q;
    // Block-0 CV words j and 4+j, adjacent in the stored order 0 4 1 5 2 6 3 7.
    const GAS u32x2* const cv0 = (const GAS u32x2*)T.cv0 + 4 * q + j;
    // Rows 0..n_steps exist (the last is idle), and every uploaded schedule
    // is followed by one more row of records and near entries, so step s can
    // load row s+2 without a bound.  Per-lane pointers to the records and
    // near entries of row s+2 and to the block-0 CV words of row s+1,
    // advanced one row per step.
    const uint32_t n_steps = T.n_steps;
    gcu4 rec2 = rec + 2 * NODES;
    const GAS uint32_t* near2 = nearp + 2 * NODES;
    const GAS u32x2* cv1 = cv0 + NODES * 4;

    // One step.  In: this step's store offset dst, the next row's records rn
    // and raw near entry hnr, where this step's quarter is in LDS (byte
    // offset qoff from the step buffers) and its block-0 CV words.  Out: the
    // same for step s+1.  Two steps per loop iteration with the in and out
    // sets swapped, so nothing is copied between steps except dst.
    auto step = [&](uint32_t s, uint32_t dst, const u32x4& rn, uint32_t hnr, uint32_t qoff,
                    uint32_t ca, uint32_t cb, uint32_t& dst_o, u32x4& rn_o, uint32_t& hnr_o,
                    uint32_t& qoff_o, uint32_t& ca_o, uint32_t& cb_o) {
        // This step's quarter: for a near predecessor the previous step's LDS
        // copy, otherwise the quarter loaded one step ago, which the lane put
        // in its own LDS slot at the end of that step: one read at a per-lane
        // address, chosen in the step before, with no execution mask.  It
        // goes first, so the global loads below are issued while it is in
        // flight.
        const u32x4 qw = *(const lds_u32x4*)(c.prevb + qoff);
        __builtin_amdgcn_sched_barrier(0);
        // Rows s+2 (records, near entries) and s+1 (block-0 CV words).
        rn_o = *rec2;
        hnr_o = *near2;
        const u32x2 cvw = *cv1;
        ca_o = cvw.x;
        cb_o = cvw.y;
        rec2 += NODES;
        near2 += NODES;
        cv1 += NODES * 4;
        __builtin_amdgcn_sched_barrier(0);
        // Step s+1's predecessors written before step s: load them now.  A
        // predecessor of the step just before is taken from LDS instead (it
        // does not exist yet); its lane loads the zero slot.
        static_assert(kZeroSlot == 0, "the zero slot is at offset 0");
        static_assert(kFarPred == 0xFFFFu, "near entries below 2^15 are lanes, kFarPred has bit 15 set");
        const uint32_t hn = __builtin_amdgcn_ubfe(hnr, hsh, 16);
        // All ones for a far predecessor, zero for a near one (bit 15 of the
        // entry, sign-extended): a mask in a VGPR, so the selects below are
        // bitwise and need no condition register.
        const uint32_t far_m = (uint32_t)__builtin_amdgcn_sbfe((int)hnr, hsh + 15, 1);
        const uint32_t pred = p1 ? rn.z : rn.y;
        const uint32_t pl = pred & far_m;
        const u32x4 far_n_q = ((gcu4)(sbase + pl))[hj];
        __builtin_amdgcn_sched_barrier(0);
        *c.wptr[0] = qw.x;
        *c.wptr[1] = qw.y;
        *c.wptr[2] = qw.z;
        *c.wptr[3] = qw.w;
        // The quad's lanes are in one wavefront, whose LDS operations are
        // performed in order: a wavefront-scope fence orders the staging
        // writes before the reads without waiting for them.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // Block 1.  Round 0's words come from the quad's registers, so the
        // staging round trip is off the step's critical path: lane j holds
        // words 4j..4j+3, its column words 2j and 2j+1 sit in lane j>>1
        // (components x, y for even j, z, w for odd j), its diagonal words
        // 8+2i and 9+2i (i = (j+3)%4) in lanes 3, 2, 2, 3 (z, w for even j,
        // x, y for odd j).  Later rounds read their four words from LDS while
        // the round before is hashing.  Reading earlier is slower: all 24 up
        // front would make round 1 wait for most of them, and two rounds ahead
        // (eight reads in flight) takes 6 % longer per step.
        uint32_t a = ca, b = cb, cc = c.iv_j, dd = c.d1;
        uint32_t mc[4], mn[4];
        {
            // Both candidates of each word are moved on every lane (a move
            // under a partial execution mask would read disabled lanes).
            const uint32_t ocx = quad_perm<kQuadCol, true>(qw.z), ocy = quad_perm<kQuadCol, true>(qw.w);
            const uint32_t odx = quad_perm<kQuadDiag, true>(qw.x), ody = quad_perm<kQuadDiag, true>(qw.y);
            const uint32_t ecx = quad_perm<kQuadCol, true>(qw.x), ecy = quad_perm<kQuadCol, true>(qw.y);
            const uint32_t edx = quad_perm<kQuadDiag, true>(qw.z), edy = quad_perm<kQuadDiag, true>(qw.w);
            mc[0] = odd ? ocx : ecx;
            mc[1] = odd ? ocy : ecy;
            mc[2] = odd ? odx : edx;
            mc[3] = odd ? ody : edy;
        }
#pragma unroll
        for (int r = 0; r < 7; ++r) {
            if (r < 6) {
#pragma unroll
                for (int k = 0; k < 4; ++k) mn[k] = *c.mptr[4 * (r + 1) + k];
            }
            if (r == 6) {
                __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt <= 0
            } else if (r > 0) {
                // One wait for this round's four words (lgkmcnt <= 4; vmcnt
                // and expcnt at their maxima, i.e. not waited on).
                __builtin_amdgcn_s_waitcnt(0xC47F);
            }
            __builtin_amdgcn_sched_barrier(0);
            if (r < 6)
                quad_round<(W >= kQuadFoldMinW)>(a, b, cc, dd, mc[0], mc[1], mc[2], mc[3]);
            else
                quad_last_round(a, b, cc, dd, mc[0], mc[1], mc[2], mc[3]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < 4; ++k) mc[k] = mn[k];
        }
        compress_quad<(W >= kQuadFoldMinW)>(a, b, c.iv_j, c.d2, m2);
        __builtin_amdgcn_sched_barrier(0);
        // sbase + a 32-bit offset: the store takes sbase from SGPRs.
        GAS uint32_t* d = (GAS uint32_t*)(sbase + (dst + 4 * j));
        d[0] = a;
        d[4] = b;
        lds_u32* const po = c.prev_out + (s & 1) * NODES * 8;
        po[0] = a;
        po[4] = b;
        *c.farq4 = far_n_q;
        __syncthreads();
        dst_o = rn.x;
        // Step s+1 reads its near predecessor from this step's copy, a far
        // one from the lane's slot.
        const uint32_t near_off = (s & 1) * NODES * 32 + (hn << 5) + c.hjs;
        qoff_o = (far_m & c.farq_off) | (~far_m & near_off);
    };

    // Rows 0 and 1; step 0's predecessors are loaded before the loop (row 0
    // has no near predecessors).
    const u32x4 r0 = rec[0];
    uint32_t dA = r0.x, hnrA = nearp[NODES];
    uint32_t qsA = c.farq_off, qsB;
    u32x4 rnA = rec[NODES];
    const u32x2 cv00 = cv0[0];
    uint32_t caA = cv00.x, cbA = cv00.y;
    *c.farq4 = ((gcu4)(sbase + (p1 ? r0.z : r0.y)))[hj];
    uint32_t dB, hnrB, caB, cbB;
    u32x4 rnB;
    // Drain these loads here.  Otherwise the wait for them, merged into the
    // loop body, also waits for the previous step's label stores.
    __builtin_amdgcn_s_waitcnt(0);
    uint32_t s = 0;
    for (; s + 2 <= n_steps; s += 2) {
        step(s, dA, rnA, hnrA, qsA, caA, cbA, dB, rnB, hnrB, qsB, caB, cbB);
        step(s + 1, dB, rnB, hnrB, qsB, caB, cbB, dA, rnA, hnrA, qsA, caA, cbA);
    }
    if (s < n_steps)
        step(s, dA, rnA, hnrA, qsA, caA, cbA, dB, rnB, hnrB, qsB, caB, cbB);
    // The finished labels, from the slot region to the caller's buffer.
    const gcu4 src = (gcu4)(sbase + T.out_off);
    const gu4 out = (gu4)pbase;
    const uint32_t n16 = T.out16;
    for (uint32_t i = tid; i < n16; i += 64 * W) out[i] = src[i];
    __syncthreads();
}

}  // namespace dev
}  // namespace pose
