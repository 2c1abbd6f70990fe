// blake3_dev.h — BLAKE3 compression for CDNA4, one node hash per lane.
//
// The PoSE node hash (common/src/cuda/blake3_node.cuh:195-230 in the reference)
// is a keyed BLAKE3 of one 192-byte chunk: three compressions, counter 0,
// block length 64, flags KEYED|CHUNK_START, KEYED, KEYED|CHUNK_END|ROOT.
// Everything here is exact 32-bit add/xor/rotate; rotates map to
// v_alignbit_b32 and a+b+m to v_add3_u32, so one G is 12 VALU instructions and
// one compression 7*8*12 + 8 = 680.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pose {
namespace dev {

constexpr uint32_t kFlagChunkStart = 1, kFlagChunkEnd = 2, kFlagRoot = 8, kFlagKeyed = 16;

// Message word index of each G input, per round (the spec's permutation
// applied r times).
constexpr uint8_t kSched[7][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15},
    {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8},
    {3, 4, 10, 12, 13, 2, 7, 14, 6, 5, 9, 0, 11, 15, 8, 1},
    {10, 7, 12, 9, 14, 3, 13, 15, 4, 0, 11, 2, 5, 8, 1, 6},
    {12, 13, 9, 11, 15, 10, 14, 8, 7, 2, 5, 3, 0, 1, 6, 4},
    {9, 14, 11, 5, 8, 12, 15, 1, 13, 3, 0, 10, 2, 6, 4, 7},
    {11, 15, 5, 0, 1, 9, 8, 6, 14, 10, 2, 12, 3, 4, 7, 13},
};

__device__ __forceinline__ uint32_t rotr(uint32_t x, uint32_t n) {
    return __builtin_amdgcn_alignbit(x, x, n);
}

#define POSE_G(a, b, c, d, x, y)                 \
    a = a + b + (x); d = rotr(d ^ a, 16);        \
    c = c + d;       b = rotr(b ^ c, 12);        \
    a = a + b + (y); d = rotr(d ^ a, 8);         \
    c = c + d;       b = rotr(b ^ c, 7);

// cv <- first 8 output words of compress(cv, m, counter 0, len 64, flags).
__device__ __forceinline__ void compress(uint32_t (&cv)[8], const uint32_t (&m)[16],
                                         uint32_t flags) {
    uint32_t v0 = cv[0], v1 = cv[1], v2 = cv[2], v3 = cv[3];
    uint32_t v4 = cv[4], v5 = cv[5], v6 = cv[6], v7 = cv[7];
    uint32_t v8 = 0x6A09E667u, v9 = 0xBB67AE85u, v10 = 0x3C6EF372u, v11 = 0xA54FF53Au;
    uint32_t v12 = 0, v13 = 0, v14 = 64, v15 = flags;
#pragma unroll
    for (int r = 0; r < 7; ++r) {
        POSE_G(v0, v4, v8, v12, m[kSched[r][0]], m[kSched[r][1]]);
        POSE_G(v1, v5, v9, v13, m[kSched[r][2]], m[kSched[r][3]]);
        POSE_G(v2, v6, v10, v14, m[kSched[r][4]], m[kSched[r][5]]);
        POSE_G(v3, v7, v11, v15, m[kSched[r][6]], m[kSched[r][7]]);
        POSE_G(v0, v5, v10, v15, m[kSched[r][8]], m[kSched[r][9]]);
        POSE_G(v1, v6, v11, v12, m[kSched[r][10]], m[kSched[r][11]]);
        POSE_G(v2, v7, v8, v13, m[kSched[r][12]], m[kSched[r][13]]);
        POSE_G(v3, v4, v9, v14, m[kSched[r][14]], m[kSched[r][15]]);
    }
    cv[0] = v0 ^ v8;  cv[1] = v1 ^ v9;  cv[2] = v2 ^ v10; cv[3] = v3 ^ v11;
    cv[4] = v4 ^ v12; cv[5] = v5 ^ v13; cv[6] = v6 ^ v14; cv[7] = v7 ^ v15;
}

// ── four lanes per compression ─────────────────────────────────────────────
//
// A quad of lanes holds one state: lane j has column j, i.e. a = v[j],
// b = v[4+j], c = v[8+j], d = v[12+j].  The column step is one G per lane.
// In the diagonal step lane k keeps its b = v[4+k] and runs the diagonal G
// that uses it, G_i with i = (k+3)%4: G(v[i], v[4+k], v[8+(k+1)%4],
// v[12+(k+2)%4]).  Rows a, c, d are rotated by 3, 1, 2 lanes inside the quad
// (DPP quad_perm) for it and rotated back afterwards.  b, the last word each
// G produces, never moves, so no DPP move sits on the serial G chain: a, c
// and d are final 6, 2 and 3 instructions before b.  msg[4r..4r+3] are this
// lane's words of round r: sched(2j), sched(2j+1), sched(8+2i), sched(9+2i).

// BC = false keeps each row move a separate v_mov_b32_dpp: the moves are
// independent and cover each other's DPP read hazard, and none of them sits
// on the G chain.  BC = true (bound_ctrl, which changes nothing for
// quad_perm) lets the compiler fold the move into the VOP2 instruction that
// reads it (v_xor_b32_dpp, v_add_u32_dpp): fewer instructions, but a DPP
// instruction waits two wait states after the VALU write of any operand it
// reads, and here that operand is on the chain.
template <int CTRL, bool BC = false>
__device__ __forceinline__ uint32_t quad_perm(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, BC);
}
// quad_perm selectors: lane i takes lane p_i, encoded p0 | p1<<2 | p2<<4 | p3<<6.
constexpr int kQuadNext1 = 0x39;   // lane j <- lane j+1
constexpr int kQuadNext2 = 0x4E;   // lane j <- lane j+2
constexpr int kQuadNext3 = 0x93;   // lane j <- lane j+3 (= j-1)
constexpr int kQuadCol = 0x50;     // lanes 0..3 <- lanes 0, 0, 1, 1
constexpr int kQuadDiag = 0xEB;    // lanes 0..3 <- lanes 3, 2, 2, 3

// One round on the quad layout: column G, rows to the diagonal layout,
// diagonal G, rows back.  m0, m1 are this lane's column words of the round,
// m2, m3 its diagonal words.  FOLD lets the c and d row moves fold into their
// readers (see quad_perm): the better choice when another wave on the SIMD
// fills the wait states.
// One G whose three row moves (a, d, c by PA, PD, PC lanes) are placed inside
// its tail, each where its source is already two instructions old: a after
// d's last rotate, d after b's last XOR, c after b's last rotate.  A DPP read
// needs two wait states after the VALU write of its source, so placed like
// this no move needs an s_nop, and each move fills part of the issue latency
// of the dependent G instruction after it.  sched_barrier pins the moves
// (the G chain is serial, so it keeps its order by itself).
template <int PA, int PD, int PC>
__device__ __forceinline__ void g_then_move(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d,
                                            uint32_t x, uint32_t y) {
    a = a + b + x;
    d = rotr(d ^ a, 16);
    c = c + d;
    b = rotr(b ^ c, 12);
    a = a + b + y;
    d = rotr(d ^ a, 8);
    __builtin_amdgcn_sched_barrier(0);
    a = quad_perm<PA>(a);
    __builtin_amdgcn_sched_barrier(0);
    c = c + d;
    b = b ^ c;
    __builtin_amdgcn_sched_barrier(0);
    d = quad_perm<PD>(d);
    __builtin_amdgcn_sched_barrier(0);
    b = rotr(b, 7);
    __builtin_amdgcn_sched_barrier(0);
    c = quad_perm<PC>(c);
    __builtin_amdgcn_sched_barrier(0);
}

template <bool FOLD>
__device__ __forceinline__ void quad_round(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d,
                                           uint32_t m0, uint32_t m1, uint32_t m2, uint32_t m3) {
    if constexpr (!FOLD) {
        g_then_move<kQuadNext3, kQuadNext2, kQuadNext1>(a, b, c, d, m0, m1);
        g_then_move<kQuadNext1, kQuadNext2, kQuadNext3>(a, b, c, d, m2, m3);
    } else {
        POSE_G(a, b, c, d, m0, m1);
        a = quad_perm<kQuadNext3>(a);
        c = quad_perm<kQuadNext1, FOLD>(c);
        d = quad_perm<kQuadNext2, FOLD>(d);
        POSE_G(a, b, c, d, m2, m3);
        a = quad_perm<kQuadNext1>(a);
        c = quad_perm<kQuadNext3, FOLD>(c);
        d = quad_perm<kQuadNext2, FOLD>(d);
    }
}

// The last round of a compression, ending in its output words: on return
// lane k holds a = v[k] ^ v[8+k] and b = v[4+k] ^ v[12+k].  After the last
// diagonal G lane k holds a = v[(k+3)%4], b = v[4+k], c = v[8+(k+1)%4],
// d = v[12+(k+2)%4], so the rows are not rotated back: t = a ^ (c of lane
// k+2) is output word (k+3)%4, which lane k-1 needs, and b ^ (d of lane k+2)
// is output word 4+k.  Two of the three moves fold into the XORs
// (v_xor_b32_dpp): three instructions where moving a, c and d back and two
// XORs take five.
__device__ __forceinline__ void quad_last_round(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d,
                                                uint32_t m0, uint32_t m1, uint32_t m2, uint32_t m3) {
    g_then_move<kQuadNext3, kQuadNext2, kQuadNext1>(a, b, c, d, m0, m1);
    a = a + b + m2;
    d = rotr(d ^ a, 16);
    c = c + d;
    b = rotr(b ^ c, 12);
    a = a + b + m3;
    d = rotr(d ^ a, 8);
    c = c + d;
    b = rotr(b ^ c, 7);
    const uint32_t t = a ^ quad_perm<kQuadNext2, true>(c);
    b ^= quad_perm<kQuadNext2, true>(d);
    a = quad_perm<kQuadNext1>(t);
}

// (a, b) in: chaining value words j and 4+j; out: the new words j and 4+j.
// c0 = IV[j]; d0 = {counter lo, counter hi, block length, flags}[j].
template <bool FOLD>
__device__ __forceinline__ void compress_quad(uint32_t& a, uint32_t& b, uint32_t c0, uint32_t d0,
                                              const uint32_t (&msg)[28]) {
    uint32_t c = c0, d = d0;
#pragma unroll
    for (int r = 0; r < 6; ++r)
        quad_round<FOLD>(a, b, c, d, msg[4 * r], msg[4 * r + 1], msg[4 * r + 2], msg[4 * r + 3]);
    quad_last_round(a, b, c, d, msg[24], msg[25], msg[26], msg[27]);
}

#undef POSE_G

// Block 0 of a node message: the node id as 8 big-endian bytes, then zeros.
// As little-endian words that is (bswap(id_hi), bswap(id_lo), 0, ...).
__device__ __forceinline__ void block0_words(uint64_t id, uint32_t (&m)[16]) {
    m[0] = __builtin_bswap32((uint32_t)(id >> 32));
    m[1] = __builtin_bswap32((uint32_t)id);
#pragma unroll
    for (int i = 2; i < 16; ++i) m[i] = 0;
}

}  // namespace dev
}  // namespace pose
