// b3_host.h — host-side BLAKE3 for the labeler's per-topology and per-chunk
// constants: the two node keys, the graph descriptor and the chunk seeds.
//
// Written from the BLAKE3 specification (single-chunk and tree modes, keyed and
// derive-key modes).  The node hash itself runs on the GPU; this file is only
// the host's share: pose_label_node_keys (common/src/hash.c:62-67),
// pose_graph_descriptor (common/src/hash.c:198-216) and pose_chunk_seed
// (common/src/hash.c:218-233) of the reference.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

namespace pose {
namespace b3 {

constexpr uint32_t kChunkStart = 1, kChunkEnd = 2, kParent = 4, kRoot = 8,
                   kKeyedHash = 16, kDeriveKeyContext = 32, kDeriveKeyMaterial = 64;

constexpr uint32_t kIV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                             0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};

// Word order of round r: the spec's message permutation applied r times.
struct Schedule {
    uint8_t w[7][16];
    constexpr Schedule() : w{} {
        constexpr uint8_t p[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
        for (int i = 0; i < 16; ++i) w[0][i] = (uint8_t)i;
        for (int r = 1; r < 7; ++r)
            for (int i = 0; i < 16; ++i) w[r][i] = w[r - 1][p[i]];
    }
};
constexpr Schedule kSchedule{};

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

inline void quarter(uint32_t* v, int a, int b, int c, int d, uint32_t x, uint32_t y) {
    v[a] += v[b] + x; v[d] = rotr(v[d] ^ v[a], 16);
    v[c] += v[d];     v[b] = rotr(v[b] ^ v[c], 12);
    v[a] += v[b] + y; v[d] = rotr(v[d] ^ v[a], 8);
    v[c] += v[d];     v[b] = rotr(v[b] ^ v[c], 7);
}

// Chaining-value form of the compression: cv <- first 8 words of the output.
inline void compress(uint32_t cv[8], const uint32_t m[16], uint64_t counter,
                     uint32_t block_len, uint32_t flags) {
    uint32_t v[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                      kIV[0], kIV[1], kIV[2], kIV[3], (uint32_t)counter,
                      (uint32_t)(counter >> 32), block_len, flags};
    for (int r = 0; r < 7; ++r) {
        const uint8_t* s = kSchedule.w[r];
        quarter(v, 0, 4, 8, 12, m[s[0]], m[s[1]]);
        quarter(v, 1, 5, 9, 13, m[s[2]], m[s[3]]);
        quarter(v, 2, 6, 10, 14, m[s[4]], m[s[5]]);
        quarter(v, 3, 7, 11, 15, m[s[6]], m[s[7]]);
        quarter(v, 0, 5, 10, 15, m[s[8]], m[s[9]]);
        quarter(v, 1, 6, 11, 12, m[s[10]], m[s[11]]);
        quarter(v, 2, 7, 8, 13, m[s[12]], m[s[13]]);
        quarter(v, 3, 4, 9, 14, m[s[14]], m[s[15]]);
    }
    for (int i = 0; i < 8; ++i) cv[i] = v[i] ^ v[i + 8];
}

inline uint32_t load_le(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
           ((uint32_t)p[3] << 24);
}
inline void store_le(uint8_t* p, uint32_t v) {
    p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); p[2] = (uint8_t)(v >> 16);
    p[3] = (uint8_t)(v >> 24);
}

// A node of the hash tree whose final compression is still pending, so the
// caller can apply either the chaining or the root flag.
struct Pending {
    uint32_t cv[8];
    uint32_t m[16];
    uint64_t counter;
    uint32_t len;
    uint32_t flags;
    void chaining_value(uint32_t out[8]) const {
        std::memcpy(out, cv, 32);
        compress(out, m, counter, len, flags);
    }
};

inline void block_words(const uint8_t* p, size_t n, uint32_t m[16]) {
    uint8_t b[64] = {0};
    if (n) std::memcpy(b, p, n);
    for (int i = 0; i < 16; ++i) m[i] = load_le(b + 4 * i);
}

// One chunk (<= 1024 bytes) with chunk counter `ctr`.
inline Pending chunk(const uint32_t key[8], uint32_t flags, const uint8_t* p,
                     size_t n, uint64_t ctr) {
    Pending o;
    std::memcpy(o.cv, key, 32);
    size_t blocks = n ? (n + 63) / 64 : 1;
    for (size_t b = 0; b + 1 < blocks; ++b) {
        uint32_t m[16];
        block_words(p + 64 * b, 64, m);
        compress(o.cv, m, ctr, 64, flags | (b == 0 ? kChunkStart : 0u));
    }
    size_t tail = n - 64 * (blocks - 1);
    block_words(p + 64 * (blocks - 1), tail, o.m);
    o.counter = ctr;
    o.len = (uint32_t)tail;
    o.flags = flags | kChunkEnd | (blocks == 1 ? kChunkStart : 0u);
    return o;
}

// Subtree over p[0..n): left side takes the largest power-of-two count of
// whole chunks that leaves the right side non-empty.
inline Pending subtree(const uint32_t key[8], uint32_t flags, const uint8_t* p,
                       size_t n, uint64_t ctr) {
    if (n <= 1024) return chunk(key, flags, p, n, ctr);
    uint64_t chunks = (n + 1023) / 1024, left = 1;
    while (2 * left < chunks) left *= 2;
    uint32_t l[8], r[8];
    subtree(key, flags, p, left * 1024, ctr).chaining_value(l);
    subtree(key, flags, p + left * 1024, n - left * 1024, ctr + left).chaining_value(r);
    Pending o;
    std::memcpy(o.cv, key, 32);
    std::memcpy(o.m, l, 32);
    std::memcpy(o.m + 8, r, 32);
    o.counter = 0;
    o.len = 64;
    o.flags = flags | kParent;
    return o;
}

// 32-byte root output of BLAKE3 with the given key words and mode flags.
inline void hash(const uint32_t key[8], uint32_t flags, const uint8_t* p, size_t n,
                 uint8_t out[32]) {
    Pending root = subtree(key, flags, p, n, 0);
    uint32_t cv[8];
    std::memcpy(cv, root.cv, 32);
    compress(cv, root.m, 0, root.len, root.flags | kRoot);
    for (int i = 0; i < 8; ++i) store_le(out + 4 * i, cv[i]);
}

// Derive-key mode, step 1: the context key as 8 little-endian words.
inline void context_key(const char* ctx, uint32_t key[8]) {
    uint8_t k[32];
    hash(kIV, kDeriveKeyContext, reinterpret_cast<const uint8_t*>(ctx), std::strlen(ctx), k);
    for (int i = 0; i < 8; ++i) key[i] = load_le(k + 4 * i);
}

inline void derive_key(const char* ctx, const uint8_t* material, size_t n, uint8_t out[32]) {
    uint32_t key[8];
    context_key(ctx, key);
    hash(key, kDeriveKeyMaterial, material, n, out);
}

}  // namespace b3

inline void put_be64(uint8_t* p, uint64_t v) {
    for (int i = 0; i < 8; ++i) p[i] = (uint8_t)(v >> (56 - 8 * i));
}

// Node keys: context keys of "pose-db/label/src" and "pose-db/label/int".
inline void node_keys(uint32_t ksrc[8], uint32_t kint[8]) {
    b3::context_key("pose-db/label/src", ksrc);
    b3::context_key("pose-db/label/int", kint);
}

// Graph descriptor: derive_key("pose-db/descriptor", be64(m) be64(n) be64(256)).
inline void graph_descriptor(uint8_t out[32], uint64_t m, uint64_t n) {
    uint8_t buf[24];
    put_be64(buf, m);
    put_be64(buf + 8, n);
    put_be64(buf + 16, 256);
    b3::derive_key("pose-db/descriptor", buf, sizeof buf, out);
}

// Chunk seed: derive_key("pose-db/chunk-seed", session_seed || be64(index)).
// The context key is hoisted so a batch of seeds costs one compression each.
struct ChunkSeeder {
    uint32_t key[8];
    std::vector<uint8_t> buf;
    size_t seed_len;
    ChunkSeeder(const uint8_t* seed, size_t len) : buf(len + 8), seed_len(len) {
        b3::context_key("pose-db/chunk-seed", key);
        if (len) std::memcpy(buf.data(), seed, len);
    }
    void seed_words(uint64_t index, uint32_t out[8]) {
        put_be64(buf.data() + seed_len, index);
        uint8_t o[32];
        b3::hash(key, b3::kDeriveKeyMaterial, buf.data(), buf.size(), o);
        for (int i = 0; i < 8; ++i) out[i] = b3::load_le(o + 4 * i);
    }
};

}  // namespace pose
