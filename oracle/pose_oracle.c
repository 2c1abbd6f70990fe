/*
 * pose_oracle.c — CPU restatement of the reference labeling path.
 *
 * TEST INFRASTRUCTURE ONLY (see pose_oracle.h): the checker for the HIP
 * labeler, never linked into the product.
 *
 * Written from the BLAKE3 specification (O'Connor, Aumasson, Neves,
 * Wilcox-O'Hearn, "BLAKE3: one function, fast everywhere", 2020) and from the
 * behaviour of the reference's common/src/hash.c and common/src/graph.c.
 * Citations are to files of the reference repository.
 */
#define _GNU_SOURCE
#include "pose_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ── BLAKE3 (spec §2) ───────────────────────────────────────────────────── */

enum {
    B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_PARENT = 4, B3_ROOT = 8,
    B3_KEYED_HASH = 16, B3_DERIVE_KEY_CONTEXT = 32, B3_DERIVE_KEY_MATERIAL = 64,
};

static const uint32_t B3_IV[8] = {
    0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
    0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u,
};

/* Message word order of each of the 7 rounds: round r uses the spec's
 * permutation applied r times to 0..15. */
static uint8_t B3_SCHED[7][16];
static void b3_init_sched(void)
{
    static const uint8_t perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
    for (int i = 0; i < 16; i++) B3_SCHED[0][i] = (uint8_t)i;
    for (int r = 1; r < 7; r++)
        for (int i = 0; i < 16; i++) B3_SCHED[r][i] = B3_SCHED[r - 1][perm[i]];
}
static pthread_once_t b3_once = PTHREAD_ONCE_INIT;

static inline uint32_t ror32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

#define B3_G(a, b, c, d, x, y) do {                      \
        s[a] = s[a] + s[b] + (x); s[d] = ror32(s[d] ^ s[a], 16); \
        s[c] = s[c] + s[d];       s[b] = ror32(s[b] ^ s[c], 12); \
        s[a] = s[a] + s[b] + (y); s[d] = ror32(s[d] ^ s[a], 8);  \
        s[c] = s[c] + s[d];       s[b] = ror32(s[b] ^ s[c], 7);  \
    } while (0)

/* Full 16-word compression output (spec §2.2). */
static void b3_compress(const uint32_t cv[8], const uint32_t m[16],
                        uint64_t counter, uint32_t block_len, uint32_t flags,
                        uint32_t out[16])
{
    uint32_t s[16];
    for (int i = 0; i < 8; i++) s[i] = cv[i];
    for (int i = 0; i < 4; i++) s[8 + i] = B3_IV[i];
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = block_len;
    s[15] = flags;
    for (int r = 0; r < 7; r++) {
        const uint8_t *o = B3_SCHED[r];
        B3_G(0, 4,  8, 12, m[o[0]],  m[o[1]]);
        B3_G(1, 5,  9, 13, m[o[2]],  m[o[3]]);
        B3_G(2, 6, 10, 14, m[o[4]],  m[o[5]]);
        B3_G(3, 7, 11, 15, m[o[6]],  m[o[7]]);
        B3_G(0, 5, 10, 15, m[o[8]],  m[o[9]]);
        B3_G(1, 6, 11, 12, m[o[10]], m[o[11]]);
        B3_G(2, 7,  8, 13, m[o[12]], m[o[13]]);
        B3_G(3, 4,  9, 14, m[o[14]], m[o[15]]);
    }
    for (int i = 0; i < 8; i++) {
        out[i]     = s[i] ^ s[i + 8];
        out[i + 8] = s[i + 8] ^ cv[i];
    }
}

static inline uint32_t le32(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
           ((uint32_t)p[3] << 24);
}

static void words_from_block(const uint8_t *blk, size_t len, uint32_t m[16])
{
    uint8_t tmp[64];
    memset(tmp, 0, sizeof(tmp));
    memcpy(tmp, blk, len);
    for (int i = 0; i < 16; i++) m[i] = le32(tmp + 4 * i);
}

static void store_words(uint8_t *out, const uint32_t *w, int n)
{
    for (int i = 0; i < n; i++) {
        out[4 * i + 0] = (uint8_t)(w[i]);
        out[4 * i + 1] = (uint8_t)(w[i] >> 8);
        out[4 * i + 2] = (uint8_t)(w[i] >> 16);
        out[4 * i + 3] = (uint8_t)(w[i] >> 24);
    }
}

/* An unfinished compression: the spec's "output" of a chunk or parent node. */
typedef struct {
    uint32_t cv[8];
    uint32_t m[16];
    uint64_t counter;
    uint32_t block_len;
    uint32_t flags;
} b3_pending;

static void b3_pending_cv(const b3_pending *p, uint32_t cv[8])
{
    uint32_t o[16];
    b3_compress(p->cv, p->m, p->counter, p->block_len, p->flags, o);
    memcpy(cv, o, 32);
}

/* Chunk i of the input (1..1024 bytes; 0 bytes only for the empty message). */
static void b3_chunk(const uint32_t key[8], uint32_t flags, const uint8_t *in,
                     size_t len, uint64_t chunk_index, b3_pending *outp)
{
    uint32_t cv[8];
    memcpy(cv, key, 32);
    size_t nblk = len == 0 ? 1 : (len + 63) / 64;
    for (size_t b = 0; b + 1 < nblk; b++) {
        uint32_t m[16], o[16];
        words_from_block(in + 64 * b, 64, m);
        b3_compress(cv, m, chunk_index, 64,
                    flags | (b == 0 ? B3_CHUNK_START : 0), o);
        memcpy(cv, o, 32);
    }
    size_t last = len - 64 * (nblk - 1);
    memcpy(outp->cv, cv, 32);
    words_from_block(in + 64 * (nblk - 1), last, outp->m);
    outp->counter   = chunk_index;
    outp->block_len = (uint32_t)last;
    outp->flags     = flags | B3_CHUNK_END | (nblk == 1 ? B3_CHUNK_START : 0);
}

static void b3_parent(const uint32_t key[8], uint32_t flags, const uint32_t l[8],
                      const uint32_t r[8], b3_pending *outp)
{
    memcpy(outp->cv, key, 32);
    memcpy(outp->m, l, 32);
    memcpy(outp->m + 8, r, 32);
    outp->counter   = 0;
    outp->block_len = 64;
    outp->flags     = flags | B3_PARENT;
}

/* Root pending node of the subtree over in[0..len) whose first chunk has index
 * chunk0 (spec §2.1: the left subtree holds the largest power-of-two number of
 * full chunks that leaves at least one byte for the right subtree). */
static void b3_subtree(const uint32_t key[8], uint32_t flags, const uint8_t *in,
                       size_t len, uint64_t chunk0, b3_pending *outp)
{
    if (len <= 1024) {
        b3_chunk(key, flags, in, len, chunk0, outp);
        return;
    }
    uint64_t nchunks = (len + 1023) / 1024;
    uint64_t left_chunks = 1;
    while (left_chunks * 2 < nchunks) left_chunks *= 2;
    size_t left_len = (size_t)left_chunks * 1024;
    b3_pending lp, rp;
    uint32_t lcv[8], rcv[8];
    b3_subtree(key, flags, in, left_len, chunk0, &lp);
    b3_subtree(key, flags, in + left_len, len - left_len, chunk0 + left_chunks, &rp);
    b3_pending_cv(&lp, lcv);
    b3_pending_cv(&rp, rcv);
    b3_parent(key, flags, lcv, rcv, outp);
}

void oracle_blake3_hash(const uint32_t *key8, uint32_t flags,
                        const uint8_t *in, size_t len, uint8_t out[32])
{
    pthread_once(&b3_once, b3_init_sched);
    const uint32_t *key = key8 ? key8 : B3_IV;
    b3_pending root;
    b3_subtree(key, flags, in, len, 0, &root);
    uint32_t o[16];
    b3_compress(root.cv, root.m, 0, root.block_len, root.flags | B3_ROOT, o);
    store_words(out, o, 8);
}

void oracle_blake3_context_key(const char *context, uint32_t key_out[8])
{
    uint8_t ck[32];
    oracle_blake3_hash(NULL, B3_DERIVE_KEY_CONTEXT, (const uint8_t *)context,
                       strlen(context), ck);
    for (int i = 0; i < 8; i++) key_out[i] = le32(ck + 4 * i);
}

void oracle_blake3_derive_key(const char *context, const uint8_t *material,
                              size_t len, uint8_t out[32])
{
    uint32_t k[8];
    oracle_blake3_context_key(context, k);
    oracle_blake3_hash(k, B3_DERIVE_KEY_MATERIAL, material, len, out);
}

/* ── PoSE node hash, keys, descriptor, chunk seed ───────────────────────── */

/* Keys: the derive-key context keys of "pose-db/label/src" / ".../int", as 8
 * little-endian words (common/src/hash.c:46-67). */
void oracle_label_node_keys(uint32_t key_src[8], uint32_t key_int[8])
{
    oracle_blake3_context_key("pose-db/label/src", key_src);
    oracle_blake3_context_key("pose-db/label/int", key_int);
}

static inline void put_be64(uint8_t *p, uint64_t v)
{
    for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (56 - 8 * i));
}

/* derive_key("pose-db/descriptor", be64(m) || be64(n) || be64(256))
 * (common/src/hash.c:198-216). */
void oracle_graph_descriptor(uint8_t out[32], uint64_t m, uint64_t n)
{
    uint8_t buf[24];
    put_be64(buf, m);
    put_be64(buf + 8, n);
    put_be64(buf + 16, 256);
    oracle_blake3_derive_key("pose-db/descriptor", buf, sizeof(buf), out);
}

/* derive_key("pose-db/chunk-seed", session_seed || be64(chunk_index))
 * (common/src/hash.c:218-233). */
void oracle_chunk_seed(uint8_t out[32], const uint8_t *session_seed,
                       size_t seed_len, uint64_t chunk_index)
{
    uint8_t *buf = (uint8_t *)malloc(seed_len + 8);
    if (!buf) { memset(out, 0, 32); return; }
    if (seed_len) memcpy(buf, session_seed, seed_len);
    put_be64(buf + seed_len, chunk_index);
    oracle_blake3_derive_key("pose-db/chunk-seed", buf, seed_len + 8, out);
    free(buf);
}

static uint32_t g_ksrc[8], g_kint[8];
static pthread_once_t keys_once = PTHREAD_ONCE_INIT;
static void keys_init(void)
{
    pthread_once(&b3_once, b3_init_sched);
    oracle_label_node_keys(g_ksrc, g_kint);
}

/* Keyed BLAKE3 of the 192-byte node message (common/src/hash.c:80-144,
 * message layout common/include/hash.h:47-57): block 0 = be64(node_index) and
 * zeros, block 1 = pred0 | pred1 (zeros when absent; a source node ignores
 * both), block 2 = seed | descriptor.  One chunk of three blocks. */
void oracle_label_node(uint8_t out[32], int is_source, uint64_t node_index,
                       const uint8_t *pred0, const uint8_t *pred1,
                       const uint8_t seed[32], const uint8_t desc[32])
{
    pthread_once(&keys_once, keys_init);
    uint8_t msg[192];
    memset(msg, 0, sizeof(msg));
    put_be64(msg, node_index);
    if (!is_source && pred0) memcpy(msg + 64, pred0, 32);
    if (!is_source && pred1) memcpy(msg + 96, pred1, 32);
    memcpy(msg + 128, seed, 32);
    memcpy(msg + 160, desc, 32);
    uint32_t cv[8], m[16], o[16];
    memcpy(cv, is_source ? g_ksrc : g_kint, 32);
    for (int b = 0; b < 3; b++) {
        uint32_t fl = B3_KEYED_HASH | (b == 0 ? B3_CHUNK_START : 0) |
                      (b == 2 ? (B3_CHUNK_END | B3_ROOT) : 0);
        words_from_block(msg + 64 * b, 64, m);
        b3_compress(cv, m, 0, 64, fl, o);
        memcpy(cv, o, 32);
    }
    store_words(out, cv, 8);
}

/* ── graph sizes ────────────────────────────────────────────────────────── */

/* n = max(0, bitlen(m-1) - 1) (common/src/graph.c:562-570). */
int oracle_graph_parameter_n(uint64_t m)
{
    if (m <= 1) return 0;
    int bits = 64 - __builtin_clzll(m - 1);
    return bits - 1 < 0 ? 0 : bits - 1;
}

/* Butterfly B(d) = (d+1)·2^d; connected C(l) = 2C(l-1) + 2B(l-1); standalone
 * S(l) = S(l-1) + B(l-1) + C(l-1); C(0) = S(0) = 1 (common/src/graph.c:33-54). */
static uint64_t bfly_count(int d) { return (uint64_t)(d + 1) << d; }
static uint64_t conn_count(int l) { return l == 0 ? 1 : 2 * conn_count(l - 1) + 2 * bfly_count(l - 1); }
static uint64_t stand_count(int l)
{
    return l == 0 ? 1 : stand_count(l - 1) + bfly_count(l - 1) + conn_count(l - 1);
}

/* Scaffold = two standalone copies of level n+1 (common/src/graph.c:647-652). */
uint64_t oracle_scaffold_node_count(uint64_t m)
{
    return 2 * stand_count(oracle_graph_parameter_n(m) + 1);
}

/* ── emitter shared by the two labeler formulations ─────────────────────── */

/*
 * The graph is defined by the order in which its components are emitted; node
 * ids are assigned in emission order (common/src/graph.c:99-520):
 *
 *   standalone(l):  standalone(l-1) -> L;  connector(l-1, L) -> C;
 *                   connected(l-1, C) -> R;  base = L ++ R
 *   connected(l, in[2^l]):  connected(l-1, in[0:h]) -> L;
 *                   merged(l-1, primary = L, ingress = in[h:]) -> C;
 *                   connected(l-1, C) -> R;  base = L ++ R
 *   connected(0, in[1]) / standalone(0): one node (1 pred / source)
 *   connector(d, in): layer 0 node(in[i]); layers k = 1..d with
 *                   bit = 2^(d-k): node(prev[i & ~bit], prev[i | bit])
 *   merged(d, P, I): connector(d, I) -> G; layer 0 node(P[i], G[i]);
 *                   layers k = 1..d as in a connector.
 *
 * (merged(0) is a one-node connector followed by one node(P[0], G[0]), the
 * reference's dim==0 special case, common/src/graph.c:265-275.)
 *
 * The operations are abstract: a "value" is a node handle (an id in edge mode,
 * a label pointer in in-place mode).  Lower-id predecessor first: within a
 * butterfly layer the lower id sits at i & ~bit; in merged layer 0 the primary
 * input was emitted before the ingress output (common/src/graph.c:660-676).
 */
typedef struct emit_ops emit_ops;
struct emit_ops {
    /* new node with 0/1/2 preds; value written to *dst (value size vbytes) */
    void (*node)(emit_ops *e, void *dst, const void *p0, const void *p1, int np);
    size_t vbytes;
    uint8_t *arena;    /* bump arena of values */
    size_t arena_off;  /* in values */
    size_t arena_cap;
};

#define V(base, i) ((uint8_t *)(base) + (size_t)(i) * e->vbytes)

static void *e_alloc(emit_ops *e, size_t n)
{
    void *p = e->arena + e->arena_off * e->vbytes;
    e->arena_off += n;
    if (e->arena_off > e->arena_cap) abort();
    return p;
}

static void e_connector(emit_ops *e, int dim, const uint8_t *in, uint8_t *out)
{
    size_t saved = e->arena_off;
    int width = 1 << dim;
    uint8_t *a = (uint8_t *)e_alloc(e, (size_t)width);
    uint8_t *b = (uint8_t *)e_alloc(e, (size_t)width);
    for (int i = 0; i < width; i++)
        e->node(e, V(a, i), V(in, i), NULL, 1);
    for (int k = 1; k <= dim; k++) {
        int bit = 1 << (dim - k);
        for (int i = 0; i < width; i++)
            e->node(e, V(b, i), V(a, i & ~bit), V(a, i | bit), 2);
        uint8_t *t = a; a = b; b = t;
    }
    memcpy(out, a, (size_t)width * e->vbytes);
    e->arena_off = saved;
}

static void e_merged(emit_ops *e, int dim, const uint8_t *primary,
                     const uint8_t *ingress, uint8_t *out)
{
    size_t saved = e->arena_off;
    int width = 1 << dim;
    uint8_t *g = (uint8_t *)e_alloc(e, (size_t)width);
    uint8_t *a = (uint8_t *)e_alloc(e, (size_t)width);
    uint8_t *b = (uint8_t *)e_alloc(e, (size_t)width);
    e_connector(e, dim, ingress, g);
    for (int i = 0; i < width; i++)
        e->node(e, V(a, i), V(primary, i), V(g, i), 2);
    for (int k = 1; k <= dim; k++) {
        int bit = 1 << (dim - k);
        for (int i = 0; i < width; i++)
            e->node(e, V(b, i), V(a, i & ~bit), V(a, i | bit), 2);
        uint8_t *t = a; a = b; b = t;
    }
    memcpy(out, a, (size_t)width * e->vbytes);
    e->arena_off = saved;
}

static void e_connected(emit_ops *e, int level, const uint8_t *in, uint8_t *base)
{
    if (level == 0) {
        e->node(e, V(base, 0), V(in, 0), NULL, 1);
        return;
    }
    int half = 1 << (level - 1);
    size_t saved = e->arena_off;
    uint8_t *l = (uint8_t *)e_alloc(e, (size_t)half);
    uint8_t *c = (uint8_t *)e_alloc(e, (size_t)half);
    uint8_t *r = (uint8_t *)e_alloc(e, (size_t)half);
    e_connected(e, level - 1, in, l);
    e_merged(e, level - 1, l, V(in, half), c);
    e_connected(e, level - 1, c, r);
    memcpy(base, l, (size_t)half * e->vbytes);
    memcpy(V(base, half), r, (size_t)half * e->vbytes);
    e->arena_off = saved;
}

static void e_standalone(emit_ops *e, int level, uint8_t *base)
{
    if (level == 0) {
        e->node(e, V(base, 0), NULL, NULL, 0);
        return;
    }
    int half = 1 << (level - 1);
    size_t saved = e->arena_off;
    uint8_t *l = (uint8_t *)e_alloc(e, (size_t)half);
    uint8_t *c = (uint8_t *)e_alloc(e, (size_t)half);
    uint8_t *r = (uint8_t *)e_alloc(e, (size_t)half);
    e_standalone(e, level - 1, l);
    e_connector(e, level - 1, l, c);
    e_connected(e, level - 1, c, r);
    memcpy(base, l, (size_t)half * e->vbytes);
    memcpy(V(base, half), r, (size_t)half * e->vbytes);
    e->arena_off = saved;
}

/* Arena bound: the reference sizes its arena at 16·half_w values
 * (common/src/graph.c:60-65); the recursion's live set peaks near 6.5·half_w. */
#define ORACLE_ARENA_FACTOR 16u

/* ── edge mode (common/src/graph.c:1044-1100) ───────────────────────────── */

typedef struct {
    emit_ops base;
    uint64_t next_id;
    uint8_t *np;
    uint32_t *p0, *p1;
} edge_emitter;

static void edge_node(emit_ops *eo, void *dst, const void *a, const void *b, int np)
{
    edge_emitter *e = (edge_emitter *)eo;
    uint64_t id = e->next_id++;
    uint32_t x = a ? *(const uint32_t *)a : 0, y = b ? *(const uint32_t *)b : 0;
    if (np == 2 && x > y) { uint32_t t = x; x = y; y = t; }
    e->np[id] = (uint8_t)np;
    e->p0[id] = np >= 1 ? x : 0;
    e->p1[id] = np >= 2 ? y : 0;
    *(uint32_t *)dst = (uint32_t)id;
}

int oracle_graph_edges(uint64_t m, uint8_t *np, uint32_t *p0, uint32_t *p1,
                       int32_t *rank)
{
    if (m == 0 || m > (1u << 20)) return -1;
    int n = oracle_graph_parameter_n(m);
    int level = n + 1;
    uint64_t half_w = (uint64_t)1 << n, retained = m - half_w;
    uint64_t total = oracle_scaffold_node_count(m);
    uint32_t *lb = (uint32_t *)malloc(2 * half_w * 4);
    uint32_t *rb = (uint32_t *)malloc(2 * half_w * 4);
    uint8_t *arena = (uint8_t *)malloc((size_t)ORACLE_ARENA_FACTOR * half_w * 4 + 64);
    if (!lb || !rb || !arena) { free(lb); free(rb); free(arena); return -1; }
    edge_emitter e;
    memset(&e, 0, sizeof(e));
    e.base.node = edge_node;
    e.base.vbytes = 4;
    e.base.arena = arena;
    e.base.arena_cap = (size_t)ORACLE_ARENA_FACTOR * half_w + 16;
    e.np = np; e.p0 = p0; e.p1 = p1;
    e_standalone(&e.base, level, (uint8_t *)lb);
    e_standalone(&e.base, level, (uint8_t *)rb);
    int rc = (e.next_id == total) ? 0 : -1;
    if (rank) {
        for (uint64_t i = 0; i < total; i++) rank[i] = -1;
        for (uint64_t i = 0; i < half_w; i++) rank[lb[half_w + i]] = (int32_t)i;
        for (uint64_t i = 0; i < retained; i++) rank[rb[half_w + i]] = (int32_t)(half_w + i);
    }
    free(lb); free(rb); free(arena);
    return rc;
}

/* ── formulation 1: id-indexed topological labeling ──────────────────────
 * Emit the edges, then label every node in id order into an id-indexed
 * buffer (the full-scratch prover, common/src/graph.c:575-645), and gather
 * the output set by challenge rank. */
int oracle_label_chunk_edges(const uint8_t cseed[32], uint64_t m, uint8_t *out)
{
    uint64_t total = oracle_scaffold_node_count(m);
    uint8_t *np = (uint8_t *)malloc(total);
    uint32_t *p0 = (uint32_t *)malloc(total * 4), *p1 = (uint32_t *)malloc(total * 4);
    int32_t *rank = (int32_t *)malloc(total * 4);
    uint8_t *lab = (uint8_t *)malloc(total * 32);
    int rc = -1;
    if (!np || !p0 || !p1 || !rank || !lab) goto done;
    if (oracle_graph_edges(m, np, p0, p1, rank) != 0) goto done;
    uint8_t desc[32];
    oracle_graph_descriptor(desc, m, (uint64_t)oracle_graph_parameter_n(m));
    for (uint64_t id = 0; id < total; id++) {
        oracle_label_node(lab + id * 32, np[id] == 0, id,
                          np[id] >= 1 ? lab + (uint64_t)p0[id] * 32 : NULL,
                          np[id] >= 2 ? lab + (uint64_t)p1[id] * 32 : NULL,
                          cseed, desc);
        if (rank[id] >= 0) memcpy(out + (uint64_t)rank[id] * 32, lab + id * 32, 32);
    }
    rc = 0;
done:
    free(np); free(p0); free(p1); free(rank); free(lab);
    return rc;
}

/* ── formulation 2: in-place arena labeling ─────────────────────────────
 * Values are the 32-byte labels themselves, held in the bump arena; only the
 * output set survives (common/src/graph.c:654-1001). */
typedef struct {
    emit_ops base;
    uint64_t next_id;
    const uint8_t *seed;
    uint8_t desc[32];
} label_emitter;

static void label_node_cb(emit_ops *eo, void *dst, const void *a, const void *b, int np)
{
    label_emitter *e = (label_emitter *)eo;
    uint64_t id = e->next_id++;
    oracle_label_node((uint8_t *)dst, np == 0, id, (const uint8_t *)a,
                      (const uint8_t *)b, e->seed, e->desc);
}

int oracle_label_chunk_inplace(const uint8_t cseed[32], uint64_t m, uint8_t *out)
{
    if (m == 0 || m > (1u << 20)) return -1;
    int n = oracle_graph_parameter_n(m);
    int level = n + 1;
    uint64_t half_w = (uint64_t)1 << n, retained = m - half_w;
    uint64_t left_total = oracle_scaffold_node_count(m) / 2;
    size_t arena_vals = (size_t)ORACLE_ARENA_FACTOR * half_w + 16;
    uint8_t *lb = (uint8_t *)malloc(2 * half_w * 32);
    uint8_t *rb = (uint8_t *)malloc(2 * half_w * 32);
    uint8_t *arena = (uint8_t *)malloc(arena_vals * 32);
    if (!lb || !rb || !arena) { free(lb); free(rb); free(arena); return -1; }
    label_emitter e;
    memset(&e, 0, sizeof(e));
    e.base.node = label_node_cb;
    e.base.vbytes = 32;
    e.base.arena = arena;
    e.base.arena_cap = arena_vals;
    e.seed = cseed;
    oracle_graph_descriptor(e.desc, m, (uint64_t)n);
    e_standalone(&e.base, level, lb);
    e.next_id = left_total;   /* right copy continues at left_total */
    e_standalone(&e.base, level, rb);
    for (uint64_t i = 0; i < half_w; i++) memcpy(out + i * 32, lb + (half_w + i) * 32, 32);
    for (uint64_t i = 0; i < retained; i++)
        memcpy(out + (half_w + i) * 32, rb + (half_w + i) * 32, 32);
    free(lb); free(rb); free(arena);
    return 0;
}

/* ── region labeler (common/src/graph.c:1604-1689) ─────────────────────── */

typedef struct {
    const uint8_t *session_seed;
    size_t seed_len;
    uint64_t region_block_offset;
    uint32_t cb;
    const uint64_t *js;      /* super-chunk indices, or NULL for 0..n-1 */
    size_t n;
    uint8_t *out;            /* n * cb * 32 */
    atomic_size_t next;
    atomic_int rc;
} region_job;

static void *region_worker(void *arg)
{
    region_job *r = (region_job *)arg;
    size_t super_bytes = (size_t)r->cb * 32;
    for (;;) {
        size_t k = atomic_fetch_add(&r->next, 1);
        if (k >= r->n) break;
        uint64_t j = r->js ? r->js[k] : k;
        uint64_t chunk_index = (r->region_block_offset + j * (uint64_t)r->cb) / r->cb;
        uint8_t cseed[32];
        oracle_chunk_seed(cseed, r->session_seed, r->seed_len, chunk_index);
        if (oracle_label_chunk_inplace(cseed, r->cb, r->out + k * super_bytes) != 0)
            atomic_store(&r->rc, -1);
    }
    return NULL;
}

static int run_region_job(region_job *r, int threads)
{
    if (threads < 1) threads = 1;
    if ((size_t)threads > r->n) threads = (int)r->n;
    pthread_t *t = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    if (!t) return -1;
    int started = 0;
    for (int i = 1; i < threads; i++)
        if (pthread_create(&t[i], NULL, region_worker, r) == 0) started = i;
        else break;
    region_worker(r);
    for (int i = 1; i <= started; i++) pthread_join(t[i], NULL);
    free(t);
    return atomic_load(&r->rc);
}

int oracle_label_region(const uint8_t *session_seed, size_t seed_len,
                        uint8_t *region, size_t region_len,
                        uint64_t region_block_offset, uint32_t chunk_blocks,
                        int threads)
{
    if (chunk_blocks == 0) chunk_blocks = 4096;
    if (!region || region_len % 32) return -1;
    size_t super_bytes = (size_t)chunk_blocks * 32;
    region_job r;
    memset(&r, 0, sizeof(r));
    r.session_seed = session_seed; r.seed_len = seed_len;
    r.region_block_offset = region_block_offset; r.cb = chunk_blocks;
    r.n = region_len / super_bytes; r.out = region;
    if (r.n == 0) return 0;
    return run_region_job(&r, threads);
}

int oracle_label_superchunks(const uint8_t *session_seed, size_t seed_len,
                             uint64_t region_block_offset, uint32_t chunk_blocks,
                             const uint64_t *js, size_t n, uint8_t *out,
                             int threads)
{
    if (chunk_blocks == 0) chunk_blocks = 4096;
    region_job r;
    memset(&r, 0, sizeof(r));
    r.session_seed = session_seed; r.seed_len = seed_len;
    r.region_block_offset = region_block_offset; r.cb = chunk_blocks;
    r.js = js; r.n = n; r.out = out;
    if (n == 0) return 0;
    return run_region_job(&r, threads);
}

/* SplitMix64 (Steele, Lea, Flood, OOPSLA 2014). */
uint64_t oracle_splitmix64_next(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
