/*
 * pose_oracle.h — CPU restatement of the reference PoSE-DB labeling path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP labeler:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product library never links or calls it.
 *
 * Every function restates a reference function; the citation is given at the
 * definition in pose_oracle.c (paths relative to the reference repository).
 *
 * Pinning (what checks each function beyond its own formula):
 *   oracle_blake3_hash          official BLAKE3 vectors (common/test/test_hash.c:54-66)
 *                               + fixtures from the reference build (keyed, derive-key,
 *                               multi-chunk lengths)
 *   oracle_label_node_keys,
 *   oracle_graph_descriptor,
 *   oracle_chunk_seed,
 *   oracle_label_node          fixtures from the reference build
 *   oracle_graph_parameter_n   common/test/test_graph.c:29-41 KAT
 *   oracle_scaffold_node_count closed form 2*S(n+1) (common/src/graph.c:33-54)
 *   oracle_graph_edges         fixture digests of (np,p0,p1,rank) from the reference
 *                              build; invariants (pred<id, 2 sources, cb ranks)
 *   oracle_label_chunk_edges   == oracle_label_chunk_inplace (two formulations),
 *   oracle_label_chunk_inplace fixture digests of whole super-chunks from the
 *                              reference build
 */
#ifndef POSE_ORACLE_H
#define POSE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Generic BLAKE3: key8 == NULL means the IV (unkeyed / derive-key context). */
void oracle_blake3_hash(const uint32_t *key8, uint32_t flags,
                        const uint8_t *in, size_t len, uint8_t out[32]);
void oracle_blake3_context_key(const char *context, uint32_t key_out[8]);
void oracle_blake3_derive_key(const char *context, const uint8_t *material,
                              size_t len, uint8_t out[32]);

void oracle_label_node_keys(uint32_t key_src[8], uint32_t key_int[8]);
void oracle_graph_descriptor(uint8_t out[32], uint64_t m, uint64_t n);
void oracle_chunk_seed(uint8_t out[32], const uint8_t *session_seed,
                       size_t seed_len, uint64_t chunk_index);
void oracle_label_node(uint8_t out[32], int is_source, uint64_t node_index,
                       const uint8_t *pred0, const uint8_t *pred1,
                       const uint8_t seed[32], const uint8_t desc[32]);

int      oracle_graph_parameter_n(uint64_t m);
uint64_t oracle_scaffold_node_count(uint64_t m);

/* Topology in id order: np[id] in {0,1,2}; p0 <= p1 (ids); rank[id] = challenge
 * rank or -1.  Arrays must hold oracle_scaffold_node_count(m) entries. */
int oracle_graph_edges(uint64_t m, uint8_t *np, uint32_t *p0, uint32_t *p1,
                       int32_t *rank);

/* One super-chunk: out[m*32] = output labels in challenge-rank order for the
 * 32-byte chunk seed `cseed`.  Two independent formulations. */
int oracle_label_chunk_inplace(const uint8_t cseed[32], uint64_t m, uint8_t *out);
int oracle_label_chunk_edges(const uint8_t cseed[32], uint64_t m, uint8_t *out);

/* Region labeler: super-chunk j of region gets chunk index
 * (region_block_offset + j*cb)/cb; threads >= 1 workers.  Tail smaller than a
 * super-chunk is left untouched.  Returns 0 / -1. */
int oracle_label_region(const uint8_t *session_seed, size_t seed_len,
                        uint8_t *region, size_t region_len,
                        uint64_t region_block_offset, uint32_t chunk_blocks,
                        int threads);

/* Label only the listed super-chunk indices j (absolute within the region);
 * out[k*cb*32] receives super-chunk js[k]. */
int oracle_label_superchunks(const uint8_t *session_seed, size_t seed_len,
                             uint64_t region_block_offset, uint32_t chunk_blocks,
                             const uint64_t *js, size_t n, uint8_t *out,
                             int threads);

/* SplitMix64 stream used to derive benchmark keys and workload draws. */
uint64_t oracle_splitmix64_next(uint64_t *state);

#ifdef __cplusplus
}
#endif

#endif
