/*
 * qcp_oracle.h -- CPU oracle for the Hilbert bulk-load + range-search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker: only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it.  The product library never links, loads or calls it.
 *
 * Every function restates the reference's CPU semantics (file:line citations
 * are relative to the reference repository root) in plain C11, with the fanout
 * F, the dimensionality D, the key mode and the key source as runtime
 * parameters instead of the reference's compile-time constants.
 *
 * Parity pins (what each function is checked against in tests/):
 *   orc_hilbert_c2i      bijection on small grids, unit-step adjacency of
 *                        consecutive keys, round trip through orc_hilbert_i2c,
 *                        agreement of the induced order with orc_hilbert_cmp
 *                        (an independent code path), 2-D order-1/2 tables.
 *   orc_sort_stable      sortedness + permutation + stable ties (closed form).
 *   orc_pack             parent box == exact min/max of children, level
 *                        counts == ceil-div chain (brute-force recomputation).
 *   orc_bvh_build        child ranges partition [0,N), leaf size <= F,
 *                        children <= F-1, split positions match a brute-force
 *                        common-prefix rule.
 *   orc_guttman_build    fill bounds (F/2..F below the root), one leaf level,
 *                        exact covers, a permutation of the inputs, and
 *                        hand-derived splits and descents at small F.
 *   searches             DFS == MPRS == brute-force closed-interval filter.
 *   partitions           split sizes against the reference's rule, every
 *                        partition's subtree and boxes against orc_packed_
 *                        topology / orc_pack of its range (Python mapping of
 *                        the depth-major ids), DFS == brute force, partition
 *                        MPRS == brute force.
 */
#ifndef QCP_ORACLE_H
#define QCP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_KEY_REF = 0, ORC_KEY_B16 = 1 };
enum { ORC_SRC_LO = 0, ORC_SRC_CENTER = 1 };

/* ---- Hilbert mapping (src/mapper/hilbert_mapper.cpp) ---------------------- */
uint64_t orc_bit_transpose(unsigned nd, unsigned nb, uint64_t in_coords);
uint64_t orc_hilbert_c2i(unsigned nd, unsigned nb, const uint64_t *coord);
void     orc_hilbert_i2c(unsigned nd, unsigned nb, uint64_t index, uint64_t *coord);
int      orc_hilbert_cmp(unsigned nd, unsigned nbytes, unsigned nb,
                         const void *c1, const void *c2);
int64_t  orc_quant_ref(float p);
uint32_t orc_quant_b16(float p);
unsigned orc_ref_bits(unsigned D);
/* keys for n objects (points: n x D, boxes: n x 2D rows [lo(D),hi(D)]).
 * Returns 0, or 1 if a REF-mode coordinate falls outside [0, 2^bits). */
int      orc_keys(const float *objs, uint64_t n, unsigned D, int is_box,
                  int key_mode, int key_src, uint64_t *keys);

/* ---- sort (src/sort/thrust_sorter.cpp, src/sort/parallel_sorter.cpp) ----- */
void     orc_sort_stable(const uint64_t *keys, uint64_t n, uint32_t *perm);

/* ---- packed (MPHR / Hilbert-packed R-tree) levels (src/tree/tree.cpp) ----- */
int      orc_level_counts(uint64_t n, unsigned F, uint64_t *counts, int max_levels);
/* sorted objects (AoS, rank order) -> node boxes for every node, level order
 * (root first), AoS 2D floats per node.  Returns the number of nodes. */
uint64_t orc_pack(const float *sobjs, uint64_t n, unsigned D, int is_box,
                  unsigned F, float *node_boxes);

/* ---- BVH topology (src/tree/tree.cpp:267-408) ---------------------------- */
/* Computes the BVH over ranks 0..n-1 in BFS order.  For node i:
 *   node_first[i]  first child node id (internal) or first rank (leaf)
 *   node_count[i]  number of children / entries
 *   node_leaf[i]   1 for a leaf
 *   node_level[i]  depth
 * Returns the node count, or -1 if cap is too small. */
int64_t  orc_bvh_topology(uint64_t n, unsigned F, int64_t cap,
                          uint32_t *node_first, uint32_t *node_count,
                          uint8_t *node_leaf, uint8_t *node_level);
/* node boxes of an explicit tree, padded-min/max rule (src/node/node.cpp:19-70);
 * F = 0 gives exact covers (no padding). */
void     orc_tree_boxes(const float *sobjs, unsigned D, int is_box, unsigned F,
                        int64_t n_nodes, const uint32_t *node_first,
                        const uint32_t *node_count, const uint8_t *node_leaf,
                        float *node_boxes);
/* explicit representation of a packed tree (same BFS arrays as the BVH). */
int64_t  orc_packed_topology(uint64_t n, unsigned F, int64_t cap,
                             uint32_t *node_first, uint32_t *node_count,
                             uint8_t *node_leaf, uint8_t *node_level);

/* ---- search ---------------------------------------------------------------- */
/* Per-ID term of the ID-set hash: id*id + 0x9E3779B97F4A7C15*id mod 2^64. */
uint64_t orc_id_hash(uint32_t id);
/* Canonical top-down DFS (src/tree/bvh.cpp:256-278) over an explicit tree.
 * ids/offsets optional (NULL).  If ids != NULL, offsets[Q+1] must be given and
 * already hold the exclusive prefix of counts (two-pass use), and ids receives
 * the 1-based rank IDs in ascending order per query.
 * int_entries / leaf_entries (optional) receive per-query sums of the entry
 * counts of visited internal / leaf nodes (algorithmic-bytes accounting). */
void     orc_search_dfs(const float *sobjs, unsigned D, int is_box,
                        int64_t n_nodes, const uint32_t *node_first,
                        const uint32_t *node_count, const uint8_t *node_leaf,
                        const float *node_boxes,
                        const float *queries, uint64_t q0, uint64_t q1,
                        uint32_t *counts, uint64_t *hashes,
                        const uint64_t *offsets, uint32_t *ids,
                        uint64_t *int_entries, uint64_t *leaf_entries);
/* Same search over nthreads POSIX threads with the reference's static query
 * chunking (src/tree/bvh.cpp:195-220).  Counts + hashes only. */
void     orc_search_dfs_mt(const float *sobjs, unsigned D, int is_box,
                           int64_t n_nodes, const uint32_t *node_first,
                           const uint32_t *node_count, const uint8_t *node_leaf,
                           const float *node_boxes,
                           const float *queries, uint64_t nq, unsigned nthreads,
                           uint32_t *counts, uint64_t *hashes);
/* Both searches from each of the roots 0 .. roots-1 in order (a partitioned
 * MPHR has P roots; one root otherwise). */
void     orc_search_dfs_roots(const float *sobjs, unsigned D, int is_box,
                              int64_t n_nodes, const uint32_t *node_first,
                              const uint32_t *node_count, const uint8_t *node_leaf,
                              const float *node_boxes, unsigned roots,
                              const float *queries, uint64_t q0, uint64_t q1,
                              uint32_t *counts, uint64_t *hashes,
                              const uint64_t *offsets, uint32_t *ids,
                              uint64_t *int_entries, uint64_t *leaf_entries);
void     orc_search_dfs_mt_roots(const float *sobjs, unsigned D, int is_box,
                                 int64_t n_nodes, const uint32_t *node_first,
                                 const uint32_t *node_count, const uint8_t *node_leaf,
                                 const float *node_boxes, unsigned roots,
                                 const float *queries, uint64_t nq, unsigned nthreads,
                                 uint32_t *counts, uint64_t *hashes);
/* MPRS restart-scanning + parent-check (src/tree/mphr.cpp:303-438) over a
 * packed tree, serial.  visits[2*q] = root visits, visits[2*q+1] = node visits. */
void     orc_search_mprs(const float *sobjs, uint64_t n, unsigned D, int is_box,
                         unsigned F, const float *node_boxes,
                         const float *queries, uint64_t nq,
                         uint32_t *counts, uint64_t *hashes,
                         const uint64_t *offsets, uint32_t *ids,
                         uint32_t *visits);
/* The build over nthreads POSIX threads (the reference's threaded CPU build:
 * key mapping chunks, a parallel sort, gather, bottom-up boxes), for the CPU
 * baseline.  Outputs equal orc_keys + orc_sort_stable + orc_pack: keys[n] are
 * the sorted keys, perm[n] the input index of each rank, sobjs[n*E] the
 * rank-ordered objects, node_boxes the packed tree's MBRs.  ms[4] = wall ms of
 * keys, sort, gather, pack.  Returns nonzero on a domain error. */
int      orc_build_mt(const float *objs, uint64_t n, unsigned D, int is_box, int key_mode,
                      int key_src, unsigned F, unsigned nthreads, uint64_t *keys,
                      uint32_t *perm, float *sobjs, float *node_boxes, double *ms);
/* orc_search_mprs over nthreads POSIX threads, static query chunks. */
void     orc_search_mprs_mt(const float *sobjs, uint64_t n, unsigned D, int is_box, unsigned F,
                            const float *node_boxes, const float *queries, uint64_t nq,
                            unsigned nthreads, uint32_t *counts, uint64_t *hashes);
/* k-means + Hilbert sort keys (this build's defined variant of the reference's
 * KmeansMapper, src/mapper/kmeans_mapper.cpp:15-259): K clusters seeded at
 * init_idx[K]; Lloyd iterations (assignment, then centre update) repeated
 * while (float)changes / (float)n > 1e-5 and loop++ < max_loop (the
 * reference's rule, kmeans_mapper.cpp:170, with its max_loop 1000); keys[n] =
 * composite sort keys, cid[n] = final clusters, *iterations (optional) = the
 * iterations run.  Follow with orc_sort_stable. */
void     orc_kmeans_keys(const float *objs, uint64_t n, unsigned D, int is_box, int key_mode,
                         int key_src, unsigned K, unsigned max_loop, const uint32_t *init_idx,
                         uint64_t *keys, uint32_t *cid, unsigned *iterations);
/* Guttman R-tree (src/tree/rtree.h; Tree::RTree_Top_Down and SetNodeIndex,
 * src/tree/tree.cpp:126-185,244-256): inserts objs[i] in index order (node
 * capacity F, minimum fill F/2, quadratic split, spherical-volume metric in
 * FP32), then writes the BFS node arrays as orc_bvh_topology does, leaves
 * holding depth-first positions, and order[p] = index of the object at
 * depth-first position p.  Returns the node count, or -1 if cap is too small.
 * Node boxes: orc_tree_boxes with F = 0 (exact covers, no padding). */
int64_t  orc_guttman_build(const float *objs, uint64_t n, unsigned D, int is_box, unsigned F,
                           int64_t cap, uint32_t *node_first, uint32_t *node_count,
                           uint8_t *node_leaf, uint8_t *node_level, uint32_t *order);
/* ---- MPHR data partitions (src/tree/mphr.cpp:60-108, 235-257, 314-333) ---- */
/* Partition p's rank range: chunk = n/P, partition 0 takes chunk + n%P. */
void     orc_partition_range(uint64_t n, unsigned P, unsigned p, uint64_t *begin,
                             uint64_t *count);
/* The P packed subtrees as one explicit tree, numbered depth-major (partition
 * p's root is node p); arrays as orc_bvh_topology.  -1 if cap is too small or
 * P > n.  Node boxes: orc_tree_boxes with F (the padded rule). */
int64_t  orc_partitioned_topology(uint64_t n, unsigned F, unsigned P, int64_t cap,
                                  uint32_t *node_first, uint32_t *node_count,
                                  uint8_t *node_leaf, uint8_t *node_level);
/* MPRS over every partition's own subtree, per-query results summed over the
 * partitions (counts, hashes, visits[2q] root / [2q+1] node visits). */
void     orc_search_mprs_part(const float *sobjs, uint64_t n, unsigned D, int is_box,
                              unsigned F, unsigned P, const float *queries, uint64_t nq,
                              uint32_t *counts, uint64_t *hashes, uint32_t *visits);
/* Closed-form brute force: every object against every query. */
void     orc_search_brute(const float *sobjs, uint64_t n, unsigned D, int is_box,
                          const float *queries, uint64_t nq,
                          uint32_t *counts, uint64_t *hashes,
                          const uint64_t *offsets, uint32_t *ids);

#ifdef __cplusplus
}
#endif
#endif
