/*
 * qcp.h -- C ABI of the MI355X-native Hilbert bulk-load + range-search library
 * (libqcp.so).  Plain C types only: pointers, sizes and status codes.
 *
 * Every entry point names the reference interface it replaces (paths are
 * relative to the reference repository root).  Semantics carried over:
 *   - objects are rank-sorted by a Hilbert key of their lower corner (or centre)
 *     and addressed by their 1-based Hilbert rank ("rank ID"), exactly as the
 *     reference overwrites each branch index with its rank
 *     (src/sort/thrust_sorter.cpp:39-47);
 *   - a query box overlaps an entry unless q.lo[d] > e.hi[d] or q.hi[d] < e.lo[d]
 *     for some d, in FP32 (src/node/node_soa.cpp:94-113);
 *   - per-query hit counts (the reference's per-block h_hit[],
 *     src/tree/mphr.cpp:244-252) plus, new here, sorted rank-ID lists and an
 *     order-independent ID-set hash: the sum over hit rank IDs of
 *     id*id + 0x9E3779B97F4A7C15*id, mod 2^64 (a function of the count, the sum
 *     and the sum of squares of the IDs; sets equal in all three collide).
 * Conventions that differ from the reference on purpose: every error is a
 * status code (nothing calls exit()), every handle owns its device memory and
 * HIP stream (no device globals, no singletons), and inputs are caller-owned.
 * A handle is used by one host thread at a time; distinct handles are
 * independent.
 */
#ifndef QCP_H
#define QCP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  QCP_OK = 0,
  QCP_EINVAL = -1,   /* bad argument or configuration                         */
  QCP_ENOMEM = -2,   /* device or host allocation failed                      */
  QCP_EHIP = -3,     /* a HIP runtime call failed                             */
  QCP_EIO = -4,      /* file open/read/write failed                           */
  QCP_EFORMAT = -5,  /* index file malformed, truncated or checksum mismatch  */
  QCP_EBUSY = -6,    /* result buffer too small / admission refused           */
  QCP_EDOMAIN = -7,  /* an object coordinate lies outside the key quantiser's domain */
  QCP_ENODEV = -8    /* no usable GPU                                         */
} qcp_status;

/* Node layouts: MPHR and RTREE are Hilbert-packed (F consecutive ranks per
 * leaf); BVH splits rank ranges at common-prefix positions; GUTTMAN inserts the
 * Hilbert-ordered objects one by one on the host (quadratic split, minimum
 * fill F/2; the reference RTree, src/tree/tree.cpp:126-185) and numbers the
 * entries depth-first (src/tree/tree.cpp:244-256), so its entry IDs are
 * depth-first positions rather than Hilbert ranks; fanout >= 4, not with
 * QCP_BUILD_GRAPH or qcp_pipeline_run. */
enum { QCP_LAYOUT_MPHR = 0, QCP_LAYOUT_RTREE = 1, QCP_LAYOUT_BVH = 2, QCP_LAYOUT_GUTTMAN = 3 };
enum { QCP_KEY_REF = 0, QCP_KEY_B16 = 1 };      /* (ll)(1e6f*p) @ 31|20 bits, or 16 bits */
enum { QCP_SRC_LO = 0, QCP_SRC_CENTER = 1 };    /* key from lower corner or box centre  */
enum { QCP_OBJ_POINTS = 0, QCP_OBJ_BOXES = 1 };
/* qcp_build_cfg.flags.  QCP_BUILD_GRAPH: run the build as a captured hipGraph.
 * QCP_BUILD_NO_PERM: the index need not keep the input index of each rank (the
 * reference keeps none: it overwrites each branch index by its rank,
 * src/sort/thrust_sorter.cpp:39-47); qcp_index_info.has_perm tells whether it did.
 * The key sort either sorts (key, input index) and then gathers every object
 * through the permutation (QCP_BUILD_SORT_INDEX), or moves the objects
 * themselves (and the input indices unless QCP_BUILD_NO_PERM) through every
 * pass as the keys' payload, so that no object is read at a random place
 * (QCP_BUILD_SORT_OBJECTS; not for GUTTMAN layouts or the k-means order).
 * Neither flag: the object form once the objects take 64 MiB or more (for
 * layouts and orders it serves).  Both forms give the same index. */
enum { QCP_BUILD_GRAPH = 1, QCP_BUILD_NO_PERM = 2, QCP_BUILD_SORT_INDEX = 4,
       QCP_BUILD_SORT_OBJECTS = 8 };
enum { QCP_ORDER_HILBERT = 0, QCP_ORDER_KMEANS_HILBERT = 1 };
#define QCP_MAX_CLUSTERS (1u << 20) /* k-means order: largest cluster count */

enum { QCP_SEARCH_PERSISTENT = 0, QCP_SEARCH_PLAIN = 1, QCP_SEARCH_HYBRID = 2 };
enum { QCP_OUT_COUNTS = 1, QCP_OUT_HASH = 2, QCP_OUT_IDS = 4, QCP_OUT_VISITS = 8 };

typedef struct {
  uint32_t dims;      /* 2 or 3                                                     */
  uint32_t fanout;    /* entries per node, 2..256 (reference: 192 MPHR, 128 RTREE/BVH/GUTTMAN) */
  uint32_t layout;    /* QCP_LAYOUT_*                                               */
  uint32_t key_mode;  /* QCP_KEY_*                                                  */
  uint32_t key_src;   /* QCP_SRC_*                                                  */
  uint32_t obj_kind;  /* QCP_OBJ_*: points are N x D floats, boxes N x 2D [lo, hi] */
  uint32_t device;    /* HIP device ordinal                                         */
  uint32_t flags;     /* QCP_BUILD_GRAPH: run the build as a captured hipGraph      */
  uint32_t order;     /* QCP_ORDER_*: Hilbert order of the objects' keys, or k-means
                         clusters ranked by their centres' keys, then Hilbert order
                         within a cluster (KmeansMapper, src/mapper/kmeans_mapper.cpp:
                         201-259; not with QCP_BUILD_GRAPH)                        */
  uint32_t clusters;  /* k-means order: clusters K (0 = 20000, the reference's
                         DFN_NumberOfClusters; at most QCP_MAX_CLUSTERS; K > n
                         means n)                                                  */
  uint32_t kmeans_iters; /* k-means order: Lloyd iterations at most (0 = 1001: the
                            reference's loop runs while the changed fraction
                            (float)changes / (float)n exceeds 1e-5 and loop++ < 1000,
                            src/mapper/kmeans_mapper.cpp:170)                      */
  uint32_t partitions;   /* MPHR data partitions P (0 or 1 = one tree; at most 65536 and
                            at most n): the ranks split as MPHR::Build does (chunk = n/P,
                            the first partition also taking n % P), one packed subtree
                            per partition (src/tree/mphr.cpp:60-108, SetNumberOfPartition
                            279-285); searched from every partition root               */
} qcp_build_cfg;

typedef struct {
  uint64_t n;                 /* objects                                  */
  uint64_t n_nodes;           /* tree nodes (root included)               */
  uint32_t height;            /* node levels (root .. leaves)             */
  uint32_t dims, fanout, layout, obj_kind, key_mode, key_src, key_bits;
  uint64_t level_count[32];   /* nodes per level, root first              */
  uint64_t tree_bytes;        /* device bytes of node boxes + objects     */
  double build_ms;            /* device time of the last build (HIP events) */
  double build_phase_ms[6];   /* keys, sort, gather, pack, topology (GUTTMAN: host
                                 insertion, wall ms), total */
  uint32_t partitions;        /* MPHR data partitions (1 = one tree); partition p's root
                                 is node p (nodes are numbered depth-major)  */
  uint32_t kmeans_iterations; /* Lloyd iterations of a k-means build (0 otherwise, and
                                 for an index loaded from a file)            */
  double kmeans_ms[2];        /* k-means build: assignment kernels and centre-update
                                 kernels (HIP events, each summed over the
                                 iterations) */
  uint32_t sort_form;         /* the last build's key sort: 0 = (key, input index) and a
                                 gather, 1 = objects carried through the passes */
  uint32_t has_perm;          /* 1: qcp_export returns the input index of each rank */
} qcp_index_info;

typedef struct {
  uint32_t mode;           /* QCP_SEARCH_*                                          */
  uint32_t outputs;        /* QCP_OUT_* mask                                        */
  uint32_t plain_batch;    /* plain form: queries per launch (reference default 128) */
  uint32_t block_waves;    /* persistent form: waves per workgroup, 1..8 (0 = 8)    */
  uint32_t grid_blocks;    /* persistent form: workgroups (0 = fill the device)     */
  uint32_t leaf_chunks;    /* persistent form: 64-entry leaf chunks in flight per wave, 2/4/8 (0 = 8) */
  uint32_t donate;         /* persistent form: subtree donation (work stealing) on/off */
  uint32_t order;          /* persistent form: 0 = process queries in Hilbert order of their
                              centres, split into one range per XCD (batches of 8192 queries
                              or more; smaller ones as given); 1 = as given */
  uint32_t max_waiters;    /* donation: waves kept waiting for donated work (0 = 256)  */
  uint32_t hunger_period;  /* donation: internal-node descents between queue checks (0 = 8) */
  uint32_t donate_height;  /* donation (packed layouts): minimum height of a donated subtree,
                              leaves having height 0 and leaf parents 1 (0 = 1) */
  uint32_t graph;          /* persistent form, counts/hash outputs: capture the whole launch
                              sequence (query ordering, buffer fills, search) in a hipGraph
                              on first use and replay it while the configuration is unchanged */
  uint32_t contain;        /* persistent form: report every child subtree whose node box lies
                              inside the query box as hits without visiting it (count from
                              its rank range, hash in closed form, IDs written as a run).
                              Same results; the stats then count the pruned traversal.  Not
                              applied to box indexes holding a box with lo > hi.  (The count
                              pass of an ID search, which only sizes the output, always uses
                              it; an ID search's stats describe its ID pass.) */
  uint32_t cpu_threads;    /* hybrid form: host threads walking the upper tree (0 = all cores) */
  uint32_t handoff_level;  /* hybrid form: tree level whose overlapping nodes the host hands
                              to the GPU (1 = the root's children; 0 or beyond = the leaves,
                              the reference's CPU-internal / GPU-leaf split)              */
  uint32_t hybrid_batch;   /* hybrid form: queries per pipeline block (0 = 16384)          */
  uint32_t plain_waves;    /* plain form: waves per workgroup, 1 (0) or 3 (the reference's
                              192-thread block, src/tree/mphr.cpp:241-243).  A partitioned
                              MPHR runs the reference's partition mode: one launch per
                              query with one workgroup per partition (235-257)           */
} qcp_search_cfg;

typedef struct {
  double kernel_ms;          /* device time of the search kernels (HIP events)  */
  double total_ms;           /* device time of the whole search, setup included */
  uint64_t total_hits;
  uint64_t internal_entries; /* entries of visited internal nodes (canonical traversal) */
  uint64_t leaf_entries;     /* entries of visited leaf nodes (canonical traversal)     */
  uint64_t donated_items;    /* subtrees handed to other waves                          */
  uint64_t launches;         /* kernel launches issued by the search                    */
  uint64_t contained_runs;   /* contain: runs of sibling subtrees reported without a visit */
  double search_ms;          /* persistent form: device time of the (first) search kernel
                                alone, without query ordering and buffer fills (whole
                                graph when graph-captured)                              */
} qcp_search_stats;

typedef struct qcp_index qcp_index;
typedef struct qcp_batch qcp_batch;

/* Status text.  Replaces the reference's cudaErrCheck/exit() reporting
 * (src/common/macro.h:100-108). */
const char *qcp_strerror(qcp_status s);

/* Number of visible HIP devices (0 when none); never initialises a context. */
int qcp_device_count(void);

/* Build an index on the GPU: Hilbert keys, stable radix sort by (key, input
 * index), rank-ordered SoA objects, bottom-up (MPHR/RTREE) or rank-prefix
 * (BVH) node boxes, all resident in HBM (GUTTMAN: the host inserts the sorted
 * objects and the tree is then uploaded in depth-first order).  Replaces tree::Tree::Build
 * (src/tree/tree.h:24) as implemented by MPHR::Build (src/tree/mphr.cpp:35-136),
 * BVH::Build (src/tree/bvh.cpp:25-63) and RTree::Build (src/tree/rtree.cpp:25-63),
 * together with the upload of manager::ChunkManager::Init/CopyNode
 * (src/manager/chunk_manager.cpp:16-40).  objs is host memory, row-major. */
qcp_status qcp_build(const qcp_build_cfg *cfg, const float *objs, uint64_t n,
                     qcp_index **out);

/* Same, with objs already in device memory of cfg->device (no PCIe copy). */
qcp_status qcp_build_device(const qcp_build_cfg *cfg, const float *d_objs, uint64_t n,
                            qcp_index **out);

/* Rebuild an existing index in place from new device-resident objects of the
 * same count and configuration (timed builds). */
qcp_status qcp_rebuild_device(qcp_index *ix, const float *d_objs);

/* Device buffers for inputs that should be resident before a timed build or
 * search (the reference's cudaMalloc + cudaMemcpy of DataSet::GetDeviceQuery,
 * src/io/dataset.cpp:82-88, as a caller-owned buffer). */
qcp_status qcp_device_alloc(uint32_t device, uint64_t bytes, void **out);
qcp_status qcp_device_free(void *p);
qcp_status qcp_copy_to_device(void *dst, const void *src, uint64_t bytes);
qcp_status qcp_copy_to_host(void *dst, const void *src, uint64_t bytes);

void qcp_index_destroy(qcp_index *ix);

/* Shape and timing of a built index.  Replaces the reference's build-time
 * logging of level counts and phase times (src/tree/tree.cpp:426-428,492). */
qcp_status qcp_index_info_get(const qcp_index *ix, qcp_index_info *out);

/* Copy the built tree back to the host for inspection and parity checks:
 *   sorted_keys[n]   Hilbert keys in rank order          (may be NULL)
 *   perm[n]          input index of each rank            (may be NULL; must be NULL,
 *                    or QCP_EINVAL, when qcp_index_info.has_perm is 0)
 *   sobjs[n*E]       objects in rank order, row-major     (may be NULL)
 *   node_boxes[n_nodes*2D] node MBRs, BFS order, [lo(D),hi(D)] (may be NULL)
 *   node_first/node_count[n_nodes], node_leaf[n_nodes]   (may be NULL)
 * Replaces the reference's tree dump (MPHR::DumpToFile, src/tree/mphr.cpp:169-194)
 * as an in-memory export. */
qcp_status qcp_export(const qcp_index *ix, uint64_t *sorted_keys, uint32_t *perm,
                      float *sobjs, float *node_boxes, uint32_t *node_first,
                      uint32_t *node_count, uint8_t *node_leaf);

/* Save / load the device-resident index (versioned header + checksum).
 * Replaces MPHR::DumpToFile / DumpFromFile (src/tree/mphr.cpp:138-194) and
 * BVH::DumpToFile / DumpFromFile (src/tree/bvh.cpp:65-163). */
qcp_status qcp_save(const qcp_index *ix, const char *path);
qcp_status qcp_load(const char *path, uint32_t device, qcp_index **out);

/* Build from objects already in rank order (entry i gets rank ID i+1): no keys
 * are computed and nothing is sorted; node boxes as in qcp_build.  For data
 * whose order was fixed elsewhere, such as the ranks of a reference dump. */
qcp_status qcp_build_ranked(const qcp_build_cfg *cfg, const float *objs, uint64_t n,
                            qcp_index **out);

/* Read a file written by the reference's MPHR::DumpToFile (src/tree/mphr.cpp:
 * 167-194) and recover its objects in the reference's rank order (its leaf
 * entries' indexes), as points (lo corner, D floats each) or boxes (2D).
 * dims and leaf_degree are the reference build's GetNumberOfDims() and
 * GetNumberOfLeafNodeDegrees() (3 and 192 as shipped).  objs == NULL returns
 * the object count only; QCP_EFORMAT if the size, a node's count or the ranks
 * do not form a valid dump.  Host only; follow with qcp_build_ranked. */
qcp_status qcp_read_reference_mphr(const char *path, uint32_t dims, uint32_t leaf_degree,
                                   uint32_t obj_kind, float *objs, uint64_t cap,
                                   uint64_t *n_out);

/* Upload a query set to HBM.  Replaces io::DataSet::GetDeviceQuery
 * (src/io/dataset.cpp:82-88); queries are Q x 2D floats [lo(D), hi(D)]. */
qcp_status qcp_batch_create(qcp_index *ix, const float *boxes, uint64_t q,
                            qcp_batch **out);
/* Same, from device memory on the index's device. */
qcp_status qcp_batch_create_device(qcp_index *ix, const float *d_boxes, uint64_t q,
                                   qcp_batch **out);
void qcp_batch_destroy(qcp_batch *b);

/* Run the range search of a resident query batch.  Replaces tree::Tree::Search
 * (src/tree/tree.h:47-48) as implemented by MPHR::Search (src/tree/mphr.cpp:196-272,
 * cfg->mode = QCP_SEARCH_PLAIN: one workgroup per query, plain_batch queries per
 * launch, a synchronous count copy after each launch) and, new here, the
 * persistent work-stealing form (QCP_SEARCH_PERSISTENT: all queries in one
 * launch).  QCP_SEARCH_HYBRID is the CPU-GPU co-processing search of
 * Hybrid::Search (src/tree/hybrid.cpp:451-969): host threads walk the upper
 * levels and hand the overlapping nodes of cfg->handoff_level to the GPU, in
 * double-buffered blocks of queries (counts and hash only).  With QCP_OUT_IDS
 * the sorted rank-ID lists are produced by a count pass, a device prefix sum
 * and an ordered write pass.  cfg may be NULL (persistent, counts+hash, tuned). */
qcp_status qcp_batch_run(qcp_batch *b, const qcp_search_cfg *cfg, qcp_search_stats *stats);

/* The sorted rank-ID lists of queries [q0, q1) of the last run with
 * QCP_OUT_IDS: offsets[q1-q0+1] relative to the first (may be NULL), ids
 * (ids_cap entries; QCP_EBUSY if too small).  For result sets too large to copy
 * whole (pagination of the reference's per-query results). */
qcp_status qcp_batch_ids(const qcp_batch *b, uint64_t q0, uint64_t q1, uint64_t *offsets,
                         uint32_t *ids, uint64_t ids_cap);

/* The hipGraph-captured build+query pipeline: one graph holding the rebuild of
 * ix from device objects d_objs (same count and configuration as its build)
 * followed by the persistent search of batch b (counts + hash), captured on
 * the first call for a given d_objs and search configuration and replayed
 * after.  New (the reference builds and searches in separate host-driven
 * steps, src/evaluator/evaluator.cpp:121-280).  QCP_EDOMAIN after the run if
 * the new objects are outside the quantiser's domain (results then invalid). */
qcp_status qcp_pipeline_run(qcp_index *ix, const float *d_objs, qcp_batch *b,
                            const qcp_search_cfg *cfg, qcp_search_stats *stats);

/* Results of the last run (host copies).  counts[q], hashes[q],
 * offsets[q+1] (exclusive prefix of counts), ids[offsets[Q]], visits[2q]
 * (root visits, node visits: plain form only).  Each pointer may be NULL. */
qcp_status qcp_batch_results(const qcp_batch *b, uint32_t *counts, uint64_t *hashes,
                             uint64_t *offsets, uint32_t *ids, uint64_t ids_cap,
                             uint32_t *visits);
uint64_t qcp_batch_total_hits(const qcp_batch *b);

/* One-shot search of host queries (upload, run, copy back).  Same semantics
 * as qcp_batch_run; counts/hashes/offsets/ids as in qcp_batch_results. */
qcp_status qcp_search(qcp_index *ix, const float *boxes, uint64_t q,
                      const qcp_search_cfg *cfg, uint32_t *counts, uint64_t *hashes,
                      uint64_t *offsets, uint32_t *ids, uint64_t ids_cap,
                      qcp_search_stats *stats);

/* ---- autotuner ---------------------------------------------------------------------
 * Replaces the reference's evaluation-mode parameter sweeps (-e,
 * src/evaluator/evaluator.cpp:171-280): builds the index at every candidate
 * fanout (cfg->fanout is ignored) from device-resident objects, and times the
 * persistent search of q device-resident sample queries for every workgroup
 * size {1, 2, 4, 8 waves}, leaf-chunk count {2, 4, 8} (packed layouts) and
 * donation {off, on}.  Returns the index built at the best fanout (owned by the
 * caller) and the best search configuration; table (optional) receives every
 * measurement. */
typedef struct {
  uint32_t fanout, block_waves, leaf_chunks, donate;
  double kernel_ms;          /* best of two timed runs of the sample */
} qcp_tune_entry;

qcp_status qcp_autotune(const qcp_build_cfg *cfg, const float *d_objs, uint64_t n,
                        const float *d_queries, uint64_t q, const uint32_t *fanouts,
                        uint32_t n_fanouts, uint32_t outputs, qcp_index **best_index,
                        qcp_search_cfg *best, qcp_tune_entry *table, uint32_t table_cap,
                        uint32_t *table_len);

/* Co-processing policy tuner.  The reference hand-tunes its CPU-GPU split:
 * the scan level (-l), the pipeline chunk (-c) and its chunk heuristic
 * (src/tree/hybrid.cpp:833-876, src/evaluator/evaluator.cpp:174-178).  This
 * times the hybrid form (QCP_SEARCH_HYBRID) of q device-resident sample
 * queries over ix for every hand-off level and every pipeline block in
 * {4096, 16384, 65536} (blocks larger than the sample are skipped).  The
 * levels are 1 .. height-2 and the leaves (0) for packed layouts, and depths
 * 1 .. min(height-2, 6) and the leaves for explicit ones.  The other fields
 * come from base (null = defaults: all host cores, counts + hash).  best
 * receives base with the (handoff_level, hybrid_batch) of the lowest wall
 * time, host traversal included; table (optional) receives every
 * measurement. */
typedef struct {
  uint32_t handoff_level;  /* 0 = the leaves (the reference's split)            */
  uint32_t hybrid_batch;   /* queries per pipeline block                        */
  double total_ms;         /* best of two timed runs after a warm-up, wall      */
  double device_ms;        /* device time of that run                           */
  uint64_t launches;       /* search launches of that run                       */
} qcp_hybrid_tune_entry;

qcp_status qcp_autotune_hybrid(qcp_index *ix, const float *d_queries, uint64_t q,
                               const qcp_search_cfg *base, qcp_search_cfg *best,
                               qcp_hybrid_tune_entry *table, uint32_t table_cap,
                               uint32_t *table_len);

/* ---- serving: continuous batching over a resident index -----------------------
 * New (the reference has no serving path; its closest interface is the
 * per-batch MPHR::Search loop, src/tree/mphr.cpp:231-258).  A server keeps the
 * index resident and runs a dispatcher thread: requests are queued as they are
 * submitted, and whenever one of `slots` in-flight batch slots is free the
 * waiting requests (up to max_batch) are taken as one batch on that slot's HIP
 * stream, so batches that arrive while others are on the GPU overlap on the
 * device.  Admission is charged the batch's result bytes against
 * result_budget (a batch alone is always admitted):
 *   - count + ID-set hash service: 12 bytes per query, checked before launch;
 *   - ID-list service (outputs has QCP_OUT_IDS): the batch's count pass runs
 *     first and sizes its sorted rank-ID lists; the ID pass is admitted when
 *     4 bytes per ID + 12 per query fit beside the batches in flight.
 * Results are kept in a table of max_requests entries until they are read:
 * qcp_server_result / _wait read (and free) a count + hash result;
 * qcp_server_result_ids reads (and frees) an ID-list result; qcp_server_release
 * frees one unread.  Submissions are refused (QCP_EBUSY) only while the table
 * is full of unread results. */
typedef struct qcp_server qcp_server;
typedef struct {
  uint32_t slots;              /* in-flight batches (HIP streams), default 4          */
  uint32_t max_batch;          /* queries per batch, default 4096                     */
  uint64_t result_budget;      /* bytes of result buffers in flight, default 64 MiB   */
  uint32_t max_requests;       /* result-table entries (unread tickets), default 1 << 20 */
  uint32_t grid_blocks;        /* per-batch grid cap (0 = fill the device)            */
  uint32_t contain;            /* as qcp_search_cfg.contain (same results)            */
  uint32_t outputs;            /* QCP_OUT_* mask: counts + hash (0 means that), or with
                                  QCP_OUT_IDS also the sorted rank-ID lists          */
} qcp_serve_cfg;
typedef struct {
  uint64_t submitted, completed, batches;
  uint64_t deferred_by_budget; /* admissions postponed because the budget was full   */
  uint64_t max_batch_seen, max_inflight_seen;
  double p50_us, p99_us, mean_us, max_us;  /* submit -> result latency (p50/p99 from a
                                              histogram of 1/128-octave bins)        */
  uint64_t bytes_peak;         /* most result bytes admitted at once                 */
  uint64_t ids_served;         /* rank IDs delivered (ID-list service)               */
} qcp_serve_stats;

qcp_status qcp_server_start(qcp_index *ix, const qcp_serve_cfg *cfg, qcp_server **out);
/* Thread-safe, non-blocking: queue one query box; *ticket identifies it.
 * QCP_EBUSY while the result table is full of unread results. */
qcp_status qcp_server_submit(qcp_server *sv, const float *box, uint64_t *ticket);
/* QCP_OK with the result once done (the entry is freed, except with ID-list
 * service, where the IDs are still to be read), QCP_EBUSY while pending,
 * QCP_EINVAL for a ticket not in the table (never issued or already read). */
qcp_status qcp_server_result(qcp_server *sv, uint64_t ticket, uint32_t *count, uint64_t *hash,
                             double *latency_us);
/* Block until the ticket is done, then as qcp_server_result. */
qcp_status qcp_server_wait(qcp_server *sv, uint64_t ticket, uint32_t *count, uint64_t *hash,
                           double *latency_us);
/* ID-list service: the sorted rank-ID list of a finished request.  ids == NULL:
 * only *n_ids; otherwise the list is copied (QCP_EBUSY if cap is too small)
 * and the entry freed. */
qcp_status qcp_server_result_ids(qcp_server *sv, uint64_t ticket, uint32_t *ids, uint64_t cap,
                                 uint64_t *n_ids);
/* Free a finished request's entry without reading it. */
qcp_status qcp_server_release(qcp_server *sv, uint64_t ticket);
qcp_status qcp_server_stats(qcp_server *sv, qcp_serve_stats *out);
/* Drain outstanding work and stop the dispatcher. */
qcp_status qcp_server_stop(qcp_server *sv);
/* Load driver: submits query i at start + arrival_s[i] (arrival_s ascending,
 * seconds) from a native thread while another reads the results out in ticket
 * order, and returns per-query results and submit-to-result latencies. */
qcp_status qcp_server_replay(qcp_server *sv, const float *boxes, const double *arrival_s,
                             uint64_t n, uint32_t *counts, uint64_t *hashes, double *latency_us);
/* The same for ID-list service: also offsets[n+1] (prefix of the counts) and
 * the lists into ids (ids_cap entries; ids == NULL reads them out without
 * copying; QCP_EBUSY at the end if they did not fit). */
qcp_status qcp_server_replay_ids(qcp_server *sv, const float *boxes, const double *arrival_s,
                                 uint64_t n, uint32_t *counts, uint64_t *hashes,
                                 double *latency_us, uint64_t *offsets, uint32_t *ids,
                                 uint64_t ids_cap);

/* ---- search without allocation --------------------------------------------------
 * No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The persistent search of qcp_batch_run
 * (cfg->mode QCP_SEARCH_PERSISTENT; counts, or counts + ID-set hash with
 * QCP_OUT_HASH; no ID lists) of q query boxes already in device memory of the
 * index's device, into caller-owned device arrays d_counts[q] and, with the
 * hash, d_hashes[q]; same configuration rules and results as qcp_batch_run.
 * Every other buffer (queue control block, donation slots, the query-order
 * sort) is carved from one fixed array of QCP_NOALLOC_SCRATCH_BYTES per GPU
 * inside the library, held until the process ends; nothing is allocated.
 * q above QCP_NOALLOC_MAX_QUERIES is refused (QCP_EINVAL).  Returns when the
 * results are in place.  Calls that overlap on one GPU share that array and
 * race: serialise them. */
#define QCP_NOALLOC_MAX_QUERIES 262144u
#define QCP_NOALLOC_SCRATCH_BYTES (7u << 20)
qcp_status qcp_search_noalloc(qcp_index *ix, const float *d_queries, uint64_t q,
                              const qcp_search_cfg *cfg, uint32_t *d_counts,
                              uint64_t *d_hashes);

/* ---- measurement support ----------------------------------------------------------
 * Read bandwidth (GB/s) of the memory level that holds a window of
 * window_bytes: every wave streams 64-float chunks with dword loads (the leaf
 * scan's load form), 8 chunks in flight, sweeping the window for `rounds`
 * rounds of 2 KiB per wave (8 workgroups of 4 waves per CU).  A window of a
 * few MiB measures the XCDs' L2, a few hundred MiB the Infinity Cache, several
 * GiB HBM.  No counterpart in the reference (roofline peaks for the search
 * figures); *ms receives the timed kernel's duration. */
qcp_status qcp_probe_read_gbps(uint32_t device, uint64_t window_bytes, uint32_t rounds,
                               double *gbps, double *ms);

/* ---- seeded generator (replaces generator/cuRAND.cpp:17-75, without cuRAND) ----
 * Philox4x32-10 (Salmon et al., SC'11).  Word i of (seed, stream) is output
 * word i%4 of philox(counter = {i/4 low, i/4 high, seed high, 0},
 * key = {seed low, stream}); it is order independent and parallel. */
void qcp_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void qcp_gen_words(uint32_t *out, uint64_t n, uint64_t seed, uint32_t stream,
                   uint64_t offset);
/* u = (w >> 8) * 2^-24 in [0, 1). */
void qcp_gen_uniform(float *out, uint64_t n, uint64_t seed, uint32_t stream,
                     uint64_t offset);
/* Standard normals, Box-Muller in double from word pairs (2i, 2i+1):
 * u1 = ((w0 >> 8) + 1) * 2^-24, u2 = (w1 >> 8) * 2^-24,
 * z[2i] = sqrt(-2 ln u1) cos(2 pi u2), z[2i+1] = sqrt(-2 ln u1) sin(2 pi u2). */
void qcp_gen_normal(double *out, uint64_t n, uint64_t seed, uint32_t stream,
                    uint64_t offset);

#ifdef __cplusplus
}
#endif
#endif /* QCP_H */
