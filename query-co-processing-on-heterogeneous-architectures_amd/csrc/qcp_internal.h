// qcp_internal.h -- host-side definitions of the index and batch handles,
// shared by the C ABI implementation files.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/qcp.h"
#include "qcp_kernels.h"

struct qcp_index {
  int device = 0;
  hipStream_t stream = nullptr;
  qcp_build_cfg cfg{};
  uint32_t D = 0, F = 0, E = 0, nb = 0;
  bool box = false, expl = false;
  uint64_t n = 0;
  int h = 0;
  std::vector<uint64_t> level_count;  // root first
  std::vector<uint64_t> level_off;    // h+1 entries
  uint64_t n_nodes = 0;
  uint64_t obj_stride = 0, node_stride = 0;
  // device
  float* d_obj = nullptr;        // [E][obj_stride]
  float* d_nbox = nullptr;       // [2D][node_stride]
  uint32_t* d_nfirst = nullptr;  // explicit layouts
  uint32_t* d_ncount = nullptr;
  uint32_t* d_nrlo = nullptr;    // explicit layouts: rank range under each node
  uint32_t* d_nrhi = nullptr;
  uint32_t* d_nparent = nullptr; // explicit layouts: parent node (roots: themselves)
  uint32_t parts = 1;            // MPHR data partitions (roots are nodes 0 .. parts-1)
  uint32_t kmeans_its = 0;       // Lloyd iterations of the last k-means build
  double kmeans_ms[2] = {0, 0};  // its assignment and centre-update device time
  uint64_t* d_keys = nullptr;    // sorted keys [n]
  uint32_t* d_perm = nullptr;    // [n] input index of each rank (when perm_valid)
  bool perm_valid = true;
  bool carry = false;            // the last build carried the objects through the sort
  // object-carrying sort: payload words of its two scratch sets
  uint32_t* d_carry = nullptr;
  uint64_t carry_words = 0;
  // build scratch
  uint64_t* d_keys2 = nullptr;
  uint32_t* d_perm2 = nullptr;
  uint32_t* d_sort_ghist = nullptr;   // digit counts of every pass + tile counters
  uint64_t* d_sort_status = nullptr;  // look-back words of two consecutive passes
  unsigned long long* d_key_or = nullptr;
  uint32_t* d_bad = nullptr;
  float* d_in = nullptr;         // device copy of host input
  hipEvent_t ev[8] = {};
  hipGraphExec_t build_exec = nullptr;      // captured build (QCP_BUILD_GRAPH)
  const float* build_exec_objs = nullptr;   // input pointer baked into build_exec
  int cus = 0;
  // every object box has lo <= hi in each dimension, so a node box inside a
  // query implies that every object under it overlaps the query
  bool contain_ok = true;
  // d_keys does not hold the sorted keys (a packed sort leaves them out);
  // qcp_export recomputes them from the rank-ordered objects
  mutable bool keys_stale = false;
  // host copies for the co-processing search: explicit topology (kept from the
  // build) and the node boxes [node][lo(D), hi(D)] (mirrored on first use,
  // cleared by every build)
  std::vector<uint32_t> h_first, h_count;
  std::vector<float> h_nbox;
  qcp_index_info info{};
};

struct qcp_batch {
  qcp_index* ix = nullptr;
  uint64_t q = 0;
  float* d_queries = nullptr;
  uint32_t* d_counts = nullptr;
  unsigned long long* d_hashes = nullptr;
  unsigned long long* d_offsets = nullptr;
  unsigned long long* d_scan_sums = nullptr;
  // spatial processing order of the queries (Hilbert key of each centre)
  uint32_t* d_qperm = nullptr;
  uint32_t* d_qperm2 = nullptr;
  uint64_t* d_qkeys = nullptr;
  uint64_t* d_qkeys2 = nullptr;
  uint32_t* d_qghist = nullptr;   // one-sweep sort: digit counts + tile counters
  uint64_t* d_qstatus = nullptr;  // one-sweep sort: look-back words
  unsigned long long* d_qor = nullptr;
  uint32_t* d_qbad = nullptr;
  uint32_t* d_ids = nullptr;
  uint64_t ids_cap = 0;
  uint32_t* d_visits = nullptr;
  std::vector<float> h_queries;   // host copy of the queries (co-processing search)
  qcp::QueueCtl* d_ctl = nullptr;
  unsigned long long* d_slots = nullptr;
  uint32_t slot_cap = 0;
  uint64_t slots_dirty = 0;       // slots written by the last donating run
  uint32_t* h_counts = nullptr;   // pinned, plain-form per-batch copies
  uint32_t* h_visits = nullptr;
  // plain form over a partitioned index: one workgroup's results per partition
  uint32_t* d_pcounts = nullptr;
  unsigned long long* d_phashes = nullptr;
  uint32_t* d_pvisits = nullptr;
  uint32_t* h_pcounts = nullptr;  // pinned
  unsigned long long* h_phashes = nullptr;
  uint32_t* h_pvisits = nullptr;
  hipEvent_t ev[6] = {};
  hipGraphExec_t exec = nullptr;            // captured search (qcp_search_cfg.graph)
  uint32_t exec_key[12] = {};               // configuration baked into exec
  const float* exec_objs = nullptr;         // rebuild input baked into exec (pipeline)
  const float* pipe_objs = nullptr;         // set while qcp_pipeline_run captures / runs
  uint32_t last_outputs = 0;
  uint32_t last_mode = 0;
  uint64_t total_hits = 0;
  bool has_offsets = false;
};

namespace qcp {
// Kernel arguments of a search over `ix` for `nq` queries at d_queries with
// the given output buffers (no queue / donation state).
SearchArgs index_search_args(const qcp_index* ix, const float* d_queries, uint64_t nq,
                             uint32_t* d_counts, unsigned long long* d_hashes);
// Workgroups that fill the device with the persistent kernel of `out` mode.
unsigned persistent_grid(const qcp_index* ix, int out, unsigned block);
// k-means + Hilbert order (QCP_ORDER_KMEANS_HILBERT), in kmeans.cpp: replaces
// the object keys in ix->d_keys by composite keys; *key_bits = significant bits;
// ix->kmeans_its = iterations run
qcp_status kmeans_keys(qcp_index* ix, const float* d_objs, uint64_t* key_bits);
uint32_t kmeans_seed_index(uint64_t word, uint64_t n);
// the batch's query order on the host: the Hilbert order of the query centres
// (spatial, batches of 8192 or more) or the given order
qcp_status batch_order_host(qcp_batch* b, bool spatial, std::vector<uint32_t>& ord);
// the co-processing search (QCP_SEARCH_HYBRID), in hybrid.cpp
qcp_status hybrid_run(qcp_batch* b, const qcp_search_cfg& cfg, qcp_search_stats* stats);
}  // namespace qcp
