// autotune.cpp -- search-configuration autotuner (see qcp_autotune in qcp.h).
//
// The reference sweeps CPU thread counts, chunk sizes and block counts by hand
// in its evaluation mode (src/evaluator/evaluator.cpp:171-280) and always
// searches with compile-time fanouts.  Here the fanout, the workgroup size, the
// number of 64-entry leaf chunks in flight per wave and subtree donation are
// chosen by timing the persistent search of a query sample on the GPU itself.
#include <string.h>

#include <new>
#include <vector>

#include "../../include/qcp.h"

namespace {

struct Candidate {
  qcp_search_cfg cfg;
  double ms;
};

// Best of two timed runs after one warm-up run.
qcp_status time_cfg(qcp_batch* b, const qcp_search_cfg& sc, double* ms) {
  qcp_search_stats st;
  qcp_status s = qcp_batch_run(b, &sc, &st);
  if (s != QCP_OK) return s;
  double best = 1e30;
  for (int r = 0; r < 2; ++r) {
    s = qcp_batch_run(b, &sc, &st);
    if (s != QCP_OK) return s;
    if (st.kernel_ms < best) best = st.kernel_ms;
  }
  *ms = best;
  return QCP_OK;
}

}  // namespace

extern "C" qcp_status qcp_autotune(const qcp_build_cfg* cfg, const float* d_objs, uint64_t n,
                                   const float* d_queries, uint64_t q, const uint32_t* fanouts,
                                   uint32_t n_fanouts, uint32_t outputs, qcp_index** best_index,
                                   qcp_search_cfg* best, qcp_tune_entry* table, uint32_t table_cap,
                                   uint32_t* table_len) try {
  if (!cfg || !best_index || !best || !fanouts || n_fanouts == 0 || (q && !d_queries) ||
      (n && !d_objs))
    return QCP_EINVAL;
  if (outputs & QCP_OUT_IDS) return QCP_EINVAL;  // tuned on the counting pass
  *best_index = nullptr;
  if (table_len) *table_len = 0;
  const bool packed = cfg->layout < QCP_LAYOUT_BVH;
  const uint32_t waves[] = {1, 2, 4, 8};
  const uint32_t chunks_packed[] = {2, 4, 8};
  const uint32_t chunks_expl[] = {4};
  double best_ms = 1e30;
  qcp_index* best_ix = nullptr;
  for (uint32_t f = 0; f < n_fanouts; ++f) {
    qcp_build_cfg c = *cfg;
    c.fanout = fanouts[f];
    qcp_index* ix = nullptr;
    qcp_status s = qcp_build_device(&c, d_objs, n, &ix);
    if (s != QCP_OK) {
      qcp_index_destroy(best_ix);
      return s;
    }
    qcp_batch* b = nullptr;
    s = qcp_batch_create_device(ix, d_queries, q, &b);
    if (s != QCP_OK) {
      qcp_index_destroy(ix);
      qcp_index_destroy(best_ix);
      return s;
    }
    bool ix_best = false;
    const uint32_t* chunks = packed ? chunks_packed : chunks_expl;
    const uint32_t nch = packed ? 3 : 1;
    for (uint32_t wv : waves) {
      for (uint32_t ci = 0; ci < nch; ++ci) {
        for (uint32_t don = 0; don < 2; ++don) {
          qcp_search_cfg sc;
          memset(&sc, 0, sizeof(sc));
          sc.mode = QCP_SEARCH_PERSISTENT;
          sc.outputs = outputs;
          sc.block_waves = wv;
          sc.leaf_chunks = chunks[ci];
          sc.donate = don;
          double ms = 0;
          s = time_cfg(b, sc, &ms);
          if (s != QCP_OK) {
            qcp_batch_destroy(b);
            qcp_index_destroy(ix);
            qcp_index_destroy(best_ix);
            return s;
          }
          if (table && table_len && *table_len < table_cap) {
            qcp_tune_entry& e = table[(*table_len)++];
            e.fanout = c.fanout;
            e.block_waves = wv;
            e.leaf_chunks = chunks[ci];
            e.donate = don;
            e.kernel_ms = ms;
          }
          if (ms < best_ms) {
            best_ms = ms;
            *best = sc;
            ix_best = true;
          }
        }
      }
    }
    qcp_batch_destroy(b);
    if (ix_best) {
      qcp_index_destroy(best_ix);
      best_ix = ix;
    } else {
      qcp_index_destroy(ix);
    }
  }
  *best_index = best_ix;
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}
