// autotune.cpp -- search-configuration autotuners (qcp_autotune and
// qcp_autotune_hybrid in qcp.h).
//
// The reference sweeps CPU thread counts, chunk sizes and block counts by hand
// in its evaluation mode (src/evaluator/evaluator.cpp:171-280) and always
// searches with compile-time fanouts.  Here the fanout, the workgroup size, the
// number of 64-entry leaf chunks in flight per wave and subtree donation are
// chosen by timing the persistent search of a query sample on the GPU itself;
// the co-processing form's split (hand-off level, pipeline block) by timing
// that form end to end, host traversal included.
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

// Best of two timed hybrid runs after one warm-up run, by wall time.
qcp_status time_hybrid(qcp_batch* b, const qcp_search_cfg& sc, qcp_search_stats* out) {
  qcp_search_stats st;
  qcp_status s = qcp_batch_run(b, &sc, &st);
  if (s != QCP_OK) return s;
  out->total_ms = 1e30;
  for (int r = 0; r < 2; ++r) {
    s = qcp_batch_run(b, &sc, &st);
    if (s != QCP_OK) return s;
    if (st.total_ms < out->total_ms) *out = st;
  }
  return QCP_OK;
}

}  // namespace

extern "C" qcp_status qcp_autotune_hybrid(qcp_index* ix, const float* d_queries, uint64_t q,
                                          const qcp_search_cfg* base, qcp_search_cfg* best,
                                          qcp_hybrid_tune_entry* table, uint32_t table_cap,
                                          uint32_t* table_len) try {
  if (!ix || !best || q == 0 || !d_queries) return QCP_EINVAL;
  if (table_len) *table_len = 0;
  qcp_search_cfg sc;
  memset(&sc, 0, sizeof(sc));
  sc.outputs = QCP_OUT_COUNTS | QCP_OUT_HASH;
  if (base) sc = *base;
  if (sc.outputs & QCP_OUT_IDS) return QCP_EINVAL;  // the hybrid form returns counts + hash
  sc.mode = QCP_SEARCH_HYBRID;
  qcp_index_info inf;
  qcp_status s = qcp_index_info_get(ix, &inf);
  if (s != QCP_OK) return s;
  const bool packed = inf.layout < QCP_LAYOUT_BVH;
  std::vector<uint32_t> levels;
  const uint32_t top = inf.height >= 2 ? inf.height - 2 : 0;
  for (uint32_t l = 1; l <= top && (packed || l <= 6); ++l) levels.push_back(l);
  levels.push_back(0);  // the leaves
  const uint32_t blocks[] = {4096, 16384, 65536};
  qcp_batch* b = nullptr;
  s = qcp_batch_create_device(ix, d_queries, q, &b);
  if (s != QCP_OK) return s;
  double best_ms = 1e30;
  for (uint32_t l : levels) {
    for (uint32_t bi = 0; bi < 3; ++bi) {
      if (bi > 0 && blocks[bi - 1] >= q) break;  // one block already holds the sample
      qcp_search_cfg c = sc;
      c.handoff_level = l;
      c.hybrid_batch = blocks[bi];
      qcp_search_stats st;
      s = time_hybrid(b, c, &st);
      if (s != QCP_OK) {
        qcp_batch_destroy(b);
        return s;
      }
      if (table && table_len && *table_len < table_cap) {
        qcp_hybrid_tune_entry& e = table[(*table_len)++];
        e.handoff_level = l;
        e.hybrid_batch = blocks[bi];
        e.total_ms = st.total_ms;
        e.device_ms = st.kernel_ms;
        e.launches = st.launches;
      }
      if (st.total_ms < best_ms) {
        best_ms = st.total_ms;
        *best = c;
      }
    }
  }
  qcp_batch_destroy(b);
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

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
