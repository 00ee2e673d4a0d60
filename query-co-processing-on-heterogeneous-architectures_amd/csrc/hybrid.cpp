// hybrid.cpp -- CPU-GPU co-processing search (QCP_SEARCH_HYBRID).
//
// The reference's Hybrid tree (src/tree/hybrid.cpp:451-969) splits every
// query between the two processors: CPU threads walk the upper tree
// (TraverseInternalNodes, 820-900) and, for each run of overlapping leaves
// they reach, launch a leaf-scan kernel on the GPU (global_ParallelScan_Leafnodes,
// 1031-1082), one small launch per chunk, with a hand-tuned chunk heuristic
// (833-876).  Here the split is a tree level: host threads walk the levels
// above `handoff_level` on a host mirror of the node boxes and emit one work
// item (query, node) per overlapping node at that level; a persistent search
// launch walks those subtrees (by default the leaves themselves, as in the
// reference).  Queries are taken in the batch's spatial order (Hilbert order
// of their centres, as the persistent form processes them) and in blocks,
// double-buffered: the host traverses block b+1 while the GPU searches block b.  Every overlapping leaf
// is searched exactly once, so counts, hashes and the traversal statistics
// equal the canonical traversal's.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

#include "qcp_internal.h"

using qcp::QueueCtl;
using qcp::SearchArgs;

#define HIP_TRY(expr)                             \
  do {                                            \
    if ((expr) != hipSuccess) return QCP_EHIP;    \
  } while (0)
#define QCP_TRY(expr)                 \
  do {                                \
    qcp_status _s = (expr);           \
    if (_s != QCP_OK) return _s;      \
  } while (0)

namespace {

constexpr uint32_t kLeafFlag = 0x80000000u;

// Closed-interval FP32 overlap test of the reference (src/node/node_soa.cpp:94-113).
inline bool overlap(const float* q, const float* box, uint32_t D) {
  for (uint32_t d = 0; d < D; ++d)
    if (q[d] > box[D + d] || q[D + d] < box[d]) return false;
  return true;
}

// Items of one query over a packed tree: nodes at level L (or leaves) whose
// parent entry overlaps the query.  ie counts the internal entries examined.
void items_packed(const qcp_index* ix, const float* q, uint32_t qi, int L,
                  std::vector<unsigned long long>& out, uint64_t& ie,
                  std::vector<std::pair<int, uint64_t>>& stack) {
  const int h = ix->h;
  const uint32_t D = ix->D, F = ix->F;
  const float* nb = ix->h_nbox.data();
  stack.clear();
  stack.push_back({0, 0});
  while (!stack.empty()) {
    const int l = stack.back().first;
    const uint64_t j = stack.back().second;
    stack.pop_back();
    const uint64_t c0 = j * F;
    const uint64_t c1 = std::min<uint64_t>(c0 + F, ix->level_count[l + 1]);
    ie += c1 - c0;
    // children in descending order onto the stack: popped in ascending order
    for (uint64_t cj = c1; cj-- > c0;) {
      const uint64_t node = ix->level_off[l + 1] + cj;
      if (!overlap(q, nb + node * 2 * D, D)) continue;
      if (l + 1 >= L || l + 1 == h - 1)
        out.push_back(((unsigned long long)qi << 32) | (uint32_t)node);
      else
        stack.push_back({l + 1, cj});
    }
  }
}

// Items below one internal root node of an explicit tree.
void items_from(const qcp_index* ix, const float* q, uint32_t qi, int L, uint32_t root,
                std::vector<unsigned long long>& out, uint64_t& ie,
                std::vector<std::pair<int, uint64_t>>& stack) {
  const uint32_t D = ix->D;
  const float* nb = ix->h_nbox.data();
  stack.clear();
  stack.push_back({0, root});
  while (!stack.empty()) {
    const int dep = stack.back().first;
    const uint64_t nd = stack.back().second;
    stack.pop_back();
    const uint32_t cf = ix->h_first[nd];
    const uint32_t cc = ix->h_count[nd] & ~kLeafFlag;
    ie += cc;
    for (uint32_t k = cc; k-- > 0;) {
      const uint64_t child = (uint64_t)cf + k;
      if (!overlap(q, nb + child * 2 * D, D)) continue;
      if ((ix->h_count[child] & kLeafFlag) || dep + 1 >= L)
        out.push_back(((unsigned long long)qi << 32) | (uint32_t)child);
      else
        stack.push_back({dep + 1, child});
    }
  }
}

// The same over an explicit tree (BVH, Guttman, or a partitioned MPHR, whose
// roots are nodes 0 .. parts-1, walked in order); L counts depth from a root.
void items_explicit(const qcp_index* ix, const float* q, uint32_t qi, int L,
                    std::vector<unsigned long long>& out, uint64_t& ie,
                    std::vector<std::pair<int, uint64_t>>& stack) {
  for (uint32_t r = 0; r < ix->parts; ++r) {
    if (ix->h_count[r] & kLeafFlag) {
      out.push_back(((unsigned long long)qi << 32) | r);  // a one-leaf partition
      continue;
    }
    items_from(ix, q, qi, L, r, out, ie, stack);
  }
}

// Host mirror of the node boxes, [node][lo(D), hi(D)] (made once per build).
qcp_status mirror_boxes(qcp_index* ix) {
  if (!ix->h_nbox.empty() || ix->n_nodes == 0) return QCP_OK;
  const uint64_t m = ix->n_nodes;
  const uint32_t B = 2 * ix->D;
  std::vector<float> soa((size_t)B * m);
  for (uint32_t e = 0; e < B; ++e)
    if (hipMemcpyAsync(soa.data() + (size_t)e * m, ix->d_nbox + (uint64_t)e * ix->node_stride,
                       4 * m, hipMemcpyDeviceToHost, ix->stream) != hipSuccess)
      return QCP_EHIP;
  if (hipStreamSynchronize(ix->stream) != hipSuccess) return QCP_EHIP;
  ix->h_nbox.resize((size_t)B * m);
  for (uint64_t i = 0; i < m; ++i)
    for (uint32_t e = 0; e < B; ++e) ix->h_nbox[i * B + e] = soa[(size_t)e * m + i];
  return QCP_OK;
}

struct PinnedItems {
  unsigned long long* h = nullptr;
  unsigned long long* d = nullptr;
  uint64_t cap = 0;
  hipEvent_t done = nullptr;
  bool used = false;
  ~PinnedItems() {
    if (h) hipHostFree(h);
    if (d) hipFree(d);
    if (done) hipEventDestroy(done);
  }
  qcp_status reserve(uint64_t n) {
    if (n <= cap) return QCP_OK;
    if (h) hipHostFree(h);
    if (d) hipFree(d);
    h = nullptr;
    d = nullptr;
    cap = 0;
    const uint64_t c = std::max<uint64_t>(n + n / 4, 1 << 16);
    if (hipHostMalloc((void**)&h, 8 * c, 0) != hipSuccess) return QCP_ENOMEM;
    if (hipMalloc((void**)&d, 8 * c) != hipSuccess) return QCP_ENOMEM;
    cap = c;
    return QCP_OK;
  }
};

}  // namespace

qcp_status qcp::hybrid_run(qcp_batch* b, const qcp_search_cfg& cfg, qcp_search_stats* stats) {
  qcp_index* ix = b->ix;
  hipStream_t s = ix->stream;
  const uint64_t q = b->q;
  if (cfg.outputs & QCP_OUT_IDS) return QCP_EINVAL;  // items of a query finish in any order
  if (cfg.block_waves > 8) return QCP_EINVAL;
  const auto t_start = std::chrono::steady_clock::now();
  qcp_search_stats st;
  memset(&st, 0, sizeof(st));
  if (q && ix->n) {
    QCP_TRY(mirror_boxes(ix));
    if (b->h_queries.size() != (size_t)2 * ix->D * q) {
      b->h_queries.resize((size_t)2 * ix->D * q);
      HIP_TRY(hipMemcpyAsync(b->h_queries.data(), b->d_queries, 4ull * b->h_queries.size(),
                             hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
    }
  }
  const int out_mode = (cfg.outputs & QCP_OUT_HASH) ? 1 : 0;
  const unsigned block = 64u * (cfg.block_waves ? cfg.block_waves : kDefaultBlockWaves);
  const int u = cfg.leaf_chunks ? (int)cfg.leaf_chunks : kDefaultLeafChunks;
  const unsigned full_grid = persistent_grid(ix, out_mode, block);
  const uint64_t qb = cfg.hybrid_batch ? cfg.hybrid_batch : 16384;
  unsigned nt = cfg.cpu_threads ? cfg.cpu_threads : std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(nt, 64u));
  // default hand-off: the leaves (the reference's CPU-internal / GPU-leaf split)
  const int L = cfg.handoff_level ? (int)cfg.handoff_level : (ix->expl ? 1 << 30 : ix->h - 1);

  PinnedItems buf[2];
  for (auto& p : buf) HIP_TRY(hipEventCreateWithFlags(&p.done, hipEventDisableTiming));
  const uint64_t nblocks = (q + qb - 1) / qb;
  QueueCtl* h_ctl = nullptr;
  if (nblocks) HIP_TRY(hipHostMalloc((void**)&h_ctl, sizeof(QueueCtl) * nblocks, 0));
  struct CtlFree {
    QueueCtl* p;
    ~CtlFree() {
      if (p) hipHostFree(p);
    }
  } ctl_free{h_ctl};
  std::vector<std::vector<unsigned long long>> tl(nt);
  std::vector<uint64_t> t_ie(nt, 0);
  // queries in the batch's spatial order: a block's items then cover
  // neighbouring queries, whose subtrees share leaves in each XCD's L2
  std::vector<uint32_t> ord;
  QCP_TRY(batch_order_host(b, cfg.order == 0, ord));

  HIP_TRY(launch_fill_u32(b->d_counts, q, 0u, s));
  HIP_TRY(launch_fill_u32((uint32_t*)b->d_hashes, 2 * q, 0u, s));
  HIP_TRY(hipEventRecord(b->ev[0], s));
  SearchArgs a = index_search_args(ix, b->d_queries, q, b->d_counts, b->d_hashes);
  a.ctl = b->d_ctl;
  a.contain = (cfg.contain && ix->contain_ok) ? 1u : 0u;
  uint64_t host_entries = 0, launches = 0;
  for (uint64_t blk = 0; blk < nblocks; ++blk) {
    const uint64_t q0 = blk * qb, q1 = std::min(q, q0 + qb);
    // host traversal of this block, static chunks over the threads
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
      const uint64_t a0 = q0 + (q1 - q0) * t / nt, a1 = q0 + (q1 - q0) * (t + 1) / nt;
      tl[t].clear();
      t_ie[t] = 0;
      if (a0 >= a1 || ix->h == 0) continue;
      th.emplace_back([&, t, a0, a1] {
        std::vector<std::pair<int, uint64_t>> stack;
        for (uint64_t j = a0; j < a1; ++j) {
          const uint64_t qi = ord[j];
          const float* qp = b->h_queries.data() + qi * 2 * ix->D;
          if (ix->h == 1 || L <= 0) {
            for (uint32_t r = 0; r < ix->parts; ++r)  // the roots themselves
              tl[t].push_back(((unsigned long long)qi << 32) | r);
          } else if (ix->expl) {
            items_explicit(ix, qp, (uint32_t)qi, L, tl[t], t_ie[t], stack);
          } else {
            items_packed(ix, qp, (uint32_t)qi, L, tl[t], t_ie[t], stack);
          }
        }
      });
    }
    for (auto& x : th) x.join();
    uint64_t n_items = 0;
    for (unsigned t = 0; t < nt; ++t) {
      n_items += tl[t].size();
      host_entries += t_ie[t];
    }
    PinnedItems& pb = buf[blk & 1];
    if (pb.used) HIP_TRY(hipEventSynchronize(pb.done));  // the launch two blocks back
    QCP_TRY(pb.reserve(n_items));
    uint64_t pos = 0;
    for (unsigned t = 0; t < nt; ++t) {
      if (!tl[t].empty()) memcpy(pb.h + pos, tl[t].data(), 8 * tl[t].size());
      pos += tl[t].size();
    }
    memset(&h_ctl[blk], 0, sizeof(QueueCtl));
    if (n_items) {
      HIP_TRY(hipMemcpyAsync(pb.d, pb.h, 8 * n_items, hipMemcpyHostToDevice, s));
      a.items = pb.d;
      a.n_items = (uint32_t)n_items;
      const unsigned wpb = block / 64u;
      const unsigned grid = (unsigned)std::max<uint64_t>(
          1, std::min<uint64_t>(full_grid, (n_items + wpb - 1) / wpb));
      HIP_TRY(launch_search(a, ix->D, ix->box, ix->expl, out_mode, u, grid, block, s));
      HIP_TRY(hipMemcpyAsync(&h_ctl[blk], b->d_ctl, sizeof(QueueCtl), hipMemcpyDeviceToHost, s));
      ++launches;
    }
    HIP_TRY(hipEventRecord(pb.done, s));
    pb.used = true;
  }
  HIP_TRY(hipEventRecord(b->ev[1], s));
  HIP_TRY(hipStreamSynchronize(s));
  float dev_ms = 0;
  hipEventElapsedTime(&dev_ms, b->ev[0], b->ev[1]);
  for (uint64_t blk = 0; blk < nblocks; ++blk)
    for (int x = 0; x < 8; ++x) {
      st.internal_entries += h_ctl[blk].internal_entries[x];
      st.leaf_entries += h_ctl[blk].leaf_entries[x];
      st.total_hits += h_ctl[blk].total_hits[x];
      st.contained_runs += h_ctl[blk].contained[x];
    }
  st.internal_entries += host_entries;
  st.launches = launches;
  st.kernel_ms = dev_ms;
  st.search_ms = dev_ms;
  st.total_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  b->total_hits = st.total_hits;
  if (stats) *stats = st;
  return QCP_OK;
}
