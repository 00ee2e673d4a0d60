// kmeans.cpp -- k-means + Hilbert ordering of the objects (QCP_ORDER_KMEANS_HILBERT).
//
// The reference's KmeansMapper (src/mapper/kmeans_mapper.cpp:15-259) clusters
// the objects with Lloyd's algorithm (GPU assignment, host centre update),
// gives each cluster centre a Hilbert key, ranks the clusters by it, and then
// orders objects by (cluster rank, own Hilbert key).  Its composite key
// cluster_key * N + object_key overflows 64 bits and it skips the last object
// (kmeans_mapper.cpp:250-253); this build defines the order as follows:
//  - seeds: K distinct objects, index (w * N) >> 32 for the Philox words w of
//    (seed 1234, stream 5) in order, repeats skipped;
//  - each iteration assigns every key point to its nearest centre (GPU,
//    k_assign) and, unless no assignment changed, moves each non-empty
//    cluster's centre to the mean of its key points: FP64 sums in index order
//    within chunks of 65536 objects, the chunk sums added in chunk order, then
//    rounded to FP32 (so any number of host threads gives the same centres);
//  - clusters are ranked by (Hilbert key of the centre under the index's
//    quantiser, cluster number);
//  - an object's sort key is rank << (key bits - s) | key >> s with
//    s = max(0, key bits + kbits - 64), kbits = ceil(log2 K) (the cluster rank
//    above the object's key, the key's lowest bits dropped only when both do
//    not fit 64 bits); the stable sort breaks ties by input index as
//    everywhere else.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "hilbert_key.h"
#include "qcp_internal.h"

#define HIP_TRY(expr)                          \
  do {                                         \
    if ((expr) != hipSuccess) return QCP_EHIP; \
  } while (0)

namespace {

constexpr uint64_t kChunk = 1 << 16;
constexpr uint64_t kSeed = 1234;
constexpr uint32_t kSeedStream = 5;

template <typename T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() {
    if (p) hipFree(p);
  }
  bool alloc(uint64_t n) { return hipMalloc((void**)&p, sizeof(T) * (n ? n : 1)) == hipSuccess; }
};

}  // namespace

uint32_t qcp::kmeans_seed_index(uint64_t word, uint64_t n) {
  return (uint32_t)((word * n) >> 32);
}

void qcp::kmeans_update(const float* pts, uint64_t n, uint32_t D, const uint32_t* cid,
                        uint32_t K, float* centers, unsigned threads) {
  const uint64_t nchunks = (n + kChunk - 1) / kChunk;
  std::vector<double> part(nchunks * K * D, 0.0);
  std::vector<uint64_t> cnt(nchunks * K, 0);
  auto work = [&](uint64_t c0, uint64_t c1) {
    for (uint64_t c = c0; c < c1; ++c) {
      double* ps = part.data() + c * K * D;
      uint64_t* cs = cnt.data() + c * K;
      const uint64_t e = std::min(n, (c + 1) * kChunk);
      for (uint64_t i = c * kChunk; i < e; ++i) {
        const uint32_t k = cid[i];
        cs[k] += 1;
        for (uint32_t d = 0; d < D; ++d) ps[k * D + d] += (double)pts[i * D + d];
      }
    }
  };
  threads = std::max(1u, std::min<unsigned>(threads, (unsigned)std::max<uint64_t>(1, nchunks)));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < threads; ++t)
    th.emplace_back(work, nchunks * t / threads, nchunks * (t + 1) / threads);
  for (auto& x : th) x.join();
  for (uint32_t k = 0; k < K; ++k) {
    uint64_t total = 0;
    double sum[3] = {0.0, 0.0, 0.0};
    for (uint64_t c = 0; c < nchunks; ++c) {
      total += cnt[c * K + k];
      for (uint32_t d = 0; d < D; ++d) sum[d] += part[(c * K + k) * D + d];
    }
    if (total)
      for (uint32_t d = 0; d < D; ++d) centers[k * D + d] = (float)(sum[d] / (double)total);
  }
}

qcp_status qcp::kmeans_keys(qcp_index* ix, const float* d_objs, uint64_t* key_bits) {
  const uint64_t n = ix->n;
  const uint32_t D = ix->D, E = ix->E;
  hipStream_t s = ix->stream;
  uint32_t K = ix->cfg.clusters ? ix->cfg.clusters : 64;
  if ((uint64_t)K > n) K = (uint32_t)n;
  const uint32_t iters = ix->cfg.kmeans_iters ? ix->cfg.kmeans_iters : 20;
  const bool center = ix->box && ix->cfg.key_src == QCP_SRC_CENTER;
  // key points on the host (lower corner or centre, the Hilbert key's source)
  std::vector<float> objs((size_t)n * E), pts((size_t)n * D);
  HIP_TRY(hipMemcpyAsync(objs.data(), d_objs, 4ull * n * E, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (uint64_t i = 0; i < n; ++i)
    for (uint32_t d = 0; d < D; ++d) {
      float p = objs[i * E + d];
      if (center) p = (p + objs[i * E + D + d]) * 0.5f;
      pts[i * D + d] = p;
    }
  // seeds
  std::vector<float> centers((size_t)K * D);
  {
    std::vector<uint8_t> taken(n, 0);
    std::vector<uint32_t> words;
    uint64_t next = 0;
    for (uint32_t k = 0; k < K;) {
      if (next >= words.size()) {
        const uint64_t m = std::max<uint64_t>(1024, 2 * (uint64_t)K);
        words.resize(next + m);
        qcp_gen_words(words.data() + next, m, kSeed, kSeedStream, next);
      }
      const uint32_t idx = kmeans_seed_index(words[next++], n);
      if (taken[idx]) continue;
      taken[idx] = 1;
      for (uint32_t d = 0; d < D; ++d) centers[k * D + d] = pts[(uint64_t)idx * D + d];
      ++k;
    }
  }
  DevBuf<float> d_c;
  DevBuf<uint32_t> d_cid, d_changes, d_rank;
  if (!d_c.alloc((uint64_t)K * D) || !d_cid.alloc(n) || !d_changes.alloc(1) || !d_rank.alloc(K))
    return QCP_ENOMEM;
  HIP_TRY(launch_fill_u32(d_cid.p, n, 0xFFFFFFFFu, s));
  std::vector<uint32_t> cid(n);
  unsigned hw = std::thread::hardware_concurrency();
  hw = std::max(1u, std::min(hw, 16u));
  for (uint32_t it = 0; it < iters; ++it) {
    uint32_t changes = 0;
    HIP_TRY(hipMemcpyAsync(d_c.p, centers.data(), 4ull * K * D, hipMemcpyHostToDevice, s));
    HIP_TRY(launch_fill_u32(d_changes.p, 1, 0u, s));
    HIP_TRY(launch_assign(d_objs, n, D, ix->box, center, d_c.p, K, d_cid.p, d_changes.p, s));
    HIP_TRY(hipMemcpyAsync(&changes, d_changes.p, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(cid.data(), d_cid.p, 4ull * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (it > 0 && changes == 0) break;
    kmeans_update(pts.data(), n, D, cid.data(), K, centers.data(), hw);
  }
  // cluster ranks by (centre key, cluster number)
  std::vector<std::pair<uint64_t, uint32_t>> ck(K);
  for (uint32_t k = 0; k < K; ++k) {
    uint32_t c[3] = {0, 0, 0};
    for (uint32_t d = 0; d < D; ++d) {
      const float p = centers[k * D + d];
      if (ix->cfg.key_mode == QCP_KEY_REF) {
        if (!quant_ref(p, ix->nb, &c[d])) c[d] = 0;
      } else {
        c[d] = quant_b16(p);
      }
    }
    ck[k] = {D == 2 ? hilbert_key<2>(c, ix->nb) : hilbert_key<3>(c, ix->nb), k};
  }
  std::sort(ck.begin(), ck.end());
  std::vector<uint32_t> rank(K);
  for (uint32_t r = 0; r < K; ++r) rank[ck[r].second] = r;
  unsigned kbits = 0;
  while ((1ull << kbits) < K) ++kbits;
  const unsigned keybits = D * ix->nb;
  const unsigned shift_key = keybits + kbits > 64 ? keybits + kbits - 64 : 0;
  const unsigned shift_rank = kbits ? keybits - shift_key : 64;
  HIP_TRY(hipMemcpyAsync(d_rank.p, rank.data(), 4ull * K, hipMemcpyHostToDevice, s));
  HIP_TRY(launch_compose(ix->d_keys, d_cid.p, d_rank.p, n, shift_key, shift_rank, s));
  HIP_TRY(hipStreamSynchronize(s));
  *key_bits = keybits - shift_key + kbits;
  return QCP_OK;
}
