// kmeans.cpp -- k-means + Hilbert ordering of the objects (QCP_ORDER_KMEANS_HILBERT).
//
// The reference's KmeansMapper (src/mapper/kmeans_mapper.cpp:15-259) clusters
// the objects with Lloyd's algorithm (GPU assignment, host centre update; here
// both run on the GPU),
// gives each cluster centre a Hilbert key, ranks the clusters by it, and then
// orders objects by (cluster rank, own Hilbert key).  Its composite key
// cluster_key * N + object_key overflows 64 bits and it skips the last object
// (kmeans_mapper.cpp:250-253); this build defines the order as follows:
//  - seeds: K distinct objects, index (w * N) >> 32 for the Philox words w of
//    (seed 1234, stream 5) in order, repeats skipped;
//  - each iteration assigns every key point to its nearest centre (k_assign)
//    and then moves each non-empty cluster's centre to the mean of its key
//    points: FP64 sums in index order within chunks of 65536 objects, the chunk
//    sums added in chunk order, then rounded to FP32 (k_km_partial /
//    k_km_combine / k_km_finish keep exactly that order, so the result does not
//    depend on how the work is spread);
//  - the loop is the reference's (kmeans_mapper.cpp:170, kmeans_macro.h:4-5):
//    delta = (float)changes / (float)n, repeat while delta > 1e-5 and
//    loop++ < max_loop, loop from 0, so at most max_loop + 1 iterations; the
//    reference's max_loop is 1000 and its K 20000, the defaults here;
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
#include <vector>

#include "hilbert_key.h"
#include "qcp_internal.h"
#include "qcp_kernels.h"

#define HIP_TRY(expr)                          \
  do {                                         \
    if ((expr) != hipSuccess) return QCP_EHIP; \
  } while (0)

namespace {

constexpr uint32_t kDefaultClusters = 20000;  // DFN_NumberOfClusters, kmeans_macro.h
constexpr uint32_t kDefaultIters = 1001;      // loop++ < 1000: kMeansLoopIteration + 1
constexpr double kDeltaMin = 0.00001;         // DFN_KmeansThreashhold
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

struct Events {
  hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
  bool create() {
    for (auto& x : e)
      if (hipEventCreate(&x) != hipSuccess) return false;
    return true;
  }
  ~Events() {
    for (auto x : e)
      if (x) hipEventDestroy(x);
  }
};

}  // namespace

uint32_t qcp::kmeans_seed_index(uint64_t word, uint64_t n) {
  return (uint32_t)((word * n) >> 32);
}

qcp_status qcp::kmeans_keys(qcp_index* ix, const float* d_objs, uint64_t* key_bits) {
  const uint64_t n = ix->n;
  const uint32_t D = ix->D;
  hipStream_t s = ix->stream;
  uint32_t K = ix->cfg.clusters ? ix->cfg.clusters : kDefaultClusters;
  if ((uint64_t)K > n) K = (uint32_t)n;
  const uint32_t max_iters = ix->cfg.kmeans_iters ? ix->cfg.kmeans_iters : kDefaultIters;
  const bool center = ix->box && ix->cfg.key_src == QCP_SRC_CENTER;
  // seeds: K distinct object indexes from the Philox words
  std::vector<uint32_t> seeds(K);
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
      seeds[k++] = idx;
    }
  }
  // the centre update works on windows of G chunks: partials G x K x D
  // (bounded to about 1 GiB), running sums K x D
  const uint64_t nchunks = (n + kKmChunk - 1) / kKmChunk;
  const uint64_t per_chunk = std::max<uint64_t>(1, (uint64_t)K * (8ull * D + 4));
  const uint64_t G = std::max<uint64_t>(1, std::min<uint64_t>(nchunks, (1ull << 30) / per_chunk));
  DevBuf<float> d_c;
  DevBuf<uint32_t> d_cid, d_changes, d_rank, d_seeds, d_cnt;
  DevBuf<double> d_part, d_sum;
  DevBuf<uint64_t> d_total;
  if (!d_c.alloc((uint64_t)K * D) || !d_cid.alloc(n) || !d_changes.alloc(1) || !d_rank.alloc(K) ||
      !d_seeds.alloc(K) || !d_part.alloc(G * K * D) || !d_cnt.alloc(G * K) ||
      !d_sum.alloc((uint64_t)K * D) || !d_total.alloc(K))
    return QCP_ENOMEM;
  HIP_TRY(hipMemcpyAsync(d_seeds.p, seeds.data(), 4ull * K, hipMemcpyHostToDevice, s));
  HIP_TRY(launch_km_seed(d_objs, D, ix->box, center, d_seeds.p, K, d_c.p, s));
  HIP_TRY(launch_fill_u32(d_cid.p, n, 0xFFFFFFFFu, s));
  const uint32_t max_loop = max_iters - 1;
  uint32_t loop = 0, its = 0;
  float delta = 0.0f;
  double assign_ms = 0.0, update_ms = 0.0;
  Events ev;
  if (!ev.create()) return QCP_EHIP;
  do {
    uint32_t changes = 0;
    HIP_TRY(launch_fill_u32(d_changes.p, 1, 0u, s));
    HIP_TRY(hipEventRecord(ev.e[0], s));
    HIP_TRY(launch_assign(d_objs, n, D, ix->box, center, d_c.p, K, d_cid.p, d_changes.p, s));
    HIP_TRY(hipEventRecord(ev.e[1], s));
    HIP_TRY(hipMemcpyAsync(&changes, d_changes.p, 4, hipMemcpyDeviceToHost, s));
    // centre update, windows of chunks in chunk order
    HIP_TRY(launch_fill_u32((uint32_t*)d_sum.p, 2ull * K * D, 0u, s));
    HIP_TRY(launch_fill_u32((uint32_t*)d_total.p, 2ull * K, 0u, s));
    for (uint64_t c0 = 0; c0 < nchunks; c0 += G) {
      const uint64_t w = std::min(G, nchunks - c0);
      HIP_TRY(launch_fill_u32((uint32_t*)d_part.p, 2ull * w * K * D, 0u, s));
      HIP_TRY(launch_fill_u32(d_cnt.p, w * K, 0u, s));
      HIP_TRY(launch_km_update(d_objs, n, D, ix->box, center, d_cid.p, K, c0, w, d_part.p,
                               d_cnt.p, d_sum.p, d_total.p, s));
    }
    HIP_TRY(launch_km_finish(d_sum.p, d_total.p, K, D, d_c.p, s));
    HIP_TRY(hipEventRecord(ev.e[2], s));
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
    assign_ms += ms;
    HIP_TRY(hipEventElapsedTime(&ms, ev.e[1], ev.e[2]));
    update_ms += ms;
    ++its;
    delta = (float)changes;
    delta /= (float)n;
  } while ((double)delta > kDeltaMin && loop++ < max_loop);
  ix->kmeans_its = its;
  ix->kmeans_ms[0] = assign_ms;
  ix->kmeans_ms[1] = update_ms;
  std::vector<float> centers((size_t)K * D);
  HIP_TRY(hipMemcpyAsync(centers.data(), d_c.p, 4ull * K * D, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
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
