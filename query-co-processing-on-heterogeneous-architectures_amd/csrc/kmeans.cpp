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
//    k_assign) and then moves each non-empty cluster's centre to the mean of
//    its key points: FP64 sums in index order within chunks of 65536 objects,
//    the chunk sums added in chunk order, then rounded to FP32 (so any number
//    of host threads gives the same centres);
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
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>
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

struct PinnedU32 {
  uint32_t* p = nullptr;
  ~PinnedU32() {
    if (p) hipHostFree(p);
  }
  bool alloc(uint64_t n) { return hipHostMalloc((void**)&p, 4 * (n ? n : 1), 0) == hipSuccess; }
};

struct Events {
  hipEvent_t e[2] = {nullptr, nullptr};
  bool create() {
    return hipEventCreate(&e[0]) == hipSuccess && hipEventCreate(&e[1]) == hipSuccess;
  }
  ~Events() {
    for (auto x : e)
      if (x) hipEventDestroy(x);
  }
};

// A fixed set of host threads that run fn(i) for i in [0, count) together with
// the calling thread; run() returns when every index is done.
class Pool {
 public:
  explicit Pool(unsigned threads) {
    try {
      for (unsigned t = 1; t < threads; ++t) th_.emplace_back([this] { loop(); });
    } catch (...) {
      stop();  // a thread that could not start: end the ones that did
      throw;
    }
  }
  ~Pool() { stop(); }
  void run(uint64_t count, const std::function<void(uint64_t)>& fn) {
    {
      std::lock_guard<std::mutex> g(mu_);
      fn_ = &fn;
      count_ = count;
      next_.store(0);
      busy_ = th_.size();
      ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> l(mu_);
    done_.wait(l, [this] { return busy_ == 0; });
    fn_ = nullptr;
  }

 private:
  void work() {
    for (;;) {
      const uint64_t i = next_.fetch_add(1);
      if (i >= count_) return;
      (*fn_)(i);
    }
  }
  void stop() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (auto& t : th_)
      if (t.joinable()) t.join();
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> l(mu_);
        cv_.wait(l, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
      }
      work();
      std::lock_guard<std::mutex> g(mu_);
      if (--busy_ == 0) done_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(uint64_t)>* fn_ = nullptr;
  uint64_t count_ = 0, gen_ = 0;
  size_t busy_ = 0;
  bool stop_ = false;
  std::atomic<uint64_t> next_{0};
};

// Centre update with the chunk rule above.  Chunks are summed in windows of
// G at a time (partials G x K x D, bounded to about 256 MiB); within a window
// each chunk's partials are formed in parallel, then added to the running sums
// in chunk order, clusters in parallel.  Only clusters present in a chunk are
// added (adding a chunk's +0.0 partial would leave the sum unchanged), and only
// their partials are cleared again, so the buffers are allocated and zeroed
// once for all iterations.
class CentreUpdate {
 public:
  CentreUpdate(unsigned threads, uint64_t n, uint32_t D, uint32_t K) : n_(n), D_(D), K_(K) {
    nchunks_ = (n + kChunk - 1) / kChunk;
    const uint64_t per_chunk = std::max<uint64_t>(1, (uint64_t)K * (8ull * D + 4));
    G_ = std::max<uint64_t>(
        1, std::min<uint64_t>({nchunks_, 2ull * threads, (256ull << 20) / per_chunk}));
    part_.assign(G_ * K * D, 0.0);
    cnt_.assign(G_ * K, 0);
    sum_.resize((size_t)K * D);
    total_.resize(K);
    // the clusters present in each chunk of the window: at most min(K, chunk)
    // each, reserved here so that the workers allocate nothing
    touched_.resize(G_);
    for (auto& t : touched_) t.reserve(std::min<uint64_t>(K, kChunk));
  }

  void run(Pool& pool, const float* pts, const uint32_t* cid, float* centers) {
    const uint64_t K = K_, D = D_, n = n_;
    std::fill(sum_.begin(), sum_.end(), 0.0);
    std::fill(total_.begin(), total_.end(), 0);
    const uint64_t kblk = 1024;
    for (uint64_t w0 = 0; w0 < nchunks_; w0 += G_) {
      const uint64_t w = std::min(G_, nchunks_ - w0);
      pool.run(w, [&](uint64_t g) {
        double* ps = part_.data() + g * K * D;
        uint32_t* cs = cnt_.data() + g * K;
        std::vector<uint32_t>& tl = touched_[g];
        const uint64_t b = (w0 + g) * kChunk, e = std::min(n, b + kChunk);
        for (uint64_t i = b; i < e; ++i) {
          const uint32_t k = cid[i];
          if (cs[k]++ == 0) tl.push_back(k);
          for (uint64_t d = 0; d < D; ++d) ps[k * D + d] += (double)pts[i * D + d];
        }
      });
      pool.run((K + kblk - 1) / kblk, [&](uint64_t kb) {
        const uint64_t k1 = std::min<uint64_t>(K, (kb + 1) * kblk);
        for (uint64_t g = 0; g < w; ++g)
          for (uint64_t k = kb * kblk; k < k1; ++k)
            if (cnt_[g * K + k]) {
              total_[k] += cnt_[g * K + k];
              for (uint64_t d = 0; d < D; ++d) sum_[k * D + d] += part_[(g * K + k) * D + d];
            }
      });
      pool.run(w, [&](uint64_t g) {
        for (const uint32_t k : touched_[g]) {
          cnt_[g * K + k] = 0;
          for (uint64_t d = 0; d < D; ++d) part_[(g * K + k) * D + d] = 0.0;
        }
        touched_[g].clear();
      });
    }
    for (uint64_t k = 0; k < K; ++k)
      if (total_[k])
        for (uint64_t d = 0; d < D; ++d)
          centers[k * D + d] = (float)(sum_[k * D + d] / (double)total_[k]);
  }

 private:
  uint64_t n_, nchunks_, G_;
  uint32_t D_, K_;
  std::vector<double> part_, sum_;
  std::vector<uint32_t> cnt_;
  std::vector<uint64_t> total_;
  std::vector<std::vector<uint32_t>> touched_;
};

}  // namespace

uint32_t qcp::kmeans_seed_index(uint64_t word, uint64_t n) {
  return (uint32_t)((word * n) >> 32);
}

qcp_status qcp::kmeans_keys(qcp_index* ix, const float* d_objs, uint64_t* key_bits) {
  const uint64_t n = ix->n;
  const uint32_t D = ix->D, E = ix->E;
  hipStream_t s = ix->stream;
  uint32_t K = ix->cfg.clusters ? ix->cfg.clusters : kDefaultClusters;
  if ((uint64_t)K > n) K = (uint32_t)n;
  const uint32_t max_iters = ix->cfg.kmeans_iters ? ix->cfg.kmeans_iters : kDefaultIters;
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
  // the clusters come back every iteration: pinned host memory
  PinnedU32 cid;
  if (!cid.alloc(n)) return QCP_ENOMEM;
  unsigned hw = std::thread::hardware_concurrency();
  hw = std::max(1u, std::min(hw, 16u));
  Pool pool(hw);
  CentreUpdate update(hw, n, D, K);
  const uint32_t max_loop = max_iters - 1;
  uint32_t loop = 0, its = 0;
  float delta = 0.0f;
  double assign_ms = 0.0, update_ms = 0.0;
  Events ev;
  if (!ev.create()) return QCP_EHIP;
  do {
    uint32_t changes = 0;
    HIP_TRY(hipMemcpyAsync(d_c.p, centers.data(), 4ull * K * D, hipMemcpyHostToDevice, s));
    HIP_TRY(launch_fill_u32(d_changes.p, 1, 0u, s));
    HIP_TRY(hipEventRecord(ev.e[0], s));
    HIP_TRY(launch_assign(d_objs, n, D, ix->box, center, d_c.p, K, d_cid.p, d_changes.p, s));
    HIP_TRY(hipEventRecord(ev.e[1], s));
    HIP_TRY(hipMemcpyAsync(&changes, d_changes.p, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(cid.p, d_cid.p, 4ull * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
    assign_ms += ms;
    ++its;
    const auto t0 = std::chrono::steady_clock::now();
    update.run(pool, pts.data(), cid.p, centers.data());
    update_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    delta = (float)changes;
    delta /= (float)n;
  } while ((double)delta > kDeltaMin && loop++ < max_loop);
  ix->kmeans_its = its;
  ix->kmeans_ms[0] = assign_ms;
  ix->kmeans_ms[1] = update_ms;
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
