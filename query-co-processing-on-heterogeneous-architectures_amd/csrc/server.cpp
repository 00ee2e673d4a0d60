// server.cpp -- continuous-batching range-search server over a resident index.
//
// A dispatcher thread owns `slots` batch slots, each with its own HIP stream,
// device buffers and pinned staging buffers.  Submitted requests wait in a FIFO;
// whenever a slot is free and the result-buffer budget allows, every waiting
// request (up to max_batch) is admitted as one batch: H2D copy of the boxes,
// the persistent search kernel (counts + ID-set hash, subtree donation on),
// D2H copy of the results, and an event.  Batches admitted while others are
// still running go to other streams, so their kernels overlap on the device.
// A request's latency is measured from submit to the moment the dispatcher sees
// its batch's event complete.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "qcp_internal.h"

using namespace qcp;

namespace {

double now_us() {
  return std::chrono::duration<double, std::micro>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

constexpr uint32_t kServerSlotCap = 1u << 12;  // donated items per batch slot

struct Slot {
  hipStream_t s = nullptr;
  float* d_q = nullptr;
  uint32_t* d_counts = nullptr;
  unsigned long long* d_hashes = nullptr;
  QueueCtl* d_ctl = nullptr;
  unsigned long long* d_slots = nullptr;
  float* h_q = nullptr;
  uint32_t* h_counts = nullptr;
  unsigned long long* h_hashes = nullptr;
  hipEvent_t done = nullptr;
  std::vector<uint64_t> tickets;
  bool busy = false;
  uint64_t bytes = 0;
};

enum : uint8_t { kQueued = 1, kInFlight = 2, kDone = 3 };

}  // namespace

struct qcp_server {
  qcp_index* ix = nullptr;
  qcp_serve_cfg cfg{};
  unsigned full_grid = 0;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_submit, cv_done;
  std::deque<uint64_t> queue;
  std::vector<float> boxes;
  std::vector<double> t_submit, t_done;
  std::vector<uint32_t> res_count;
  std::vector<unsigned long long> res_hash;
  std::vector<uint8_t> state;
  uint64_t next_ticket = 0;
  std::vector<Slot> slots;
  uint64_t bytes_in_flight = 0;
  bool deferring = false;
  std::atomic<bool> stop{false};
  qcp_serve_stats st{};
  qcp_status err = QCP_OK;
};

static void slot_free(Slot& s) {
  if (s.s) hipStreamSynchronize(s.s);
  if (s.d_q) hipFree(s.d_q);
  if (s.d_counts) hipFree(s.d_counts);
  if (s.d_hashes) hipFree(s.d_hashes);
  if (s.d_ctl) hipFree(s.d_ctl);
  if (s.d_slots) hipFree(s.d_slots);
  if (s.h_q) hipHostFree(s.h_q);
  if (s.h_counts) hipHostFree(s.h_counts);
  if (s.h_hashes) hipHostFree(s.h_hashes);
  if (s.done) hipEventDestroy(s.done);
  if (s.s) hipStreamDestroy(s.s);
  s = Slot();
}

static qcp_status slot_alloc(Slot& s, uint32_t D, uint32_t max_batch) {
  const size_t qb = (size_t)2 * D * sizeof(float) * max_batch;
  if (hipStreamCreateWithFlags(&s.s, hipStreamNonBlocking) != hipSuccess) return QCP_EHIP;
  if (hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) return QCP_EHIP;
  if (hipMalloc((void**)&s.d_q, qb) != hipSuccess ||
      hipMalloc((void**)&s.d_counts, 4ull * max_batch) != hipSuccess ||
      hipMalloc((void**)&s.d_hashes, 8ull * max_batch) != hipSuccess ||
      hipMalloc((void**)&s.d_ctl, sizeof(QueueCtl)) != hipSuccess ||
      hipMalloc((void**)&s.d_slots, 8ull * kServerSlotCap) != hipSuccess)
    return QCP_ENOMEM;
  if (hipHostMalloc((void**)&s.h_q, qb, 0) != hipSuccess ||
      hipHostMalloc((void**)&s.h_counts, 4ull * max_batch, 0) != hipSuccess ||
      hipHostMalloc((void**)&s.h_hashes, 8ull * max_batch, 0) != hipSuccess)
    return QCP_ENOMEM;
  s.tickets.reserve(max_batch);
  return QCP_OK;
}

// Enqueue one admitted batch on the slot's stream.
static qcp_status slot_launch(qcp_server* sv, Slot& s, uint32_t n) {
  const qcp_index* ix = sv->ix;
  const uint32_t D = ix->D;
  hipStream_t st = s.s;
  if (hipMemcpyAsync(s.d_q, s.h_q, (size_t)2 * D * 4 * n, hipMemcpyHostToDevice, st) != hipSuccess)
    return QCP_EHIP;
  SearchArgs a = index_search_args(ix, s.d_q, n, s.d_counts, s.d_hashes);
  a.ctl = s.d_ctl;
  a.slots = s.d_slots;
  a.slot_cap = kServerSlotCap;
  a.donate = 1;
  a.max_waiters = 64;
  a.hunger_period = 8;
  a.donate_min_height = 1;
  a.contain = (sv->cfg.contain && ix->contain_ok) ? 1u : 0u;
  if (launch_fill_u32(s.d_counts, n, 0u, st) != hipSuccess ||
      launch_fill_u32((uint32_t*)s.d_hashes, 2ull * n, 0u, st) != hipSuccess ||
      launch_fill_u32((uint32_t*)s.d_slots, 2ull * kServerSlotCap, 0xFFFFFFFFu, st) != hipSuccess)
    return QCP_EHIP;
  unsigned grid = std::max<unsigned>(16u, (n + 1) / 2);
  grid = std::min(grid, sv->full_grid);
  if (sv->cfg.grid_blocks) grid = std::min(grid, sv->cfg.grid_blocks);
  if (launch_search(a, D, ix->box, ix->expl, 1, 4, grid, 256, st) != hipSuccess) return QCP_EHIP;
  if (hipMemcpyAsync(s.h_counts, s.d_counts, 4ull * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(s.h_hashes, s.d_hashes, 8ull * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipEventRecord(s.done, st) != hipSuccess)
    return QCP_EHIP;
  return QCP_OK;
}

static void dispatcher(qcp_server* sv) {
  hipSetDevice(sv->ix->device);
  const uint32_t D = sv->ix->D;
  while (true) {
    bool did = false;
    // retire finished batches
    for (Slot& s : sv->slots) {
      if (!s.busy) continue;
      hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) continue;
      const double t = now_us();
      {
        std::lock_guard<std::mutex> g(sv->mu);
        if (q != hipSuccess && sv->err == QCP_OK) sv->err = QCP_EHIP;
        for (size_t i = 0; i < s.tickets.size(); ++i) {
          const uint64_t tk = s.tickets[i];
          sv->res_count[tk] = s.h_counts[i];
          sv->res_hash[tk] = s.h_hashes[i];
          sv->t_done[tk] = t;
          sv->state[tk] = kDone;
        }
        sv->st.completed += s.tickets.size();
        sv->bytes_in_flight -= s.bytes;
        s.busy = false;
        s.tickets.clear();
      }
      sv->cv_done.notify_all();
      did = true;
    }
    // admit waiting requests into free slots
    for (Slot& s : sv->slots) {
      if (s.busy) continue;
      uint32_t n = 0;
      {
        std::lock_guard<std::mutex> g(sv->mu);
        if (sv->queue.empty()) break;
        n = (uint32_t)std::min<size_t>(sv->queue.size(), sv->cfg.max_batch);
        const uint64_t bytes = 12ull * n;  // count + hash per query
        if (sv->bytes_in_flight > 0 && sv->bytes_in_flight + bytes > sv->cfg.result_budget) {
          if (!sv->deferring) sv->st.deferred_by_budget++;
          sv->deferring = true;
          break;
        }
        sv->deferring = false;
        for (uint32_t i = 0; i < n; ++i) {
          const uint64_t tk = sv->queue.front();
          sv->queue.pop_front();
          memcpy(s.h_q + (size_t)2 * D * i, sv->boxes.data() + (size_t)2 * D * tk,
                 sizeof(float) * 2 * D);
          sv->state[tk] = kInFlight;
          s.tickets.push_back(tk);
        }
        s.bytes = bytes;
        sv->bytes_in_flight += bytes;
        sv->st.batches++;
        sv->st.max_batch_seen = std::max<uint64_t>(sv->st.max_batch_seen, n);
      }
      qcp_status rc = slot_launch(sv, s, n);
      if (rc != QCP_OK) {
        std::lock_guard<std::mutex> g(sv->mu);
        if (sv->err == QCP_OK) sv->err = rc;
      }
      s.busy = true;
      did = true;
      uint64_t inflight = 0;
      for (Slot& x : sv->slots) inflight += x.busy ? 1 : 0;
      std::lock_guard<std::mutex> g(sv->mu);
      sv->st.max_inflight_seen = std::max<uint64_t>(sv->st.max_inflight_seen, inflight);
    }
    if (!did) {
      bool any_busy = false;
      for (Slot& s : sv->slots) any_busy |= s.busy;
      std::unique_lock<std::mutex> lk(sv->mu);
      if (sv->stop.load() && sv->queue.empty() && !any_busy) break;
      if (!any_busy && sv->queue.empty())
        sv->cv_submit.wait_for(lk, std::chrono::microseconds(500));
      else if (any_busy) {
        lk.unlock();
        std::this_thread::yield();
      }
    }
  }
}

extern "C" {

qcp_status qcp_server_start(qcp_index* ix, const qcp_serve_cfg* cfg_in, qcp_server** out) try {
  if (!ix || !out) return QCP_EINVAL;
  *out = nullptr;
  qcp_serve_cfg cfg;
  memset(&cfg, 0, sizeof(cfg));
  if (cfg_in) cfg = *cfg_in;
  if (!cfg.slots) cfg.slots = 4;
  if (!cfg.max_batch) cfg.max_batch = 4096;
  if (!cfg.result_budget) cfg.result_budget = 64ull << 20;
  if (!cfg.max_requests) cfg.max_requests = 1u << 22;
  if (cfg.slots > 64 || cfg.max_batch > (1u << 24)) return QCP_EINVAL;
  if (hipSetDevice(ix->device) != hipSuccess) return QCP_EHIP;
  qcp_server* sv = new (std::nothrow) qcp_server();
  if (!sv) return QCP_ENOMEM;
  sv->ix = ix;
  sv->cfg = cfg;
  sv->full_grid = persistent_grid(ix, 1, 256);
  const size_t R = cfg.max_requests;
  try {
    sv->boxes.resize(R * 2 * ix->D);
    sv->t_submit.resize(R);
    sv->t_done.resize(R);
    sv->res_count.resize(R);
    sv->res_hash.resize(R);
    sv->state.assign(R, 0);
  } catch (...) {
    delete sv;
    return QCP_ENOMEM;
  }
  sv->slots.resize(cfg.slots);
  for (Slot& s : sv->slots) {
    qcp_status rc = slot_alloc(s, ix->D, cfg.max_batch);
    if (rc != QCP_OK) {
      for (Slot& x : sv->slots) slot_free(x);
      delete sv;
      return rc;
    }
  }
  sv->worker = std::thread(dispatcher, sv);
  *out = sv;
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

qcp_status qcp_server_submit(qcp_server* sv, const float* box, uint64_t* ticket) try {
  if (!sv || !box || !ticket) return QCP_EINVAL;
  const uint32_t D = sv->ix->D;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    if (sv->stop.load()) return QCP_EINVAL;
    if (sv->next_ticket >= sv->cfg.max_requests) return QCP_EBUSY;
    const uint64_t tk = sv->next_ticket++;
    memcpy(sv->boxes.data() + (size_t)2 * D * tk, box, sizeof(float) * 2 * D);
    sv->t_submit[tk] = now_us();
    sv->state[tk] = kQueued;
    sv->queue.push_back(tk);
    sv->st.submitted++;
    *ticket = tk;
  }
  sv->cv_submit.notify_one();
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

qcp_status qcp_server_result(qcp_server* sv, uint64_t ticket, uint32_t* count, uint64_t* hash,
                             double* latency_us) {
  if (!sv) return QCP_EINVAL;
  std::lock_guard<std::mutex> g(sv->mu);
  if (ticket >= sv->next_ticket) return QCP_EINVAL;
  if (sv->err != QCP_OK) return sv->err;
  if (sv->state[ticket] != kDone) return QCP_EBUSY;
  if (count) *count = sv->res_count[ticket];
  if (hash) *hash = sv->res_hash[ticket];
  if (latency_us) *latency_us = sv->t_done[ticket] - sv->t_submit[ticket];
  return QCP_OK;
}

qcp_status qcp_server_wait(qcp_server* sv, uint64_t ticket, uint32_t* count, uint64_t* hash,
                           double* latency_us) {
  if (!sv) return QCP_EINVAL;
  {
    std::unique_lock<std::mutex> lk(sv->mu);
    if (ticket >= sv->next_ticket) return QCP_EINVAL;
    sv->cv_done.wait(lk, [&] { return sv->state[ticket] == kDone || sv->err != QCP_OK; });
  }
  return qcp_server_result(sv, ticket, count, hash, latency_us);
}

qcp_status qcp_server_stats(qcp_server* sv, qcp_serve_stats* out) {
  if (!sv || !out) return QCP_EINVAL;
  std::vector<double> lat;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    *out = sv->st;
    lat.reserve(sv->st.completed);
    for (uint64_t t = 0; t < sv->next_ticket; ++t)
      if (sv->state[t] == kDone) lat.push_back(sv->t_done[t] - sv->t_submit[t]);
  }
  if (!lat.empty()) {
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
      size_t i = (size_t)std::min<double>(lat.size() - 1, std::ceil(p * lat.size()) - 1);
      return lat[i];
    };
    out->p50_us = pct(0.50);
    out->p99_us = pct(0.99);
    out->max_us = lat.back();
    double s = 0;
    for (double x : lat) s += x;
    out->mean_us = s / lat.size();
  }
  return QCP_OK;
}

qcp_status qcp_server_stop(qcp_server* sv) {
  if (!sv) return QCP_EINVAL;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stop.store(true);
  }
  sv->cv_submit.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  hipSetDevice(sv->ix->device);
  for (Slot& s : sv->slots) slot_free(s);
  qcp_status e = sv->err;
  delete sv;
  return e;
}

qcp_status qcp_server_replay(qcp_server* sv, const float* boxes, const double* arrival_s,
                             uint64_t n, uint32_t* counts, uint64_t* hashes, double* latency_us) try {
  if (!sv || (n && (!boxes || !arrival_s))) return QCP_EINVAL;
  const uint32_t D = sv->ix->D;
  std::vector<uint64_t> tk(n);
  const double t0 = now_us();
  for (uint64_t i = 0; i < n; ++i) {
    const double target = t0 + arrival_s[i] * 1e6;
    while (now_us() < target) {
    }
    qcp_status rc = qcp_server_submit(sv, boxes + (size_t)2 * D * i, &tk[i]);
    if (rc != QCP_OK) return rc;
  }
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t c = 0;
    uint64_t h = 0;
    double l = 0;
    qcp_status rc = qcp_server_wait(sv, tk[i], &c, &h, &l);
    if (rc != QCP_OK) return rc;
    if (counts) counts[i] = c;
    if (hashes) hashes[i] = h;
    if (latency_us) latency_us[i] = l;
  }
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

}  // extern "C"
