// server.cpp -- continuous-batching range-search server over a resident index.
//
// A dispatcher thread owns `slots` batch slots, each with its own HIP stream,
// device buffers and pinned staging buffers.  Submitted requests wait in a FIFO;
// whenever a slot is free, the waiting requests (up to max_batch) are taken as
// one batch.  Batches taken while others are still running go to other
// streams, so their kernels overlap on the device.
//
// Count + hash service: the batch is admitted when its result bytes (12 per
// query) fit the result-buffer budget; one persistent search (subtree
// donation on) and a D2H copy of the results follow.
//
// ID-list service (QCP_OUT_IDS): the batch's count pass (containment shortcut,
// donation) runs first and sizes its sorted ID lists; the batch is then
// admitted to the ID pass only when those bytes (4 per ID + 12 per query) fit
// the budget beside the batches already in flight, so admission is charged
// the real result memory.  The ID pass writes the lists at the prefix-sum
// offsets and they are copied to the host.
//
// Results live in a ring of max_requests entries: ticket t uses entry
// t % max_requests, and an entry is reused once its result has been read
// (qcp_server_result / _wait for count + hash service, qcp_server_result_ids
// for ID-list service, or qcp_server_release).  Latencies (submit to the
// moment the dispatcher sees the batch's last event) go into a log-linear
// histogram (1/128 of an octave per bin), so statistics never rescan tickets.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
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
constexpr uint32_t kScanBlocks = 1024;         // block sums of the per-batch offset scan

// Latency histogram in integer nanoseconds: exact below 256 ns, then 128
// linear bins per power of two (relative width 1/128).
struct LatHist {
  static constexpr int kSub = 7;
  static constexpr int kBins = 256 + (64 - 8) * (1 << kSub);
  std::vector<uint64_t> bins = std::vector<uint64_t>(kBins, 0);
  uint64_t n = 0;
  double sum_us = 0.0, max_us = 0.0;
  static int bin_of(uint64_t ns) {
    if (ns < 256) return (int)ns;
    const int e = 63 - __builtin_clzll(ns);  // >= 8
    const int sub = (int)((ns >> (e - kSub)) & ((1u << kSub) - 1));
    return 256 + (e - 8) * (1 << kSub) + sub;
  }
  // midpoint of a bin, in microseconds
  static double value_of(int b) {
    if (b < 256) return b * 1e-3;
    const int e = (b - 256) / (1 << kSub) + 8;
    const int sub = (b - 256) % (1 << kSub);
    const double lo = (double)(((1ull << kSub) + sub) << (e - kSub));
    const double w = (double)(1ull << (e - kSub));
    return (lo + 0.5 * w) * 1e-3;
  }
  void add(double us) {
    const double ns = us * 1e3;
    const uint64_t v = ns <= 0 ? 0 : (ns >= 1.8e19 ? ~0ull >> 1 : (uint64_t)ns);
    bins[bin_of(v)]++;
    n++;
    sum_us += us;
    max_us = std::max(max_us, us);
  }
  double pct(double p) const {
    if (!n) return 0.0;
    const uint64_t rank = (uint64_t)std::max(1.0, std::ceil(p * (double)n));
    uint64_t c = 0;
    for (int b = 0; b < kBins; ++b) {
      c += bins[b];
      if (c >= rank) return std::min(value_of(b), max_us);
    }
    return max_us;
  }
};

enum : uint8_t { kFree = 0, kQueued = 1, kInFlight = 2, kDone = 3 };

// One result-table entry (ticket % capacity).
struct Entry {
  uint64_t ticket = 0;
  uint8_t state = kFree;
  float box[6] = {};
  double t_submit = 0.0, t_done = 0.0;
  uint32_t count = 0;
  unsigned long long hash = 0;
  // ID-list service: the batch's host block and this request's place in it
  std::shared_ptr<std::vector<uint32_t>> ids;
  uint64_t id_off = 0;
};

enum SlotPhase : uint8_t { kIdle = 0, kCounting = 1, kSized = 2, kWriting = 3 };

struct Slot {
  hipStream_t s = nullptr;
  float* d_q = nullptr;
  uint32_t* d_counts = nullptr;
  unsigned long long* d_hashes = nullptr;
  unsigned long long* d_offsets = nullptr;
  unsigned long long* d_scan = nullptr;
  uint32_t* d_ids = nullptr;
  uint64_t ids_cap = 0;
  QueueCtl* d_ctl = nullptr;
  unsigned long long* d_slots = nullptr;
  float* h_q = nullptr;
  uint32_t* h_counts = nullptr;
  unsigned long long* h_hashes = nullptr;
  unsigned long long* h_total = nullptr;
  uint32_t* h_ids = nullptr;
  uint64_t h_ids_cap = 0;
  hipEvent_t done = nullptr;
  std::vector<uint64_t> tickets;
  SlotPhase phase = kIdle;
  uint64_t seq = 0;        // batch number (sized batches are admitted in this order)
  uint64_t bytes = 0;      // result bytes charged to the budget
  uint64_t total_ids = 0;
};

}  // namespace

struct qcp_server {
  qcp_index* ix = nullptr;
  qcp_serve_cfg cfg{};
  bool ids = false;
  unsigned full_grid = 0;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_submit, cv_done;
  std::deque<uint64_t> queue;
  std::vector<Entry> ring;
  uint64_t next_ticket = 0;
  uint64_t next_seq = 0;
  std::vector<Slot> slots;
  uint64_t bytes_in_flight = 0;
  bool deferring = false;
  std::atomic<bool> stop{false};
  qcp_serve_stats st{};
  LatHist hist;
  qcp_status err = QCP_OK;
};

static void slot_free(Slot& s) {
  if (s.s) hipStreamSynchronize(s.s);
  for (void* p : {(void*)s.d_q, (void*)s.d_counts, (void*)s.d_hashes, (void*)s.d_offsets,
                  (void*)s.d_scan, (void*)s.d_ids, (void*)s.d_ctl, (void*)s.d_slots})
    if (p) hipFree(p);
  for (void* p : {(void*)s.h_q, (void*)s.h_counts, (void*)s.h_hashes, (void*)s.h_total,
                  (void*)s.h_ids})
    if (p) hipHostFree(p);
  if (s.done) hipEventDestroy(s.done);
  if (s.s) hipStreamDestroy(s.s);
  s = Slot();
}

static qcp_status slot_alloc(Slot& s, uint32_t D, uint32_t max_batch, bool ids) {
  const size_t qb = (size_t)2 * D * sizeof(float) * max_batch;
  if (hipStreamCreateWithFlags(&s.s, hipStreamNonBlocking) != hipSuccess) return QCP_EHIP;
  if (hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) return QCP_EHIP;
  if (hipMalloc((void**)&s.d_q, qb) != hipSuccess ||
      hipMalloc((void**)&s.d_counts, 4ull * max_batch) != hipSuccess ||
      hipMalloc((void**)&s.d_hashes, 8ull * max_batch) != hipSuccess ||
      hipMalloc((void**)&s.d_ctl, sizeof(QueueCtl)) != hipSuccess ||
      hipMalloc((void**)&s.d_slots, 8ull * kServerSlotCap) != hipSuccess)
    return QCP_ENOMEM;
  if (ids && (hipMalloc((void**)&s.d_offsets, 8ull * (max_batch + 1)) != hipSuccess ||
              hipMalloc((void**)&s.d_scan, 8ull * (kScanBlocks + 1)) != hipSuccess))
    return QCP_ENOMEM;
  if (hipHostMalloc((void**)&s.h_q, qb, 0) != hipSuccess ||
      hipHostMalloc((void**)&s.h_counts, 4ull * max_batch, 0) != hipSuccess ||
      hipHostMalloc((void**)&s.h_hashes, 8ull * max_batch, 0) != hipSuccess ||
      hipHostMalloc((void**)&s.h_total, 8, 0) != hipSuccess)
    return QCP_ENOMEM;
  s.tickets.reserve(max_batch);
  return QCP_OK;
}

static SearchArgs slot_args(const qcp_server* sv, const Slot& s, uint32_t n) {
  const qcp_index* ix = sv->ix;
  SearchArgs a = index_search_args(ix, s.d_q, n, s.d_counts, s.d_hashes);
  a.ctl = s.d_ctl;
  a.slots = s.d_slots;
  a.slot_cap = kServerSlotCap;
  return a;
}

// The count (+ hash) pass of one batch on the slot's stream, with donation;
// for ID-list service it always takes the containment shortcut (same counts
// and hashes) and ends with the offsets scan and a copy of the ID total.
static qcp_status slot_launch_count(qcp_server* sv, Slot& s, uint32_t n) {
  const qcp_index* ix = sv->ix;
  const uint32_t D = ix->D;
  hipStream_t st = s.s;
  if (hipMemcpyAsync(s.d_q, s.h_q, (size_t)2 * D * 4 * n, hipMemcpyHostToDevice, st) != hipSuccess)
    return QCP_EHIP;
  SearchArgs a = slot_args(sv, s, n);
  a.donate = 1;
  a.max_waiters = 64;
  a.hunger_period = 8;
  a.donate_min_height = 1;
  a.contain = ((sv->cfg.contain || sv->ids) && ix->contain_ok) ? 1u : 0u;
  if (launch_fill_u32(s.d_counts, n, 0u, st) != hipSuccess ||
      launch_fill_u32((uint32_t*)s.d_hashes, 2ull * n, 0u, st) != hipSuccess ||
      launch_fill_u32((uint32_t*)s.d_slots, 2ull * kServerSlotCap, 0xFFFFFFFFu, st) != hipSuccess)
    return QCP_EHIP;
  unsigned grid = std::max<unsigned>(16u, (n + 1) / 2);
  grid = std::min(grid, sv->full_grid);
  if (sv->cfg.grid_blocks) grid = std::min(grid, sv->cfg.grid_blocks);
  if (launch_search(a, D, ix->box, ix->expl, 1, 4, grid, 256, st) != hipSuccess) return QCP_EHIP;
  if (sv->ids) {
    if (scan_u32_u64(s.d_counts, n, s.d_offsets, s.d_scan, kScanBlocks, st) != hipSuccess ||
        hipMemcpyAsync(s.h_total, s.d_offsets + n, 8, hipMemcpyDeviceToHost, st) != hipSuccess)
      return QCP_EHIP;
  }
  if (hipMemcpyAsync(s.h_counts, s.d_counts, 4ull * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(s.h_hashes, s.d_hashes, 8ull * n, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipEventRecord(s.done, st) != hipSuccess)
    return QCP_EHIP;
  return QCP_OK;
}

// The ID pass of a sized batch: buffers grown to its total, sorted lists
// written at the scanned offsets, copied to the pinned host block.
static qcp_status slot_launch_ids(qcp_server* sv, Slot& s) {
  const qcp_index* ix = sv->ix;
  hipStream_t st = s.s;
  const uint32_t n = (uint32_t)s.tickets.size();
  const uint64_t total = s.total_ids;
  if (total + 64 > s.ids_cap) {  // grow geometrically (the stream is idle here)
    const uint64_t cap = std::max<uint64_t>(total + 64, 2 * s.ids_cap);
    if (s.d_ids) hipFree(s.d_ids);
    s.d_ids = nullptr;
    s.ids_cap = 0;
    if (hipMalloc((void**)&s.d_ids, 4 * cap) != hipSuccess) return QCP_ENOMEM;
    s.ids_cap = cap;
  }
  if (total > s.h_ids_cap) {
    const uint64_t cap = std::max<uint64_t>(std::max<uint64_t>(total, 1u << 20), 2 * s.h_ids_cap);
    if (s.h_ids) hipHostFree(s.h_ids);
    s.h_ids = nullptr;
    s.h_ids_cap = 0;
    if (hipHostMalloc((void**)&s.h_ids, 4 * cap, 0) != hipSuccess) return QCP_ENOMEM;
    s.h_ids_cap = cap;
  }
  if (total) {
    SearchArgs a = slot_args(sv, s, n);
    a.offsets = s.d_offsets;
    a.ids = s.d_ids;
    a.contain = (sv->cfg.contain && ix->contain_ok) ? 1u : 0u;
    unsigned grid = std::min<unsigned>(sv->full_grid, std::max<unsigned>(1u, (n + 3) / 4));
    if (sv->cfg.grid_blocks) grid = std::min(grid, sv->cfg.grid_blocks);
    if (launch_search(a, ix->D, ix->box, ix->expl, 2, 4, grid, 256, st) != hipSuccess)
      return QCP_EHIP;
    if (hipMemcpyAsync(s.h_ids, s.d_ids, 4 * total, hipMemcpyDeviceToHost, st) != hipSuccess)
      return QCP_EHIP;
  }
  if (hipEventRecord(s.done, st) != hipSuccess) return QCP_EHIP;
  return QCP_OK;
}

// Hand a finished batch's results to its tickets (caller holds the lock).
static void slot_retire_locked(qcp_server* sv, Slot& s, double t) {
  std::shared_ptr<std::vector<uint32_t>> block;
  if (sv->ids) {
    block = std::make_shared<std::vector<uint32_t>>(s.h_ids, s.h_ids + s.total_ids);
    sv->st.ids_served += s.total_ids;
  }
  const uint64_t R = sv->ring.size();
  uint64_t off = 0;
  for (size_t i = 0; i < s.tickets.size(); ++i) {
    Entry& e = sv->ring[s.tickets[i] % R];
    e.count = s.h_counts[i];
    e.hash = s.h_hashes[i];
    e.t_done = t;
    e.state = kDone;
    if (block) {
      e.ids = block;
      e.id_off = off;
      off += e.count;
    }
    sv->hist.add(t - e.t_submit);
  }
  sv->st.completed += s.tickets.size();
  sv->bytes_in_flight -= s.bytes;
  s.phase = kIdle;
  s.bytes = 0;
  s.tickets.clear();
}

static void set_err(qcp_server* sv, qcp_status rc) {
  std::lock_guard<std::mutex> g(sv->mu);
  if (sv->err == QCP_OK) sv->err = rc;
}

static void dispatcher(qcp_server* sv) {
  hipSetDevice(sv->ix->device);
  const uint32_t D = sv->ix->D;
  while (true) {
    bool did = false;
    // 1. events: retire finished batches; a finished count pass of ID-list
    //    service learns its ID total and waits for admission
    for (Slot& s : sv->slots) {
      if (s.phase != kCounting && s.phase != kWriting) continue;
      hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) set_err(sv, QCP_EHIP);
      did = true;
      if (sv->ids && s.phase == kCounting) {
        s.total_ids = *s.h_total;
        s.bytes = 4 * s.total_ids + 12ull * s.tickets.size();
        s.phase = kSized;
        continue;
      }
      const double t = now_us();
      {
        std::lock_guard<std::mutex> g(sv->mu);
        slot_retire_locked(sv, s, t);
      }
      sv->cv_done.notify_all();
    }
    // 2. ID-list service: admit sized batches, oldest first, while their
    //    result bytes fit the budget (a batch alone is always admitted)
    while (sv->ids) {
      Slot* next = nullptr;
      for (Slot& s : sv->slots)
        if (s.phase == kSized && (!next || s.seq < next->seq)) next = &s;
      if (!next) break;
      {
        std::lock_guard<std::mutex> g(sv->mu);
        if (sv->bytes_in_flight > 0 && sv->bytes_in_flight + next->bytes > sv->cfg.result_budget) {
          if (!sv->deferring) sv->st.deferred_by_budget++;
          sv->deferring = true;
          break;
        }
        sv->deferring = false;
        sv->bytes_in_flight += next->bytes;
        sv->st.bytes_peak = std::max<uint64_t>(sv->st.bytes_peak, sv->bytes_in_flight);
      }
      qcp_status rc = slot_launch_ids(sv, *next);
      if (rc != QCP_OK) set_err(sv, rc);
      next->phase = kWriting;
      did = true;
    }
    // 3. take waiting requests into free slots
    for (Slot& s : sv->slots) {
      if (s.phase != kIdle) continue;
      uint32_t n = 0;
      {
        std::lock_guard<std::mutex> g(sv->mu);
        if (sv->queue.empty()) break;
        n = (uint32_t)std::min<size_t>(sv->queue.size(), sv->cfg.max_batch);
        if (!sv->ids) {
          const uint64_t bytes = 12ull * n;  // count + hash per query
          if (sv->bytes_in_flight > 0 && sv->bytes_in_flight + bytes > sv->cfg.result_budget) {
            if (!sv->deferring) sv->st.deferred_by_budget++;
            sv->deferring = true;
            break;
          }
          sv->deferring = false;
          s.bytes = bytes;
          sv->bytes_in_flight += bytes;
          sv->st.bytes_peak = std::max<uint64_t>(sv->st.bytes_peak, sv->bytes_in_flight);
        }
        const uint64_t R = sv->ring.size();
        for (uint32_t i = 0; i < n; ++i) {
          const uint64_t tk = sv->queue.front();
          sv->queue.pop_front();
          Entry& e = sv->ring[tk % R];
          memcpy(s.h_q + (size_t)2 * D * i, e.box, sizeof(float) * 2 * D);
          e.state = kInFlight;
          s.tickets.push_back(tk);
        }
        s.seq = sv->next_seq++;
        sv->st.batches++;
        sv->st.max_batch_seen = std::max<uint64_t>(sv->st.max_batch_seen, n);
      }
      qcp_status rc = slot_launch_count(sv, s, n);
      if (rc != QCP_OK) set_err(sv, rc);
      s.phase = kCounting;
      did = true;
      uint64_t inflight = 0;
      for (Slot& x : sv->slots) inflight += x.phase != kIdle ? 1 : 0;
      std::lock_guard<std::mutex> g(sv->mu);
      sv->st.max_inflight_seen = std::max<uint64_t>(sv->st.max_inflight_seen, inflight);
    }
    if (!did) {
      bool any_busy = false;
      for (Slot& s : sv->slots) any_busy |= s.phase != kIdle;
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

// The entry of a ticket that is still in the table, or nullptr (caller holds the lock).
static Entry* entry_of(qcp_server* sv, uint64_t ticket) {
  if (ticket >= sv->next_ticket) return nullptr;
  Entry& e = sv->ring[ticket % sv->ring.size()];
  if (e.state == kFree || e.ticket != ticket) return nullptr;
  return &e;
}

static void retire_entry(Entry& e) {
  e.state = kFree;
  e.ids.reset();
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
  if (!cfg.max_requests) cfg.max_requests = 1u << 20;
  if (cfg.slots > 64 || cfg.max_batch > (1u << 24)) return QCP_EINVAL;
  if (cfg.outputs & ~(uint32_t)(QCP_OUT_COUNTS | QCP_OUT_HASH | QCP_OUT_IDS)) return QCP_EINVAL;
  if (hipSetDevice(ix->device) != hipSuccess) return QCP_EHIP;
  qcp_server* sv = new (std::nothrow) qcp_server();
  if (!sv) return QCP_ENOMEM;
  sv->ix = ix;
  sv->cfg = cfg;
  sv->ids = (cfg.outputs & QCP_OUT_IDS) != 0;
  sv->full_grid = persistent_grid(ix, 1, 256);
  try {
    sv->ring.resize(cfg.max_requests);
  } catch (...) {
    delete sv;
    return QCP_ENOMEM;
  }
  sv->slots.resize(cfg.slots);
  for (Slot& s : sv->slots) {
    qcp_status rc = slot_alloc(s, ix->D, cfg.max_batch, sv->ids);
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
    const uint64_t tk = sv->next_ticket;
    Entry& e = sv->ring[tk % sv->ring.size()];
    if (e.state != kFree) return QCP_EBUSY;  // the table is full of unread results
    sv->next_ticket++;
    e.ticket = tk;
    memcpy(e.box, box, sizeof(float) * 2 * D);
    e.t_submit = now_us();
    e.state = kQueued;
    e.ids.reset();
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
  Entry* e = entry_of(sv, ticket);
  if (!e) return QCP_EINVAL;
  if (sv->err != QCP_OK) return sv->err;
  if (e->state != kDone) return QCP_EBUSY;
  if (count) *count = e->count;
  if (hash) *hash = e->hash;
  if (latency_us) *latency_us = e->t_done - e->t_submit;
  if (!sv->ids) retire_entry(*e);  // an ID-list result stays until its IDs are read
  return QCP_OK;
}

qcp_status qcp_server_wait(qcp_server* sv, uint64_t ticket, uint32_t* count, uint64_t* hash,
                           double* latency_us) {
  if (!sv) return QCP_EINVAL;
  {
    std::unique_lock<std::mutex> lk(sv->mu);
    if (!entry_of(sv, ticket)) return QCP_EINVAL;
    sv->cv_done.wait(lk, [&] {
      Entry* e = entry_of(sv, ticket);
      return !e || e->state == kDone || sv->err != QCP_OK;
    });
  }
  return qcp_server_result(sv, ticket, count, hash, latency_us);
}

qcp_status qcp_server_result_ids(qcp_server* sv, uint64_t ticket, uint32_t* ids, uint64_t cap,
                                 uint64_t* n_ids) {
  if (!sv) return QCP_EINVAL;
  if (!sv->ids) return QCP_EINVAL;
  std::lock_guard<std::mutex> g(sv->mu);
  Entry* e = entry_of(sv, ticket);
  if (!e) return QCP_EINVAL;
  if (sv->err != QCP_OK) return sv->err;
  if (e->state != kDone) return QCP_EBUSY;
  if (n_ids) *n_ids = e->count;
  if (!ids) return QCP_OK;
  if (cap < e->count) return QCP_EBUSY;
  if (e->count) memcpy(ids, e->ids->data() + e->id_off, 4ull * e->count);
  retire_entry(*e);
  return QCP_OK;
}

qcp_status qcp_server_release(qcp_server* sv, uint64_t ticket) {
  if (!sv) return QCP_EINVAL;
  std::lock_guard<std::mutex> g(sv->mu);
  Entry* e = entry_of(sv, ticket);
  if (!e) return QCP_EINVAL;
  if (e->state != kDone) return QCP_EBUSY;
  retire_entry(*e);
  return QCP_OK;
}

qcp_status qcp_server_stats(qcp_server* sv, qcp_serve_stats* out) {
  if (!sv || !out) return QCP_EINVAL;
  std::lock_guard<std::mutex> g(sv->mu);
  *out = sv->st;
  const LatHist& h = sv->hist;
  if (h.n) {
    out->p50_us = h.pct(0.50);
    out->p99_us = h.pct(0.99);
    out->max_us = h.max_us;
    out->mean_us = h.sum_us / (double)h.n;
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

// Submits query i at start + arrival_s[i] from this thread while a consumer
// thread waits for the results in ticket order and reads them out (so the
// table only holds what is in flight); with ID-list service the lists go to
// ids at the prefix offsets of the counts (offsets[n+1]; ids == NULL: not
// copied, the count still is).
static qcp_status replay_impl(qcp_server* sv, const float* boxes, const double* arrival_s,
                              uint64_t n, uint32_t* counts, uint64_t* hashes, double* latency_us,
                              uint64_t* offsets, uint32_t* ids, uint64_t ids_cap) {
  const uint32_t D = sv->ix->D;
  std::vector<uint64_t> tk(n);
  std::atomic<uint64_t> submitted{0};
  std::atomic<bool> abort{false};
  qcp_status crc = QCP_OK;
  std::thread consumer([&] {
    uint64_t pos = 0;
    if (offsets) offsets[0] = 0;
    for (uint64_t i = 0; i < n; ++i) {
      while (submitted.load(std::memory_order_acquire) <= i) {
        if (abort.load()) return;
        std::this_thread::yield();
      }
      uint32_t c = 0;
      uint64_t h = 0;
      double l = 0;
      qcp_status rc = qcp_server_wait(sv, tk[i], &c, &h, &l);
      if (rc == QCP_OK && sv->ids) {
        if (ids && pos + c <= ids_cap) {
          rc = qcp_server_result_ids(sv, tk[i], ids + pos, c, nullptr);
        } else {
          if (ids && crc == QCP_OK) crc = QCP_EBUSY;  // caller's buffer too small
          rc = qcp_server_release(sv, tk[i]);
        }
      }
      if (rc != QCP_OK) {
        crc = rc;
        abort.store(true);
        return;
      }
      pos += c;
      if (counts) counts[i] = c;
      if (hashes) hashes[i] = h;
      if (latency_us) latency_us[i] = l;
      if (offsets) offsets[i + 1] = pos;
    }
  });
  qcp_status src = QCP_OK;
  const double t0 = now_us();
  for (uint64_t i = 0; i < n && !abort.load(); ++i) {
    const double target = t0 + arrival_s[i] * 1e6;
    while (now_us() < target) {
    }
    qcp_status rc;
    // a full table (results not yet read out) delays the submission
    while ((rc = qcp_server_submit(sv, boxes + (size_t)2 * D * i, &tk[i])) == QCP_EBUSY &&
           !abort.load())
      std::this_thread::yield();
    if (rc != QCP_OK) {
      src = rc;
      break;
    }
    submitted.store(i + 1, std::memory_order_release);
  }
  if (src != QCP_OK) abort.store(true);
  consumer.join();
  return src != QCP_OK ? src : crc;
}

qcp_status qcp_server_replay(qcp_server* sv, const float* boxes, const double* arrival_s,
                             uint64_t n, uint32_t* counts, uint64_t* hashes, double* latency_us) try {
  if (!sv || (n && (!boxes || !arrival_s))) return QCP_EINVAL;
  return replay_impl(sv, boxes, arrival_s, n, counts, hashes, latency_us, nullptr, nullptr, 0);
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

qcp_status qcp_server_replay_ids(qcp_server* sv, const float* boxes, const double* arrival_s,
                                 uint64_t n, uint32_t* counts, uint64_t* hashes,
                                 double* latency_us, uint64_t* offsets, uint32_t* ids,
                                 uint64_t ids_cap) try {
  if (!sv || (n && (!boxes || !arrival_s))) return QCP_EINVAL;
  if (!sv->ids) return QCP_EINVAL;
  return replay_impl(sv, boxes, arrival_s, n, counts, hashes, latency_us, offsets, ids, ids_cap);
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

}  // extern "C"
