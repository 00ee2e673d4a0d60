// noalloc.hip -- the persistent search with no device allocation: every buffer
// beyond the caller's inputs and outputs lives in one fixed array, `scratch`,
// of QCP_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends).
//
// Layout of the array (each part 256-byte aligned):
//   queue control block | donation slots | query keys x2 | query order x2 |
//   sort digit counts + tile counters | sort look-back words | key OR + flags
// The host never uses the array's name as an address: on the first call on a
// device, with that device current, hipGetSymbolAddress gives the device's
// copy, kept per device and passed to the kernels like any other buffer.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "../../include/qcp.h"
#include "qcp_internal.h"
#include "qcp_kernels.h"

namespace qcp {
namespace {

constexpr uint64_t kNaSlots = 1ull << 16;      // donation slots
constexpr uint64_t kNaMaxQ = QCP_NOALLOC_MAX_QUERIES;
constexpr uint64_t kNaOrderMin = 8192;          // smaller batches are searched as given
constexpr uint64_t kNaStatusWords = 2ull * ((kNaMaxQ + 4095) / 4096) * 256;

constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }
constexpr uint64_t kOffCtl = 0;
constexpr uint64_t kOffSlots = align256(kOffCtl + sizeof(QueueCtl));
constexpr uint64_t kOffKeys = align256(kOffSlots + 8 * kNaSlots);
constexpr uint64_t kOffKeys2 = align256(kOffKeys + 8 * kNaMaxQ);
constexpr uint64_t kOffPerm = align256(kOffKeys2 + 8 * kNaMaxQ);
constexpr uint64_t kOffPerm2 = align256(kOffPerm + 4 * kNaMaxQ);
constexpr uint64_t kOffGhist = align256(kOffPerm2 + 4 * kNaMaxQ);
constexpr uint64_t kOffStatus = align256(kOffGhist + 4 * kSortGhistWords);
constexpr uint64_t kOffOr = align256(kOffStatus + 8 * kNaStatusWords);
constexpr uint64_t kOffBad = kOffOr + 8;
constexpr uint64_t kScratchUsed = align256(kOffBad + 4);
static_assert(kScratchUsed <= QCP_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds its array");

constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[QCP_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};

// The device address of this device's copy of `scratch` (resolved once per
// device, with that device current).
qcp_status scratch_ptr(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return QCP_EINVAL;
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return QCP_EHIP;
    g_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_ptr[dev];
  return QCP_OK;
}

}  // namespace
}  // namespace qcp

using namespace qcp;

extern "C" qcp_status qcp_search_noalloc(qcp_index* ix, const float* d_queries, uint64_t q,
                                         const qcp_search_cfg* cfg_in, uint32_t* d_counts,
                                         uint64_t* d_hashes) try {
  if (!ix || (q && (!d_queries || !d_counts))) return QCP_EINVAL;
  qcp_search_cfg cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.mode = QCP_SEARCH_PERSISTENT;
  cfg.outputs = QCP_OUT_COUNTS | QCP_OUT_HASH;
  if (cfg_in) cfg = *cfg_in;
  if (cfg.mode != QCP_SEARCH_PERSISTENT || (cfg.outputs & (QCP_OUT_IDS | QCP_OUT_VISITS)))
    return QCP_EINVAL;
  const bool hash = (cfg.outputs & QCP_OUT_HASH) != 0;
  if (hash && q && !d_hashes) return QCP_EINVAL;
  if (q > kNaMaxQ) return QCP_EINVAL;  // beyond the fixed scratch array
  if (cfg.block_waves > 8) return QCP_EINVAL;
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) return QCP_ENODEV;
  if (hipSetDevice(ix->device) != hipSuccess) return QCP_EHIP;
  unsigned char* base = nullptr;
  const qcp_status ps = scratch_ptr(ix->device, &base);
  if (ps != QCP_OK) return ps;
  if (q == 0) return QCP_OK;
  hipStream_t s = ix->stream;
  QueueCtl* ctl = reinterpret_cast<QueueCtl*>(base + kOffCtl);
  unsigned long long* slots = reinterpret_cast<unsigned long long*>(base + kOffSlots);
  uint64_t* qkeys = reinterpret_cast<uint64_t*>(base + kOffKeys);
  uint64_t* qkeys2 = reinterpret_cast<uint64_t*>(base + kOffKeys2);
  uint32_t* qperm = reinterpret_cast<uint32_t*>(base + kOffPerm);
  uint32_t* qperm2 = reinterpret_cast<uint32_t*>(base + kOffPerm2);
  uint32_t* ghist = reinterpret_cast<uint32_t*>(base + kOffGhist);
  uint64_t* status = reinterpret_cast<uint64_t*>(base + kOffStatus);
  unsigned long long* qor = reinterpret_cast<unsigned long long*>(base + kOffOr);
  uint32_t* qbad = reinterpret_cast<uint32_t*>(base + kOffBad);
  if (sort_status_words(q) > kNaStatusWords) return QCP_EINVAL;

  // the same launch sequence and configuration rules as qcp_batch_run's
  // persistent form, with the scratch array's parts as its buffers
  const unsigned block = 64u * (cfg.block_waves ? cfg.block_waves : kDefaultBlockWaves);
  const int out_mode = hash ? 1 : 0;
  const int u = cfg.leaf_chunks ? (int)cfg.leaf_chunks : kDefaultLeafChunks;
  const bool donate = cfg.donate != 0;
  const unsigned wpb = block / 64u;
  unsigned grid = cfg.grid_blocks;
  if (!grid) {
    const unsigned full = persistent_grid(ix, out_mode, block);
    grid = donate ? full
                  : (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(full, (q + wpb - 1) / wpb));
  }
  SearchArgs a = index_search_args(ix, d_queries, q, d_counts,
                                   reinterpret_cast<unsigned long long*>(d_hashes));
  a.ctl = ctl;
  a.contain = (cfg.contain && ix->contain_ok) ? 1u : 0u;
  if (donate) {
    a.slots = slots;
    a.slot_cap = (uint32_t)kNaSlots;
    a.donate = 1;
    a.max_waiters = cfg.max_waiters ? cfg.max_waiters : 256;
    a.hunger_period = cfg.hunger_period ? cfg.hunger_period : 8;
    a.donate_min_height = cfg.donate_height ? cfg.donate_height : 1;
    // per-item results are accumulated atomically; donated slots start empty
    // (the array keeps the last call's contents)
    if (launch_fill_u32(d_counts, q, 0u, s) != hipSuccess ||
        (hash && launch_fill_u32((uint32_t*)d_hashes, 2 * q, 0u, s) != hipSuccess) ||
        launch_fill_u32((uint32_t*)slots, 2 * kNaSlots, 0xFFFFFFFFu, s) != hipSuccess)
      return QCP_EHIP;
  }
  if (cfg.order == 0 && q >= kNaOrderMin) {
    // spatial processing order: the 16-bit Hilbert key of each query box's
    // centre, radix-sorted with the query index (an even pass count, so the
    // order lands in qperm)
    const unsigned passes = (16 * ix->D + 7) / 8;
    if (passes & 1u) return QCP_EINVAL;
    if (launch_fill_u32((uint32_t*)qor, 2, 0u, s) != hipSuccess ||
        launch_fill_u32(qbad, 1, 0u, s) != hipSuccess ||
        launch_keys(d_queries, q, ix->D, true, false, true, 16, qkeys, nullptr, qor, qbad, s) !=
            hipSuccess ||
        launch_sort_keys(qkeys, qperm, qkeys2, qperm2, q, passes, nullptr, ghist, status, s) !=
            hipSuccess)
      return QCP_EHIP;
    a.qperm = qperm;
  }
  if (launch_search(a, ix->D, ix->box, ix->expl, out_mode, u, grid, block, s) != hipSuccess)
    return QCP_EHIP;
  if (hipStreamSynchronize(s) != hipSuccess) return QCP_EHIP;
  return QCP_OK;
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}
