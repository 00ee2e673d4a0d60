// qcp_api.cpp -- host runtime and C ABI of libqcp.so.
//
// An index handle owns its device (HBM) tree and a HIP stream; there are no
// device globals and no singletons (the reference keeps the one tree of the
// process in a ChunkManager singleton plus a __device__ pointer,
// src/manager/chunk_manager.cpp:11-50).  Every launch goes to the handle's
// stream; host synchronisation happens only where a size must be known on the
// host (sort pass count, ID buffer size) and at the end of a call.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <new>
#include <vector>

#include "../../include/qcp.h"
#include "bvh_topology.h"
#include "guttman.h"
#include "partition_topology.h"
#include "qcp_internal.h"
#include "qcp_kernels.h"

using namespace qcp;

#define HIP_TRY(expr)                       \
  do {                                      \
    hipError_t _e = (expr);                 \
    if (_e != hipSuccess) {                 \
      last_hip_error(_e, #expr, __LINE__);  \
      return QCP_EHIP;                      \
    }                                       \
  } while (0)

#define QCP_TRY(expr)                 \
  do {                                \
    qcp_status _s = (expr);           \
    if (_s != QCP_OK) return _s;      \
  } while (0)

static void last_hip_error(hipError_t e, const char* what, int line) {
  if (getenv("QCP_DEBUG"))
    fprintf(stderr, "qcp: HIP error %s (%d) at line %d: %s\n", hipGetErrorString(e), (int)e,
            line, what);
}

namespace {

constexpr uint64_t kObjSlack = 64 * 16;   // entries read past the end by masked chunks
constexpr uint64_t kNodeSlack = 256 + 64;
constexpr uint32_t kMaxScanBlocks = 4096;
constexpr uint32_t kSlotCap = 1u << 21;   // donated work items per launch
constexpr uint64_t kOrderMinQueries = 8192;  // smaller batches are searched as given

template <typename T>
qcp_status dalloc(T** p, uint64_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  if (hipMalloc((void**)p, sizeof(T) * count) != hipSuccess) {
    *p = nullptr;
    return QCP_ENOMEM;
  }
  return QCP_OK;
}
template <typename T>
void dfree(T*& p) {
  if (p) hipFree((void*)p);
  p = nullptr;
}

}  // namespace

static qcp_status set_device(int dev) {
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) return QCP_ENODEV;
  if (dev < 0 || dev >= cnt) return QCP_EINVAL;
  HIP_TRY(hipSetDevice(dev));
  return QCP_OK;
}

static void index_free_scratch(qcp_index* ix) {
  dfree(ix->d_keys2);
  dfree(ix->d_perm2);
  dfree(ix->d_carry);
  ix->carry_words = 0;
  dfree(ix->d_sort_ghist);
  dfree(ix->d_sort_status);
}

static void index_free(qcp_index* ix) {
  if (!ix) return;
  hipSetDevice(ix->device);
  if (ix->stream) hipStreamSynchronize(ix->stream);
  index_free_scratch(ix);
  dfree(ix->d_obj);
  dfree(ix->d_nbox);
  dfree(ix->d_nfirst);
  dfree(ix->d_ncount);
  dfree(ix->d_nrlo);
  dfree(ix->d_nrhi);
  dfree(ix->d_nparent);
  dfree(ix->d_keys);
  dfree(ix->d_perm);
  dfree(ix->d_key_or);
  dfree(ix->d_bad);
  dfree(ix->d_in);
  if (ix->build_exec) hipGraphExecDestroy(ix->build_exec);
  for (auto& e : ix->ev)
    if (e) hipEventDestroy(e);
  if (ix->stream) hipStreamDestroy(ix->stream);
  delete ix;
}

static qcp_status validate_cfg(const qcp_build_cfg* cfg) {
  if (!cfg) return QCP_EINVAL;
  if (cfg->dims != 2 && cfg->dims != 3) return QCP_EINVAL;
  if (cfg->fanout < 2 || cfg->fanout > 256) return QCP_EINVAL;
  if (cfg->layout > QCP_LAYOUT_GUTTMAN) return QCP_EINVAL;
  if (cfg->layout == QCP_LAYOUT_GUTTMAN && (cfg->flags & QCP_BUILD_GRAPH)) return QCP_EINVAL;
  // a minimum fill of 1 (F < 4) lets single-child chains grow without bound
  if (cfg->layout == QCP_LAYOUT_GUTTMAN && cfg->fanout < 4) return QCP_EINVAL;
  if (cfg->key_mode > QCP_KEY_B16 || cfg->key_src > QCP_SRC_CENTER) return QCP_EINVAL;
  if (cfg->obj_kind > QCP_OBJ_BOXES) return QCP_EINVAL;
  if (cfg->key_src == QCP_SRC_CENTER && cfg->obj_kind != QCP_OBJ_BOXES) return QCP_EINVAL;
  if (cfg->order > QCP_ORDER_KMEANS_HILBERT || cfg->clusters > QCP_MAX_CLUSTERS) return QCP_EINVAL;
  if (cfg->order == QCP_ORDER_KMEANS_HILBERT && (cfg->flags & QCP_BUILD_GRAPH)) return QCP_EINVAL;
  if (cfg->partitions > 1 && (cfg->layout != QCP_LAYOUT_MPHR || cfg->partitions > 65536))
    return QCP_EINVAL;  // partitions are MPHR's (src/tree/mphr.cpp:60-108)
  if (cfg->flags & ~(uint32_t)(QCP_BUILD_GRAPH | QCP_BUILD_NO_PERM | QCP_BUILD_SORT_INDEX |
                               QCP_BUILD_SORT_OBJECTS))
    return QCP_EINVAL;
  if ((cfg->flags & QCP_BUILD_SORT_INDEX) && (cfg->flags & QCP_BUILD_SORT_OBJECTS))
    return QCP_EINVAL;
  // the Guttman insertion and the k-means composite keys read the sorted
  // (key, input index) pairs on the host
  if ((cfg->flags & QCP_BUILD_SORT_OBJECTS) &&
      (cfg->layout == QCP_LAYOUT_GUTTMAN || cfg->order == QCP_ORDER_KMEANS_HILBERT))
    return QCP_EINVAL;
  return QCP_OK;
}

// Bit width of the input index in a packed sort element (indices < n).
static unsigned index_bits(uint64_t n) {
  unsigned pbits = 1;
  while (pbits < 32 && (1ull << pbits) < n) ++pbits;
  return pbits;
}

// First pass of the index form whose output packs the key bits still to be
// sorted with the input index into one 64-bit word (-1: none).
static int pack_pass_for(unsigned bits, unsigned pbits) {
  const unsigned passes = (bits + 7) / 8;
  for (unsigned j = 0; j + 1 < passes; ++j)
    if (bits - 8 * (j + 1) + pbits <= 64) return (int)j;
  return -1;
}

// Default choice between the two key-sort forms for `bits`-bit keys, by a
// model of their bytes a key at the rates the build's kernels reach on
// MI355X (one-sweep passes 3.0 TB/s in the index form and 3.4 TB/s in the
// object form, whose payload rounds overlap; the gather's raw HBM rate
// 5.6 TB/s; streamed leaf boxes 4.6 TB/s), which orders the two forms as
// measured on the 16M and 33.5M-point 3-D, 67M-box and 134M-point 2-D
// builds.
// Index form: pass 0 reads the key (the index is the position); a pass
// writes key + index (12 bytes), or 8 from the pass that packs them on, and
// the next reads what it wrote; then the gather reads the permutation and
// each object (a whole 128-byte line once the objects outgrow the 256 MiB
// Infinity Cache) and writes it, reducing the leaf boxes at fanout 32 / 64.
// Object form: a pass moves the remainder (8 bytes while more than 32 bits
// remain, else 4; the odd payload word's 4 bytes then ride in an 8-byte slot
// or alone) and the W = E (+1) payload words; pass 0 reads the key and the E
// object words; the last pass writes the W words alone; then the leaf boxes
// read the objects again.
static bool carry_cheaper(const qcp_index* ix, unsigned bits, bool keep_perm) {
  const unsigned passes = (bits + 7) / 8;
  if (passes == 0) return false;
  const double E4 = 4.0 * ix->E;
  const double W4 = E4 + (keep_perm ? 4.0 : 0.0);
  const int pk = pack_pass_for(bits, index_bits(ix->n));
  double idx_sort = 0.0, obj_sort = 0.0;
  for (unsigned p = 0; p < passes; ++p) {
    const bool packed_in = pk >= 0 && (int)p > pk, packed_out = pk >= 0 && (int)p >= pk;
    idx_sort += (p == 0 ? 8.0 : (packed_in ? 8.0 : 12.0)) + (packed_out ? 8.0 : 12.0);
    const bool last = p + 1 == passes;
    const double in = p == 0 ? 8.0 + E4 : (bits - 8 * p > 32 ? 8.0 : 4.0) + W4;
    const double out = last ? W4 : (bits - 8 * (p + 1) > 32 ? 8.0 : 4.0) + W4;
    obj_sort += in + out;
  }
  const bool big = (double)ix->n * E4 > 256.0 * 1048576.0;
  const double gather = 8.0 + (big ? 128.0 : E4) + E4;
  const bool fused = !ix->expl && (ix->F == 32 || ix->F == 64);
  const double t_index = idx_sort / 3.0 + gather / 5.6;
  const double t_object = obj_sort / 3.4 + (fused ? E4 / 4.6 : 0.0);
  return t_object < t_index;
}

// The key sort of a build over keys of `bits` bits: objects carried through
// the passes, or (key, input index) pairs and a gather.
static bool use_carry_sort(const qcp_index* ix, unsigned bits) {
  const qcp_build_cfg& c = ix->cfg;
  if (c.layout == QCP_LAYOUT_GUTTMAN || c.order == QCP_ORDER_KMEANS_HILBERT) return false;
  if (c.flags & QCP_BUILD_SORT_INDEX) return false;
  if (c.flags & QCP_BUILD_SORT_OBJECTS) return true;
  return carry_cheaper(ix, bits, !(c.flags & QCP_BUILD_NO_PERM));
}

// The object-carrying sort keeps the input indices unless the caller waived them.
static bool carry_keeps_perm(const qcp_index* ix) { return !(ix->cfg.flags & QCP_BUILD_NO_PERM); }

// Payload row stride of the object-carrying sort's scratch sets.
static uint64_t carry_stride(uint64_t n) { return ((n + 63) / 64) * 64; }

// Scratch of the object-carrying sort: two sets of E (+1) payload word arrays
// (the key remainders use d_keys2 and d_keys).
static qcp_status ensure_carry_scratch(qcp_index* ix) {
  const uint64_t W = ix->E + (carry_keeps_perm(ix) ? 1 : 0);
  const uint64_t words = 2 * W * carry_stride(ix->n > 0 ? ix->n : 1);
  if (ix->carry_words >= words) return QCP_OK;
  dfree(ix->d_carry);
  ix->carry_words = 0;
  QCP_TRY(dalloc(&ix->d_carry, words));
  ix->carry_words = words;
  return QCP_OK;
}

// The sort form of the next build over keys of `bits` bits, and its scratch
// (allocated before any graph capture of the build).
static qcp_status prepare_build_form(qcp_index* ix, unsigned bits) {
  ix->carry = use_carry_sort(ix, bits);
  ix->perm_valid = !ix->carry || carry_keeps_perm(ix);
  if (ix->carry && ix->n) QCP_TRY(ensure_carry_scratch(ix));
  return QCP_OK;
}

// The sort form of a build about to be captured: the captured sort plans its
// passes on the device for any key width, and its form is chosen here from
// the keys of the objects at capture time (one eager key pass); replays keep
// that form.
static qcp_status prepare_graph_form(qcp_index* ix, const float* d_objs) {
  if (ix->n == 0) return prepare_build_form(ix, 0);
  hipStream_t s = ix->stream;
  HIP_TRY(launch_fill_u32((uint32_t*)ix->d_key_or, 2, 0u, s));
  HIP_TRY(launch_fill_u32(ix->d_bad, 1, 0u, s));
  HIP_TRY(launch_keys(d_objs, ix->n, ix->D, ix->box, ix->cfg.key_mode == QCP_KEY_REF,
                      ix->cfg.key_src == QCP_SRC_CENTER, ix->nb, ix->d_keys, nullptr,
                      ix->d_key_or, ix->d_bad, s));
  ix->keys_stale = true;  // d_keys now holds the unsorted keys
  uint64_t key_or = 0;
  HIP_TRY(hipMemcpyAsync(&key_or, ix->d_key_or, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return prepare_build_form(ix, key_or ? 64u - (unsigned)__builtin_clzll(key_or) : 0u);
}

// Arguments of the object-carrying sort over the index's buffers: the keys
// are in d_keys; the key remainders alternate between d_keys2 (set 0) and
// d_keys (set 1, free once pass 0 has read the keys).
static CarryArgs carry_args(qcp_index* ix, const float* d_objs, unsigned bits,
                            const unsigned long long* key_or) {
  const bool idx = carry_keeps_perm(ix);
  const uint64_t cs = carry_stride(ix->n);
  CarryArgs a{};
  a.keys = ix->d_keys;
  a.objs = d_objs;
  a.krem[0] = ix->d_keys2;
  a.krem[1] = ix->d_keys;
  a.pay[0] = ix->d_carry;
  a.pay[1] = ix->d_carry + (ix->E + (idx ? 1 : 0)) * cs;
  a.cs = cs;
  a.obj = ix->d_obj;
  a.ostride = ix->obj_stride;
  a.perm = idx ? ix->d_perm : nullptr;
  a.n = ix->n;
  a.key_or = key_or;
  a.bits = bits;
  a.ghist = ix->d_sort_ghist;
  a.status = ix->d_sort_status;
  a.words = (uint64_t)sort_tiles(ix->n) * 256;
  return a;
}

// Level sizes of a packed tree, root first (Tree::GetLevelNodeCount,
// src/tree/tree.cpp:654-667); N == 1 is one leaf root, N == 0 no levels.
static std::vector<uint64_t> packed_levels(uint64_t n, uint32_t F) {
  std::vector<uint64_t> lv;
  if (n == 0) return lv;
  if (n == 1) return {1};
  uint64_t cur = n;
  while (cur > 1) {
    cur = cur / F + ((cur % F) ? 1 : 0);
    lv.insert(lv.begin(), cur);
  }
  return lv;
}

// Explicit layouts: adopt a BFS topology (first child / first position, count
// with kLeafBit on leaves, BFS offset of each depth).  The node arrays are
// (re)allocated when the node count outgrows them; the rank range under each
// node is derived here: leaves hold [first, first+count), a parent spans its
// first child's start to its last child's end (children follow their parent
// in BFS order, so one reverse pass suffices).
static qcp_status attach_topology(qcp_index* ix, std::vector<uint32_t>&& first,
                                  std::vector<uint32_t>&& count,
                                  const std::vector<uint64_t>& level_off) {
  const uint64_t m = first.size();
  if (count.size() != m || level_off.empty() || level_off.back() != m) return QCP_EINVAL;
  const uint64_t stride = ((m + kNodeSlack + 63) / 64) * 64;
  qcp_status s;
  if (!ix->d_nbox || stride > ix->node_stride) {
    for (uint32_t** a : {&ix->d_nfirst, &ix->d_ncount, &ix->d_nrlo, &ix->d_nrhi, &ix->d_nparent})
      dfree(*a);
    dfree(ix->d_nbox);
    ix->node_stride = stride;
    if ((s = dalloc(&ix->d_nbox, (uint64_t)2 * ix->D * stride)) != QCP_OK) return s;
    if (launch_fill_f32(ix->d_nbox, (uint64_t)2 * ix->D * stride, 0.0f, ix->stream) != hipSuccess)
      return QCP_EHIP;
    for (uint32_t** a : {&ix->d_nfirst, &ix->d_ncount, &ix->d_nrlo, &ix->d_nrhi, &ix->d_nparent}) {
      if ((s = dalloc(a, stride)) != QCP_OK) return s;
      if (launch_fill_u32(*a, stride, 0u, ix->stream) != hipSuccess) return QCP_EHIP;
    }
  }
  std::vector<uint32_t> rlo(m), rhi(m), parent(m);
  for (uint64_t i = m; i-- > 0;) {
    parent[i] = (uint32_t)i;  // a root stays its own parent
    if (count[i] & kLeafBit) {
      rlo[i] = first[i];
      rhi[i] = first[i] + (count[i] & ~kLeafBit);
    } else {
      rlo[i] = rlo[first[i]];
      rhi[i] = rhi[first[i] + count[i] - 1];
    }
  }
  for (uint64_t i = 0; i < m; ++i)
    if (!(count[i] & kLeafBit))
      for (uint32_t k = 0; k < count[i]; ++k) parent[first[i] + k] = (uint32_t)i;
  uint32_t* dst[5] = {ix->d_nfirst, ix->d_ncount, ix->d_nrlo, ix->d_nrhi, ix->d_nparent};
  const uint32_t* src[5] = {first.data(), count.data(), rlo.data(), rhi.data(), parent.data()};
  for (int i = 0; i < 5 && m; ++i)
    HIP_TRY(hipMemcpyAsync(dst[i], src[i], 4 * m, hipMemcpyHostToDevice, ix->stream));
  HIP_TRY(hipStreamSynchronize(ix->stream));
  ix->n_nodes = m;
  ix->h = (int)level_off.size() - 1;
  ix->level_off.assign(level_off.begin(), level_off.end());
  ix->level_count.clear();
  for (int l = 0; l < ix->h; ++l) ix->level_count.push_back(level_off[l + 1] - level_off[l]);
  ix->h_first = std::move(first);
  ix->h_count = std::move(count);
  ix->h_nbox.clear();
  return QCP_OK;
}

// Allocate the index and its topology for n objects.
static qcp_status index_alloc(const qcp_build_cfg* cfg, uint64_t n, qcp_index** out) {
  QCP_TRY(validate_cfg(cfg));
  if (n >= 0xFFFFFE00ull) return QCP_EINVAL;  // ranks (+ one fanout) and node ids are 32-bit
  if (cfg->partitions > 1 && n > 0 && cfg->partitions > n) return QCP_EINVAL;  // an empty partition
  QCP_TRY(set_device((int)cfg->device));
  qcp_index* ix = new (std::nothrow) qcp_index();
  if (!ix) return QCP_ENOMEM;
  ix->device = (int)cfg->device;
  ix->cfg = *cfg;
  ix->D = cfg->dims;
  ix->F = cfg->fanout;
  ix->box = cfg->obj_kind == QCP_OBJ_BOXES;
  ix->E = ix->box ? 2 * ix->D : ix->D;
  // MPHR data partitions (an empty index is one empty tree)
  ix->parts = (cfg->partitions > 1 && n > 0) ? cfg->partitions : 1u;
  ix->expl = cfg->layout == QCP_LAYOUT_BVH || cfg->layout == QCP_LAYOUT_GUTTMAN || ix->parts > 1;
  ix->nb = cfg->key_mode == QCP_KEY_REF ? (ix->D > 2 ? 20u : 31u) : 16u;
  ix->n = n;
  auto fail = [&](qcp_status s) {
    index_free(ix);
    return s;
  };
  if (hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(QCP_EHIP);
  for (auto& e : ix->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(QCP_EHIP);
  hipDeviceGetAttribute(&ix->cus, hipDeviceAttributeMultiprocessorCount, ix->device);
  if (ix->cus <= 0) ix->cus = 256;

  qcp_status s;
  if (!ix->expl) {
    ix->level_count = packed_levels(n, ix->F);
    ix->h = (int)ix->level_count.size();
    if (ix->h > kMaxLevels) return fail(QCP_EINVAL);
    ix->level_off.assign(ix->h + 1, 0);
    for (int l = 0; l < ix->h; ++l) ix->level_off[l + 1] = ix->level_off[l] + ix->level_count[l];
    ix->n_nodes = ix->level_off[ix->h];
    ix->node_stride = ((ix->n_nodes + kNodeSlack + 63) / 64) * 64;
    if ((s = dalloc(&ix->d_nbox, (uint64_t)2 * ix->D * ix->node_stride)) != QCP_OK) return fail(s);
    if (launch_fill_f32(ix->d_nbox, (uint64_t)2 * ix->D * ix->node_stride, 0.0f, ix->stream) != hipSuccess)
      return fail(QCP_EHIP);
  } else if (ix->parts > 1) {
    PartitionTopology t;
    if (!partition_topology(n, ix->F, ix->parts, kMaxDepth, &t)) return fail(QCP_EINVAL);
    if ((s = attach_topology(ix, std::move(t.first), std::move(t.count), t.level_off)) != QCP_OK)
      return fail(s);
  } else if (cfg->layout == QCP_LAYOUT_BVH) {
    BvhTopology t;
    if (!bvh_topology(n, ix->F, kMaxDepth, &t)) return fail(QCP_EINVAL);
    std::vector<uint64_t> lo(t.level_off.begin(), t.level_off.end());
    if ((s = attach_topology(ix, std::move(t.first), std::move(t.count), lo)) != QCP_OK) return fail(s);
  } else {
    // Guttman: the topology follows from the objects; empty until the build
    if ((s = attach_topology(ix, {}, {}, {0})) != QCP_OK) return fail(s);
  }
  ix->obj_stride = ((n + kObjSlack + 63) / 64) * 64;
  if ((s = dalloc(&ix->d_obj, (uint64_t)ix->E * ix->obj_stride)) != QCP_OK) return fail(s);
  if ((s = dalloc(&ix->d_keys, n)) != QCP_OK) return fail(s);
  if ((s = dalloc(&ix->d_perm, n)) != QCP_OK) return fail(s);
  if ((s = dalloc(&ix->d_key_or, 1)) != QCP_OK) return fail(s);
  if ((s = dalloc(&ix->d_bad, 1)) != QCP_OK) return fail(s);
  if (launch_fill_f32(ix->d_obj, (uint64_t)ix->E * ix->obj_stride, 0.0f, ix->stream) != hipSuccess)
    return fail(QCP_EHIP);
  // sort scratch
  if ((s = dalloc(&ix->d_keys2, n)) != QCP_OK) return fail(s);
  if ((s = dalloc(&ix->d_perm2, n)) != QCP_OK) return fail(s);
  if ((s = dalloc(&ix->d_sort_ghist, kSortGhistWords)) != QCP_OK) return fail(s);
  if ((s = dalloc(&ix->d_sort_status, sort_status_words(n))) != QCP_OK) return fail(s);
  *out = ix;
  return QCP_OK;
}

// Shape fields of the index info (timings cleared).
static void set_info_shape(qcp_index* ix) {
  qcp_index_info& in = ix->info;
  memset(&in, 0, sizeof(in));
  in.n = ix->n;
  in.n_nodes = ix->n_nodes;
  in.height = (uint32_t)ix->h;
  in.dims = ix->D;
  in.fanout = ix->F;
  in.layout = ix->cfg.layout;
  in.obj_kind = ix->cfg.obj_kind;
  in.key_mode = ix->cfg.key_mode;
  in.key_src = ix->cfg.key_src;
  in.key_bits = ix->nb * ix->D;
  for (int l = 0; l < ix->h && l < 32; ++l) in.level_count[l] = ix->level_count[l];
  in.tree_bytes = (uint64_t)ix->E * 4 * ix->n + (uint64_t)2 * ix->D * 4 * ix->n_nodes;
  in.partitions = ix->parts;
  in.kmeans_iterations = ix->kmeans_its;
  in.kmeans_ms[0] = ix->kmeans_ms[0];
  in.kmeans_ms[1] = ix->kmeans_ms[1];
  in.sort_form = ix->carry ? 1u : 0u;
  in.has_perm = ix->perm_valid ? 1u : 0u;
}

// Guttman layout: insert the objects on the host in their sorted order
// (sobjs: n rows of E floats, perm / keys of each sorted position), adopt the
// tree, and store perm / keys in depth-first order.  *host_ms: insertion wall
// time.
static qcp_status guttman_adopt(qcp_index* ix, const float* sobjs, const uint32_t* perm,
                                const uint64_t* keys, double* host_ms) {
  const uint64_t n = ix->n;
  const auto t0 = std::chrono::steady_clock::now();
  GuttmanTree t;
  if (!guttman_build(sobjs, n, ix->D, ix->box, ix->F, &t)) return QCP_EINVAL;
  if (t.level_off.size() > (size_t)kMaxDepth) return QCP_EINVAL;  // deeper than a wave's stack
  std::vector<uint32_t> perm2(n);
  std::vector<uint64_t> keys2(n);
  for (uint64_t p = 0; p < n; ++p) {
    perm2[p] = perm[t.order[p]];
    keys2[p] = keys[t.order[p]];
  }
  *host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (n) {
    HIP_TRY(hipMemcpyAsync(ix->d_perm, perm2.data(), 4 * n, hipMemcpyHostToDevice, ix->stream));
    HIP_TRY(hipMemcpyAsync(ix->d_keys, keys2.data(), 8 * n, hipMemcpyHostToDevice, ix->stream));
  }
  return attach_topology(ix, std::move(t.first), std::move(t.count), t.level_off);
}

// The same from the device state after the key sort (d_perm / d_keys sorted).
static qcp_status guttman_from_sorted(qcp_index* ix, const float* d_objs, double* host_ms) {
  const uint64_t n = ix->n;
  const uint32_t E = ix->E;
  hipStream_t s = ix->stream;
  std::vector<float> objs((size_t)E * n), sobjs((size_t)E * n);
  std::vector<uint32_t> perm(n);
  std::vector<uint64_t> keys(n);
  HIP_TRY(hipMemcpyAsync(objs.data(), d_objs, 4ull * E * n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(perm.data(), ix->d_perm, 4 * n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(keys.data(), ix->d_keys, 8 * n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (uint64_t r = 0; r < n; ++r)
    memcpy(sobjs.data() + r * E, objs.data() + (uint64_t)perm[r] * E, 4ull * E);
  return guttman_adopt(ix, sobjs.data(), perm.data(), keys.data(), host_ms);
}

// Node boxes of the rank-ordered objects in d_obj: bottom-up over the packed
// levels, or level by level of an explicit tree.
// Fanout whose leaf boxes the build's gather writes itself (0: none).
static uint32_t fused_leaf_fanout(const qcp_index* ix) {
  return (!ix->expl && ix->h >= 1 && (ix->F == 32 || ix->F == 64)) ? ix->F : 0;
}

// The build's gather (see launch_gather_objs) into the index's arrays.
static hipError_t enqueue_gather(qcp_index* ix, const float* d_objs, const uint64_t* packed,
                                 unsigned pbits) {
  const uint32_t fl = fused_leaf_fanout(ix);
  return launch_gather_objs(d_objs, packed ? nullptr : ix->d_perm, packed, ix->n, pbits, ix->D,
                            ix->box, ix->d_obj, ix->obj_stride, packed ? ix->d_perm : nullptr,
                            fl, ix->d_nbox, ix->node_stride, fl ? ix->level_off[ix->h - 1] : 0,
                            ix->stream);
}

// leaves_done: the gather already wrote the leaf level (fused_leaf_fanout).
static qcp_status enqueue_boxes(qcp_index* ix, bool leaves_done = false) {
  hipStream_t s = ix->stream;
  if (!ix->expl) {
    const int h = ix->h;
    if (!leaves_done)
      HIP_TRY(launch_leaf_boxes(ix->d_obj, ix->obj_stride, ix->n, ix->D, ix->box, ix->F,
                                ix->level_count[h - 1], ix->d_nbox, ix->node_stride,
                                ix->level_off[h - 1], s));
    for (int l = h - 2; l >= 0; --l)
      HIP_TRY(launch_level_boxes(ix->d_nbox, ix->node_stride, ix->D, ix->level_off[l + 1],
                                 ix->level_count[l + 1], ix->F, ix->level_off[l],
                                 ix->level_count[l], s));
  } else {
    for (int l = ix->h - 1; l >= 0; --l)
      HIP_TRY(launch_tree_boxes(ix->d_obj, ix->obj_stride, ix->d_nfirst, ix->d_ncount, ix->D,
                                ix->box, ix->F, ix->cfg.layout == QCP_LAYOUT_GUTTMAN ? 0 : ix->F,
                                ix->level_off[l], ix->level_off[l + 1], ix->d_nbox,
                                ix->node_stride, s));
  }
  return QCP_OK;
}

// Sort passes of a graph-captured build: every byte a key can hold (an even
// count for every mode).  Passes above the keys' actual width end at once on
// the device, so the result lies where the last reordering pass left it:
// d_keys/d_perm after an even number of reordering passes, d_keys2/d_perm2
// after an odd number, packed or not (launch_gather_objs_planned reads the
// same plan from the OR of the keys and finds it).
static unsigned graph_sort_passes(const qcp_index* ix) { return (ix->D * ix->nb + 7) / 8; }

// The whole build as one asynchronous sequence (no host synchronisation), for
// hipGraph capture: keys, a fixed number of sort passes, gather, node boxes.
static qcp_status enqueue_build(qcp_index* ix, const float* d_objs, unsigned passes) {
  hipStream_t s = ix->stream;
  const uint64_t n = ix->n;
  // the sort's plan (passes that reorder, packing) is made on the device, so
  // d_keys may end packed: exports recompute the keys from the objects (the
  // graph launch sites set keys_stale again for every replay)
  ix->keys_stale = true;
  HIP_TRY(launch_fill_u32((uint32_t*)ix->d_key_or, 2, 0u, s));
  HIP_TRY(launch_fill_u32(ix->d_bad, 1, 0u, s));
  // the key kernel also counts the sort's digits (graph builds have no
  // k-means keys)
  HIP_TRY(launch_fill_u32(ix->d_sort_ghist, kSortGhistWords, 0u, s));
  HIP_TRY(launch_keys(d_objs, n, ix->D, ix->box, ix->cfg.key_mode == QCP_KEY_REF,
                      ix->cfg.key_src == QCP_SRC_CENTER, ix->nb, ix->d_keys, nullptr,
                      ix->d_key_or, ix->d_bad, s, ix->d_sort_ghist, ix->d_sort_status,
                      (uint64_t)sort_tiles(n) * 256));
  if (ix->carry) {
    // every pass a key can need; the last one that reorders writes the index
    const unsigned cpasses = (ix->D * ix->nb + 7) / 8;
    HIP_TRY(launch_sort_carry(carry_args(ix, d_objs, 0, ix->d_key_or), ix->D, ix->box,
                              carry_keeps_perm(ix), cpasses, s, true));
    HIP_TRY(launch_carry_identity(d_objs, n, ix->d_key_or, ix->D, ix->box, ix->d_obj,
                                  ix->obj_stride, carry_keeps_perm(ix) ? ix->d_perm : nullptr, s));
    return enqueue_boxes(ix, false);
  }
  if (passes & 1u) return QCP_EINVAL;  // odd pass counts are not captured
  const unsigned pbits = index_bits(n);
  HIP_TRY(launch_sort_keys(ix->d_keys, ix->d_perm, ix->d_keys2, ix->d_perm2, n, passes,
                           ix->d_key_or, ix->d_sort_ghist, ix->d_sort_status, s, -1, pbits,
                           true));
  const uint32_t fl = fused_leaf_fanout(ix);
  HIP_TRY(launch_gather_objs_planned(d_objs, ix->d_perm, ix->d_perm2, ix->d_keys, ix->d_keys2,
                                     n, pbits, passes, ix->d_key_or, ix->D, ix->box, ix->d_obj,
                                     ix->obj_stride, fl, ix->d_nbox, ix->node_stride,
                                     fl ? ix->level_off[ix->h - 1] : 0, s));
  return enqueue_boxes(ix, fl != 0);
}

// Graph-captured build: captured once per input pointer, replayed after.
static qcp_status index_run_build_graph(qcp_index* ix, const float* d_objs) {
  hipStream_t s = ix->stream;
  ix->h_nbox.clear();
  if (ix->n == 0) return QCP_OK;
  if (!ix->build_exec || ix->build_exec_objs != d_objs) QCP_TRY(prepare_graph_form(ix, d_objs));
  if (!ix->build_exec || ix->build_exec_objs != d_objs) {
    if (ix->build_exec) {
      hipGraphExecDestroy(ix->build_exec);
      ix->build_exec = nullptr;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    qcp_status st = enqueue_build(ix, d_objs, graph_sort_passes(ix));
    hipError_t e = hipStreamEndCapture(s, &g);
    if (st != QCP_OK) {
      if (g) hipGraphDestroy(g);
      return st;
    }
    HIP_TRY(e);
    e = hipGraphInstantiate(&ix->build_exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    HIP_TRY(e);
    ix->build_exec_objs = d_objs;
  }
  HIP_TRY(hipEventRecord(ix->ev[0], s));
  HIP_TRY(hipGraphLaunch(ix->build_exec, s));
  ix->keys_stale = true;  // each replay may leave packed elements in d_keys
  HIP_TRY(hipEventRecord(ix->ev[4], s));
  uint32_t bad = 0;
  HIP_TRY(hipMemcpyAsync(&bad, ix->d_bad, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  float tot = 0;
  hipEventElapsedTime(&tot, ix->ev[0], ix->ev[4]);
  set_info_shape(ix);
  ix->info.build_ms = tot;
  ix->info.build_phase_ms[5] = tot;
  ix->contain_ok = !(bad & kBadInverted);
  return (bad & kBadDomain) ? QCP_EDOMAIN : QCP_OK;
}

// The build pipeline on the index's stream.
static qcp_status index_run_build(qcp_index* ix, const float* d_objs) {
  hipStream_t s = ix->stream;
  ix->h_nbox.clear();
  const uint64_t n = ix->n;
  const bool guttman = ix->cfg.layout == QCP_LAYOUT_GUTTMAN;
  double host_ms = 0.0;
  HIP_TRY(hipEventRecord(ix->ev[0], s));
  uint64_t key_or = 0;
  uint32_t bad = 0;
  if (n > 0) {
    ix->keys_stale = false;
    // the key kernel also counts the sort's digits, unless k-means replaces
    // the keys it writes
    const bool counted = ix->cfg.order != QCP_ORDER_KMEANS_HILBERT;
    HIP_TRY(launch_fill_u32((uint32_t*)ix->d_key_or, 2, 0u, s));
    HIP_TRY(launch_fill_u32(ix->d_bad, 1, 0u, s));
    if (counted) HIP_TRY(launch_fill_u32(ix->d_sort_ghist, kSortGhistWords, 0u, s));
    HIP_TRY(launch_keys(d_objs, n, ix->D, ix->box, ix->cfg.key_mode == QCP_KEY_REF,
                        ix->cfg.key_src == QCP_SRC_CENTER, ix->nb, ix->d_keys, nullptr,
                        ix->d_key_or, ix->d_bad, s, counted ? ix->d_sort_ghist : nullptr,
                        counted ? ix->d_sort_status : nullptr,
                        counted ? (uint64_t)sort_tiles(n) * 256 : 0));
    HIP_TRY(hipMemcpyAsync(&key_or, ix->d_key_or, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&bad, ix->d_bad, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipEventRecord(ix->ev[1], s));
    HIP_TRY(hipStreamSynchronize(s));
    if (bad & kBadDomain) return QCP_EDOMAIN;
    ix->contain_ok = !(bad & kBadInverted);
    // LSD radix sort over the significant bytes of the keys
    unsigned bits = key_or ? 64 - __builtin_clzll(key_or) : 0;
    if (ix->cfg.order == QCP_ORDER_KMEANS_HILBERT) {
      uint64_t kb = 0;
      QCP_TRY(kmeans_keys(ix, d_objs, &kb));
      bits = (unsigned)kb;
    }
    const unsigned passes = (bits + 7) / 8;
    QCP_TRY(prepare_build_form(ix, bits));
    if (ix->carry) {
      // objects carried through the passes: the last pass writes the index
      uint32_t* perm_out = carry_keeps_perm(ix) ? ix->d_perm : nullptr;
      if (passes == 0) {  // all keys equal: the input order is the rank order
        HIP_TRY(launch_gather_objs(d_objs, nullptr, nullptr, n, 0, ix->D, ix->box, ix->d_obj,
                                   ix->obj_stride, perm_out, 0, nullptr, 0, 0, s));
      } else {
        HIP_TRY(launch_sort_carry(carry_args(ix, d_objs, bits, nullptr), ix->D, ix->box,
                                  perm_out != nullptr, passes, s, counted));
      }
      ix->keys_stale = true;  // d_keys holds key remainders
      HIP_TRY(hipEventRecord(ix->ev[2], s));
      HIP_TRY(hipEventRecord(ix->ev[3], s));
      QCP_TRY(enqueue_boxes(ix, false));
    } else {
    // Packed sort elements from the first pass after which the key bits still
    // to be sorted and the input index fit 64 bits; the gather then recomputes
    // the keys from the objects.  Not for k-means keys (not a function of the
    // object alone) or Guttman builds (their host step reads the sorted keys).
    const unsigned pbits = index_bits(n);
    const int pack_pass = (!guttman && ix->cfg.order != QCP_ORDER_KMEANS_HILBERT)
                              ? pack_pass_for(bits, pbits) : -1;
    HIP_TRY(launch_sort_keys(ix->d_keys, ix->d_perm, ix->d_keys2, ix->d_perm2, n, passes,
                             nullptr, ix->d_sort_ghist, ix->d_sort_status, s, pack_pass, pbits,
                             counted));
    if (passes == 0) HIP_TRY(launch_iota_u32(ix->d_perm, n, s));  // all keys equal
    if (pack_pass < 0 && (passes & 1u)) {
      std::swap(ix->d_keys, ix->d_keys2);
      std::swap(ix->d_perm, ix->d_perm2);
    }
    HIP_TRY(hipEventRecord(ix->ev[2], s));
    if (guttman) {
      QCP_TRY(guttman_from_sorted(ix, d_objs, &host_ms));
      HIP_TRY(hipEventRecord(ix->ev[6], s));
    }
    if (pack_pass >= 0) {
      const uint64_t* packed = (passes & 1u) ? ix->d_keys2 : ix->d_keys;
      HIP_TRY(enqueue_gather(ix, d_objs, packed, pbits));
      ix->keys_stale = true;
    } else {
      HIP_TRY(enqueue_gather(ix, d_objs, nullptr, 0));
    }
    HIP_TRY(hipEventRecord(ix->ev[3], s));
    QCP_TRY(enqueue_boxes(ix, fused_leaf_fanout(ix) != 0));
    }
  } else {
    if (guttman) QCP_TRY(guttman_adopt(ix, nullptr, nullptr, nullptr, &host_ms));
    HIP_TRY(hipEventRecord(ix->ev[1], s));
    HIP_TRY(hipEventRecord(ix->ev[2], s));
    HIP_TRY(hipEventRecord(ix->ev[3], s));
  }
  HIP_TRY(hipEventRecord(ix->ev[4], s));
  HIP_TRY(hipStreamSynchronize(s));
  float ms[4] = {0, 0, 0, 0}, tot = 0;
  hipEventElapsedTime(&ms[0], ix->ev[0], ix->ev[1]);
  hipEventElapsedTime(&ms[1], ix->ev[1], ix->ev[2]);
  hipEventElapsedTime(&ms[2], ix->ev[guttman && n ? 6 : 2], ix->ev[3]);
  hipEventElapsedTime(&ms[3], ix->ev[3], ix->ev[4]);
  hipEventElapsedTime(&tot, ix->ev[0], ix->ev[4]);
  set_info_shape(ix);
  qcp_index_info& in = ix->info;
  in.build_ms = tot;
  for (int i = 0; i < 4; ++i) in.build_phase_ms[i] = ms[i];
  in.build_phase_ms[4] = host_ms;
  in.build_phase_ms[5] = tot;
  return QCP_OK;
}

extern "C" {

const char* qcp_strerror(qcp_status s) {
  switch (s) {
    case QCP_OK: return "ok";
    case QCP_EINVAL: return "invalid argument";
    case QCP_ENOMEM: return "out of memory";
    case QCP_EHIP: return "HIP runtime error";
    case QCP_EIO: return "I/O error";
    case QCP_EFORMAT: return "malformed index file";
    case QCP_EBUSY: return "buffer too small or admission refused";
    case QCP_EDOMAIN: return "object coordinate outside the key quantiser's domain";
    case QCP_ENODEV: return "no GPU";
  }
  return "unknown status";
}

int qcp_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

qcp_status qcp_build_device(const qcp_build_cfg* cfg, const float* d_objs, uint64_t n,
                            qcp_index** out) try {
  if (!out || (n && !d_objs)) return QCP_EINVAL;
  *out = nullptr;
  qcp_index* ix = nullptr;
  QCP_TRY(index_alloc(cfg, n, &ix));
  qcp_status s = (cfg->flags & QCP_BUILD_GRAPH) ? index_run_build_graph(ix, d_objs) : index_run_build(ix, d_objs);
  if (s != QCP_OK) {
    index_free(ix);
    return s;
  }
  *out = ix;
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

qcp_status qcp_build_ranked(const qcp_build_cfg* cfg, const float* objs, uint64_t n,
                            qcp_index** out) try {
  if (!out || (n && !objs)) return QCP_EINVAL;
  *out = nullptr;
  qcp_index* ix = nullptr;
  QCP_TRY(index_alloc(cfg, n, &ix));
  auto fail = [&](qcp_status st) {
    index_free(ix);
    return st;
  };
  hipStream_t s = ix->stream;
  for (uint64_t i = 0; i < (uint64_t)ix->E * n; ++i)
    if (objs[i] != objs[i]) return fail(QCP_EDOMAIN);  // NaN objects are rejected
  double host_ms = 0.0;
  HIP_TRY(hipEventRecord(ix->ev[0], s));
  if (n) {
    // the given order is the rank order: keys and permutation are the identity
    std::vector<uint64_t> keys(n);
    std::vector<uint32_t> perm(n);
    for (uint64_t i = 0; i < n; ++i) {
      keys[i] = i;
      perm[i] = (uint32_t)i;
    }
    qcp_status st = dalloc(&ix->d_in, (uint64_t)ix->E * n);
    if (st != QCP_OK) return fail(st);
    if (ix->cfg.layout == QCP_LAYOUT_GUTTMAN) {
      // the given order is the insertion order
      if ((st = guttman_adopt(ix, objs, perm.data(), keys.data(), &host_ms)) != QCP_OK) return fail(st);
    } else if (hipMemcpyAsync(ix->d_keys, keys.data(), 8 * n, hipMemcpyHostToDevice, s) != hipSuccess ||
               hipMemcpyAsync(ix->d_perm, perm.data(), 4 * n, hipMemcpyHostToDevice, s) != hipSuccess) {
      return fail(QCP_EHIP);
    }
    if (hipMemcpyAsync(ix->d_in, objs, (uint64_t)ix->E * 4 * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        enqueue_gather(ix, ix->d_in, nullptr, 0) != hipSuccess)
      return fail(QCP_EHIP);
    if ((st = enqueue_boxes(ix, fused_leaf_fanout(ix) != 0)) != QCP_OK) return fail(st);
    if (hipStreamSynchronize(s) != hipSuccess) return fail(QCP_EHIP);
    dfree(ix->d_in);
  }
  if (hipEventRecord(ix->ev[4], s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail(QCP_EHIP);
  float tot = 0;
  hipEventElapsedTime(&tot, ix->ev[0], ix->ev[4]);
  set_info_shape(ix);
  ix->info.build_ms = tot;
  ix->info.build_phase_ms[4] = host_ms;
  ix->info.build_phase_ms[5] = tot;
  if (ix->box)
    for (uint64_t r = 0; r < n && ix->contain_ok; ++r)
      for (uint32_t d = 0; d < ix->D; ++d)
        if (objs[r * ix->E + d] > objs[r * ix->E + ix->D + d]) ix->contain_ok = false;
  *out = ix;
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

qcp_status qcp_build(const qcp_build_cfg* cfg, const float* objs, uint64_t n, qcp_index** out) try {
  if (!out || (n && !objs)) return QCP_EINVAL;
  *out = nullptr;
  qcp_index* ix = nullptr;
  QCP_TRY(index_alloc(cfg, n, &ix));
  const uint64_t bytes = (uint64_t)ix->E * 4 * n;
  qcp_status s = dalloc(&ix->d_in, n ? (uint64_t)ix->E * n : 1);
  if (s == QCP_OK && n &&
      (hipMemcpyAsync(ix->d_in, objs, bytes, hipMemcpyHostToDevice, ix->stream) != hipSuccess))
    s = QCP_EHIP;
  if (s == QCP_OK)
    s = (cfg->flags & QCP_BUILD_GRAPH) ? index_run_build_graph(ix, ix->d_in) : index_run_build(ix, ix->d_in);
  dfree(ix->d_in);
  if (s != QCP_OK) {
    index_free(ix);
    return s;
  }
  *out = ix;
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

qcp_status qcp_rebuild_device(qcp_index* ix, const float* d_objs) try {
  if (!ix || (ix->n && !d_objs)) return QCP_EINVAL;
  QCP_TRY(set_device(ix->device));
  if (!ix->d_keys2) return QCP_EINVAL;
  if (ix->cfg.flags & QCP_BUILD_GRAPH) return index_run_build_graph(ix, d_objs);
  return index_run_build(ix, d_objs);
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

void qcp_index_destroy(qcp_index* ix) { index_free(ix); }

qcp_status qcp_device_alloc(uint32_t device, uint64_t bytes, void** out) {
  if (!out) return QCP_EINVAL;
  *out = nullptr;
  QCP_TRY(set_device((int)device));
  if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) {
    *out = nullptr;
    return QCP_ENOMEM;
  }
  return QCP_OK;
}

qcp_status qcp_device_free(void* p) {
  if (p) HIP_TRY(hipFree(p));
  return QCP_OK;
}

qcp_status qcp_copy_to_device(void* dst, const void* src, uint64_t bytes) {
  if (bytes && (!dst || !src)) return QCP_EINVAL;
  if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return QCP_OK;
}

qcp_status qcp_copy_to_host(void* dst, const void* src, uint64_t bytes) {
  if (bytes && (!dst || !src)) return QCP_EINVAL;
  if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return QCP_OK;
}

qcp_status qcp_index_info_get(const qcp_index* ix, qcp_index_info* out) {
  if (!ix || !out) return QCP_EINVAL;
  *out = ix->info;
  return QCP_OK;
}

qcp_status qcp_export(const qcp_index* ix, uint64_t* sorted_keys, uint32_t* perm, float* sobjs,
                      float* node_boxes, uint32_t* node_first, uint32_t* node_count,
                      uint8_t* node_leaf) try {
  if (!ix) return QCP_EINVAL;
  if (perm && !ix->perm_valid) return QCP_EINVAL;  // the build kept no permutation
  QCP_TRY(set_device(ix->device));
  const uint64_t n = ix->n;
  hipStream_t s = ix->stream;
  if (sorted_keys && n && ix->keys_stale) {
    HIP_TRY(launch_keys_sorted(ix->d_obj, ix->obj_stride, n, ix->D, ix->box,
                               ix->cfg.key_mode == QCP_KEY_REF, ix->cfg.key_src == QCP_SRC_CENTER,
                               ix->nb, ix->d_keys, s));
    ix->keys_stale = false;
  }
  if (sorted_keys && n) HIP_TRY(hipMemcpyAsync(sorted_keys, ix->d_keys, 8 * n, hipMemcpyDeviceToHost, s));
  if (perm && n) HIP_TRY(hipMemcpyAsync(perm, ix->d_perm, 4 * n, hipMemcpyDeviceToHost, s));
  std::vector<float> soa;
  if (sobjs && n) {
    soa.resize((size_t)ix->E * n);
    for (uint32_t e = 0; e < ix->E; ++e)
      HIP_TRY(hipMemcpyAsync(soa.data() + (size_t)e * n, ix->d_obj + (uint64_t)e * ix->obj_stride,
                             4 * n, hipMemcpyDeviceToHost, s));
  }
  std::vector<float> nb;
  const uint64_t m = ix->n_nodes;
  if (node_boxes && m) {
    nb.resize((size_t)2 * ix->D * m);
    for (uint32_t e = 0; e < 2 * ix->D; ++e)
      HIP_TRY(hipMemcpyAsync(nb.data() + (size_t)e * m, ix->d_nbox + (uint64_t)e * ix->node_stride,
                             4 * m, hipMemcpyDeviceToHost, s));
  }
  std::vector<uint32_t> cnt;
  if (ix->expl && m) {
    if (node_first) HIP_TRY(hipMemcpyAsync(node_first, ix->d_nfirst, 4 * m, hipMemcpyDeviceToHost, s));
    cnt.resize(m);
    HIP_TRY(hipMemcpyAsync(cnt.data(), ix->d_ncount, 4 * m, hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  if (sobjs)
    for (uint64_t r = 0; r < n; ++r)
      for (uint32_t e = 0; e < ix->E; ++e) sobjs[r * ix->E + e] = soa[(size_t)e * n + r];
  if (node_boxes)
    for (uint64_t i = 0; i < m; ++i)
      for (uint32_t e = 0; e < 2 * ix->D; ++e) node_boxes[i * 2 * ix->D + e] = nb[(size_t)e * m + i];
  if (ix->expl) {
    for (uint64_t i = 0; i < m; ++i) {
      if (node_count) node_count[i] = cnt[i] & 0x7FFFFFFFu;
      if (node_leaf) node_leaf[i] = (cnt[i] & 0x80000000u) ? 1 : 0;
    }
  } else {
    for (int l = 0; l < ix->h; ++l) {
      const bool leaf = l == ix->h - 1;
      const uint64_t below = leaf ? n : ix->level_count[l + 1];
      for (uint64_t j = 0; j < ix->level_count[l]; ++j) {
        const uint64_t id = ix->level_off[l] + j;
        const uint64_t f = j * ix->F;
        const uint64_t c = std::min<uint64_t>(ix->F, below - f);
        if (node_first) node_first[id] = (uint32_t)(leaf ? f : ix->level_off[l + 1] + f);
        if (node_count) node_count[id] = (uint32_t)c;
        if (node_leaf) node_leaf[id] = leaf ? 1 : 0;
      }
    }
  }
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

// ---------------------------------------------------------------------------
// batches

void qcp_batch_destroy(qcp_batch* b) {
  if (!b) return;
  hipSetDevice(b->ix->device);
  hipStreamSynchronize(b->ix->stream);
  dfree(b->d_queries);
  dfree(b->d_counts);
  dfree(b->d_hashes);
  dfree(b->d_offsets);
  dfree(b->d_scan_sums);
  dfree(b->d_qperm);
  dfree(b->d_qperm2);
  dfree(b->d_qkeys);
  dfree(b->d_qkeys2);
  dfree(b->d_qghist);
  dfree(b->d_qstatus);
  dfree(b->d_qor);
  dfree(b->d_qbad);
  dfree(b->d_ids);
  dfree(b->d_visits);
  dfree(b->d_ctl);
  dfree(b->d_slots);
  if (b->h_counts) hipHostFree(b->h_counts);
  if (b->h_visits) hipHostFree(b->h_visits);
  dfree(b->d_pcounts);
  dfree(b->d_phashes);
  dfree(b->d_pvisits);
  for (void* h : {(void*)b->h_pcounts, (void*)b->h_phashes, (void*)b->h_pvisits})
    if (h) hipHostFree(h);
  if (b->exec) hipGraphExecDestroy(b->exec);
  for (auto& e : b->ev)
    if (e) hipEventDestroy(e);
  delete b;
}

static qcp_status batch_alloc(qcp_index* ix, uint64_t q, qcp_batch** out) {
  if (q >= 0xFFFFFFF0ull) return QCP_EINVAL;
  qcp_batch* b = new (std::nothrow) qcp_batch();
  if (!b) return QCP_ENOMEM;
  b->ix = ix;
  b->q = q;
  auto fail = [&](qcp_status s) {
    qcp_batch_destroy(b);
    return s;
  };
  qcp_status s;
  if ((s = dalloc(&b->d_queries, (uint64_t)2 * ix->D * q)) != QCP_OK) return fail(s);
  if ((s = dalloc(&b->d_counts, q)) != QCP_OK) return fail(s);
  if ((s = dalloc(&b->d_hashes, q)) != QCP_OK) return fail(s);
  if ((s = dalloc(&b->d_offsets, q + 1)) != QCP_OK) return fail(s);
  if ((s = dalloc(&b->d_scan_sums, kMaxScanBlocks + 1)) != QCP_OK) return fail(s);
  {
    if ((s = dalloc(&b->d_qperm, q)) != QCP_OK) return fail(s);
    if ((s = dalloc(&b->d_qperm2, q)) != QCP_OK) return fail(s);
    if ((s = dalloc(&b->d_qkeys, q)) != QCP_OK) return fail(s);
    if ((s = dalloc(&b->d_qkeys2, q)) != QCP_OK) return fail(s);
    if ((s = dalloc(&b->d_qghist, kSortGhistWords)) != QCP_OK) return fail(s);
    if ((s = dalloc(&b->d_qstatus, sort_status_words(q))) != QCP_OK) return fail(s);
    if ((s = dalloc(&b->d_qor, 1)) != QCP_OK) return fail(s);
    if ((s = dalloc(&b->d_qbad, 1)) != QCP_OK) return fail(s);
  }
  if ((s = dalloc(&b->d_visits, 2 * q)) != QCP_OK) return fail(s);
  if ((s = dalloc(&b->d_ctl, 1)) != QCP_OK) return fail(s);
  for (auto& e : b->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(QCP_EHIP);
  *out = b;
  return QCP_OK;
}

qcp_status qcp_batch_create(qcp_index* ix, const float* boxes, uint64_t q, qcp_batch** out) try {
  if (!ix || !out || (q && !boxes)) return QCP_EINVAL;
  *out = nullptr;
  QCP_TRY(set_device(ix->device));
  qcp_batch* b = nullptr;
  QCP_TRY(batch_alloc(ix, q, &b));
  if (q) {
    if (hipMemcpyAsync(b->d_queries, boxes, (uint64_t)2 * ix->D * 4 * q, hipMemcpyHostToDevice,
                       ix->stream) != hipSuccess ||
        hipStreamSynchronize(ix->stream) != hipSuccess) {
      qcp_batch_destroy(b);
      return QCP_EHIP;
    }
  }
  *out = b;
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

qcp_status qcp_batch_create_device(qcp_index* ix, const float* d_boxes, uint64_t q,
                                   qcp_batch** out) try {
  if (!ix || !out || (q && !d_boxes)) return QCP_EINVAL;
  *out = nullptr;
  QCP_TRY(set_device(ix->device));
  qcp_batch* b = nullptr;
  QCP_TRY(batch_alloc(ix, q, &b));
  if (q) {
    if (hipMemcpyAsync(b->d_queries, d_boxes, (uint64_t)2 * ix->D * 4 * q,
                       hipMemcpyDeviceToDevice, ix->stream) != hipSuccess ||
        hipStreamSynchronize(ix->stream) != hipSuccess) {
      qcp_batch_destroy(b);
      return QCP_EHIP;
    }
  }
  *out = b;
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

}  // extern "C"

SearchArgs qcp::index_search_args(const qcp_index* ix, const float* d_queries, uint64_t nq,
                                 uint32_t* d_counts, unsigned long long* d_hashes) {
  SearchArgs a;
  memset(&a, 0, sizeof(a));
  a.obj = ix->d_obj;
  a.nbox = ix->d_nbox;
  a.nfirst = ix->d_nfirst;
  a.ncount = ix->d_ncount;
  a.nrlo = ix->d_nrlo;
  a.nrhi = ix->d_nrhi;
  a.nparent = ix->d_nparent;
  a.parts = ix->parts;
  a.queries = d_queries;
  a.obj_stride = ix->obj_stride;
  a.node_stride = ix->node_stride;
  a.n = ix->n;
  a.nq = (uint32_t)nq;
  a.F = ix->F;
  a.h = ix->h;
  a.n_nodes = (uint32_t)ix->n_nodes;
  if (!ix->expl) {
    for (int l = 0; l <= ix->h && l <= kMaxLevels; ++l) a.lvl_off[l] = (uint32_t)ix->level_off[l];
    for (int l = 0; l < ix->h && l < kMaxLevels; ++l) a.lvl_cnt[l] = (uint32_t)ix->level_count[l];
    uint64_t sp = ix->F;
    for (int l = ix->h - 1; l >= 0; --l) {
      a.span[l] = sp;
      sp = (sp > (1ull << 40)) ? sp : sp * ix->F;
    }
  }
  a.counts = d_counts;
  a.hashes = d_hashes;
  return a;
}

static SearchArgs make_args(const qcp_batch* b) {
  SearchArgs a = index_search_args(b->ix, b->d_queries, b->q, b->d_counts, b->d_hashes);
  a.offsets = b->d_offsets;
  a.ids = b->d_ids;
  a.visits = b->d_visits;
  a.ctl = b->d_ctl;
  a.slots = b->d_slots;
  a.slot_cap = b->slot_cap;
  return a;
}

unsigned qcp::persistent_grid(const qcp_index* ix, int out, unsigned block) {
  const int mw = (int)((ix->F + 63) / 64);
  const void* k = search_kernel_ptr(ix->D, ix->box, ix->expl, out, mw);
  int per_cu = 0;
  if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, k, (int)block, search_lds_bytes(ix->expl, mw, out, block)) != hipSuccess ||
      per_cu <= 0)
    per_cu = 2;
  return (unsigned)(ix->cus * per_cu);
}

extern "C" {

// Spatial processing order of a batch: the 16-bit Hilbert key of each query
// box's centre, radix-sorted (stable, one-sweep) with the query index as
// payload; 2D passes, an even count, so the order lands in d_qperm.  All on
// the index's stream, no host synchronisation (graph-capturable).
static qcp_status batch_order(qcp_batch* b) {
  qcp_index* ix = b->ix;
  hipStream_t s = ix->stream;
  const uint64_t q = b->q;
  HIP_TRY(launch_fill_u32((uint32_t*)b->d_qor, 2, 0u, s));
  HIP_TRY(launch_fill_u32(b->d_qbad, 1, 0u, s));
  HIP_TRY(launch_keys(b->d_queries, q, ix->D, true, false, true, 16, b->d_qkeys, nullptr,
                      b->d_qor, b->d_qbad, s));
  const unsigned passes = (16 * ix->D + 7) / 8;
  if (passes & 1u) return QCP_EINVAL;
  HIP_TRY(launch_sort_keys(b->d_qkeys, b->d_qperm, b->d_qkeys2, b->d_qperm2, q, passes, nullptr,
                           b->d_qghist, b->d_qstatus, s));
  return QCP_OK;
}

}  // extern "C"

// The batch's spatial order on the host (the co-processing form walks its
// queries in it, so that the work items of neighbouring queries reach the GPU
// together): ord[j] = the j-th query, the identity for small batches or
// order = 1.
qcp_status qcp::batch_order_host(qcp_batch* b, bool spatial, std::vector<uint32_t>& ord) {
  const uint64_t q = b->q;
  ord.resize(q);
  if (!spatial || q < kOrderMinQueries) {
    for (uint64_t i = 0; i < q; ++i) ord[i] = (uint32_t)i;
    return QCP_OK;
  }
  QCP_TRY(batch_order(b));
  hipStream_t s = b->ix->stream;
  HIP_TRY(hipMemcpyAsync(ord.data(), b->d_qperm, 4ull * q, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return QCP_OK;
}

extern "C" {

// The reference's partition mode of MPHR::Search (src/tree/mphr.cpp:235-257):
// one launch per query with one workgroup per partition, each running MPRS
// over its partition's subtree from its root (g_root_offset[bid], 314-333);
// the per-workgroup counts, hashes and visit counters are copied back
// synchronously after every launch and summed on the host, then stored as the
// batch's per-query results.
static qcp_status plain_partitions(qcp_batch* b, uint32_t waves, qcp_search_stats* st) {
  qcp_index* ix = b->ix;
  hipStream_t s = ix->stream;
  const uint64_t q = b->q;
  const uint32_t P = ix->parts;
  if (!b->d_pcounts) {
    QCP_TRY(dalloc(&b->d_pcounts, P));
    QCP_TRY(dalloc(&b->d_phashes, P));
    QCP_TRY(dalloc(&b->d_pvisits, 2ull * P));
    HIP_TRY(hipHostMalloc((void**)&b->h_pcounts, 4ull * P, 0));
    HIP_TRY(hipHostMalloc((void**)&b->h_phashes, 8ull * P, 0));
    HIP_TRY(hipHostMalloc((void**)&b->h_pvisits, 8ull * P, 0));
  }
  std::vector<uint32_t> counts(q), visits(2 * q);
  std::vector<unsigned long long> hashes(q);
  SearchArgs a = make_args(b);
  a.counts = b->d_pcounts;
  a.hashes = b->d_phashes;
  a.visits = b->d_pvisits;
  HIP_TRY(hipEventRecord(b->ev[0], s));
  uint64_t total = 0;
  for (uint64_t qi = 0; qi < q; ++qi) {
    a.q_begin = (uint32_t)qi;
    HIP_TRY(launch_mprs_part(a, ix->D, ix->box, P, waves, s));
    HIP_TRY(hipMemcpyAsync(b->h_pcounts, b->d_pcounts, 4ull * P, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(b->h_phashes, b->d_phashes, 8ull * P, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(b->h_pvisits, b->d_pvisits, 8ull * P, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (uint32_t p = 0; p < P; ++p) {
      counts[qi] += b->h_pcounts[p];
      hashes[qi] += b->h_phashes[p];
      visits[2 * qi] += b->h_pvisits[2 * p];
      visits[2 * qi + 1] += b->h_pvisits[2 * p + 1];
    }
    total += counts[qi];
    st->launches++;
  }
  HIP_TRY(hipEventRecord(b->ev[1], s));
  HIP_TRY(hipStreamSynchronize(s));
  float ms = 0;
  hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
  if (q) {
    HIP_TRY(hipMemcpyAsync(b->d_counts, counts.data(), 4 * q, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(b->d_hashes, hashes.data(), 8 * q, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(b->d_visits, visits.data(), 8 * q, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  st->kernel_ms = ms;
  st->total_ms = ms;
  st->total_hits = total;
  b->total_hits = total;
  return QCP_OK;
}

qcp_status qcp_batch_run(qcp_batch* b, const qcp_search_cfg* cfg_in, qcp_search_stats* stats) try {
  if (!b) return QCP_EINVAL;
  qcp_search_cfg cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.mode = QCP_SEARCH_PERSISTENT;
  cfg.outputs = QCP_OUT_COUNTS | QCP_OUT_HASH;
  if (cfg_in) cfg = *cfg_in;
  if (cfg.mode > QCP_SEARCH_HYBRID) return QCP_EINVAL;
  qcp_index* ix = b->ix;
  QCP_TRY(set_device(ix->device));
  if (cfg.mode == QCP_SEARCH_HYBRID) {
    b->has_offsets = false;
    b->last_mode = cfg.mode;
    b->last_outputs = cfg.outputs;
    return hybrid_run(b, cfg, stats);
  }
  hipStream_t s = ix->stream;
  const uint64_t q = b->q;
  qcp_search_stats st;
  memset(&st, 0, sizeof(st));
  b->has_offsets = false;
  b->last_mode = cfg.mode;
  b->last_outputs = cfg.outputs;
  if (cfg.mode == QCP_SEARCH_PLAIN) {
    if (cfg.outputs & QCP_OUT_IDS) return QCP_EINVAL;
    const uint32_t waves = cfg.plain_waves ? cfg.plain_waves : 1;
    if (waves != 1 && waves != 3) return QCP_EINVAL;
    if (ix->parts > 1) {
      QCP_TRY(plain_partitions(b, waves, &st));
      if (stats) *stats = st;
      return QCP_OK;
    }
    if (ix->expl) return QCP_EINVAL;  // MPRS needs the packed level layout
    const uint32_t batch = cfg.plain_batch ? cfg.plain_batch : 128;
    if (!b->h_counts && q) {
      HIP_TRY(hipHostMalloc((void**)&b->h_counts, 4 * q, 0));
      HIP_TRY(hipHostMalloc((void**)&b->h_visits, 8 * q, 0));
    }
    SearchArgs a = make_args(b);
    HIP_TRY(hipEventRecord(b->ev[0], s));
    uint64_t total = 0;
    for (uint64_t q0 = 0; q0 < q; q0 += batch) {
      const uint32_t nbk = (uint32_t)std::min<uint64_t>(batch, q - q0);
      a.q_begin = (uint32_t)q0;
      HIP_TRY(launch_mprs(a, ix->D, ix->box, nbk, waves, s));
      // the reference copies the per-block results back synchronously after
      // every launch (src/tree/mphr.cpp:244-246)
      HIP_TRY(hipMemcpyAsync(b->h_counts + q0, b->d_counts + q0, 4ull * nbk, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipMemcpyAsync(b->h_visits + 2 * q0, b->d_visits + 2 * q0, 8ull * nbk,
                             hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      for (uint32_t i = 0; i < nbk; ++i) total += b->h_counts[q0 + i];
      st.launches++;
    }
    HIP_TRY(hipEventRecord(b->ev[1], s));
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0;
    hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
    st.kernel_ms = ms;
    st.total_ms = ms;
    st.total_hits = total;
    b->total_hits = total;
    if (stats) *stats = st;
    return QCP_OK;
  }
  // persistent form
  if (q == 0) {
    HIP_TRY(launch_fill_u32((uint32_t*)b->d_offsets, 2, 0u, s));
    HIP_TRY(hipStreamSynchronize(s));
    b->total_hits = 0;
    b->has_offsets = true;
    if (stats) *stats = st;
    return QCP_OK;
  }
  if (cfg.block_waves > 8) return QCP_EINVAL;
  const unsigned block = 64u * (cfg.block_waves ? cfg.block_waves : kDefaultBlockWaves);
  const int out_mode = (cfg.outputs & QCP_OUT_HASH) ? 1 : 0;
  const int u = cfg.leaf_chunks ? (int)cfg.leaf_chunks : kDefaultLeafChunks;
  const bool donate = cfg.donate != 0;
  // Small batches: a wave per query is enough (no donation) and a spatial
  // order buys nothing, so the grid is sized to the batch and ordering skipped.
  const unsigned wpb = block / 64u;
  auto fit_grid = [&](unsigned full) -> unsigned {
    if (cfg.grid_blocks) return cfg.grid_blocks;
    if (donate) return full;
    const uint64_t need = (q + wpb - 1) / wpb;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(full, need));
  };
  const bool order_queries = cfg.order == 0 && q >= kOrderMinQueries;
  unsigned grid = fit_grid(persistent_grid(ix, out_mode, block));
  SearchArgs a = make_args(b);
  const bool want_ids = (cfg.outputs & QCP_OUT_IDS) != 0;
  // The count pass of an ID search only sizes the output: the containment
  // shortcut gives the same counts and hashes with fewer reads, so it is always
  // on there; the traversal statistics then come from the ID pass, which
  // follows the caller's choice.
  a.contain = ((cfg.contain || want_ids) && ix->contain_ok) ? 1u : 0u;
  if (donate) {
    if (!b->d_slots) {
      b->slot_cap = kSlotCap;
      QCP_TRY(dalloc(&b->d_slots, b->slot_cap));
      b->slots_dirty = b->slot_cap;
    }
    a.slots = b->d_slots;
    a.slot_cap = b->slot_cap;
    a.donate = 1;
    a.max_waiters = cfg.max_waiters ? cfg.max_waiters : 256;
    a.hunger_period = cfg.hunger_period ? cfg.hunger_period : 8;
    a.donate_min_height = cfg.donate_height ? cfg.donate_height : 1;
  }
  // the device work of one run, all on the stream (graph-capturable)
  // timed: HIP events around the search kernel alone (not inside a capture)
  auto enqueue = [&](uint64_t slots_to_clear, bool timed) -> qcp_status {
    if (donate) {
      // per-item results are accumulated atomically; donated slots start empty
      HIP_TRY(launch_fill_u32(b->d_counts, q, 0u, s));
      HIP_TRY(launch_fill_u32((uint32_t*)b->d_hashes, 2 * q, 0u, s));
      HIP_TRY(launch_fill_u32((uint32_t*)b->d_slots, 2ull * slots_to_clear, 0xFFFFFFFFu, s));
    }
    if (order_queries) {
      QCP_TRY(batch_order(b));
      a.qperm = b->d_qperm;
    }
    if (timed) HIP_TRY(hipEventRecord(b->ev[4], s));
    HIP_TRY(launch_search(a, ix->D, ix->box, ix->expl, out_mode, u, grid, block, s));
    if (timed) HIP_TRY(hipEventRecord(b->ev[5], s));
    return QCP_OK;
  };
  const bool use_graph = cfg.graph && !(cfg.outputs & QCP_OUT_IDS);
  if (use_graph) {
    const uint32_t key[12] = {1u, (uint32_t)out_mode, (uint32_t)u, grid, donate ? 1u : 0u,
                              cfg.order, a.max_waiters, a.hunger_period, a.donate_min_height,
                              block, a.contain, b->pipe_objs ? 1u : 0u};
    if (!b->exec || memcmp(key, b->exec_key, sizeof(key)) != 0 || b->exec_objs != b->pipe_objs) {
      if (b->exec) {
        hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
      }
      if (b->pipe_objs && ix->n) QCP_TRY(prepare_graph_form(ix, b->pipe_objs));
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      qcp_status cs = QCP_OK;
      // build+query pipeline: the index's rebuild from device objects first
      if (b->pipe_objs && ix->n) cs = enqueue_build(ix, b->pipe_objs, graph_sort_passes(ix));
      if (cs == QCP_OK) cs = enqueue(b->slot_cap, false);
      hipError_t e = hipStreamEndCapture(s, &g);
      if (cs != QCP_OK) {
        if (g) hipGraphDestroy(g);
        return cs;
      }
      HIP_TRY(e);
      e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      HIP_TRY(e);
      memcpy(b->exec_key, key, sizeof(key));
      b->exec_objs = b->pipe_objs;
    }
    HIP_TRY(hipEventRecord(b->ev[0], s));
    HIP_TRY(hipGraphLaunch(b->exec, s));
    if (b->pipe_objs && ix->n) ix->keys_stale = true;  // the replayed rebuild's sort
    HIP_TRY(hipEventRecord(b->ev[1], s));
  } else {
    HIP_TRY(hipEventRecord(b->ev[0], s));
    QCP_TRY(enqueue(b->slots_dirty, true));
    HIP_TRY(hipEventRecord(b->ev[1], s));
  }
  st.launches = 1;
  // queue statistics of the count pass (the ID pass resets the control block)
  QueueCtl ctl, ctl_ids;
  HIP_TRY(hipMemcpyAsync(&ctl, b->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost, s));
  if (cfg.outputs & QCP_OUT_IDS) {
    HIP_TRY(scan_u32_u64(b->d_counts, q, b->d_offsets, b->d_scan_sums, kMaxScanBlocks, s));
    unsigned long long total = 0;
    HIP_TRY(hipMemcpyAsync(&total, b->d_offsets + q, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (total > b->ids_cap || !b->d_ids) {
      dfree(b->d_ids);
      b->ids_cap = 0;
      QCP_TRY(dalloc(&b->d_ids, total + 64));
      b->ids_cap = total + 64;
    }
    a.ids = b->d_ids;
    a.donate = 0;
    a.contain = (cfg.contain && ix->contain_ok) ? 1u : 0u;
    unsigned grid2 = fit_grid(persistent_grid(ix, 2, block));
    HIP_TRY(hipEventRecord(b->ev[2], s));
    HIP_TRY(launch_search(a, ix->D, ix->box, ix->expl, 2, u >= 8 ? 8 : 4, grid2, block, s));
    HIP_TRY(hipEventRecord(b->ev[3], s));
    HIP_TRY(hipMemcpyAsync(&ctl_ids, b->d_ctl, sizeof(ctl_ids), hipMemcpyDeviceToHost, s));
    st.launches = 2;
    b->has_offsets = true;
  }
  HIP_TRY(hipStreamSynchronize(s));
  float k1 = 0, k2 = 0, tot = 0;
  hipEventElapsedTime(&k1, b->ev[0], b->ev[1]);
  if (cfg.outputs & QCP_OUT_IDS) {
    hipEventElapsedTime(&k2, b->ev[2], b->ev[3]);
    hipEventElapsedTime(&tot, b->ev[0], b->ev[3]);
  } else {
    tot = k1;
  }
  st.kernel_ms = k1 + k2;
  st.total_ms = tot;
  st.search_ms = k1;
  if (!use_graph) {
    float ks = 0;
    hipEventElapsedTime(&ks, b->ev[4], b->ev[5]);
    st.search_ms = ks;
  }
  for (int x = 0; x < 8; ++x) {
    const QueueCtl& tc = want_ids ? ctl_ids : ctl;  // the traversal the caller chose
    st.internal_entries += tc.internal_entries[x];
    st.leaf_entries += tc.leaf_entries[x];
    st.total_hits += tc.total_hits[x];
    st.contained_runs += tc.contained[x];
  }
  st.donated_items = ctl.donated;
  b->total_hits = st.total_hits;
  if (donate) b->slots_dirty = std::min<uint64_t>(b->slot_cap, std::max<uint64_t>(ctl.tail, 1));
  if (stats) *stats = st;
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

qcp_status qcp_pipeline_run(qcp_index* ix, const float* d_objs, qcp_batch* b,
                            const qcp_search_cfg* cfg_in, qcp_search_stats* stats) try {
  if (!ix || !b || b->ix != ix || (ix->n && !d_objs)) return QCP_EINVAL;
  qcp_search_cfg cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.outputs = QCP_OUT_COUNTS | QCP_OUT_HASH;
  if (cfg_in) cfg = *cfg_in;
  if (cfg.mode != QCP_SEARCH_PERSISTENT || (cfg.outputs & QCP_OUT_IDS)) return QCP_EINVAL;
  if (cfg.contain && ix->box) return QCP_EINVAL;  // box validity is known only after the build
  if (ix->cfg.layout == QCP_LAYOUT_GUTTMAN) return QCP_EINVAL;  // the build has a host step
  QCP_TRY(set_device(ix->device));
  cfg.graph = 1;
  b->pipe_objs = d_objs;
  qcp_status st = qcp_batch_run(b, &cfg, stats);
  b->pipe_objs = nullptr;
  if (st != QCP_OK) return st;
  uint32_t bad = 0;
  HIP_TRY(hipMemcpy(&bad, ix->d_bad, 4, hipMemcpyDeviceToHost));
  ix->h_nbox.clear();
  ix->contain_ok = !(bad & kBadInverted);
  return (bad & kBadDomain) ? QCP_EDOMAIN : QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

uint64_t qcp_batch_total_hits(const qcp_batch* b) { return b ? b->total_hits : 0; }

qcp_status qcp_batch_results(const qcp_batch* b, uint32_t* counts, uint64_t* hashes,
                             uint64_t* offsets, uint32_t* ids, uint64_t ids_cap,
                             uint32_t* visits) {
  if (!b) return QCP_EINVAL;
  const qcp_index* ix = b->ix;
  QCP_TRY(set_device(ix->device));
  hipStream_t s = ix->stream;
  const uint64_t q = b->q;
  if (counts && q) HIP_TRY(hipMemcpyAsync(counts, b->d_counts, 4 * q, hipMemcpyDeviceToHost, s));
  if (hashes && q) HIP_TRY(hipMemcpyAsync(hashes, b->d_hashes, 8 * q, hipMemcpyDeviceToHost, s));
  if (visits && q) {
    if (b->last_mode != QCP_SEARCH_PLAIN) return QCP_EINVAL;
    HIP_TRY(hipMemcpyAsync(visits, b->d_visits, 8 * q, hipMemcpyDeviceToHost, s));
  }
  if (offsets || ids) {
    if (!b->has_offsets) return QCP_EINVAL;
    unsigned long long total = 0;
    HIP_TRY(hipMemcpyAsync(&total, b->d_offsets + q, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (offsets) HIP_TRY(hipMemcpyAsync(offsets, b->d_offsets, 8 * (q + 1), hipMemcpyDeviceToHost, s));
    if (ids) {
      if (ids_cap < total) return QCP_EBUSY;
      if (total) HIP_TRY(hipMemcpyAsync(ids, b->d_ids, 4 * total, hipMemcpyDeviceToHost, s));
    }
  }
  HIP_TRY(hipStreamSynchronize(s));
  return QCP_OK;
}

qcp_status qcp_batch_ids(const qcp_batch* b, uint64_t q0, uint64_t q1, uint64_t* offsets,
                         uint32_t* ids, uint64_t ids_cap) try {
  if (!b || q0 > q1 || q1 > b->q) return QCP_EINVAL;
  if (!b->has_offsets) return QCP_EINVAL;
  const qcp_index* ix = b->ix;
  QCP_TRY(set_device(ix->device));
  hipStream_t s = ix->stream;
  std::vector<unsigned long long> off(q1 - q0 + 1);
  HIP_TRY(hipMemcpyAsync(off.data(), b->d_offsets + q0, 8 * off.size(), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const uint64_t total = off.back() - off.front();
  if (offsets)
    for (size_t i = 0; i < off.size(); ++i) offsets[i] = off[i] - off.front();
  if (ids) {
    if (ids_cap < total) return QCP_EBUSY;
    if (total)
      HIP_TRY(hipMemcpyAsync(ids, b->d_ids + off.front(), 4 * total, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

qcp_status qcp_search(qcp_index* ix, const float* boxes, uint64_t q, const qcp_search_cfg* cfg,
                      uint32_t* counts, uint64_t* hashes, uint64_t* offsets, uint32_t* ids,
                      uint64_t ids_cap, qcp_search_stats* stats) try {
  qcp_batch* b = nullptr;
  QCP_TRY(qcp_batch_create(ix, boxes, q, &b));
  qcp_status s = qcp_batch_run(b, cfg, stats);
  if (s == QCP_OK) {
    const bool want_ids = cfg && (cfg->outputs & QCP_OUT_IDS);
    s = qcp_batch_results(b, counts, hashes, want_ids ? offsets : nullptr,
                          want_ids ? ids : nullptr, ids_cap, nullptr);
  }
  qcp_batch_destroy(b);
  return s;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

// Device state of an index from host arrays (the payload of a saved index).
// The topology is recomputed from (n, cfg) and must equal the stored one.
__attribute__((visibility("hidden"))) qcp_status qcp_internal_restore(
    const qcp_build_cfg* cfg, uint64_t n, const uint64_t* keys, const uint32_t* perm,
    const float* objs, const float* boxes, const uint32_t* first, const uint32_t* count,
    uint64_t n_nodes, qcp_index** out) {
  qcp_index* ix = nullptr;
  QCP_TRY(index_alloc(cfg, n, &ix));
  auto fail = [&](qcp_status s) {
    index_free(ix);
    return s;
  };
  hipStream_t s = ix->stream;
  if (cfg->layout == QCP_LAYOUT_GUTTMAN) {
    // the stored topology is the tree: check its shape, then adopt it
    std::vector<uint64_t> lo;
    if (!first || !count || !guttman_check(first, count, n_nodes, n, ix->F, kMaxDepth, &lo))
      return fail(QCP_EFORMAT);
    qcp_status st = attach_topology(ix, std::vector<uint32_t>(first, first + n_nodes),
                                    std::vector<uint32_t>(count, count + n_nodes), lo);
    if (st != QCP_OK) return fail(st);
  }
  if (ix->n_nodes != n_nodes) return fail(QCP_EFORMAT);
  if ((ix->cfg.layout == QCP_LAYOUT_BVH || ix->parts > 1) && n_nodes) {
    std::vector<uint32_t> f(n_nodes), c(n_nodes);
    if (hipMemcpy(f.data(), ix->d_nfirst, 4 * n_nodes, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(c.data(), ix->d_ncount, 4 * n_nodes, hipMemcpyDeviceToHost) != hipSuccess)
      return fail(QCP_EHIP);
    if (memcmp(f.data(), first, 4 * n_nodes) || memcmp(c.data(), count, 4 * n_nodes))
      return fail(QCP_EFORMAT);
  }
  std::vector<float> soa;
  if (n) {
    if (hipMemcpyAsync(ix->d_keys, keys, 8 * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        (perm && hipMemcpyAsync(ix->d_perm, perm, 4 * n, hipMemcpyHostToDevice, s) != hipSuccess))
      return fail(QCP_EHIP);
    soa.resize((size_t)ix->E * n);
    for (uint64_t r = 0; r < n; ++r)
      for (uint32_t e = 0; e < ix->E; ++e) soa[(size_t)e * n + r] = objs[r * ix->E + e];
    for (uint32_t e = 0; e < ix->E; ++e)
      if (hipMemcpyAsync(ix->d_obj + (uint64_t)e * ix->obj_stride, soa.data() + (size_t)e * n,
                         4 * n, hipMemcpyHostToDevice, s) != hipSuccess)
        return fail(QCP_EHIP);
    if (hipStreamSynchronize(s) != hipSuccess) return fail(QCP_EHIP);
  }
  std::vector<float> nb((size_t)2 * ix->D * n_nodes);
  for (uint64_t i = 0; i < n_nodes; ++i)
    for (uint32_t e = 0; e < 2 * ix->D; ++e) nb[(size_t)e * n_nodes + i] = boxes[i * 2 * ix->D + e];
  for (uint32_t e = 0; e < 2 * ix->D && n_nodes; ++e)
    if (hipMemcpyAsync(ix->d_nbox + (uint64_t)e * ix->node_stride, nb.data() + (size_t)e * n_nodes,
                       4 * n_nodes, hipMemcpyHostToDevice, s) != hipSuccess)
      return fail(QCP_EHIP);
  if (hipStreamSynchronize(s) != hipSuccess) return fail(QCP_EHIP);
  if (ix->box)
    for (uint64_t r = 0; r < n && ix->contain_ok; ++r)
      for (uint32_t d = 0; d < ix->D; ++d)
        if (objs[r * ix->E + d] > objs[r * ix->E + ix->D + d]) ix->contain_ok = false;
  ix->perm_valid = perm != nullptr || n == 0;
  set_info_shape(ix);
  *out = ix;
  return QCP_OK;
}

}  // extern "C"
