// index_io.cpp -- save / load of a device-resident index.
//
// Replaces MPHR::DumpToFile / DumpFromFile (src/tree/mphr.cpp:138-194) and the
// BVH dump (src/tree/bvh.cpp:65-163), whose files carry no header, version or
// checksum.  Format (little-endian):
//   magic "QCPIDX01" | u32 version | u32 header bytes | qcp_build_cfg |
//   u64 n | u64 n_nodes | i32 h | u32 pad | u64 level_off[h+1] |
//   u64 payload bytes | u64 FNV-1a 64 of the payload | payload
// payload = sorted keys (8n) | perm (4n; absent when the stored cfg.flags has
//           QCP_BUILD_NO_PERM: the index kept none) | objects SoA (4En) | node boxes SoA
//           (8D n_nodes) | [explicit layouts and partitioned MPHR] node first
//           (4 n_nodes) | node count (4 n_nodes)
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <new>
#include <vector>

#include "../../include/qcp.h"

namespace {

constexpr char kMagic[8] = {'Q', 'C', 'P', 'I', 'D', 'X', '0', '1'};
constexpr uint32_t kVersion = 2;  // 2: qcp_build_cfg carries the MPHR partition count

uint64_t fnv1a(const uint8_t* p, size_t n, uint64_t h) {
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001B3ull;
  }
  return h;
}

struct FileCloser {
  FILE* f;
  ~FileCloser() {
    if (f) fclose(f);
  }
};

}  // namespace

extern "C" {

qcp_status qcp_save(const qcp_index* ix, const char* path) try {
  if (!ix || !path) return QCP_EINVAL;
  qcp_index_info in;
  qcp_status s = qcp_index_info_get(ix, &in);
  if (s != QCP_OK) return s;
  const uint64_t n = in.n, m = in.n_nodes;
  const uint32_t D = in.dims;
  const uint32_t E = in.obj_kind == QCP_OBJ_BOXES ? 2 * D : D;
  const bool expl =
      in.layout == QCP_LAYOUT_BVH || in.layout == QCP_LAYOUT_GUTTMAN || in.partitions > 1;
  const bool has_perm = in.has_perm != 0;
  std::vector<uint64_t> keys(n);
  std::vector<uint32_t> perm(has_perm ? n : 0);
  std::vector<float> objs((size_t)E * n), boxes((size_t)2 * D * m);
  std::vector<uint32_t> first(m), count(m);
  std::vector<uint8_t> leaf(m);
  s = qcp_export(ix, keys.data(), has_perm ? perm.data() : nullptr, objs.data(), boxes.data(),
                 first.data(), count.data(), leaf.data());
  if (s != QCP_OK) return s;
  std::vector<uint8_t> payload;
  auto put = [&](const void* p, size_t bytes) {
    const uint8_t* b = (const uint8_t*)p;
    payload.insert(payload.end(), b, b + bytes);
  };
  put(keys.data(), 8 * n);
  if (has_perm) put(perm.data(), 4 * n);
  put(objs.data(), 4 * objs.size());
  put(boxes.data(), 4 * boxes.size());
  if (expl) {
    for (uint64_t i = 0; i < m; ++i) count[i] |= leaf[i] ? 0x80000000u : 0u;
    put(first.data(), 4 * m);
    put(count.data(), 4 * m);
  }
  qcp_build_cfg cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.dims = D;
  cfg.fanout = in.fanout;
  cfg.layout = in.layout;
  cfg.key_mode = in.key_mode;
  cfg.key_src = in.key_src;
  cfg.obj_kind = in.obj_kind;
  cfg.partitions = in.partitions;
  cfg.flags = has_perm ? 0u : (uint32_t)QCP_BUILD_NO_PERM;
  int32_t h = (int32_t)in.height;
  std::vector<uint64_t> level_off(h + 1, 0);
  for (int l = 0; l < h; ++l) level_off[l + 1] = level_off[l] + in.level_count[l];
  FileCloser fc{fopen(path, "wb")};
  if (!fc.f) return QCP_EIO;
  const uint32_t hdr = (uint32_t)(8 + 4 + 4 + sizeof(cfg) + 8 + 8 + 4 + 4 + 8 * (h + 1) + 8 + 8);
  const uint64_t pbytes = payload.size();
  const uint64_t sum = fnv1a(payload.data(), payload.size(), 0xCBF29CE484222325ull);
  const uint32_t pad = 0;
  bool ok = fwrite(kMagic, 8, 1, fc.f) == 1 && fwrite(&kVersion, 4, 1, fc.f) == 1 &&
            fwrite(&hdr, 4, 1, fc.f) == 1 && fwrite(&cfg, sizeof(cfg), 1, fc.f) == 1 &&
            fwrite(&n, 8, 1, fc.f) == 1 && fwrite(&m, 8, 1, fc.f) == 1 &&
            fwrite(&h, 4, 1, fc.f) == 1 && fwrite(&pad, 4, 1, fc.f) == 1 &&
            fwrite(level_off.data(), 8, h + 1, fc.f) == (size_t)(h + 1) &&
            fwrite(&pbytes, 8, 1, fc.f) == 1 && fwrite(&sum, 8, 1, fc.f) == 1 &&
            (pbytes == 0 || fwrite(payload.data(), 1, pbytes, fc.f) == pbytes);
  return ok ? QCP_OK : QCP_EIO;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

// Implemented in qcp_api.cpp: rebuild an index's device state from exported arrays.
__attribute__((visibility("hidden"))) qcp_status qcp_internal_restore(const qcp_build_cfg* cfg, uint64_t n, const uint64_t* keys,
                                const uint32_t* perm, const float* objs, const float* boxes,
                                const uint32_t* first, const uint32_t* count, uint64_t n_nodes,
                                qcp_index** out);

qcp_status qcp_load(const char* path, uint32_t device, qcp_index** out) try {
  if (!path || !out) return QCP_EINVAL;
  *out = nullptr;
  FileCloser fc{fopen(path, "rb")};
  if (!fc.f) return QCP_EIO;
  char magic[8];
  uint32_t ver = 0, hdr = 0, pad = 0;
  qcp_build_cfg cfg;
  uint64_t n = 0, m = 0, pbytes = 0, sum = 0;
  int32_t h = 0;
  if (fread(magic, 8, 1, fc.f) != 1 || memcmp(magic, kMagic, 8) != 0) return QCP_EFORMAT;
  if (fread(&ver, 4, 1, fc.f) != 1 || ver != kVersion) return QCP_EFORMAT;
  if (fread(&hdr, 4, 1, fc.f) != 1) return QCP_EFORMAT;
  if (fread(&cfg, sizeof(cfg), 1, fc.f) != 1) return QCP_EFORMAT;
  if (fread(&n, 8, 1, fc.f) != 1 || fread(&m, 8, 1, fc.f) != 1) return QCP_EFORMAT;
  if (fread(&h, 4, 1, fc.f) != 1 || fread(&pad, 4, 1, fc.f) != 1) return QCP_EFORMAT;
  if (h < 0 || h > 64 || cfg.dims < 2 || cfg.dims > 3 || n >= 0xFFFFFFF0ull || m > 2 * n + 2)
    return QCP_EFORMAT;
  std::vector<uint64_t> level_off(h + 1);
  if (fread(level_off.data(), 8, h + 1, fc.f) != (size_t)(h + 1)) return QCP_EFORMAT;
  if (fread(&pbytes, 8, 1, fc.f) != 1 || fread(&sum, 8, 1, fc.f) != 1) return QCP_EFORMAT;
  const uint32_t D = cfg.dims;
  const uint32_t E = cfg.obj_kind == QCP_OBJ_BOXES ? 2 * D : D;
  if (cfg.partitions > 65536) return QCP_EFORMAT;
  const bool expl =
      cfg.layout == QCP_LAYOUT_BVH || cfg.layout == QCP_LAYOUT_GUTTMAN || cfg.partitions > 1;
  if (cfg.flags & ~(uint32_t)QCP_BUILD_NO_PERM) return QCP_EFORMAT;
  const bool has_perm = !(cfg.flags & QCP_BUILD_NO_PERM);
  const uint64_t expect =
      8 * n + (has_perm ? 4 * n : 0) + 4ull * E * n + 8ull * D * m + (expl ? 8 * m : 0);
  if (pbytes != expect) return QCP_EFORMAT;
  std::vector<uint8_t> payload(pbytes);
  if (pbytes && fread(payload.data(), 1, pbytes, fc.f) != pbytes) return QCP_EFORMAT;
  if (fnv1a(payload.data(), payload.size(), 0xCBF29CE484222325ull) != sum) return QCP_EFORMAT;
  const uint8_t* p = payload.data();
  const uint64_t* keys = (const uint64_t*)p;
  p += 8 * n;
  const uint32_t* perm = has_perm ? (const uint32_t*)p : nullptr;
  if (has_perm) p += 4 * n;
  const float* objs = (const float*)p;
  p += 4ull * E * n;
  const float* boxes = (const float*)p;
  p += 8ull * D * m;
  const uint32_t* first = expl ? (const uint32_t*)p : nullptr;
  const uint32_t* count = expl ? (const uint32_t*)(p + 4 * m) : nullptr;
  cfg.device = device;
  return qcp_internal_restore(&cfg, n, keys, perm, objs, boxes, first, count, m, out);
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}

}  // extern "C"

// ---- the reference's MPHR dump ------------------------------------------------
// MPHR::DumpToFile (src/tree/mphr.cpp:167-194) writes u32 partitions, i64 root
// offset per partition, u32 node count, then the Node_SOA array as raw memory
// (src/node/node_soa.h:40-55): points[2*D*F] floats (dimension-major, lo
// dimensions then hi), index[F] and child_offset[F] as i64, then node_type,
// level (i32) and branch_count (u32), padded to 8 bytes; D and F are the
// reference's compile-time GetNumberOfDims() and GetNumberOfLeafNodeDegrees()
// (src/common/config.h:4-8).  Leaf nodes (node_type 1) hold the entries, and
// an entry's index is its 1-based Hilbert rank (src/sort/thrust_sorter.cpp:
// 39-47) in every partition, so the objects are recovered in rank order.
extern "C" qcp_status qcp_read_reference_mphr(const char* path, uint32_t dims,
                                              uint32_t leaf_degree, uint32_t obj_kind,
                                              float* objs, uint64_t cap, uint64_t* n_out) try {
  if (!path || !n_out || (dims != 2 && dims != 3) || leaf_degree == 0 ||
      leaf_degree > (1u << 16) || obj_kind > QCP_OBJ_BOXES)
    return QCP_EINVAL;
  const uint32_t D = dims, F = leaf_degree;
  const uint32_t E = obj_kind == QCP_OBJ_BOXES ? 2 * D : D;
  const uint64_t raw = (uint64_t)2 * D * F * 4 + 16ull * F + 12;
  const uint64_t node_bytes = (raw + 7) / 8 * 8;
  FILE* f = fopen(path, "rb");
  if (!f) return QCP_EIO;
  FileCloser fc{f};
  if (fseek(f, 0, SEEK_END) != 0) return QCP_EIO;
  const long fsize = ftell(f);
  if (fsize < 8 || fseek(f, 0, SEEK_SET) != 0) return QCP_EFORMAT;
  uint32_t parts = 0, count = 0;
  if (fread(&parts, 4, 1, f) != 1 || parts == 0 || parts > (1u << 16)) return QCP_EFORMAT;
  if (fseek(f, 8l * parts, SEEK_CUR) != 0 || fread(&count, 4, 1, f) != 1) return QCP_EFORMAT;
  if ((uint64_t)fsize != 8ull + 8ull * parts + node_bytes * count) return QCP_EFORMAT;
  std::vector<uint8_t> node(node_bytes);
  const float* pts = reinterpret_cast<const float*>(node.data());
  // first pass: the number of leaf entries (= N)
  const long body = ftell(f);
  uint64_t n = 0;
  for (uint32_t i = 0; i < count; ++i) {
    if (fread(node.data(), 1, node_bytes, f) != node_bytes) return QCP_EFORMAT;
    int32_t type;
    uint32_t bc;
    memcpy(&type, node.data() + raw - 12, 4);
    memcpy(&bc, node.data() + raw - 4, 4);
    if (bc > F) return QCP_EFORMAT;
    if (type == 1) n += bc;
  }
  *n_out = n;
  if (!objs) return QCP_OK;
  if (cap < n) return QCP_EBUSY;
  // second pass: every entry to its rank, each rank exactly once
  std::vector<uint8_t> seen(n, 0);
  if (fseek(f, body, SEEK_SET) != 0) return QCP_EIO;
  for (uint32_t i = 0; i < count; ++i) {
    if (fread(node.data(), 1, node_bytes, f) != node_bytes) return QCP_EFORMAT;
    int32_t type;
    uint32_t bc;
    memcpy(&type, node.data() + raw - 12, 4);
    memcpy(&bc, node.data() + raw - 4, 4);
    if (type != 1) continue;
    const int64_t* idx = reinterpret_cast<const int64_t*>(node.data() + (uint64_t)2 * D * F * 4);
    for (uint32_t k = 0; k < bc; ++k) {
      const int64_t r = idx[k];
      if (r < 1 || (uint64_t)r > n || seen[r - 1]) return QCP_EFORMAT;
      seen[r - 1] = 1;
      for (uint32_t e = 0; e < E; ++e) objs[(uint64_t)(r - 1) * E + e] = pts[(uint64_t)e * F + k];
    }
  }
  return QCP_OK;
} catch (const std::bad_alloc&) {
  return QCP_ENOMEM;  // host allocation failed
} catch (...) {
  return QCP_EINVAL;  // no exception crosses the C ABI
}
