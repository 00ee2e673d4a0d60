// partition_topology.cpp -- the partitioned MPHR's node structure (see the header).
#include "partition_topology.h"

#include <algorithm>

namespace qcp {

void partition_range(uint64_t n, uint32_t P, uint32_t p, uint64_t* begin, uint64_t* count) {
  const uint64_t chunk = n / P, first = chunk + n % P;
  *begin = p == 0 ? 0 : first + (uint64_t)(p - 1) * chunk;
  *count = p == 0 ? first : chunk;
}

bool partition_topology(uint64_t n, uint32_t F, uint32_t P, uint32_t max_depth,
                        PartitionTopology* out) {
  if (P == 0 || n < P || F < 2) return false;
  // level sizes of one packed subtree of c objects, root first (c >= 1)
  auto levels = [&](uint64_t c) {
    std::vector<uint64_t> lv;
    if (c == 1) return std::vector<uint64_t>{1};
    for (uint64_t cur = c; cur > 1;) {
      cur = (cur + F - 1) / F;
      lv.insert(lv.begin(), cur);
    }
    return lv;
  };
  // two shapes: partition 0 and the (equal) others
  uint64_t b0, c0, b1, c1;
  partition_range(n, P, 0, &b0, &c0);
  const std::vector<uint64_t> la = levels(c0);
  std::vector<uint64_t> lb;
  if (P > 1) {
    partition_range(n, P, 1, &b1, &c1);
    lb = levels(c1);
  }
  const size_t H = std::max(la.size(), lb.size());
  if (H > max_depth) return false;
  auto cnt = [&](uint32_t p, size_t d) -> uint64_t {
    const std::vector<uint64_t>& l = p == 0 ? la : lb;
    return d < l.size() ? l[d] : 0;
  };
  out->level_off.assign(H + 1, 0);
  for (size_t d = 0; d < H; ++d)
    out->level_off[d + 1] = out->level_off[d] + cnt(0, d) + (uint64_t)(P - 1) * cnt(1, d);
  const uint64_t m = out->level_off[H];
  if (m >= 0x7FFFFFFFull) return false;
  out->first.assign(m, 0);
  out->count.assign(m, 0);
  // first node id of partition p at depth d
  auto start = [&](uint32_t p, size_t d) -> uint64_t {
    return out->level_off[d] + (p ? cnt(0, d) + (uint64_t)(p - 1) * cnt(1, d) : 0);
  };
  for (uint32_t p = 0; p < P; ++p) {
    uint64_t b, c;
    partition_range(n, P, p, &b, &c);
    const std::vector<uint64_t>& l = p == 0 ? la : lb;
    const size_t h = l.size();
    for (size_t d = 0; d < h; ++d) {
      const bool leaf = d + 1 == h;
      const uint64_t below = leaf ? c : l[d + 1];
      const uint64_t s = start(p, d);
      for (uint64_t j = 0; j < l[d]; ++j) {
        const uint64_t f = j * F;
        const uint32_t k = (uint32_t)std::min<uint64_t>(F, below - f);
        out->first[s + j] = (uint32_t)(leaf ? b + f : start(p, d + 1) + f);
        out->count[s + j] = leaf ? (k | 0x80000000u) : k;
      }
    }
  }
  return true;
}

}  // namespace qcp
