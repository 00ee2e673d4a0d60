// bvh_topology.cpp -- BFS construction of the rank-prefix BVH (see the header).
#include "bvh_topology.h"

#include <algorithm>
#include <deque>

namespace qcp {

namespace {

// Highest offset in [start, end) whose 1-based rank shares more leading bits
// with rank(start) than rank(end) does, found by the reference's halving
// search (src/tree/tree.cpp:278-301); 0 doubles as "no split"
// (src/tree/tree.cpp:274-276), as in the reference.
uint32_t split_offset(uint32_t start, uint32_t end) {
  const uint64_t first = (uint64_t)start + 1, last = (uint64_t)end + 1;
  if (first >= last) return 0;
  const int prefix = __builtin_clzll(first ^ last);
  uint32_t split = start;
  uint32_t step = end - start;
  do {
    step = (step + 1) >> 1;
    const uint32_t cand = split + step;
    if (cand < end && __builtin_clzll(first ^ ((uint64_t)cand + 1)) > prefix) split = cand;
  } while (step > 1);
  return split;
}

// Sorted split positions of [start, end] (src/tree/tree.cpp:304-350).
void split_positions(uint32_t start, uint32_t end, uint32_t F, std::vector<uint32_t>& pos) {
  pos.clear();
  pos.push_back(start);
  pos.push_back(end);
  std::deque<std::pair<uint32_t, uint32_t>> work;
  work.emplace_back(start, end);
  while (!work.empty() && pos.size() < F) {
    const auto r = work.front();
    work.pop_front();
    const uint32_t s = split_offset(r.first, r.second);
    if (!s) continue;
    pos.push_back(s);
    if (s - r.first >= F) work.emplace_back(r.first, s);
    if (r.second - s + 1 >= F) work.emplace_back(s + 1, r.second);
  }
  std::sort(pos.begin(), pos.end());
}

}  // namespace

bool bvh_topology(uint64_t n, uint32_t F, uint32_t max_depth, BvhTopology* out) {
  out->first.clear();
  out->count.clear();
  out->level_off.clear();
  if (n == 0) {
    out->level_off.push_back(0);
    return true;
  }
  struct Item { uint32_t s, e, depth; };
  std::vector<Item> q;
  q.reserve((size_t)(n / (F > 1 ? F / 2 : 1)) + 16);
  q.push_back({0, (uint32_t)(n - 1), 0});
  std::vector<uint32_t> pos;
  uint32_t cur_depth = 0;
  out->level_off.push_back(0);
  for (size_t id = 0; id < q.size(); ++id) {
    const Item it = q[id];
    if (it.depth != cur_depth) {
      cur_depth = it.depth;
      if (cur_depth >= max_depth) return false;
      out->level_off.push_back((uint32_t)id);
    }
    const uint64_t entries = (uint64_t)(it.e - it.s) + 1;
    if (entries <= F) {
      out->first.push_back(it.s);
      out->count.push_back((uint32_t)entries | 0x80000000u);
      continue;
    }
    split_positions(it.s, it.e, F, pos);
    out->first.push_back((uint32_t)q.size());
    out->count.push_back((uint32_t)(pos.size() - 1));
    for (size_t c = 0; c + 1 < pos.size(); ++c) {
      q.push_back({pos[c], pos[c + 1], it.depth + 1});
      pos[c + 1] += 1;
    }
  }
  out->level_off.push_back((uint32_t)q.size());
  return true;
}

}  // namespace qcp
