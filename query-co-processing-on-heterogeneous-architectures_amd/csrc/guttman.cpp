// guttman.cpp -- host insertion build of the reference's Guttman R-tree.
//
// Semantics follow src/tree/rtree.h (RTree<float,float,D,float,128>):
//  - volume metric: the volume of the rectangle's bounding sphere,
//    r = sqrt(sum of squared half extents), V = r*r*unit (2-D) or
//    r*r*r*unit (3-D), every step in FP32 in that order
//    (RectSphericalVolume, rtree.h:1360-1389), with the unit-ball volumes
//    3.141593f and 4.188790f (UNIT_SPHERE_VOLUMES, rtree.h:636-648);
//  - descent: the entry whose rectangle grows least, ties to the smaller
//    rectangle, first index otherwise (PickBranch, rtree.h:1245-1280);
//  - a full node (F entries) plus the new entry is split quadratically
//    (SplitNode / ChoosePartition / PickSeeds / Classify / LoadNodes,
//    rtree.h:1305-1625) with minimum fill F/2: seeds are the pair wasting the
//    most volume, then the entry with the largest growth difference joins the
//    group it grows less (ties: the group with fewer entries), until one group
//    must take the rest; the first group stays in the old node and the second
//    becomes a new sibling appended to the parent;
//  - a split root grows a new root over [old root, new node] (InsertRect,
//    rtree.h:1122-1160); upper rectangles are exact covers.
// Where the reference leaves a choice undefined (seed or chosen entry never
// assigned because every comparison is false, i.e. NaN volumes), this build
// takes seeds (0, 1) and the first unassigned entry for the first group.
//
// The tree is built with explicit index arithmetic on flat arrays and an
// insertion path stack instead of the reference's recursive node objects.
#include "guttman.h"

#include <math.h>

#include <algorithm>

namespace qcp {
namespace {

constexpr uint32_t kLeaf = 0x80000000u;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr float kUnitBall2 = 3.141593f;
constexpr float kUnitBall3 = 4.188790f;

template <int D>
struct Rect {
  float lo[D], hi[D];
};

template <int D>
inline Rect<D> combine(const Rect<D>& a, const Rect<D>& b) {
  Rect<D> r;
  for (int d = 0; d < D; ++d) {
    r.lo[d] = std::min(a.lo[d], b.lo[d]);
    r.hi[d] = std::max(a.hi[d], b.hi[d]);
  }
  return r;
}

template <int D>
inline float volume(const Rect<D>& r) {
  float ss = 0.0f;
  for (int d = 0; d < D; ++d) {
    const float h = (r.hi[d] - r.lo[d]) * 0.5f;
    ss += h * h;
  }
  const float rad = sqrtf(ss);
  return D == 3 ? rad * rad * rad * kUnitBall3 : rad * rad * kUnitBall2;
}

template <int D>
class Builder {
 public:
  explicit Builder(uint32_t F) : F_(F), fill_(F / 2) {
    buf_.resize(F + 1);
    bref_.resize(F + 1);
    area_.resize(F + 1);
    group_.resize(F + 1);
    root_ = alloc(0);
  }

  void insert(const Rect<D>& r, uint32_t id) {
    path_.clear();
    uint32_t nd = root_;
    while (level_[nd] > 0) {
      const uint32_t i = pick(nd, r);
      path_.push_back({nd, i});
      nd = ref_[(uint64_t)nd * F_ + i];
    }
    uint32_t extra = add(nd, r, id);
    for (size_t k = path_.size(); k-- > 0;) {
      const uint32_t p = path_[k].node, i = path_[k].slot;
      Rect<D>& pr = rect_[(uint64_t)p * F_ + i];
      if (extra == kNone) {
        pr = combine(r, pr);
        continue;
      }
      pr = cover(ref_[(uint64_t)p * F_ + i]);
      const Rect<D> nc = cover(extra);
      extra = add(p, nc, extra);
    }
    if (extra != kNone) {
      const uint32_t old = root_;
      const Rect<D> c0 = cover(old), c1 = cover(extra);
      root_ = alloc(level_[old] + 1);
      add(root_, c0, old);
      add(root_, c1, extra);
    }
  }

  void transpose(uint64_t n, GuttmanTree* out) const {
    std::vector<uint32_t> bfs;
    bfs.reserve(level_.size());
    bfs.push_back(root_);
    out->first.clear();
    out->count.clear();
    out->order.clear();
    out->order.reserve(n);
    out->level_off.assign(1, 0);
    int cur = level_[root_];
    for (size_t h = 0; h < bfs.size(); ++h) {
      const uint32_t nd = bfs[h];
      if (level_[nd] != cur) {  // depth changes only between BFS runs
        out->level_off.push_back(h);
        cur = level_[nd];
      }
      const uint64_t base = (uint64_t)nd * F_;
      if (level_[nd] > 0) {
        out->first.push_back((uint32_t)bfs.size());
        out->count.push_back(count_[nd]);
        for (uint32_t i = 0; i < count_[nd]; ++i) bfs.push_back(ref_[base + i]);
      } else {
        out->first.push_back((uint32_t)out->order.size());
        out->count.push_back(count_[nd] | kLeaf);
        for (uint32_t i = 0; i < count_[nd]; ++i) out->order.push_back(ref_[base + i]);
      }
    }
    out->level_off.push_back(bfs.size());
  }

 private:
  struct Step {
    uint32_t node, slot;
  };

  uint32_t alloc(int level) {
    const uint32_t id = (uint32_t)level_.size();
    level_.push_back(level);
    count_.push_back(0);
    rect_.resize(rect_.size() + F_);
    ref_.resize(ref_.size() + F_);
    return id;
  }

  Rect<D> cover(uint32_t nd) const {
    const uint64_t base = (uint64_t)nd * F_;
    Rect<D> c = rect_[base];
    for (uint32_t i = 1; i < count_[nd]; ++i) c = combine(c, rect_[base + i]);
    return c;
  }

  uint32_t pick(uint32_t nd, const Rect<D>& r) const {
    const uint64_t base = (uint64_t)nd * F_;
    uint32_t best = 0;
    float best_inc = -1.0f, best_area = 0.0f;
    for (uint32_t i = 0; i < count_[nd]; ++i) {
      const float area = volume(rect_[base + i]);
      const float inc = volume(combine(r, rect_[base + i])) - area;
      if (i == 0 || inc < best_inc || (inc == best_inc && area < best_area)) {
        best = i;
        best_inc = inc;
        best_area = area;
      }
    }
    return best;
  }

  // Appends (r, ref) to node nd, splitting a full node; returns the new
  // sibling's id, or kNone when nd did not split.
  uint32_t add(uint32_t nd, const Rect<D>& r, uint32_t ref) {
    if (count_[nd] < F_) {
      const uint64_t at = (uint64_t)nd * F_ + count_[nd]++;
      rect_[at] = r;
      ref_[at] = ref;
      return kNone;
    }
    return split(nd, r, ref);
  }

  void assign(uint32_t i, int g) {
    group_[i] = (int8_t)g;
    gcover_[g] = cnt_[g] == 0 ? buf_[i] : combine(buf_[i], gcover_[g]);
    garea_[g] = volume(gcover_[g]);
    ++cnt_[g];
  }

  uint32_t split(uint32_t nd, const Rect<D>& r, uint32_t ref) {
    const uint32_t total = F_ + 1;
    const uint64_t base = (uint64_t)nd * F_;
    for (uint32_t i = 0; i < F_; ++i) {
      buf_[i] = rect_[base + i];
      bref_[i] = ref_[base + i];
    }
    buf_[F_] = r;
    bref_[F_] = ref;
    Rect<D> all = buf_[0];
    for (uint32_t i = 1; i < total; ++i) all = combine(all, buf_[i]);
    const float all_area = volume(all);
    // seeds: the pair that wastes the most volume together
    for (uint32_t i = 0; i < total; ++i) area_[i] = volume(buf_[i]);
    float worst = -all_area - 1.0f;
    uint32_t s0 = 0, s1 = 1;
    for (uint32_t a = 0; a + 1 < total; ++a)
      for (uint32_t b = a + 1; b < total; ++b) {
        const float waste = volume(combine(buf_[a], buf_[b])) - area_[a] - area_[b];
        if (waste > worst) {
          worst = waste;
          s0 = a;
          s1 = b;
        }
      }
    std::fill(group_.begin(), group_.end(), (int8_t)-1);
    cnt_[0] = cnt_[1] = 0;
    garea_[0] = garea_[1] = 0.0f;
    assign(s0, 0);
    assign(s1, 1);
    const uint32_t cap = total - fill_;
    while (cnt_[0] + cnt_[1] < total && cnt_[0] < cap && cnt_[1] < cap) {
      float biggest = -1.0f;
      uint32_t chosen = kNone;
      int better = 0;
      for (uint32_t i = 0; i < total; ++i) {
        if (group_[i] >= 0) continue;
        const float g0 = volume(combine(buf_[i], gcover_[0])) - garea_[0];
        const float g1 = volume(combine(buf_[i], gcover_[1])) - garea_[1];
        float diff = g1 - g0;
        int g = 0;
        if (!(diff >= 0.0f)) {
          g = 1;
          diff = -diff;
        }
        if (diff > biggest) {
          biggest = diff;
          chosen = i;
          better = g;
        } else if (diff == biggest && cnt_[g] < cnt_[better]) {
          chosen = i;
          better = g;
        }
      }
      if (chosen == kNone) {  // every difference NaN
        for (uint32_t i = 0; i < total && chosen == kNone; ++i)
          if (group_[i] < 0) chosen = i;
        better = 0;
      }
      assign(chosen, better);
    }
    if (cnt_[0] + cnt_[1] < total) {
      const int g = cnt_[0] >= cap ? 1 : 0;
      for (uint32_t i = 0; i < total; ++i)
        if (group_[i] < 0) assign(i, g);
    }
    // the first group stays in nd, the second fills a new node of the same level
    const uint32_t sib = alloc(level_[nd]);
    count_[nd] = 0;
    for (uint32_t i = 0; i < total; ++i) add(group_[i] == 0 ? nd : sib, buf_[i], bref_[i]);
    return sib;
  }

  uint32_t F_, fill_;
  uint32_t root_ = 0;
  std::vector<int> level_;
  std::vector<uint32_t> count_;
  std::vector<Rect<D>> rect_;
  std::vector<uint32_t> ref_;
  std::vector<Step> path_;
  // split state
  std::vector<Rect<D>> buf_;
  std::vector<uint32_t> bref_;
  std::vector<float> area_;
  std::vector<int8_t> group_;
  Rect<D> gcover_[2];
  float garea_[2];
  uint32_t cnt_[2];
};

template <int D>
void build_d(const float* objs, uint64_t n, bool box, uint32_t F, GuttmanTree* out) {
  Builder<D> b(F);
  const uint32_t E = box ? 2 * D : D;
  for (uint64_t i = 0; i < n; ++i) {
    Rect<D> r;
    for (int d = 0; d < D; ++d) {
      r.lo[d] = objs[i * E + d];
      r.hi[d] = objs[i * E + (box ? D : 0) + d];
    }
    b.insert(r, (uint32_t)i);
  }
  b.transpose(n, out);
}

}  // namespace

bool guttman_build(const float* objs, uint64_t n, uint32_t D, bool box, uint32_t F,
                   GuttmanTree* out) {
  // a minimum fill of 1 (F < 4) lets single-child chains grow without bound
  if (F < 4 || (D != 2 && D != 3) || n >= 0xFFFFFFFFull) return false;
  if (n == 0) {
    out->first.clear();
    out->count.clear();
    out->order.clear();
    out->level_off.assign(1, 0);
    return true;
  }
  if (D == 2)
    build_d<2>(objs, n, box, F, out);
  else
    build_d<3>(objs, n, box, F, out);
  return true;
}

bool guttman_check(const uint32_t* first, const uint32_t* count, uint64_t m, uint64_t n,
                   uint32_t F, uint32_t max_depth, std::vector<uint64_t>* level_off) {
  level_off->assign(1, 0);
  if (m == 0) return n == 0;
  std::vector<uint32_t> depth(m, 0);
  uint64_t next_child = 1, next_pos = 0;
  uint32_t leaf_depth = kNone;
  for (uint64_t i = 0; i < m; ++i) {
    if (i > 0 && i >= next_child) return false;  // not referenced by an earlier node
    if (i > 0 && depth[i] < depth[i - 1]) return false;
    if (i > 0 && depth[i] != depth[i - 1]) level_off->push_back(i);
    if (depth[i] > max_depth) return false;
    const uint32_t c = count[i] & ~kLeaf;
    if (c == 0 || c > F) return false;
    if (count[i] & kLeaf) {
      if (first[i] != next_pos) return false;
      if (leaf_depth == kNone) leaf_depth = depth[i];
      if (depth[i] != leaf_depth) return false;
      next_pos += c;
    } else {
      if (first[i] != next_child || next_child + c > m) return false;
      for (uint32_t k = 0; k < c; ++k) depth[next_child + k] = depth[i] + 1;
      next_child += c;
    }
  }
  level_off->push_back(m);
  return next_child == m && next_pos == n;
}

}  // namespace qcp
