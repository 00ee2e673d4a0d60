// bvh_topology.h -- the reference BVH's node structure over N Hilbert-ranked
// entries, in BFS order.
//
// The reference builds the BVH top-down from the rank-sorted branches
// (Tree::CreateNode, src/tree/tree.cpp:352-408): a node with <= F entries is a
// leaf; otherwise GetSplitPosition (304-350) runs a BFS of binary splits, each
// at the highest position whose rank shares more leading bits with the first
// rank than the last rank does (GetSplitOffset, 267-302), until F split points
// exist.  Because the branch index is the 1-based rank by then, the topology is
// a function of (N, F) alone; only the MBRs depend on the data.  The same split
// rule is implemented here (the rule itself must match for the trees to be
// identical), emitted in BFS order so that siblings are contiguous.
#pragma once
#include <stdint.h>

#include <vector>

namespace qcp {

struct BvhTopology {
  std::vector<uint32_t> first;   // child node id (internal) or first rank (leaf)
  std::vector<uint32_t> count;   // entries, with 0x80000000 set for leaves
  std::vector<uint32_t> level_off;  // BFS offset of each depth, plus the total
};

// Returns false if the tree would be deeper than max_depth.
bool bvh_topology(uint64_t n, uint32_t F, uint32_t max_depth, BvhTopology* out);

}  // namespace qcp
