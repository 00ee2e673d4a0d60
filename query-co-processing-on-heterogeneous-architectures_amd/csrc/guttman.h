// guttman.h -- the reference R-tree's insertion build (Guttman, quadratic
// split), producing the explicit BFS node arrays the search kernels read.
//
// The reference inserts the Hilbert-sorted branches one at a time into
// RTree<float, float, D, float, 128> (Tree::RTree_Top_Down,
// src/tree/tree.cpp:126-185; RTree::Insert / InsertRectRec / InsertRect /
// AddBranch / PickBranch / SplitNode / ChoosePartition / LoadNodes / PickSeeds /
// Classify, src/tree/rtree.h:659-681,1061-1280,1305-1625), transposes the
// result breadth-first into node::Node[] (Transpose, rtree.h:163-221) and
// numbers the leaf entries 1..N depth-first (Tree::SetNodeIndex,
// src/tree/tree.cpp:244-256).  All leaves sit on one level, so the BFS order
// of the leaves is their left-to-right order and the depth-first number of an
// entry is its position in the concatenated leaves.
#pragma once
#include <stdint.h>

#include <vector>

namespace qcp {

struct GuttmanTree {
  std::vector<uint32_t> first;      // child node id (internal) or first position (leaf)
  std::vector<uint32_t> count;      // entries, with 0x80000000 set for leaves
  std::vector<uint64_t> level_off;  // BFS offset of each depth, plus the total
  std::vector<uint32_t> order;      // depth-first position -> insertion index
};

// Inserts objs[i] (n rows of E floats: points [p(D)], boxes [lo(D), hi(D)]) in
// index order into a tree of node capacity F (minimum fill F/2).  n == 0
// gives an empty tree.  Returns false for D other than 2 or 3, or F < 4.
bool guttman_build(const float* objs, uint64_t n, uint32_t D, bool box, uint32_t F,
                   GuttmanTree* out);

// Checks that (first, count) over m nodes is a BFS tree whose leaves hold
// positions [0, n) in order and whose nodes hold at most F entries, with all
// leaves on one level; fills level_off.  Used for trees read from a file.
bool guttman_check(const uint32_t* first, const uint32_t* count, uint64_t m, uint64_t n,
                   uint32_t F, uint32_t max_depth, std::vector<uint64_t>* level_off);

}  // namespace qcp
