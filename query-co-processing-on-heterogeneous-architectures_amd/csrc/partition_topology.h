// partition_topology.h -- MPHR data partitions as one explicit node structure.
//
// MPHR::Build (src/tree/mphr.cpp:60-108) splits the rank-sorted branches into
// P contiguous ranges -- chunk = N / P, the first range also taking N % P --
// and builds each range's tree bottom-up on its own (Bottom_Up, packed F per
// node), recording each subtree's root offset (root_offset[p], 94-97).  Here
// the P packed subtrees become one explicit tree: nodes are numbered
// depth-major (every partition's root first, so partition p's root is node p;
// then every partition's depth-1 nodes, partitions in order, and so on), so
// children stay contiguous and every depth is one contiguous node range for
// the bottom-up box kernels.  A leaf's `first` is its first global rank.
#pragma once
#include <stdint.h>

#include <vector>

namespace qcp {

struct PartitionTopology {
  std::vector<uint32_t> first;      // child node id (internal) or first rank (leaf)
  std::vector<uint32_t> count;      // entries, with 0x80000000 set for leaves
  std::vector<uint64_t> level_off;  // node offset of each depth, plus the total
};

// Rank range [*begin, *begin + *count) of partition p (MPHR::Build's split).
void partition_range(uint64_t n, uint32_t P, uint32_t p, uint64_t* begin, uint64_t* count);

// False if a partition would be empty (P > n) or deeper than max_depth.
bool partition_topology(uint64_t n, uint32_t F, uint32_t P, uint32_t max_depth,
                        PartitionTopology* out);

}  // namespace qcp
