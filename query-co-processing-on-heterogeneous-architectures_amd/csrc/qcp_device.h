// qcp_device.h -- device-side helpers and kernel argument blocks shared by the
// build and search kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qcp {

constexpr int kWave = 64;
constexpr int kMaxLevels = 16;    // packed levels: F >= 16 and N < 2^32 give <= 8
constexpr int kMaxDepth = 48;     // explicit (BVH) trees: DFS frames per wave
constexpr uint32_t kLeafBit = 0x80000000u;

// Device-wide work queue state of one persistent search launch.
struct QueueCtl {
  uint32_t head;      // next donated-slot ticket
  uint32_t tail;      // donated slots claimed
  uint32_t pending;   // items not finished (root items + donations)
  uint32_t done;      // set once pending reaches zero
  uint32_t xhead[8];  // next root ticket of each XCD's query range
  unsigned long long donated;
  // statistics, one slot per XCD (summed on the host): spreads the end-of-
  // kernel atomics over eight addresses
  unsigned long long internal_entries[8];
  unsigned long long leaf_entries[8];
  unsigned long long total_hits[8];
  unsigned long long contained[8];  // contained subtree runs reported without a visit
};

// Tree + query batch as seen by the search kernels.
struct SearchArgs {
  const float* obj;          // [E][obj_stride] rank-ordered objects, SoA
  const float* nbox;         // [2D][node_stride] node MBRs, BFS order, SoA
  const uint32_t* nfirst;    // explicit layouts: first child node / first rank
  const uint32_t* ncount;    // explicit layouts: count | kLeafBit for leaves
  const uint32_t* nrlo;      // explicit layouts: first rank under each node
  const uint32_t* nrhi;      // explicit layouts: end rank (exclusive) under each node
  const uint32_t* nparent;   // explicit layouts: parent node (a root: itself)
  uint32_t parts;            // roots 0 .. parts-1 (MPHR data partitions; 1 otherwise)
  const float* queries;      // [nq][2D] row-major [lo(D), hi(D)]
  const uint32_t* qperm;     // processing order of the queries (nullptr = as given)
  uint64_t obj_stride;
  uint64_t node_stride;
  uint64_t n;                // objects
  uint32_t nq;
  uint32_t F;
  int32_t h;                 // node levels (packed); 0 = empty tree
  uint32_t n_nodes;
  uint32_t lvl_off[kMaxLevels + 1];
  uint32_t lvl_cnt[kMaxLevels];
  uint64_t span[kMaxLevels];   // ranks under one full node of each level
  uint32_t* counts;          // [nq]
  unsigned long long* hashes;// [nq]
  const unsigned long long* offsets;  // [nq+1] (ID write pass)
  uint32_t* ids;             // ID write pass output
  uint32_t* visits;          // plain form: [2*nq]
  QueueCtl* ctl;
  unsigned long long* slots; // donated items (q << 32 | node)
  uint32_t slot_cap;
  uint32_t donate;
  uint32_t max_waiters;      // waves kept waiting for donated work
  uint32_t hunger_period;    // internal-node descents between queue checks
  uint32_t donate_min_height;  // packed: donate subtrees this far above leaf parents
  uint32_t q_begin;          // plain form: first query of the launch
  uint32_t contain;          // report subtrees whose box lies inside the query without a visit
  // optional root items (query << 32 | subtree root node) instead of one root
  // item per query; results then accumulate (the caller zeroes them)
  const unsigned long long* items;
  uint32_t n_items;
};

// Order-independent fingerprint of a set of rank IDs: sum of squares plus
// kHashK times the sum, mod 2^64 (the sum over IDs of id^2 + kHashK*id).
// Kernels accumulate the two power sums and combine them at the end.  (A cube
// term would separate sets equal in count, sum and sum of squares, but its
// per-chunk 64-bit multiplies cost 13% of the 16M-point search; ID lists are
// compared instead wherever that matters.)
constexpr uint64_t kHashK = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ uint64_t set_hash(uint64_t s1, uint64_t s2) { return s2 + kHashK * s1; }

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// the wave's predicate mask straight from the compare (no int round trip)
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// popcount of the ballot bits below this lane
__device__ __forceinline__ uint32_t rank_below(uint64_t b) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
}

__device__ __forceinline__ uint32_t uni(uint32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    float o = __shfl_xor(v, off, kWave);
    v = (v > o) ? o : v;
  }
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    float o = __shfl_xor(v, off, kWave);
    v = (v < o) ? o : v;
  }
  return v;
}

// Box inside the query (closed): q.lo[d] <= lo[d] and hi[d] <= q.hi[d] for all
// d.  With lo <= hi per dimension for every object under a node, every object
// under a node whose box is inside the query overlaps the query.
template <int D>
__device__ __forceinline__ bool inside(const float* ql, const float* qh, const float* lo,
                                       const float* hi) {
  bool ok = true;
#pragma unroll
  for (int d = 0; d < D; ++d) ok = ok & (ql[d] <= lo[d]) & (hi[d] <= qh[d]);
  return ok;
}

// Power sums 1 + .. + m and 1 + .. + m^2 (mod 2^64) for m < 2^32 - 512 (the
// rank bound), all divisions exact and in 32 bits: one of m, m+1 is even, and
// 3 divides m, m+1 or 2m+1 (for m = 3t+1, (2m+1)/3 = 2t+1).
__device__ __forceinline__ uint64_t psum1(uint32_t m) {
  return (m & 1u) ? (uint64_t)m * ((m + 1u) >> 1) : (uint64_t)(m >> 1) * (m + 1u);
}
__device__ __forceinline__ uint64_t psum2(uint32_t m) {
  uint32_t a = m, b = m + 1u;
  if (m & 1u) b >>= 1; else a >>= 1;
  const uint32_t t = m / 3u, r = m - 3u * t;
  uint64_t c = 2ull * m + 1ull;
  if (r == 0u) a /= 3u; else if (r == 2u) b /= 3u; else c = 2ull * t + 1ull;
  return (uint64_t)a * b * c;
}

// The reference's closed-interval test: rejected iff q.lo > e.hi or q.hi < e.lo
// in some dimension (src/node/node_soa.cpp:94-113).
template <int D>
__device__ __forceinline__ bool overlaps(const float* ql, const float* qh,
                                         const float* lo, const float* hi) {
  bool ok = true;  // bitwise: no short-circuit branches
#pragma unroll
  for (int d = 0; d < D; ++d) ok = ok & !(ql[d] > hi[d]) & !(qh[d] < lo[d]);
  return ok;
}

// The same test as overlaps(), in vector arithmetic instead of per-dimension
// compare masks (which the scalar unit would have to combine): an entry is
// rejected iff max_d(max(q.lo[d] - hi[d], lo[d] - q.hi[d])) > 0.  With IEEE
// subtraction and denormals kept, x - y > 0 exactly when x > y (the difference
// of distinct floats never rounds to zero, and its sign is exact), and fmaxf
// drops a NaN operand, so a NaN coordinate rejects nothing, as in overlaps().
template <int D>
__device__ __forceinline__ bool overlap_test(const float* ql, const float* qh, const float* lo,
                                             const float* hi) {
  // one running maximum, so that each further dimension is a single v_max3
  // (fmaxf ignores NaN operands in any grouping, so the order is free)
  float t = fmaxf(ql[0] - hi[0], lo[0] - qh[0]);
#pragma unroll
  for (int d = 1; d < D; ++d) t = fmaxf(fmaxf(t, ql[d] - hi[d]), lo[d] - qh[d]);
  return !(t > 0.0f);
}

// Point v inside the closed box [pl, ph] given canonical bounds (no NaN):
// med3(v[d], pl[d], ph[d]) returns v[d] itself when pl[d] <= v[d] <= ph[d]
// and the bound it passed otherwise, so the point is inside iff every
// difference med3 - v is zero; the differences of distinct floats are never
// zero (denormals kept), and one D-operand max of their magnitudes decides
// all dimensions.  A difference is NaN only for an infinite coordinate equal
// to its bound (inf - inf), which is inside; fmaxf ignores it, and the final
// !(t > 0) accepts a max that is NaN.  With pl[d] = ql[d] (or -inf for a NaN
// ql[d]) and ph[d] = qh[d] (or +inf for a NaN qh[d]) and pl <= ph, this is the
// reference's test for a point (a lower and upper corner both equal to v):
// rejected iff ql[d] > v[d] or qh[d] < v[d] for some d
// (src/node/node_soa.cpp:94-113), in D med3, D subtractions, a max and a
// compare where overlap_test takes 2D subtractions, D maxes and a compare.
// A query with pl[d] > ph[d] rejects every point; callers handle it apart.
template <int D>
__device__ __forceinline__ bool point_test(const float* pl, const float* ph, const float* v) {
  float t = fabsf(__builtin_amdgcn_fmed3f(v[0], pl[0], ph[0]) - v[0]);
#pragma unroll
  for (int d = 1; d < D; ++d)
    t = fmaxf(t, fabsf(__builtin_amdgcn_fmed3f(v[d], pl[d], ph[d]) - v[d]));
  return !(t > 0.0f);
}

}  // namespace qcp
