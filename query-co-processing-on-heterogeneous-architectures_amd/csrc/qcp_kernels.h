// qcp_kernels.h -- host-callable launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qcp_device.h"

namespace qcp {

// Launch shape of the persistent search when the caller leaves it at 0: eight
// waves a workgroup and eight 64-entry leaf chunks in flight a wave, the shape
// the autotuner picks on the 16.8M-point 3-D, 67M-box (R-tree) and 134M-point
// 2-D configurations (each 2 to 8% faster than four and four).
constexpr unsigned kDefaultBlockWaves = 8;
constexpr int kDefaultLeafChunks = 8;

// build
// flag bits of the key kernel's `bad` word
constexpr uint32_t kBadDomain = 1u;    // NaN, or outside the quantiser's domain
constexpr uint32_t kBadInverted = 2u;  // a box with lo > hi in some dimension
// keys[i] of each object and, when idx is non-null, idx[i] = i; the OR of all
// keys into *key_or and domain / inverted-box flags into *bad.  With ghist
// (kSortGhistWords words, zero at entry) the kernel also counts the digits of
// every byte a key can hold and zeroes the sort's first look-back words
// (status0, status_words = sort_tiles(n) * 256), so that the sort of these
// keys runs with `counted` set.
hipError_t launch_keys(const float* objs, uint64_t n, unsigned D, bool box, bool ref,
                       bool center, unsigned nb, uint64_t* keys, uint32_t* idx,
                       unsigned long long* key_or, uint32_t* bad, hipStream_t s,
                       uint32_t* ghist = nullptr, uint64_t* status0 = nullptr,
                       uint64_t status_words = 0);
uint32_t sort_tiles(uint64_t n);
hipError_t launch_assign(const float* objs, uint64_t n, unsigned D, bool box, bool center,
                         const float* centers, uint32_t K, uint32_t* cid, uint32_t* changes,
                         hipStream_t s);
hipError_t launch_compose(uint64_t* keys, const uint32_t* cid, const uint32_t* rank, uint64_t n,
                          unsigned shift_key, unsigned shift_rank, hipStream_t s);
// k-means on the device.  launch_km_seed: centers[k] = key point of object
// idx[k].  launch_km_update: the centre update of chunks [c0, c0 + nc) of
// kKmChunk objects (FP64 sums in index order within a chunk, into part
// [nc][K][D] / cnt [nc][K], zero at entry), then their partials added in chunk
// order into sum [K][D] / total [K].  launch_km_finish: centres = sum / total
// for every cluster with members.
constexpr uint64_t kKmChunk = 65536;
hipError_t launch_km_seed(const float* objs, unsigned D, bool box, bool center,
                          const uint32_t* idx, uint32_t K, float* centers, hipStream_t s);
hipError_t launch_km_update(const float* objs, uint64_t n, unsigned D, bool box, bool center,
                            const uint32_t* cid, uint32_t K, uint64_t c0, uint64_t nc,
                            double* part, uint32_t* cnt, double* sum, uint64_t* total,
                            hipStream_t s);
hipError_t launch_km_finish(const double* sum, const uint64_t* total, uint32_t K, unsigned D,
                            float* centers, hipStream_t s);
// Whole LSD radix sort of (keys, vals) over `passes` 8-bit digits, one-sweep
// form; the result is in keys/vals for an even pass count, keys2/vals2 for an
// odd one.  ghist: kMaxSortPasses*256 + kMaxSortPasses words; status:
// sort_status_words(n) words.
// pack_pass >= 0 (key_or null): pass pack_pass writes packed elements
// ((key >> 8*(pack_pass+1)) << pbits) | val, which the remaining passes move
// alone; the result is then packed, in keys (even pass count) or keys2, and
// the caller unpacks it in launch_gather_objs.  Needs
// key bits - 8*(pack_pass+1) + pbits <= 64 and vals < 2^pbits.
// key_or non-null (graph-captured builds): the device word holding the OR of
// all keys decides on the device how many passes reorder (the rest end at
// once, leaving the result where the last reordering pass put it) and, when
// pbits > 0, from which pass the elements are packed; launch_gather_objs_planned
// finds the result and unpacks it.
uint64_t sort_status_words(uint64_t n);
// counted: the digit counts are already in ghist (launch_keys with ghist), and
// ghist's tile counters and the first look-back words are zero.
hipError_t launch_sort_keys(uint64_t* keys, uint32_t* vals, uint64_t* keys2, uint32_t* vals2,
                            uint64_t n, unsigned passes, const unsigned long long* key_or,
                            uint32_t* ghist, uint64_t* status, hipStream_t s, int pack_pass = -1,
                            unsigned pbits = 0, bool counted = false);
constexpr uint32_t kSortGhistWords = 8 * 256 + 8;
// Object-carrying LSD radix sort: every pass moves the key bits still to be
// sorted and the objects (plus, with idx, their input indices) themselves, and
// the last pass writes the rank-ordered SoA objects (and the permutation), so
// the build needs no gather.  Scratch set j holds the key remainders
// (krem[j], n 8-byte words, read as 4-byte words once at most 32 bits remain)
// and the payload words (pay[j], [E (+1 with idx)][cs] 4-byte words).
// key_or null: `bits` is the key width and passes = ceil(bits / 8); otherwise
// the device word holding the OR of all keys decides how many passes reorder
// (the others end at once) and which one is last (graph-captured builds,
// which also launch launch_carry_identity for the all-zero case).
struct CarryArgs {
  const uint64_t* keys;    // pass 0: the keys
  const float* objs;       // pass 0: the objects, row-major
  uint64_t* krem[2];
  uint32_t* pay[2];
  uint64_t cs;             // payload row stride of the scratch sets
  float* obj;              // last pass: objects [E][ostride]
  uint64_t ostride;
  uint32_t* perm;          // last pass (idx): input index of each rank
  uint64_t n;
  const unsigned long long* key_or;
  unsigned bits;
  uint32_t* ghist;         // kSortGhistWords words
  uint64_t* status;        // sort_status_words(n) words
  uint64_t words;          // look-back words of one pass: sort_tiles(n) * 256
};
hipError_t launch_sort_carry(const CarryArgs& a, unsigned D, bool box, bool idx,
                             unsigned passes, hipStream_t s, bool counted = false);
// objects in input order (and perm_out = identity) when *key_or == 0; nothing otherwise
hipError_t launch_carry_identity(const float* objs, uint64_t n, const unsigned long long* key_or,
                                 unsigned D, bool box, float* obj, uint64_t stride,
                                 uint32_t* perm_out, hipStream_t s);
hipError_t scan_u32_u64(const uint32_t* in, uint64_t n, unsigned long long* out,
                        unsigned long long* block_sums, uint32_t max_blocks, hipStream_t s);
// The build's gather: obj[e][r] = objs[perm[r]][e], with the permutation
// either given (perm; null = the identity) or unpacked from packed sort elements (packed, pbits;
// written to perm_out; the sorted keys are then left to launch_keys_sorted).
// fuse_leaves = 32 or 64 (packed layouts of that fanout): also writes the
// leaf level's boxes (nbox at leaf_off), as launch_leaf_boxes would.
hipError_t launch_gather_objs(const float* objs, const uint32_t* perm, const uint64_t* packed,
                              uint64_t n, unsigned pbits, unsigned D, bool box, float* obj,
                              uint64_t stride, uint32_t* perm_out, uint32_t fuse_leaves,
                              float* nbox, uint64_t nstride, uint64_t leaf_off, hipStream_t s);
// launch_gather_objs after a device-planned sort (launch_sort_keys with key_or
// over keys/packed, keys2/packed2 and vals/perm, vals2/perm2): the plan says
// whether the result is packed or a permutation and in which buffer; the
// permutation ends in perm either way.
hipError_t launch_gather_objs_planned(const float* objs, uint32_t* perm, const uint32_t* perm2,
                                      const uint64_t* packed, const uint64_t* packed2,
                                      uint64_t n, unsigned pbits, unsigned passes,
                                      const unsigned long long* key_or, unsigned D, bool box,
                                      float* obj, uint64_t stride, uint32_t fuse_leaves,
                                      float* nbox, uint64_t nstride, uint64_t leaf_off,
                                      hipStream_t s);
// Sorted keys from the rank-ordered SoA objects (the key of each rank's object).
hipError_t launch_keys_sorted(const float* obj, uint64_t stride, uint64_t n, unsigned D,
                              bool box, bool ref, bool center, unsigned nb, uint64_t* keys,
                              hipStream_t s);
hipError_t launch_leaf_boxes(const float* obj, uint64_t ostride, uint64_t n, unsigned D,
                             bool box, uint32_t F, uint64_t nleaves, float* nbox,
                             uint64_t nstride, uint64_t leaf_off, hipStream_t s);
hipError_t launch_level_boxes(float* nbox, uint64_t nstride, unsigned D, uint64_t child_off,
                              uint64_t child_cnt, uint32_t F, uint64_t parent_off,
                              uint64_t parent_cnt, hipStream_t s);
// F sizes the per-node loop (counts <= F); nodes with fewer than pad entries
// are padded 1.0f / 0.0f (pad = F for the BVH, 0 for exact covers).
hipError_t launch_tree_boxes(const float* obj, uint64_t ostride, const uint32_t* nfirst,
                             const uint32_t* ncount, unsigned D, bool box, uint32_t F,
                             uint32_t pad, uint64_t lo, uint64_t hi, float* nbox,
                             uint64_t nstride, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
hipError_t launch_iota_u32(uint32_t* p, uint64_t n, hipStream_t s);  // p[i] = i
hipError_t launch_fill_f32(float* p, uint64_t n, float v, hipStream_t s);

// search
// out: 0 counts, 1 counts + ID-set hash, 2 sorted ID lists (needs offsets);
// u: 64-entry leaf chunks in flight per wave (2, 4 or 8)
hipError_t launch_search(const SearchArgs& a, unsigned D, bool box, bool expl, int out, int u,
                         unsigned grid, unsigned block, hipStream_t s);
// the plain form over a partitioned MPHR: one workgroup of `waves` waves per
// partition (parts of them) for query a.q_begin; per-workgroup results in
// a.counts[p], a.hashes[p], a.visits[2p..2p+1]
hipError_t launch_mprs_part(const SearchArgs& a, unsigned D, bool box, unsigned parts,
                            unsigned waves, hipStream_t s);
// the plain form: nblocks workgroups of `waves` (1 or 3) waves, one per query
hipError_t launch_mprs(const SearchArgs& a, unsigned D, bool box, unsigned nblocks,
                       unsigned waves, hipStream_t s);
const void* search_kernel_ptr(unsigned D, bool box, bool expl, int out, int mw);
// dynamic LDS bytes of a persistent-search workgroup of `block` threads
size_t search_lds_bytes(bool expl, int mw, int out, unsigned block);

}  // namespace qcp
