// build_kernels.hip -- GPU bulk-load pipeline for gfx950:
//   Hilbert keys (LDS transition table) -> stable one-sweep LSD radix sort
//   (8-bit digits, wave-level multisplit ranking, decoupled look-back) ->
//   gather into rank-ordered SoA objects -> node MBRs (packed levels
//   bottom-up, or explicit BVH / Guttman levels deepest first).  The same
//   sort orders query batches by their centres' keys.
//
// Reference steps replaced: Tree::AssignHilbertIndexToBranches
// (src/tree/tree.cpp:596-639), Sorter::Sort + ReassignHilbertIndexes
// (src/sort/sorter.cpp:11-29, src/sort/thrust_sorter.cpp:12-47),
// CopyBranchToLeafNode + Transformer::Transform (src/tree/tree.cpp:683-745,
// src/transformer/transformer.cpp:13-88), global_BottomUpBuild_ILP
// (src/tree/tree.cpp:1025-1133) and Node::GetMBB (src/node/node.cpp:19-70).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "hilbert_key.h"
#include "qcp_device.h"
#include "qcp_kernels.h"

namespace qcp {

// ---------------------------------------------------------------------------
// Hilbert keys.  Key source is the lower corner (the reference reads only the
// first D entries of the branch, src/tree/tree.cpp:600) or the box centre
// (lo+hi)*0.5f.  The walk's transitions over L levels, indexed by each
// coordinate's L bits, and over the top nb % L levels from the initial state
// (hilbert_key.h) are evaluated at compile time into constant memory, in the
// form with the parity of the bits above (so the walk itself yields the
// inverse Gray code); each workgroup copies them into LDS (8 / 6 KB), and a
// key then takes one start lookup, nb/L table lookups and one constant XOR,
// with no bit interleaving.  (Computing the 1536 / 2048 entries in every
// workgroup cost about 47 ns of the device per workgroup, 0.39 ms for the key
// kernel's 8192 workgroups whatever the number of keys.)
template <int D>
struct alignas(16) KeyTabConst {
  uint16_t tab[HilbertTable<D>::PSIZE];
  constexpr KeyTabConst() : tab() {
    for (uint32_t e = 0; e < HilbertTable<D>::PSIZE; ++e) tab[e] = hilbert_ptable_entry_cm<D>(e);
  }
};
// start table of the top TOP levels (a placeholder entry when TOP = 0)
template <int D, int TOP>
struct alignas(16) KeyStartConst {
  uint16_t start[1u << (D * (TOP ? TOP : 1))];
  constexpr KeyStartConst() : start() {
    if (TOP > 0)
      for (uint32_t e = 0; e < (1u << (D * TOP)); ++e) start[e] = hilbert_pstart_entry<D>(TOP, e);
  }
};
__constant__ const KeyTabConst<2> kKeyTab2 = KeyTabConst<2>();
__constant__ const KeyTabConst<3> kKeyTab3 = KeyTabConst<3>();
// the only key widths with nb % L != 0: 2-D REF keys (31 bits, L = 4)
__constant__ const KeyStartConst<2, 3> kKeyStart2r = KeyStartConst<2, 3>();

template <int D, int NB>
struct alignas(4) KeyTables {
  static constexpr int TOP = NB % HilbertTable<D>::L;
  uint16_t tab[HilbertTable<D>::PSIZE];
  uint16_t start[1u << (D * (TOP ? TOP : 1))];
};

template <int D, int NB>
__device__ __forceinline__ void fill_key_tables(KeyTables<D, NB>& t) {
  constexpr int TOP = KeyTables<D, NB>::TOP;
  static_assert(TOP == 0 || (D == 2 && TOP == 3), "start tables exist for 2-D 31-bit keys");
  const uint32_t* src = reinterpret_cast<const uint32_t*>(D == 2 ? kKeyTab2.tab : kKeyTab3.tab);
  uint32_t* dst = reinterpret_cast<uint32_t*>(t.tab);
  for (uint32_t e = threadIdx.x; e < HilbertTable<D>::PSIZE / 2; e += blockDim.x) dst[e] = src[e];
  if (TOP > 0)
    for (uint32_t e = threadIdx.x; e < (1u << (D * TOP)); e += blockDim.x)
      t.start[e] = kKeyStart2r.start[e];
  __syncthreads();
}

// Bits a coordinate of the keys: the reference's 31 (2-D) / 20 (3-D) for REF
// keys, 16 for b16 keys.
template <int D, bool REF>
constexpr unsigned key_nb() {
  return REF ? (D == 2 ? 31u : 20u) : 16u;
}

// Point c of one object v[E]; `bad` set for a coordinate outside the REF
// quantiser's domain.
template <int D, bool BOX, bool REF, bool CENTER>
__device__ __forceinline__ void object_point(const float* v, uint32_t (&c)[D], bool& bad) {
  constexpr unsigned nb = key_nb<D, REF>();
#pragma unroll
  for (int d = 0; d < D; ++d) {
    float p = v[d];
    if (CENTER && BOX) p = (v[d] + v[D + d]) * 0.5f;
    if (REF) {
      bad |= !quant_ref(p, nb, &c[d]);
    } else {
      c[d] = quant_b16(p);
    }
  }
}

// Keys of kKeysPerThread objects a thread (objects base + r*256 + tid of each
// 256 x kKeysPerThread block of the grid-stride loop), their walks interleaved.
constexpr int kKeysPerThread = 4;

// HIST: also the digit counts of every byte a key of this width can hold,
// for the radix sort that follows (ghist[p][256], zero at entry; the sort
// then skips its own counting pass, which read every key again), and zero the
// sort's first look-back words (status0[0 .. status_words)).  A byte whose
// digit all of a wave's keys share is added once for the wave (the zero bytes
// above the keys' width, and clustered keys' high bytes), others one atomic a
// key, in LDS; each workgroup adds its counts to ghist once at the end.
template <int D, bool BOX, bool REF, bool CENTER, bool HIST>
__global__ __launch_bounds__(256) void k_keys(const float* __restrict__ objs, uint64_t n,
                                              uint64_t* __restrict__ keys,
                                              uint32_t* __restrict__ idx,
                                              unsigned long long* key_or, uint32_t* bad,
                                              uint32_t* __restrict__ ghist,
                                              uint64_t* __restrict__ status0,
                                              uint64_t status_words) {
  constexpr int E = BOX ? 2 * D : D;
  constexpr int R = kKeysPerThread;
  constexpr int NB = key_nb<D, REF>();
  constexpr int NBYTES = (D * NB + 7) / 8;
  __shared__ KeyTables<D, NB> kt;
  __shared__ uint32_t hh[HIST ? NBYTES : 1][256];
  if (HIST) {
#pragma unroll
    for (int p = 0; p < NBYTES; ++p) hh[p][threadIdx.x] = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < status_words;
         i += (uint64_t)gridDim.x * 256)
      status0[i] = 0;
  }
  fill_key_tables<D, NB>(kt);  // ends with a barrier
  uint64_t acc = 0;
  bool any_bad = false, any_inv = false;
  for (uint64_t base = (uint64_t)blockIdx.x * (256 * R); base < n;
       base += (uint64_t)gridDim.x * (256 * R)) {
    uint32_t c[R][D];
    float v[R][E];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t i = base + (uint64_t)r * 256 + threadIdx.x;
      const uint64_t j = i < n ? i : n - 1;  // tail lanes read a valid object, store nothing
#pragma unroll
      for (int e = 0; e < E; ++e) v[r][e] = objs[j * E + e];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      // a tail lane repeats object n - 1, which a valid lane also checks
#pragma unroll
      for (int e = 0; e < E; ++e) any_bad |= (v[r][e] != v[r][e]);  // NaN objects are rejected
      if (BOX) {
#pragma unroll
        for (int d = 0; d < D; ++d) any_inv |= v[r][d] > v[r][D + d];
      }
      object_point<D, BOX, REF, CENTER>(v[r], c[r], any_bad);
    }
    uint64_t k[R];
    hilbert_keys_pcm<D, NB, R>(c, kt.tab, kt.start, k);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t i = base + (uint64_t)r * 256 + threadIdx.x;
      if (i < n) {
        keys[i] = k[r];
        if (idx != nullptr) idx[i] = (uint32_t)i;  // callers whose sort reads input indices
        acc |= k[r];
      }
    }
    if (HIST) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const bool valid = base + (uint64_t)r * 256 + threadIdx.x < n;
        const uint64_t vmask = ballot(valid);
        if (vmask == 0) continue;
        const int first = __builtin_ctzll(vmask);
        const uint32_t lo = (uint32_t)k[r], hi = (uint32_t)(k[r] >> 32);
        const uint32_t xl = lo ^ (uint32_t)__builtin_amdgcn_readlane((int)lo, first);
        const uint32_t xh = hi ^ (uint32_t)__builtin_amdgcn_readlane((int)hi, first);
#pragma unroll
        for (int p = 0; p < NBYTES; ++p) {
          const uint32_t word = p < 4 ? lo : hi, x = p < 4 ? xl : xh;
          const int sh = 8 * (p % 4);
          const uint32_t dig = (word >> sh) & 0xFFu;
          if (ballot(valid && ((x >> sh) & 0xFFu) != 0) == 0) {
            if ((int)lane_id() == first) atomicAdd(&hh[p][dig], (uint32_t)__popcll(vmask));
          } else if (valid) {
            atomicAdd(&hh[p][dig], 1u);
          }
        }
      }
    }
  }
  if (HIST) {
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NBYTES; ++p) {
      const uint32_t v = hh[p][threadIdx.x];
      if (v) atomicAdd(&ghist[p * 256 + threadIdx.x], v);
    }
  }
  // wave OR, then one atomic per wave, and only when it adds bits: every wave
  // of the grid ORing into one word serialises at the L2 (32768 such atomics
  // took 0.33 ms), while the OR of the keys holds all its bits after a few
  // waves.  Bits only ever join the word, so a wave whose bits are all there
  // has nothing to add.
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc |= __shfl_xor(acc, off, kWave);
  const bool wave_bad = ballot(any_bad) != 0, wave_inv = ballot(any_inv) != 0;
  if (lane_id() == 0) {
    const unsigned long long ko =
        __hip_atomic_load(key_or, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((ko | acc) != ko) atomicOr(key_or, (unsigned long long)acc);
    const uint32_t flags = (wave_bad ? kBadDomain : 0u) | (wave_inv ? kBadInverted : 0u);
    if (flags) {
      const uint32_t b = __hip_atomic_load(bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((b | flags) != b) atomicOr(bad, flags);
    }
  }
}

// ---------------------------------------------------------------------------
// Radix sort.  Tile = 8192 consecutive keys (512 threads x 16); a pass is
// stable because each tile is ranked in index order and tiles are placed in
// tile order.  Threads 0..255 own one digit each in the per-tile steps.
constexpr int kSortThreads = 512;
constexpr int kSortKPT = 16;
constexpr int kSortTile = kSortThreads * kSortKPT;
// the digit-count kernel's workgroup and work unit
constexpr int kHistThreads = 256;
constexpr int kHistKPT = 16;
constexpr int kHistTile = kHistThreads * kHistKPT;


// ---------------------------------------------------------------------------
// Device-wide exclusive scan (three phases).  Each block owns `per_block`
// consecutive elements, processed 2048 at a time.
constexpr int kScanThreads = 256;
constexpr int kScanIPT = 8;
constexpr int kScanChunk = kScanThreads * kScanIPT;

template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* sh, T* total) {
  // sh has kScanThreads entries
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    T t = tid >= off ? sh[tid - off] : (T)0;
    __syncthreads();
    sh[tid] += t;
    __syncthreads();
  }
  T incl = sh[tid];
  *total = sh[kScanThreads - 1];
  __syncthreads();
  return incl - v;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const TI* __restrict__ in,
                                                              uint64_t n, uint64_t per_block,
                                                              TO* __restrict__ sums) {
  __shared__ TO sh[kScanThreads];
  const uint64_t b0 = (uint64_t)blockIdx.x * per_block;
  const uint64_t b1 = b0 + per_block < n ? b0 + per_block : n;
  TO acc = 0;
  for (uint64_t i = b0 + threadIdx.x; i < b1; i += kScanThreads) acc += (TO)in[i];
  TO tot;
  block_excl_scan<TO>(acc, sh, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// single block: exclusive scan of nb <= kScanThreads*64 block sums, in place;
// total written to sums[nb].
template <typename TO>
__global__ __launch_bounds__(kScanThreads) void k_scan_top(TO* sums, uint32_t nb) {
  __shared__ TO sh[kScanThreads];
  const uint32_t per = (nb + kScanThreads - 1) / kScanThreads;
  const uint32_t s0 = threadIdx.x * per;
  TO acc = 0;
  for (uint32_t i = 0; i < per; ++i)
    if (s0 + i < nb) acc += sums[s0 + i];
  TO tot;
  TO ex = block_excl_scan<TO>(acc, sh, &tot);
  for (uint32_t i = 0; i < per; ++i) {
    if (s0 + i < nb) {
      TO t = sums[s0 + i];
      sums[s0 + i] = ex;
      ex += t;
    }
  }
  if (threadIdx.x == 0) sums[nb] = tot;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const TI* __restrict__ in,
                                                             uint64_t n, uint64_t per_block,
                                                             const TO* __restrict__ sums,
                                                             TO* __restrict__ out) {
  __shared__ TO sh[kScanThreads];
  const uint64_t b0 = (uint64_t)blockIdx.x * per_block;
  const uint64_t b1 = b0 + per_block < n ? b0 + per_block : n;
  TO carry = sums[blockIdx.x];
  for (uint64_t c0 = b0; c0 < b1; c0 += kScanChunk) {
    // each thread owns kScanIPT consecutive elements of this chunk
    const uint64_t t0 = c0 + (uint64_t)threadIdx.x * kScanIPT;
    TO v[kScanIPT];
    TO loc = 0;
#pragma unroll
    for (int j = 0; j < kScanIPT; ++j) {
      v[j] = (t0 + j < b1) ? (TO)in[t0 + j] : (TO)0;
      loc += v[j];
    }
    TO tot;
    TO ex = block_excl_scan<TO>(loc, sh, &tot) + carry;
#pragma unroll
    for (int j = 0; j < kScanIPT; ++j) {
      if (t0 + j < b1) out[t0 + j] = ex;
      ex += v[j];
    }
    carry += tot;
  }
  // the last block also writes the grand total after the end (out[n]) when
  // the caller sized out for n+1 elements
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0 && sums != nullptr) {
    // sums[gridDim.x] holds the grand total (k_scan_top)
    out[n] = sums[gridDim.x];
  }
}

// ---------------------------------------------------------------------------
// The passes of a sort whose plan is made on the device from the OR of all
// keys (graph-captured builds, which run a pass for every byte a key can
// hold): `real` passes reorder, and the elements are packed from pass `pack`
// on (-1: never; pbits = 0 disables packing).
struct SortPlan {
  unsigned real;
  int pack;
};

__device__ __forceinline__ SortPlan sort_plan(unsigned passes, unsigned pbits,
                                              const unsigned long long* key_or) {
  const unsigned long long ko = *key_or;
  const unsigned bits = ko ? 64u - (unsigned)__clzll((long long)ko) : 0u;
  SortPlan pl;
  pl.real = (bits + 7) / 8 < passes ? (bits + 7) / 8 : passes;
  pl.pack = -1;
  if (pbits)
    for (unsigned j = 0; j + 1 < pl.real; ++j)
      if (bits - 8 * (j + 1) + pbits <= 64) {
        pl.pack = (int)j;
        break;
      }
  return pl;
}

// An object of E floats loaded as the widest aligned loads its size allows.
// VEC: the input pointer is 16-byte aligned (the launchers check), so objects
// of 2, 4 or 6 floats load as float2 / float4.
template <int E, bool VEC>
struct ObjVec {
  float v[E];
  __device__ __forceinline__ void load(const float* __restrict__ o) {
    if constexpr (VEC && E % 4 == 0) {
#pragma unroll
      for (int e = 0; e < E; e += 4) {
        const float4 t = *reinterpret_cast<const float4*>(o + e);
        v[e] = t.x; v[e + 1] = t.y; v[e + 2] = t.z; v[e + 3] = t.w;
      }
    } else if constexpr (VEC && E % 2 == 0) {
#pragma unroll
      for (int e = 0; e < E; e += 2) {
        const float2 t = *reinterpret_cast<const float2*>(o + e);
        v[e] = t.x; v[e + 1] = t.y;
      }
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = o[e];
    }
  }
};


// Boxes of the packed leaves of fanout FL (16, 32 or 64) that a wave's
// 64-rank chunk covers: lane l holds rank r = chunk start + l (object v,
// `valid` when r < n); each FL-lane group reduces its objects by shuffles, and
// its first lane writes the leaf's box with the padding rule of k_leaf_boxes.
template <int D, bool BOX, int FL>
__device__ __forceinline__ void leaf_box_reduce(const float (&v)[BOX ? 2 * D : D], bool valid,
                                                uint64_t r, uint64_t n, float* __restrict__ nbox,
                                                uint64_t nstride, uint64_t leaf_off) {
  static_assert(FL == 16 || FL == 32 || FL == 64, "leaves of one wave's chunk");
  const uint32_t lane = lane_id();
  float lo[D], hi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    lo[d] = valid ? v[d] : __builtin_huge_valf();
    hi[d] = valid ? v[BOX ? D + d : d] : -__builtin_huge_valf();
  }
#pragma unroll
  for (int off = FL / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float l = __shfl_xor(lo[d], off, kWave);
      const float h = __shfl_xor(hi[d], off, kWave);
      lo[d] = (lo[d] > l) ? l : lo[d];
      hi[d] = (hi[d] < h) ? h : hi[d];
    }
  }
  if ((lane & (FL - 1)) == 0 && valid) {
    const uint64_t leaf = r / FL;  // r is this leaf's first rank
    const bool pad = n - r < (uint64_t)FL;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      float l = lo[d], h = hi[d];
      if (pad) {
        l = (l > 1.0f) ? 1.0f : l;
        h = (h < 0.0f) ? 0.0f : h;
      }
      nbox[(uint64_t)d * nstride + leaf_off + leaf] = l;
      nbox[(uint64_t)(D + d) * nstride + leaf_off + leaf] = h;
    }
  }
}

// The build's gather: one 64-rank chunk per wave and iteration (a
// wave-uniform loop).  PACKED: the sort ended in packed elements (rank r holds
// ((key >> dropped bits) << pbits) | input index); the permutation is unpacked
// into perm_out, and the sorted keys are not written (the packing dropped
// their low bytes, and like the reference's build, which replaces them by
// ranks, nothing after the build needs them; k_keys_sorted recomputes them on
// demand).  FL = 32 or 64 (packed layouts whose fanout is FL): the wave also
// reduces the boxes of the leaves its chunk covers, with the padding rule of
// k_leaf_boxes, so the leaf level does not read the objects again.
template <int D, bool BOX, bool VEC, bool PACKED, int FL>
__device__ __forceinline__ void gather_objs_body(
    const float* __restrict__ objs, const uint32_t* __restrict__ perm,
    const uint64_t* __restrict__ packed, uint64_t n, unsigned pbits, float* __restrict__ obj,
    uint64_t stride, uint32_t* __restrict__ perm_out, float* __restrict__ nbox,
    uint64_t nstride, uint64_t leaf_off) {
  constexpr int E = BOX ? 2 * D : D;
  const uint32_t lane = lane_id();
  const uint64_t mask = PACKED ? (1ull << pbits) - 1ull : 0ull;
  const uint64_t T = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t wb = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(uint32_t)(kWave - 1));
       wb < n; wb += T) {
    const uint64_t r = wb + lane;
    const bool valid = r < n;
    ObjVec<E, VEC> o;
    if (valid) {
      const uint32_t p = PACKED ? (uint32_t)(packed[r] & mask)
                                : (perm != nullptr ? perm[r] : (uint32_t)r);
      o.load(objs + (uint64_t)p * E);
#pragma unroll
      for (int e = 0; e < E; ++e) obj[(uint64_t)e * stride + r] = o.v[e];
      if (PACKED || perm_out != nullptr) perm_out[r] = p;
    }
    if constexpr (FL > 0) leaf_box_reduce<D, BOX, FL>(o.v, valid, r, n, nbox, nstride, leaf_off);
  }
}

template <int D, bool BOX, bool VEC, bool PACKED, int FL>
__global__ __launch_bounds__(256) void k_gather_objs(
    const float* __restrict__ objs, const uint32_t* __restrict__ perm,
    const uint64_t* __restrict__ packed, uint64_t n, unsigned pbits, float* __restrict__ obj,
    uint64_t stride, uint32_t* __restrict__ perm_out, float* __restrict__ nbox,
    uint64_t nstride, uint64_t leaf_off) {
  gather_objs_body<D, BOX, VEC, PACKED, FL>(objs, perm, packed, n, pbits, obj, stride, perm_out,
                                            nbox, nstride, leaf_off);
}

// The same after a device-planned sort (k_sort_onesweep_planned over
// `passes`): the plan, made from the OR of all keys, says whether the result
// is packed or a permutation, and where the last reordering pass left it
// (packed / perm for an even count of such passes, packed2 / perm2 for an odd
// one).  The permutation always ends in perm_out (= perm).
template <int D, bool BOX, bool VEC, int FL>
__global__ __launch_bounds__(256) void k_gather_objs_planned(
    const float* __restrict__ objs, const uint32_t* perm, const uint32_t* perm2,
    const uint64_t* __restrict__ packed, const uint64_t* __restrict__ packed2, uint64_t n,
    unsigned pbits, float* __restrict__ obj, uint64_t stride, uint32_t* perm_out,
    float* __restrict__ nbox, uint64_t nstride, uint64_t leaf_off, unsigned passes,
    const unsigned long long* __restrict__ key_or) {
  const SortPlan pl = sort_plan(passes, pbits, key_or);
  const bool second = (pl.real & 1u) != 0;
  if (pl.real == 0)  // no pass reordered: the permutation is the identity
    gather_objs_body<D, BOX, VEC, false, FL>(objs, nullptr, nullptr, n, pbits, obj, stride,
                                             perm_out, nbox, nstride, leaf_off);
  else if (pl.pack >= 0)
    gather_objs_body<D, BOX, VEC, true, FL>(objs, nullptr, second ? packed2 : packed, n, pbits,
                                            obj, stride, perm_out, nbox, nstride, leaf_off);
  else
    gather_objs_body<D, BOX, VEC, false, FL>(objs, second ? perm2 : perm, nullptr, n, pbits, obj,
                                             stride, second ? perm_out : nullptr, nbox, nstride,
                                             leaf_off);
}

// Sorted keys recomputed from the rank-ordered SoA objects (obj[e][r]) by the
// same walk as k_keys, four a thread.
template <int D, bool BOX, bool REF, bool CENTER>
__global__ __launch_bounds__(256) void k_keys_sorted(const float* __restrict__ obj,
                                                     uint64_t stride, uint64_t n,
                                                     uint64_t* __restrict__ keys) {
  constexpr int E = BOX ? 2 * D : D;
  constexpr int R = kKeysPerThread;
  constexpr int NB = key_nb<D, REF>();
  __shared__ KeyTables<D, NB> kt;
  fill_key_tables<D, NB>(kt);
  for (uint64_t base = (uint64_t)blockIdx.x * (256 * R); base < n;
       base += (uint64_t)gridDim.x * (256 * R)) {
    uint32_t c[R][D];
    bool bad = false;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t i = base + (uint64_t)r * 256 + threadIdx.x;
      const uint64_t j = i < n ? i : n - 1;
      float v[E];
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = obj[(uint64_t)e * stride + j];
      object_point<D, BOX, REF, CENTER>(v, c[r], bad);
    }
    uint64_t k[R];
    hilbert_keys_pcm<D, NB, R>(c, kt.tab, kt.start, k);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t i = base + (uint64_t)r * 256 + threadIdx.x;
      if (i < n) keys[i] = k[r];
    }
  }
}

// Box of `cnt` entries with the reference's padding rule: when cnt < F the
// reduction also sees lower 1.0f / upper 0.0f (src/tree/tree.cpp:1070-1075);
// F = 0 gives the exact cover.
// One wave per node; lane handles entries lane + 64*w.
template <int D, int MW>
__device__ __forceinline__ void wave_box(const float* __restrict__ src, uint64_t stride,
                                         int hi_row, uint64_t first, uint32_t cnt, uint32_t F,
                                         float* __restrict__ dst, uint64_t dstride,
                                         uint64_t dst_idx) {
  const uint32_t lane = lane_id();
  float lo[D], hi[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    lo[d] = __builtin_huge_valf();
    hi[d] = -__builtin_huge_valf();
  }
#pragma unroll
  for (int w = 0; w < MW; ++w) {
    uint32_t k = w * kWave + lane;
    if (k < cnt) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        float l = src[(uint64_t)d * stride + first + k];
        float h = src[(uint64_t)(hi_row + d) * stride + first + k];
        lo[d] = (lo[d] > l) ? l : lo[d];
        hi[d] = (hi[d] < h) ? h : hi[d];
      }
    }
  }
#pragma unroll
  for (int d = 0; d < D; ++d) {
    float l = wave_min(lo[d]);
    float h = wave_max(hi[d]);
    if (cnt < F) {
      l = (l > 1.0f) ? 1.0f : l;
      h = (h < 0.0f) ? 0.0f : h;
    }
    lo[d] = l;
    hi[d] = h;
  }
  if (lane == 0) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      dst[(uint64_t)d * dstride + dst_idx] = lo[d];
      dst[(uint64_t)(D + d) * dstride + dst_idx] = hi[d];
    }
  }
}

// Packed leaves: leaf j holds ranks [jF, min(n, jF+F)).
template <int D, bool BOX, int MW>
__global__ __launch_bounds__(256) void k_leaf_boxes(const float* __restrict__ obj,
                                                    uint64_t ostride, uint64_t n, uint32_t F,
                                                    uint64_t nleaves, float* __restrict__ nbox,
                                                    uint64_t nstride, uint64_t leaf_off) {
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (wave >= nleaves) return;
  const uint64_t first = wave * F;
  const uint32_t cnt = (uint32_t)((n - first) < F ? (n - first) : F);
  wave_box<D, MW>(obj, ostride, BOX ? D : 0, first, cnt, F, nbox, nstride, leaf_off + wave);
}

// Packed leaves of fanout F (16 to 256, a power of two) from the
// rank-ordered objects, streamed: lane l of a wave takes the F/16 consecutive
// ranks from wave start + l*F/16 (one or a few vector loads a coordinate, min
// and max in rank order), so that one leaf is one 16-lane row, reduced by
// four DPP row rotations without LDS traffic.  Row g's first lane writes leaf
// (wave start / F) + g with the padding rule of k_leaf_boxes.  The order of
// the min/max pairings differs from the one-wave-a-leaf kernel, which only
// matters for a -0.0f / +0.0f tie (either may stay; no other value depends on
// the order).
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}
// row_ror:8, 4, 2, 1 (DPP controls 0x128, 0x124, 0x122, 0x121): every lane
// of a 16-lane row ends with the row's minimum / maximum
__device__ __forceinline__ float row_min16(float x) {
  float y = dpp_f32<0x128>(x);
  x = (x > y) ? y : x;
  y = dpp_f32<0x124>(x);
  x = (x > y) ? y : x;
  y = dpp_f32<0x122>(x);
  x = (x > y) ? y : x;
  y = dpp_f32<0x121>(x);
  return (x > y) ? y : x;
}
__device__ __forceinline__ float row_max16(float x) {
  float y = dpp_f32<0x128>(x);
  x = (x < y) ? y : x;
  y = dpp_f32<0x124>(x);
  x = (x < y) ? y : x;
  y = dpp_f32<0x122>(x);
  x = (x < y) ? y : x;
  y = dpp_f32<0x121>(x);
  return (x < y) ? y : x;
}

// K consecutive floats from p[r0] (K-aligned), those below rank n
template <int K>
__device__ __forceinline__ void load_run(const float* __restrict__ p, uint64_t r0, uint64_t n,
                                         float (&v)[K], bool (&ok)[K]) {
  if (r0 + K <= n) {
    if constexpr (K % 4 == 0) {
#pragma unroll
      for (int k = 0; k < K; k += 4) {
        const float4 t = *reinterpret_cast<const float4*>(p + r0 + k);
        v[k] = t.x; v[k + 1] = t.y; v[k + 2] = t.z; v[k + 3] = t.w;
      }
    } else if constexpr (K == 2) {
      const float2 t = *reinterpret_cast<const float2*>(p + r0);
      v[0] = t.x; v[1] = t.y;
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = p[r0 + k];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) ok[k] = true;
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      ok[k] = r0 + k < n;
      v[k] = ok[k] ? p[r0 + k] : 0.0f;
    }
  }
}

template <int D, bool BOX, int F>
__global__ __launch_bounds__(256) void k_leaf_boxes_row(const float* __restrict__ obj,
                                                        uint64_t ostride, uint64_t n,
                                                        float* __restrict__ nbox, uint64_t nstride,
                                                        uint64_t leaf_off) {
  static_assert(F >= 16 && F <= 256 && (F & (F - 1)) == 0, "a leaf is a 16-lane row");
  constexpr int K = F / 16;  // ranks a lane
  const uint32_t lane = lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
  for (uint64_t wb = ((uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * 4 * F;
       wb < n; wb += waves * 4 * F) {
    const uint64_t r0 = wb + (uint64_t)lane * K;
    float lo[D], hi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      float v[K];
      bool ok[K];
      load_run<K>(obj + (uint64_t)d * ostride, r0, n, v, ok);
      float l = __builtin_huge_valf(), h = -__builtin_huge_valf();
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (ok[k]) {
          l = (l > v[k]) ? v[k] : l;
          if (!BOX) h = (h < v[k]) ? v[k] : h;
        }
      }
      if (BOX) {
        load_run<K>(obj + (uint64_t)(D + d) * ostride, r0, n, v, ok);
#pragma unroll
        for (int k = 0; k < K; ++k)
          if (ok[k]) h = (h < v[k]) ? v[k] : h;
      }
      lo[d] = row_min16(l);
      hi[d] = row_max16(h);
    }
    if ((lane & 15u) == 0 && r0 < n) {
      const uint64_t leaf = r0 / F;  // r0 is this leaf's first rank
      const bool pad = n - r0 < (uint64_t)F;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        float l = lo[d], h = hi[d];
        if (pad) {
          l = (l > 1.0f) ? 1.0f : l;
          h = (h < 0.0f) ? 0.0f : h;
        }
        nbox[(uint64_t)d * nstride + leaf_off + leaf] = l;
        nbox[(uint64_t)(D + d) * nstride + leaf_off + leaf] = h;
      }
    }
  }
}

// Packed internal level: parent j boxes children [jF, min(child_cnt, jF+F)).
template <int D, int MW>
__global__ __launch_bounds__(256) void k_level_boxes(float* __restrict__ nbox,
                                                     uint64_t nstride, uint64_t child_off,
                                                     uint64_t child_cnt, uint32_t F,
                                                     uint64_t parent_off,
                                                     uint64_t parent_cnt) {
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (wave >= parent_cnt) return;
  const uint64_t first = wave * F;
  const uint32_t cnt = (uint32_t)((child_cnt - first) < F ? (child_cnt - first) : F);
  wave_box<D, MW>(nbox, nstride, D, child_off + first, cnt, F, nbox, nstride,
                  parent_off + wave);
}

// Explicit tree (BVH, Guttman): nodes [lo, hi) of one depth; leaves box their objects,
// internal nodes their children (which live at the next depth, already done).
template <int D, bool BOX, int MW>
__global__ __launch_bounds__(256) void k_tree_boxes(const float* __restrict__ obj,
                                                    uint64_t ostride,
                                                    const uint32_t* __restrict__ nfirst,
                                                    const uint32_t* __restrict__ ncount,
                                                    uint32_t pad, uint64_t lo, uint64_t hi,
                                                    float* __restrict__ nbox,
                                                    uint64_t nstride) {
  const uint64_t node = lo + ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (node >= hi) return;
  const uint32_t first = nfirst[node];
  const uint32_t cw = ncount[node];
  const uint32_t cnt = cw & ~kLeafBit;
  if (cw & kLeafBit)
    wave_box<D, MW>(obj, ostride, BOX ? D : 0, first, cnt, pad, nbox, nstride, node);
  else
    wave_box<D, MW>(nbox, nstride, D, first, cnt, pad, nbox, nstride, node);
}

// Fill helpers (kernels, so that build and search stay graph-capturable).
__global__ void k_iota_u32(uint32_t* p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}
__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_f32(float* p, uint64_t n, float v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// host-side launchers
static inline unsigned grid_for(uint64_t n, unsigned threads, unsigned cap = 65535u * 4) {
  uint64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

hipError_t launch_keys(const float* objs, uint64_t n, unsigned D, bool box, bool ref,
                       bool center, unsigned nb, uint64_t* keys, uint32_t* idx,
                       unsigned long long* key_or, uint32_t* bad, hipStream_t s,
                       uint32_t* ghist, uint64_t* status0, uint64_t status_words) {
  if (nb != (ref ? (D == 2 ? 31u : 20u) : 16u)) return hipErrorInvalidValue;
  if (ghist == nullptr && status_words != 0) return hipErrorInvalidValue;
  // counting workgroups also add their digit counts to ghist once each: fewer
  // of them, each walking more keys
  unsigned g = grid_for((n + kKeysPerThread - 1) / kKeysPerThread, 256,
                        ghist != nullptr ? 2048 : 8192);
#define QCP_KEYS(DD, BB, RR, CC)                                                                \
  if (ghist != nullptr)                                                                         \
    k_keys<DD, BB, RR, CC, true><<<g, 256, 0, s>>>(objs, n, keys, idx, key_or, bad, ghist,      \
                                                   status0, status_words);                     \
  else                                                                                          \
    k_keys<DD, BB, RR, CC, false><<<g, 256, 0, s>>>(objs, n, keys, idx, key_or, bad, nullptr,   \
                                                    nullptr, 0)
  if (D == 2) {
    if (box) {
      if (ref) { if (center) QCP_KEYS(2, true, true, true); else QCP_KEYS(2, true, true, false); }
      else { if (center) QCP_KEYS(2, true, false, true); else QCP_KEYS(2, true, false, false); }
    } else {
      if (ref) QCP_KEYS(2, false, true, false); else QCP_KEYS(2, false, false, false);
    }
  } else if (D == 3) {
    if (box) {
      if (ref) { if (center) QCP_KEYS(3, true, true, true); else QCP_KEYS(3, true, true, false); }
      else { if (center) QCP_KEYS(3, true, false, true); else QCP_KEYS(3, true, false, false); }
    } else {
      if (ref) QCP_KEYS(3, false, true, false); else QCP_KEYS(3, false, false, false);
    }
  } else {
    return hipErrorInvalidValue;
  }
#undef QCP_KEYS
  return hipGetLastError();
}

uint32_t sort_tiles(uint64_t n) { return (uint32_t)((n + kSortTile - 1) / kSortTile); }

template <typename TI, typename TO>
static hipError_t scan_impl(const TI* in, uint64_t n, TO* out, TO* block_sums,
                            uint32_t max_blocks, hipStream_t s) {
  // out must hold n+1 elements; block_sums max_blocks+1
  if (n == 0) {
    k_fill_u32<<<1, 64, 0, s>>>((uint32_t*)out, sizeof(TO) / 4, 0u);
    return hipGetLastError();
  }
  uint64_t per_block = kScanChunk;
  while ((n + per_block - 1) / per_block > max_blocks) per_block += kScanChunk;
  uint32_t nb = (uint32_t)((n + per_block - 1) / per_block);
  k_scan_reduce<TI, TO><<<nb, kScanThreads, 0, s>>>(in, n, per_block, block_sums);
  k_scan_top<TO><<<1, kScanThreads, 0, s>>>(block_sums, nb);
  k_scan_apply<TI, TO><<<nb, kScanThreads, 0, s>>>(in, n, per_block, block_sums, out);
  return hipGetLastError();
}

hipError_t scan_u32_u64(const uint32_t* in, uint64_t n, unsigned long long* out,
                        unsigned long long* block_sums, uint32_t max_blocks, hipStream_t s) {
  return scan_impl<uint32_t, unsigned long long>(in, n, out, block_sums, max_blocks, s);
}


// ---------------------------------------------------------------------------
// One-sweep LSD radix sort (index builds and query ordering).  One kernel reads
// the keys once and counts the digits of every pass; each pass is then a
// single kernel: a workgroup takes the next tile in start order (an atomic
// tile counter), ranks it into LDS in (digit, index) order (a wave multisplit
// of 8 ballots per key against wave-private counters), writes each digit's
// keys as one contiguous run, publishes its per-digit counts, and learns how
// many keys of each digit earlier tiles hold by looking back over their
// published words (decoupled look-back): a word is 0 (not yet published),
// kStatusAgg | count (that tile's own count) or kStatusPrefix | count (the
// count over tiles 0..t inclusive).  Every tile publishes its own count before
// it waits on any other tile, and a tile's predecessors hold workgroups that
// started before it, so every wait ends.
constexpr int kMaxSortPasses = 8;
constexpr uint64_t kStatusAgg = 1ull << 62;
constexpr uint64_t kStatusPrefix = 1ull << 63;

// Passes that reorder: those below the byte width of the OR of all keys (all
// of them when key_or is null).
__device__ __forceinline__ unsigned real_passes(unsigned passes,
                                                const unsigned long long* key_or) {
  if (key_or == nullptr) return passes;
  const unsigned long long ko = *key_or;
  const unsigned bits = ko ? 64u - (unsigned)__clzll((long long)ko) : 0u;
  const unsigned real = (bits + 7) / 8;
  return real < passes ? real : passes;
}

// Digit counts of every pass that reorders (ghist[p][256], zeroed by the
// caller) from one read of the keys; also zeroes the first pass's look-back
// words.  A wave whose keys share one digit adds them with one atomic (the
// high digits of clustered keys), otherwise each key adds its own.
__global__ __launch_bounds__(kHistThreads) void k_sort_hist_all(
    const uint64_t* __restrict__ keys, uint64_t n, unsigned passes,
    const unsigned long long* __restrict__ key_or, uint32_t* __restrict__ ghist,
    uint64_t* __restrict__ status0, uint64_t status_words) {
  __shared__ uint32_t h[kMaxSortPasses][256];
  const int tid = threadIdx.x;
  passes = real_passes(passes, key_or);
#pragma unroll
  for (int p = 0; p < kMaxSortPasses; ++p) h[p][tid] = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kHistThreads + tid; i < status_words;
       i += (uint64_t)gridDim.x * kHistThreads)
    status0[i] = 0;
  __syncthreads();
  const uint32_t nt = (uint32_t)((n + kHistTile - 1) / kHistTile);
  for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
    const uint64_t base = (uint64_t)t * kHistTile;
    uint64_t k[kHistKPT];
#pragma unroll
    for (int it = 0; it < kHistKPT; ++it) {
      const uint64_t i = base + (uint64_t)it * kHistThreads + tid;
      k[it] = i < n ? keys[i] : 0;
    }
#pragma unroll
    for (int it = 0; it < kHistKPT; ++it) {
      const bool valid = base + (uint64_t)it * kHistThreads + tid < n;
      const uint64_t vmask = ballot(valid);
      if (vmask == 0) continue;
      // the low digits of Hilbert keys are spread out: one atomic per key
      const unsigned hi = passes > 2 ? passes - 2 : 0;
      if (valid)
        for (unsigned p = 0; p < hi; ++p) atomicAdd(&h[p][(k[it] >> (8 * p)) & 0xFF], 1u);
      // the top two may be shared by a whole wave (clustered objects)
      const int first = __builtin_ctzll(vmask);
      const uint64_t k0 = ((uint64_t)__builtin_amdgcn_readlane((int)(k[it] >> 32), first) << 32) |
                          (uint32_t)__builtin_amdgcn_readlane((int)k[it], first);
      const uint64_t x = k[it] ^ k0;
      for (unsigned p = hi; p < passes; ++p) {
        if (ballot(valid && ((x >> (8 * p)) & 0xFF)) == 0) {
          if ((int)lane_id() == first) atomicAdd(&h[p][(k0 >> (8 * p)) & 0xFF], (uint32_t)__popcll(vmask));
        } else if (valid) {
          atomicAdd(&h[p][(k[it] >> (8 * p)) & 0xFF], 1u);
        }
      }
    }
  }
  __syncthreads();
  for (unsigned p = 0; p < passes; ++p)
    if (h[p][tid]) atomicAdd(&ghist[p * 256 + tid], h[p][tid]);
}

// LDS of one pass: 77.5 KB, so that two workgroups share a CU's 160 KB (and
// at most 128 VGPRs a lane, four waves a SIMD, for the same reason).  The
// tile's keys (or packed words) and then, for passes that move a separate
// index array, its indices pass through `sk` / `sv`; dg keeps each slot's
// digit.
struct SortLds {
  union {
    uint64_t sk[kSortTile];
    uint32_t sv[kSortTile];
  };
  uint8_t dg[kSortTile];
  uint32_t wsum[kSortThreads / kWave];            // wave totals of sort_scan
  uint32_t gstart[256];                           // global position of this tile's bucket
  uint16_t lstart[256];                           // the bucket's start inside the tile
  uint16_t wcnt[kSortThreads / kWave][256];       // per wave: keys per digit, then its offsets
  uint32_t s_tile;
};

// Exclusive scan of one value per thread over the sort workgroup (wave
// prefix by shuffles, then the wave totals); *total = the sum.
__device__ __forceinline__ uint32_t sort_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
  constexpr int W = kSortThreads / kWave;
  const int w = threadIdx.x / kWave;
  const uint32_t lane = lane_id();
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t t = __shfl_up(x, off, kWave);
    if ((int)lane >= off) x += t;
  }
  if (lane == kWave - 1) wsum[w] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int ww = 0; ww < W; ++ww) {
    const uint32_t t = wsum[ww];
    if (ww < w) pre += t;
    tot += t;
  }
  *total = tot;
  __syncthreads();
  return pre + x - v;
}

// Per-tile steps of the object-carrying passes (the same steps as
// sort_pass_body, which keeps its own copy for its register allocation).
// tile_begin: take the next tile in start order, zero this workgroup's row of
// the next pass's look-back words, clear the wave counters; returns, to
// threads 0..255, the global start of their digit's bucket.
__device__ __forceinline__ uint32_t tile_begin(SortLds& L, unsigned pass,
                                               const uint32_t* __restrict__ ghist,
                                               uint64_t* __restrict__ status_next,
                                               uint32_t* tile_ctr) {
  constexpr int W = kSortThreads / kWave;
  const int tid = threadIdx.x;
  const bool owner = tid < 256;  // owns digit tid
  if (tid == 0) L.s_tile = atomicAdd(&tile_ctr[pass], 1u);
  if (status_next != nullptr && owner) status_next[(uint64_t)blockIdx.x * 256 + tid] = 0;
  uint32_t tot;
  const uint32_t gbase = sort_scan(owner ? ghist[pass * 256 + tid] : 0u, L.wsum, &tot);
  if (owner) {
#pragma unroll
    for (int ww = 0; ww < W; ++ww) L.wcnt[ww][tid] = 0;
  }
  __syncthreads();
  return gbase;
}

// tile_rank: rank of each of this wave's keys among the wave's earlier keys of
// the same digit (rk[c] for chunk c of 64 keys at wbase + 64c; digit(c) is
// this lane's digit in chunk c), counting each digit's keys into L.wcnt[wave].
template <class DigitFn>
__device__ __forceinline__ void tile_rank(SortLds& L, uint64_t wbase, uint64_t n, DigitFn digit,
                                          uint32_t (&rk)[kSortKPT]) {
  const int w = threadIdx.x / kWave;
  const uint32_t lane = lane_id();
#pragma unroll
  for (int c = 0; c < kSortKPT; ++c) {
    const bool valid = wbase + (uint64_t)c * kWave + lane < n;
    const uint32_t dig = digit(c);
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bb = ballot((dig >> b) & 1u);
      peers &= ((dig >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t r = rank_below(peers);
    // every lane reads its digit's count before the group's first lane adds
    // the group (one wave: the LDS serves its accesses in program order)
    const uint32_t before = valid ? (uint32_t)L.wcnt[w][dig] : 0u;
    rk[c] = before + r;
    __builtin_amdgcn_wave_barrier();
    if (valid && r == 0) L.wcnt[w][dig] = (uint16_t)(before + (uint32_t)__popcll(peers));
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
}

// tile_place: the tile's count of each digit, published and then looked back
// over earlier tiles (decoupled look-back); leaves the bucket's global start
// in L.gstart, its start inside the tile in L.lstart, and in L.wcnt[wave] each
// wave's first slot of each digit inside the tile.
__device__ __forceinline__ void tile_place(SortLds& L, uint32_t tile, uint32_t gbase,
                                           uint64_t* status) {
  constexpr int W = kSortThreads / kWave;
  const int tid = threadIdx.x;
  const bool owner = tid < 256;
  // digit tid: this tile's count, and each wave's offset among the tile's keys
  uint32_t cnt = 0;
  if (owner) {
#pragma unroll
    for (int ww = 0; ww < W; ++ww) {
      const uint32_t c = L.wcnt[ww][tid];
      L.wcnt[ww][tid] = (uint16_t)cnt;
      cnt += c;
    }
  }
  uint64_t* const mine = status + (uint64_t)tile * 256 + tid;
  if (owner)
    __hip_atomic_store(mine, (tile == 0 ? kStatusPrefix : kStatusAgg) | (uint64_t)cnt,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t tot_tile;
  const uint32_t ls = sort_scan(cnt, L.wsum, &tot_tile);
  uint32_t excl = 0;
  if (owner && tile > 0) {
    // kLookback predecessors' words per round trip, consumed in order up to
    // the first inclusive prefix; a word not yet published ends the round
    // (sleep, then reload from there).  Tiles below 0 read as a zero prefix.
    constexpr int kLookback = 4;
    int32_t t = (int32_t)tile - 1;
    for (;;) {
      uint64_t wv[kLookback];
#pragma unroll
      for (int j = 0; j < kLookback; ++j)
        wv[j] = t - j >= 0 ? __hip_atomic_load(status + (uint64_t)(t - j) * 256 + tid,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           : kStatusPrefix;
      bool stop = false, done = false;
      int used = 0;
#pragma unroll
      for (int j = 0; j < kLookback; ++j) {
        if (!stop) {
          if (!(wv[j] & (kStatusAgg | kStatusPrefix))) {
            stop = true;
          } else {
            excl += (uint32_t)wv[j];
            ++used;
            if (wv[j] & kStatusPrefix) stop = done = true;
          }
        }
      }
      if (done) break;
      t -= used;
      if (used == 0) __builtin_amdgcn_s_sleep(1);
    }
    __hip_atomic_store(mine, kStatusPrefix | (uint64_t)(excl + cnt), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  if (owner) {
    L.gstart[tid] = gbase + excl;
    L.lstart[tid] = (uint16_t)ls;
#pragma unroll
    for (int ww = 0; ww < W; ++ww) L.wcnt[ww][tid] = (uint16_t)(L.wcnt[ww][tid] + ls);
  }
  __syncthreads();
}

// The work of one pass on one tile (see k_sort_onesweep).
template <bool INP, bool OUTP>
__device__ __forceinline__ void sort_pass_body(
    SortLds& L, const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, uint64_t n, unsigned pass,
    unsigned shift, unsigned pack_shift, unsigned pbits, const uint32_t* __restrict__ ghist,
    uint64_t* status, uint64_t* __restrict__ status_next, uint32_t* tile_ctr) {
  static_assert(OUTP || !INP, "a packed pass writes packed elements");
  constexpr int W = kSortThreads / kWave;
  constexpr int CH = kSortKPT;  // chunks of 64 keys per wave
  const int tid = threadIdx.x;
  const bool owner = tid < 256;  // owns digit tid
  const int w = tid / kWave;
  const uint32_t lane = lane_id();
  if (tid == 0) L.s_tile = atomicAdd(&tile_ctr[pass], 1u);
  if (status_next != nullptr && owner) status_next[(uint64_t)blockIdx.x * 256 + tid] = 0;
  uint32_t tot;
  const uint32_t gbase = sort_scan(owner ? ghist[pass * 256 + tid] : 0u, L.wsum, &tot);
  if (owner) {
#pragma unroll
    for (int ww = 0; ww < W; ++ww) L.wcnt[ww][tid] = 0;
  }
  __syncthreads();
  const uint32_t tile = L.s_tile;
  const uint64_t wbase = (uint64_t)tile * kSortTile + (uint64_t)w * CH * kWave;
  uint64_t k[CH];
  uint32_t v[CH], rk[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const uint64_t i = wbase + (uint64_t)c * kWave + lane;
    k[c] = i < n ? kin[i] : 0;
    // the first pass has no index array: an element's index is its position
    v[c] = (!INP && i < n) ? (vin != nullptr ? vin[i] : (uint32_t)i) : 0;
  }
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const bool valid = wbase + (uint64_t)c * kWave + lane < n;
    const uint32_t dig = (uint32_t)(k[c] >> shift) & 0xFF;
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bb = ballot((dig >> b) & 1u);
      peers &= ((dig >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t r = rank_below(peers);
    // every lane reads its digit's count before the group's first lane adds
    // the group (one wave: the LDS serves its accesses in program order)
    const uint32_t before = valid ? (uint32_t)L.wcnt[w][dig] : 0u;
    rk[c] = before + r;
    __builtin_amdgcn_wave_barrier();
    if (valid && r == 0) L.wcnt[w][dig] = (uint16_t)(before + (uint32_t)__popcll(peers));
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // digit tid: this tile's count, and each wave's offset among the tile's keys
  uint32_t cnt = 0;
  if (owner) {
#pragma unroll
    for (int ww = 0; ww < W; ++ww) {
      const uint32_t c = L.wcnt[ww][tid];
      L.wcnt[ww][tid] = (uint16_t)cnt;
      cnt += c;
    }
  }
  uint64_t* const mine = status + (uint64_t)tile * 256 + tid;
  if (owner)
    __hip_atomic_store(mine, (tile == 0 ? kStatusPrefix : kStatusAgg) | (uint64_t)cnt,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t tot_tile;
  const uint32_t ls = sort_scan(cnt, L.wsum, &tot_tile);
  uint32_t excl = 0;
  if (owner && tile > 0) {
    // kLookback predecessors' words per round trip, consumed in order up to
    // the first inclusive prefix; a word not yet published ends the round
    // (sleep, then reload from there).  Tiles below 0 read as a zero prefix.
    constexpr int kLookback = 4;
    int32_t t = (int32_t)tile - 1;
    for (;;) {
      uint64_t wv[kLookback];
#pragma unroll
      for (int j = 0; j < kLookback; ++j)
        wv[j] = t - j >= 0 ? __hip_atomic_load(status + (uint64_t)(t - j) * 256 + tid,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           : kStatusPrefix;
      bool stop = false, done = false;
      int used = 0;
#pragma unroll
      for (int j = 0; j < kLookback; ++j) {
        if (!stop) {
          if (!(wv[j] & (kStatusAgg | kStatusPrefix))) {
            stop = true;
          } else {
            excl += (uint32_t)wv[j];
            ++used;
            if (wv[j] & kStatusPrefix) stop = done = true;
          }
        }
      }
      if (done) break;
      t -= used;
      if (used == 0) __builtin_amdgcn_s_sleep(1);
    }
    __hip_atomic_store(mine, kStatusPrefix | (uint64_t)(excl + cnt), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  if (owner) {
    L.gstart[tid] = gbase + excl;
    L.lstart[tid] = (uint16_t)ls;
#pragma unroll
    for (int ww = 0; ww < W; ++ww) L.wcnt[ww][tid] = (uint16_t)(L.wcnt[ww][tid] + ls);
  }
  __syncthreads();
  // the tile in (digit, index) order: keys (the packing pass packs them with
  // their indices first) and their digits
  uint32_t pos[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const uint32_t dig = (uint32_t)(k[c] >> shift) & 0xFF;
    pos[c] = L.wcnt[w][dig] + rk[c];
    if (wbase + (uint64_t)c * kWave + lane < n) {
      L.sk[pos[c]] = (OUTP && !INP) ? (((k[c] >> pack_shift) << pbits) | (uint64_t)v[c]) : k[c];
      L.dg[pos[c]] = (uint8_t)dig;
    }
  }
  __syncthreads();
  const uint64_t base = (uint64_t)tile * kSortTile;
  const uint32_t tcnt = (uint32_t)(n - base < (uint64_t)kSortTile ? n - base : kSortTile);
  for (uint32_t p = tid; p < tcnt; p += kSortThreads) {
    const uint32_t dig = L.dg[p];
    kout[L.gstart[dig] + (p - L.lstart[dig])] = L.sk[p];
  }
  if (!OUTP) {
    // the indices, through the same LDS after every key has left it
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CH; ++c)
      if (wbase + (uint64_t)c * kWave + lane < n) L.sv[pos[c]] = v[c];
    __syncthreads();
    for (uint32_t p = tid; p < tcnt; p += kSortThreads) {
      const uint32_t dig = L.dg[p];
      vout[L.gstart[dig] + (p - L.lstart[dig])] = L.sv[p];
    }
  }
}

// One pass (digit `pass` of `passes`), its digit at bit `shift` of the input
// element.  status: this pass's look-back words (ntiles x 256, zero at entry);
// status_next: the next pass's, zeroed here (row blockIdx.x; null for the last
// pass); tile_ctr[pass] zero at entry.
// Packed elements: once the key bits still to be sorted and the index fit one
// 64-bit word, a pass writes ((key >> pack_shift) << pbits) | index (OUTP) and
// later passes move that word alone (INP): 8 bytes per key each way instead of
// 8 + 4.
template <bool INP, bool OUTP>
__global__ __launch_bounds__(kSortThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_sort_onesweep(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, uint64_t n, unsigned pass,
    unsigned shift, unsigned pack_shift, unsigned pbits, const uint32_t* __restrict__ ghist,
    uint64_t* status, uint64_t* __restrict__ status_next, uint32_t* tile_ctr) {
  __shared__ SortLds L;
  sort_pass_body<INP, OUTP>(L, kin, vin, kout, vout, n, pass, shift, pack_shift, pbits, ghist,
                            status, status_next, tile_ctr);
}

// One pass of a device-planned sort.  Passes at or above `real` end at once,
// so the result stays where the last reordering pass put it: in the first
// buffers for an even `real`, the second for an odd one (packed when the plan
// packs); k_gather_objs_planned reads it there.
__global__ __launch_bounds__(kSortThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_sort_onesweep_planned(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, uint64_t n, unsigned pass,
    unsigned passes, unsigned pbits, const uint32_t* __restrict__ ghist, uint64_t* status,
    uint64_t* __restrict__ status_next, uint32_t* tile_ctr,
    const unsigned long long* __restrict__ key_or) {
  __shared__ SortLds L;
  const SortPlan pl = sort_plan(passes, pbits, key_or);
  if (pass >= pl.real) return;
  if (pl.pack < 0 || (int)pass < pl.pack)
    sort_pass_body<false, false>(L, kin, vin, kout, vout, n, pass, 8 * pass, 0, 0, ghist, status,
                                 status_next, tile_ctr);
  else if ((int)pass == pl.pack)
    sort_pass_body<false, true>(L, kin, vin, kout, vout, n, pass, 8 * pass, 8 * (pass + 1),
                                pbits, ghist, status, status_next, tile_ctr);
  else
    sort_pass_body<true, true>(L, kin, vin, kout, vout, n, pass,
                               pbits + 8 * (pass - (unsigned)pl.pack - 1), 0, pbits, ghist,
                               status, status_next, tile_ctr);
}

// ---------------------------------------------------------------------------
// Object-carrying sort (CarryArgs in qcp_kernels.h).  Elements are the key
// bits still to be sorted plus W payload words: the object's E floats as raw
// words, then its input index when the index keeps its permutation.  A pass
// ranks by the low byte of the key remainder, then writes the remainder shifted
// down by 8 bits and the payload to their places, one array a round through
// the tile's LDS, 8 bytes an element where it can: payload words go as pairs
// (2i, 2i+1), and an odd last word rides in the high half of the remainder's
// 8-byte slot once at most 32 key bits remain (while more remain, the
// remainder takes the whole slot and the odd word a 4-byte array of its own;
// the last pass writes no remainder).  The last pass writes the objects
// straight into the index's rank-ordered SoA array (and the input indices
// into the permutation), so nothing reads an object at a random place: the
// index form's gather reads each object through the permutation, a whole
// 128-byte line for 8 to 24 bytes.
// Scratch set j: the remainder slots in krem[j] (8 bytes, or 4 when they hold
// a remainder alone), the pairs as 8-byte arrays at pay[j] + 2i*cs and the odd
// word as a 4-byte array at pay[j] + 2*(W/2)*cs.
// SRC = 0: pass 0, which reads the keys and the caller's row-major objects;
// SRC = 1: a later pass whose remainders in exceed 32 bits (8-byte slots, the
// odd word apart); SRC = 2: at most 32 bits (4-byte slots, or 8-byte slots
// holding the odd word).  Pass p reads set (p - 1) & 1 and writes set p & 1.
template <int E, bool IDX, int SRC>
__device__ __forceinline__ void carry_pass_body(SortLds& L, const CarryArgs& a, unsigned pass,
                                                unsigned bits, bool last) {
  constexpr int W = E + (IDX ? 1 : 0);
  constexpr int NP = W / 2;        // payload pairs
  constexpr bool ODD = (W & 1) != 0;
  constexpr int CH = kSortKPT;
  // the remainder slot read: 8 bytes for pass 0 and SRC 1, and for SRC 2 when
  // it carries the odd word
  constexpr bool KIN64 = SRC != 2 || ODD;
  using KT = typename std::conditional<KIN64, uint64_t, uint32_t>::type;
  const unsigned kout_words = last ? 0u : (bits - 8 * (pass + 1) > 32 ? 2u : 1u);
  const bool fuse_out = ODD && kout_words == 1;  // odd word in the slot's high half
  const uint64_t n = a.n;
  const int tid = threadIdx.x;
  const int w = tid / kWave;
  const uint32_t lane = lane_id();
  uint64_t* st = a.status + (uint64_t)(pass & 1u) * a.words;
  uint64_t* st_next = last ? nullptr : a.status + (uint64_t)((pass + 1) & 1u) * a.words;
  const uint32_t gbase = tile_begin(L, pass, a.ghist, st_next, a.ghist + kMaxSortPasses * 256);
  const uint32_t tile = L.s_tile;
  // the wave's first element (wave-uniform, so that loads take a scalar base
  // and a 32-bit lane offset)
  const uint64_t wbase = uni64((uint64_t)tile * kSortTile + (uint64_t)w * CH * kWave);
  const unsigned src = (pass - 1u) & 1u, dstset = pass & 1u;
  const KT* __restrict__ kin = (SRC == 0 ? reinterpret_cast<const KT*>(a.keys)
                                         : reinterpret_cast<const KT*>(a.krem[src])) + wbase;
  const float* __restrict__ oin = a.objs + wbase * E;  // pass 0
  // payload word wd of this lane's element in chunk c, pass 0
  auto word0 = [&](int wd, int c) -> uint32_t {
    const uint32_t j = (uint32_t)c * kWave + lane;
    if (wd < E) return __float_as_uint(oin[j * E + wd]);
    return (uint32_t)(wbase + j);  // the input index is the position
  };
  // payload pair i of this lane's element in chunk c
  auto pair = [&](int i, int c) -> uint64_t {
    const uint32_t j = (uint32_t)c * kWave + lane;
    if (SRC == 0) return ((uint64_t)word0(2 * i + 1, c) << 32) | word0(2 * i, c);
    return (reinterpret_cast<const uint64_t*>(a.pay[src] + (uint64_t)2 * i * a.cs) + wbase)[j];
  };
  KT k[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const uint32_t j = (uint32_t)c * kWave + lane;
    k[c] = wbase + j < n ? kin[j] : (KT)0;
  }
  // the key remainder and the odd word of chunk c
  auto krem = [&](int c) -> uint64_t {
    return (SRC == 2 && ODD) ? (uint64_t)(uint32_t)k[c] : (uint64_t)k[c];
  };
  auto odd = [&](int c) -> uint32_t {
    if (SRC == 0) return word0(W - 1, c);
    if (SRC == 2) return (uint32_t)((uint64_t)k[c] >> 32);
    return (a.pay[src] + (uint64_t)2 * NP * a.cs + wbase)[(uint32_t)c * kWave + lane];
  };
  auto digit = [&](int c) { return (uint32_t)krem(c) & 0xFFu; };
  uint32_t rk[CH];
  tile_rank(L, wbase, n, digit, rk);
  tile_place(L, tile, gbase, st);
  // slots in (digit, index) order: the digits, and the remainders (with the
  // odd word when it rides along)
  uint32_t pos[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const uint32_t dig = digit(c);
    pos[c] = L.wcnt[w][dig] + rk[c];
    if (wbase + (uint64_t)c * kWave + lane < n) {
      L.dg[pos[c]] = (uint8_t)dig;
      if (kout_words == 2)
        L.sk[pos[c]] = krem(c) >> 8;
      else if (fuse_out)
        L.sk[pos[c]] = (krem(c) >> 8) | ((uint64_t)odd(c) << 32);
      else if (kout_words == 1)
        L.sv[pos[c]] = (uint32_t)(krem(c) >> 8);
    }
  }
  __syncthreads();
  const uint64_t base = (uint64_t)tile * kSortTile;
  const uint32_t tcnt = (uint32_t)(n - base < (uint64_t)kSortTile ? n - base : kSortTile);
  // destination of slot p
  auto dst = [&](uint32_t p) {
    const uint32_t dig = L.dg[p];
    return L.gstart[dig] + (p - L.lstart[dig]);
  };
  if (kout_words != 0) {
    for (uint32_t p = tid; p < tcnt; p += kSortThreads) {
      if (kout_words == 2 || fuse_out)
        a.krem[dstset][dst(p)] = L.sk[p];
      else
        reinterpret_cast<uint32_t*>(a.krem[dstset])[dst(p)] = L.sv[p];
    }
    __syncthreads();
  }
  // the last pass writes word wd to the index
  auto out_word = [&](int wd) -> uint32_t* {
    return wd < E ? reinterpret_cast<uint32_t*>(a.obj) + (uint64_t)wd * a.ostride : a.perm;
  };
  // the odd word on its own (the last pass, or a remainder of 8 bytes)
  if (ODD && !fuse_out) {
#pragma unroll
    for (int c = 0; c < CH; ++c)
      if (wbase + (uint64_t)c * kWave + lane < n) L.sv[pos[c]] = odd(c);
    __syncthreads();
    uint32_t* out = last ? out_word(W - 1) : a.pay[dstset] + (uint64_t)2 * NP * a.cs;
    for (uint32_t p = tid; p < tcnt; p += kSortThreads) out[dst(p)] = L.sv[p];
    if (NP > 0) __syncthreads();
  }
  // the pairs, one a round; the next pair's loads are issued before this
  // round's LDS traffic
  uint64_t cur[CH];
  if (NP > 0) {
#pragma unroll
    for (int c = 0; c < CH; ++c)
      cur[c] = wbase + (uint64_t)c * kWave + lane < n ? pair(0, c) : 0ull;
  }
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    uint64_t nxt[CH];
    if (i + 1 < NP) {
#pragma unroll
      for (int c = 0; c < CH; ++c)
        nxt[c] = wbase + (uint64_t)c * kWave + lane < n ? pair(i + 1, c) : 0ull;
    }
#pragma unroll
    for (int c = 0; c < CH; ++c)
      if (wbase + (uint64_t)c * kWave + lane < n) L.sk[pos[c]] = cur[c];
    __syncthreads();
    if (last) {
      uint32_t* o0 = out_word(2 * i);
      uint32_t* o1 = out_word(2 * i + 1);
      for (uint32_t p = tid; p < tcnt; p += kSortThreads) {
        const uint32_t d = dst(p);
        const uint64_t v = L.sk[p];
        o0[d] = (uint32_t)v;
        o1[d] = (uint32_t)(v >> 32);
      }
    } else {
      uint64_t* out = reinterpret_cast<uint64_t*>(a.pay[dstset] + (uint64_t)2 * i * a.cs);
      for (uint32_t p = tid; p < tcnt; p += kSortThreads) out[dst(p)] = L.sk[p];
    }
    if (i + 1 < NP) {
      __syncthreads();
#pragma unroll
      for (int c = 0; c < CH; ++c) cur[c] = nxt[c];
    }
  }
}

// One pass.  SRC = 0, 1, 2 as for carry_pass_body, fixed by the host (eager
// builds, whose key width the host knows); SRC = -1: chosen here from the key
// width (device-planned sorts).
template <int E, bool IDX, int SRC>
__global__ __launch_bounds__(kSortThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_sort_carry(
    CarryArgs a, unsigned pass) {
  __shared__ SortLds L;
  unsigned bits = a.bits;
  if (a.key_or != nullptr) {
    const unsigned long long ko = *a.key_or;
    bits = ko ? 64u - (unsigned)__clzll((long long)ko) : 0u;
  }
  const unsigned real = (bits + 7) / 8;
  if (pass >= real) return;  // a device-planned sort: this byte of every key is zero
  const bool last = pass + 1 == real;
  if (SRC >= 0)
    carry_pass_body<E, IDX, (SRC < 0 ? 0 : SRC)>(L, a, pass, bits, last);
  else if (pass == 0)
    carry_pass_body<E, IDX, 0>(L, a, pass, bits, last);
  else if (bits - 8 * pass > 32)
    carry_pass_body<E, IDX, 1>(L, a, pass, bits, last);
  else
    carry_pass_body<E, IDX, 2>(L, a, pass, bits, last);
}

// A device-planned object-carrying sort whose keys are all equal (no pass
// reorders): the objects go to the index in input order, the permutation is
// the identity.  Ends at once when some pass reorders.
template <int D, bool BOX, bool VEC>
__global__ __launch_bounds__(256) void k_carry_identity(const float* __restrict__ objs, uint64_t n,
                                                        const unsigned long long* __restrict__ key_or,
                                                        float* __restrict__ obj, uint64_t stride,
                                                        uint32_t* __restrict__ perm_out) {
  if (*key_or != 0) return;
  gather_objs_body<D, BOX, VEC, false, 0>(objs, nullptr, nullptr, n, 0, obj, stride, perm_out,
                                          nullptr, 0, 0);
}

uint64_t sort_status_words(uint64_t n) { return 2ull * sort_tiles(n > 0 ? n : 1) * 256; }

hipError_t launch_sort_keys(uint64_t* keys, uint32_t* vals, uint64_t* keys2, uint32_t* vals2,
                            uint64_t n, unsigned passes, const unsigned long long* key_or,
                            uint32_t* ghist, uint64_t* status, hipStream_t s, int pack_pass,
                            unsigned pbits, bool counted) {
  if (passes > (unsigned)kMaxSortPasses) return hipErrorInvalidValue;
  if (pack_pass >= 0 && (key_or != nullptr || pack_pass + 1 >= (int)passes || pbits == 0))
    return hipErrorInvalidValue;
  if (pbits > 32) return hipErrorInvalidValue;
  if (n == 0 || passes == 0) return hipSuccess;
  const uint32_t nt = sort_tiles(n);
  const uint64_t words = (uint64_t)nt * 256;
  uint32_t* tile_ctr = ghist + kMaxSortPasses * 256;
  if (!counted) {
    hipError_t e = launch_fill_u32(ghist, kMaxSortPasses * 256 + kMaxSortPasses, 0u, s);
    if (e != hipSuccess) return e;
    const uint32_t ht = (uint32_t)((n + kHistTile - 1) / kHistTile);
    const unsigned g = ht < 2048 ? ht : 2048;
    k_sort_hist_all<<<g, kHistThreads, 0, s>>>(keys, n, passes, key_or, ghist, status, words);
  }
  uint64_t* kin = keys;
  uint32_t* vin = nullptr;  // the first pass takes positions as the input indices
  uint64_t* kout = keys2;
  uint32_t* vout = vals2;
  for (unsigned p = 0; p < passes; ++p) {
    uint64_t* st = status + (uint64_t)(p & 1u) * words;
    uint64_t* st_next = p + 1 < passes ? status + (uint64_t)((p + 1) & 1u) * words : nullptr;
    if (key_or != nullptr) {
      k_sort_onesweep_planned<<<nt, kSortThreads, 0, s>>>(kin, vin, kout, vout, n, p, passes,
                                                          pbits, ghist, st, st_next, tile_ctr,
                                                          key_or);
    } else if (pack_pass < 0 || (int)p < pack_pass) {
      k_sort_onesweep<false, false><<<nt, kSortThreads, 0, s>>>(
          kin, vin, kout, vout, n, p, 8 * p, 0, 0, ghist, st, st_next, tile_ctr);
    } else if ((int)p == pack_pass) {
      k_sort_onesweep<false, true><<<nt, kSortThreads, 0, s>>>(
          kin, vin, kout, vout, n, p, 8 * p, 8 * (p + 1), pbits, ghist, st, st_next, tile_ctr);
    } else {
      k_sort_onesweep<true, true><<<nt, kSortThreads, 0, s>>>(
          kin, nullptr, kout, nullptr, n, p, pbits + 8 * (p - pack_pass - 1), 0, pbits, ghist,
          st, st_next, tile_ctr);
    }
    uint64_t* tk = kin;
    kin = kout;
    kout = tk;
    uint32_t* tv = p == 0 ? vals : vin;
    vin = vout;
    vout = tv;
  }
  return hipGetLastError();
}



hipError_t launch_sort_carry(const CarryArgs& a, unsigned D, bool box, bool idx,
                             unsigned passes, hipStream_t s, bool counted) {
  if (passes > (unsigned)kMaxSortPasses || (D != 2 && D != 3)) return hipErrorInvalidValue;
  if (a.key_or == nullptr && (a.bits > 64 || (a.bits + 7) / 8 != passes))
    return hipErrorInvalidValue;
  if (a.n == 0 || passes == 0) return hipSuccess;
  const uint32_t nt = sort_tiles(a.n);
  if (a.words != (uint64_t)nt * 256) return hipErrorInvalidValue;
  if (!counted) {
    hipError_t e = launch_fill_u32(a.ghist, kMaxSortPasses * 256 + kMaxSortPasses, 0u, s);
    if (e != hipSuccess) return e;
    const uint32_t ht = (uint32_t)((a.n + kHistTile - 1) / kHistTile);
    k_sort_hist_all<<<ht < 2048 ? ht : 2048, kHistThreads, 0, s>>>(a.keys, a.n, passes, a.key_or,
                                                                    a.ghist, a.status, a.words);
  }
  const unsigned E = box ? 2 * D : D;
  for (unsigned p = 0; p < passes; ++p) {
    // the pass's source kind, when the host knows the key width
    const int src = a.key_or != nullptr ? -1 : (p == 0 ? 0 : (a.bits - 8 * p > 32 ? 1 : 2));
#define QCP_SC2(EE, II)                                                                   \
  switch (src) {                                                                          \
    case 0: k_sort_carry<EE, II, 0><<<nt, kSortThreads, 0, s>>>(a, p); break;             \
    case 1: k_sort_carry<EE, II, 1><<<nt, kSortThreads, 0, s>>>(a, p); break;             \
    case 2: k_sort_carry<EE, II, 2><<<nt, kSortThreads, 0, s>>>(a, p); break;             \
    default: k_sort_carry<EE, II, -1><<<nt, kSortThreads, 0, s>>>(a, p); break;           \
  }
#define QCP_SC(EE) \
  if (idx) { QCP_SC2(EE, true) } else { QCP_SC2(EE, false) }
    switch (E) {
      case 2: QCP_SC(2); break;
      case 3: QCP_SC(3); break;
      case 4: QCP_SC(4); break;
      default: QCP_SC(6); break;
    }
#undef QCP_SC
#undef QCP_SC2
  }
  return hipGetLastError();
}

hipError_t launch_carry_identity(const float* objs, uint64_t n, const unsigned long long* key_or,
                                 unsigned D, bool box, float* obj, uint64_t stride,
                                 uint32_t* perm_out, hipStream_t s) {
  if (key_or == nullptr || (D != 2 && D != 3)) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  const unsigned g = grid_for(n, 256, 16384);
  const bool vec = ((uintptr_t)objs & 15u) == 0;
#define QCP_CI(DD, BB)                                                                       \
  if (vec) k_carry_identity<DD, BB, true><<<g, 256, 0, s>>>(objs, n, key_or, obj, stride, perm_out); \
  else k_carry_identity<DD, BB, false><<<g, 256, 0, s>>>(objs, n, key_or, obj, stride, perm_out)
  if (D == 2) {
    if (box) { QCP_CI(2, true); } else { QCP_CI(2, false); }
  } else {
    if (box) { QCP_CI(3, true); } else { QCP_CI(3, false); }
  }
#undef QCP_CI
  return hipGetLastError();
}

hipError_t launch_gather_objs(const float* objs, const uint32_t* perm, const uint64_t* packed,
                              uint64_t n, unsigned pbits, unsigned D, bool box, float* obj,
                              uint64_t stride, uint32_t* perm_out, uint32_t fuse_leaves,
                              float* nbox, uint64_t nstride, uint64_t leaf_off, hipStream_t s) {
  const bool pk = packed != nullptr;
  if (pk && (pbits == 0 || pbits > 32 || perm_out == nullptr)) return hipErrorInvalidValue;
  if (fuse_leaves != 0 && fuse_leaves != 32 && fuse_leaves != 64) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  unsigned g = grid_for(n, 256, 16384);
  const bool vec = ((uintptr_t)objs & 15u) == 0;
#define QCP_GO3(DD, BB, VV, PP, FF)                                                        \
  k_gather_objs<DD, BB, VV, PP, FF><<<g, 256, 0, s>>>(objs, perm, packed, n, pbits, obj,   \
                                                      stride, perm_out, nbox, nstride,     \
                                                      leaf_off)
#define QCP_GO2(DD, BB, VV, PP)                                      \
  switch (fuse_leaves) {                                             \
    case 32: QCP_GO3(DD, BB, VV, PP, 32); break;                     \
    case 64: QCP_GO3(DD, BB, VV, PP, 64); break;                     \
    default: QCP_GO3(DD, BB, VV, PP, 0); break;                      \
  }
#define QCP_GO1(DD, BB)                                              \
  if (vec) { if (pk) { QCP_GO2(DD, BB, true, true) } else { QCP_GO2(DD, BB, true, false) } } \
  else { if (pk) { QCP_GO2(DD, BB, false, true) } else { QCP_GO2(DD, BB, false, false) } }
  if (D == 2) {
    if (box) { QCP_GO1(2, true) } else { QCP_GO1(2, false) }
  } else if (D == 3) {
    if (box) { QCP_GO1(3, true) } else { QCP_GO1(3, false) }
  } else {
    return hipErrorInvalidValue;
  }
#undef QCP_GO1
#undef QCP_GO2
#undef QCP_GO3
  return hipGetLastError();
}

hipError_t launch_gather_objs_planned(const float* objs, uint32_t* perm, const uint32_t* perm2,
                                      const uint64_t* packed, const uint64_t* packed2,
                                      uint64_t n, unsigned pbits, unsigned passes,
                                      const unsigned long long* key_or, unsigned D, bool box,
                                      float* obj, uint64_t stride, uint32_t fuse_leaves,
                                      float* nbox, uint64_t nstride, uint64_t leaf_off,
                                      hipStream_t s) {
  if (perm == nullptr || perm2 == nullptr || packed == nullptr || packed2 == nullptr ||
      key_or == nullptr || pbits > 32)
    return hipErrorInvalidValue;
  if (fuse_leaves != 0 && fuse_leaves != 32 && fuse_leaves != 64) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  unsigned g = grid_for(n, 256, 16384);
  const bool vec = ((uintptr_t)objs & 15u) == 0;
#define QCP_GP3(DD, BB, VV, FF)                                                              \
  k_gather_objs_planned<DD, BB, VV, FF><<<g, 256, 0, s>>>(objs, perm, perm2, packed, packed2, \
                                                          n, pbits, obj, stride, perm, nbox,   \
                                                          nstride, leaf_off, passes, key_or)
#define QCP_GP2(DD, BB, VV)                          \
  switch (fuse_leaves) {                             \
    case 32: QCP_GP3(DD, BB, VV, 32); break;         \
    case 64: QCP_GP3(DD, BB, VV, 64); break;         \
    default: QCP_GP3(DD, BB, VV, 0); break;          \
  }
#define QCP_GP1(DD, BB) \
  if (vec) { QCP_GP2(DD, BB, true) } else { QCP_GP2(DD, BB, false) }
  if (D == 2) {
    if (box) { QCP_GP1(2, true) } else { QCP_GP1(2, false) }
  } else if (D == 3) {
    if (box) { QCP_GP1(3, true) } else { QCP_GP1(3, false) }
  } else {
    return hipErrorInvalidValue;
  }
#undef QCP_GP1
#undef QCP_GP2
#undef QCP_GP3
  return hipGetLastError();
}

hipError_t launch_keys_sorted(const float* obj, uint64_t stride, uint64_t n, unsigned D,
                              bool box, bool ref, bool center, unsigned nb, uint64_t* keys,
                              hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (nb != (ref ? (D == 2 ? 31u : 20u) : 16u)) return hipErrorInvalidValue;
  unsigned g = grid_for((n + kKeysPerThread - 1) / kKeysPerThread, 256, 4096);
#define QCP_KS(DD, BB, RR, CC) \
  k_keys_sorted<DD, BB, RR, CC><<<g, 256, 0, s>>>(obj, stride, n, keys)
  if (D == 2) {
    if (box) {
      if (ref) { if (center) QCP_KS(2, true, true, true); else QCP_KS(2, true, true, false); }
      else { if (center) QCP_KS(2, true, false, true); else QCP_KS(2, true, false, false); }
    } else {
      if (ref) QCP_KS(2, false, true, false); else QCP_KS(2, false, false, false);
    }
  } else if (D == 3) {
    if (box) {
      if (ref) { if (center) QCP_KS(3, true, true, true); else QCP_KS(3, true, true, false); }
      else { if (center) QCP_KS(3, true, false, true); else QCP_KS(3, true, false, false); }
    } else {
      if (ref) QCP_KS(3, false, true, false); else QCP_KS(3, false, false, false);
    }
  } else {
    return hipErrorInvalidValue;
  }
#undef QCP_KS
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// k-means + Hilbert order (the assignment step of KmeansMapper::cuda_kmeans,
// find_nearest_cluster, src/mapper/kmeans_mapper.cpp:100-126,280-370): each
// object's key point goes to the nearest centre by FP32 squared distance
// summed over d = 0..D-1 (no contraction), ties to the lower cluster.  Any K:
// the centres pass through LDS in tiles of kAssignTile (centre k's distance is
// computed exactly as in a single sequential scan, and the tiles are visited
// in cluster order, so the strict `<` keeps the first minimum); every thread
// holds R objects in registers so that one broadcast LDS read of a centre
// serves R distance evaluations.  `changes` counts objects whose cluster
// changed.
constexpr uint32_t kAssignTile = 1024;

template <int D>
struct CentreVec;
template <>
struct CentreVec<2> {
  using T = float2;
};
template <>
struct CentreVec<3> {
  using T = float4;
};

template <int D, bool BOX, bool CENTER, int R>
__global__ __launch_bounds__(256) void k_assign(const float* __restrict__ objs, uint64_t n,
                                                const float* __restrict__ centers, uint32_t K,
                                                uint32_t* __restrict__ cid, uint32_t* changes) {
  constexpr int E = BOX ? 2 * D : D;
  using V = typename CentreVec<D>::T;
  __shared__ V tile[kAssignTile];
  const uint64_t base = (uint64_t)blockIdx.x * (256 * R) + threadIdx.x;
  float p[R][D];
  float best[R];
  uint32_t bk[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint64_t i = base + (uint64_t)r * 256;
    const uint64_t j = i < n ? i : n - 1;  // tail lanes evaluate a valid object, store nothing
#pragma unroll
    for (int d = 0; d < D; ++d) {
      p[r][d] = objs[j * E + d];
      if (CENTER && BOX) p[r][d] = (p[r][d] + objs[j * E + D + d]) * 0.5f;
    }
    // cluster 0 starts the scan (the sequential rule's `k == 0` case)
    float dd = 0.0f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float t = p[r][d] - centers[d];
      dd = dd + t * t;
    }
    best[r] = dd;
    bk[r] = 0;
  }
  for (uint32_t k0 = 0; k0 < K; k0 += kAssignTile) {
    const uint32_t m = K - k0 < kAssignTile ? K - k0 : kAssignTile;
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < m; t += 256) {
      V c;
      c.x = centers[(uint64_t)(k0 + t) * D];
      c.y = centers[(uint64_t)(k0 + t) * D + 1];
      if constexpr (D == 3) {
        c.z = centers[(uint64_t)(k0 + t) * D + 2];
        c.w = 0.0f;
      }
      tile[t] = c;
    }
    __syncthreads();
#pragma unroll 4
    for (uint32_t t = k0 == 0 ? 1u : 0u; t < m; ++t) {
      const V c = tile[t];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float t0 = p[r][0] - c.x;
        const float t1 = p[r][1] - c.y;
        float dd = 0.0f;
        dd = dd + t0 * t0;
        dd = dd + t1 * t1;
        if constexpr (D == 3) {
          const float t2 = p[r][2] - c.z;
          dd = dd + t2 * t2;
        }
        if (dd < best[r]) {
          best[r] = dd;
          bk[r] = k0 + t;
        }
      }
    }
  }
  uint32_t changed = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint64_t i = base + (uint64_t)r * 256;
    if (i < n && cid[i] != bk[r]) {
      cid[i] = bk[r];
      ++changed;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) changed += __shfl_xor(changed, off, kWave);
  if (lane_id() == 0 && changed) atomicAdd(changes, changed);
}

// keys[i] = cluster rank << shift_rank | keys[i] >> shift_key
__global__ void k_compose(uint64_t* keys, const uint32_t* __restrict__ cid,
                          const uint32_t* __restrict__ rank, uint64_t n, unsigned shift_key,
                          unsigned shift_rank) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t hi = shift_rank < 64 ? ((uint64_t)rank[cid[i]] << shift_rank) : 0ull;
    keys[i] = hi | (keys[i] >> shift_key);
  }
}

hipError_t launch_assign(const float* objs, uint64_t n, unsigned D, bool box, bool center,
                         const float* centers, uint32_t K, uint32_t* cid, uint32_t* changes,
                         hipStream_t s) {
  if (n == 0 || K == 0) return hipSuccess;
  // R objects a thread, the largest of 4, 2, 1 that still leaves 2048
  // workgroups (eight a CU)
  const unsigned R = n >= 256ull * 4 * 2048 ? 4u : (n >= 256ull * 2 * 2048 ? 2u : 1u);
  const uint64_t g64 = (n + 256ull * R - 1) / (256ull * R);
  if (g64 > 0x7FFFFFFFull) return hipErrorInvalidValue;
  const unsigned g = (unsigned)g64;
#define QCP_ASSIGN(DD, BB, CC)                                                        \
  do {                                                                              \
    if (R == 4)                                                                     \
      k_assign<DD, BB, CC, 4><<<g, 256, 0, s>>>(objs, n, centers, K, cid, changes); \
    else if (R == 2)                                                                \
      k_assign<DD, BB, CC, 2><<<g, 256, 0, s>>>(objs, n, centers, K, cid, changes); \
    else                                                                            \
      k_assign<DD, BB, CC, 1><<<g, 256, 0, s>>>(objs, n, centers, K, cid, changes); \
  } while (0)
  if (D == 2) {
    if (box) { if (center) QCP_ASSIGN(2, true, true); else QCP_ASSIGN(2, true, false); }
    else QCP_ASSIGN(2, false, false);
  } else if (D == 3) {
    if (box) { if (center) QCP_ASSIGN(3, true, true); else QCP_ASSIGN(3, true, false); }
    else QCP_ASSIGN(3, false, false);
  } else {
    return hipErrorInvalidValue;
  }
#undef QCP_ASSIGN
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// k-means centre update on the device, in the exact order of the defined rule
// (FP64 sums of each cluster's key points in index order within a chunk of
// kKmChunk objects, chunk sums added in chunk order).  One workgroup per
// chunk walks it in sub-blocks of kKmSub objects, in order: a sub-block's keys
// (cluster << 11 | position) are bitonic-sorted in LDS and its key points are
// staged in LDS, so every cluster's members come out in index order, and the
// thread at the start of a cluster's run adds them one by one onto the chunk's
// running sums (read, add each member, write back: the same sequence of FP64
// additions as a serial loop).  The barrier between sub-blocks orders one
// sub-block's sums before the next.
constexpr uint32_t kKmSub = 2048;  // 11-bit positions: keys stay below the sentinel

template <int D, bool BOX, bool CENTER>
__device__ __forceinline__ float km_point(const float* __restrict__ objs, uint64_t o, int d) {
  constexpr int E = BOX ? 2 * D : D;
  float p = objs[o * E + d];
  if (CENTER && BOX) p = (p + objs[o * E + D + d]) * 0.5f;
  return p;
}

constexpr uint32_t kKmThreads = 512;

template <int D, bool BOX, bool CENTER>
__global__ __launch_bounds__(kKmThreads) void k_km_partial(
    const float* __restrict__ objs, uint64_t n, const uint32_t* __restrict__ cid, uint32_t K,
    uint64_t c0, double* __restrict__ part, uint32_t* __restrict__ cnt) {
  constexpr uint32_t T = kKmThreads;
  constexpr int SLOTS = kKmSub / T;  // sorted positions a thread owns: tid + T j
  __shared__ uint32_t sk[kKmSub];
  __shared__ float kp[D][kKmSub];  // the sub-block's key points
  const uint32_t tid = threadIdx.x;
  const uint64_t base = (c0 + blockIdx.x) * kKmChunk;
  double* pc = part + (uint64_t)blockIdx.x * K * D;
  uint32_t* nc = cnt + (uint64_t)blockIdx.x * K;
  for (uint32_t sb = 0; sb < kKmChunk / kKmSub; ++sb) {
    const uint64_t b0 = base + (uint64_t)sb * kKmSub;
    if (b0 >= n) break;  // uniform
    const uint32_t m = n - b0 < kKmSub ? (uint32_t)(n - b0) : kKmSub;
    for (uint32_t t = tid; t < kKmSub; t += T) {
      sk[t] = t < m ? ((cid[b0 + t] << 11) | t) : 0xFFFFFFFFu;
      if (t < m) {
#pragma unroll
        for (int d = 0; d < D; ++d) kp[d][t] = km_point<D, BOX, CENTER>(objs, b0 + t, d);
      }
    }
    __syncthreads();
    // bitonic sort: a thread's compare-exchanges of one stage are issued
    // together (independent pairs), then the stage's barrier
    for (uint32_t k = 2; k <= kKmSub; k <<= 1)
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        uint32_t a[SLOTS], b[SLOTS];
#pragma unroll
        for (int r = 0; r < SLOTS; ++r) {
          const uint32_t i = tid + T * r, x = i ^ j;
          a[r] = sk[i];
          b[r] = x > i ? sk[x] : 0u;
        }
#pragma unroll
        for (int r = 0; r < SLOTS; ++r) {
          const uint32_t i = tid + T * r, x = i ^ j;
          if (x > i && ((a[r] > b[r]) == ((i & k) == 0))) {
            sk[i] = b[r];
            sk[x] = a[r];
          }
        }
        __syncthreads();
      }
    // Keys are unique (positions differ), so this is the (cluster, position)
    // order.  The thread owning a run's first slot adds the run; a thread's
    // runs read their running sums together, then add, then write back.
    uint32_t kk[SLOTS];
    bool start[SLOTS];
    double acc[SLOTS][D];
    uint32_t len[SLOTS];
#pragma unroll
    for (int j = 0; j < SLOTS; ++j) {
      const uint32_t p = tid + T * j;
      kk[j] = sk[p] >> 11;
      start[j] = p < m && (p == 0 || (sk[p - 1] >> 11) != kk[j]);
      len[j] = 0;
    }
#pragma unroll
    for (int j = 0; j < SLOTS; ++j)
      if (start[j]) {
#pragma unroll
        for (int d = 0; d < D; ++d) acc[j][d] = pc[(uint64_t)kk[j] * D + d];
      }
#pragma unroll
    for (int j = 0; j < SLOTS; ++j)
      if (start[j]) {
        for (uint32_t q = tid + T * j; q < m && (sk[q] >> 11) == kk[j]; ++q) {
          const uint32_t o = sk[q] & (kKmSub - 1);
#pragma unroll
          for (int d = 0; d < D; ++d) acc[j][d] += (double)kp[d][o];
          ++len[j];
        }
      }
#pragma unroll
    for (int j = 0; j < SLOTS; ++j)
      if (start[j]) {
#pragma unroll
        for (int d = 0; d < D; ++d) pc[(uint64_t)kk[j] * D + d] = acc[j][d];
        nc[kk[j]] += len[j];
      }
    __syncthreads();
  }
}

// Cluster k: the chunks' partials added in chunk order (a chunk without
// members of k adds nothing, as in the serial rule).
template <int D>
__global__ void k_km_combine(const double* __restrict__ part, const uint32_t* __restrict__ cnt,
                             uint64_t nc, uint32_t K, double* __restrict__ sum,
                             uint64_t* __restrict__ total) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  double s[D];
#pragma unroll
  for (int d = 0; d < D; ++d) s[d] = sum[k * D + d];
  uint64_t t = total[k];
  for (uint64_t c = 0; c < nc; ++c) {
    const uint32_t m = cnt[c * K + k];
    if (m) {
      t += m;
#pragma unroll
      for (int d = 0; d < D; ++d) s[d] += part[(c * K + k) * D + d];
    }
  }
#pragma unroll
  for (int d = 0; d < D; ++d) sum[k * D + d] = s[d];
  total[k] = t;
}

template <int D>
__global__ void k_km_finish(const double* __restrict__ sum, const uint64_t* __restrict__ total,
                            uint32_t K, float* __restrict__ centers) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K || total[k] == 0) return;
#pragma unroll
  for (int d = 0; d < D; ++d) centers[k * D + d] = (float)(sum[k * D + d] / (double)total[k]);
}

template <int D, bool BOX, bool CENTER>
__global__ void k_km_seed(const float* __restrict__ objs, const uint32_t* __restrict__ idx,
                          uint32_t K, float* __restrict__ centers) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
#pragma unroll
  for (int d = 0; d < D; ++d) centers[k * D + d] = km_point<D, BOX, CENTER>(objs, idx[k], d);
}

#define QCP_KM_DISPATCH(CALL)                                                    \
  do {                                                                         \
    if (D == 2) {                                                              \
      if (box) { if (center) CALL(2, true, true); else CALL(2, true, false); }  \
      else CALL(2, false, false);                                              \
    } else if (D == 3) {                                                       \
      if (box) { if (center) CALL(3, true, true); else CALL(3, true, false); }  \
      else CALL(3, false, false);                                              \
    } else {                                                                   \
      return hipErrorInvalidValue;                                             \
    }                                                                          \
  } while (0)

hipError_t launch_km_seed(const float* objs, unsigned D, bool box, bool center,
                          const uint32_t* idx, uint32_t K, float* centers, hipStream_t s) {
  if (K == 0) return hipSuccess;
  const unsigned g = (K + 255) / 256;
#define QCP_KM_SEED(DD, BB, CC) k_km_seed<DD, BB, CC><<<g, 256, 0, s>>>(objs, idx, K, centers)
  QCP_KM_DISPATCH(QCP_KM_SEED);
#undef QCP_KM_SEED
  return hipGetLastError();
}

hipError_t launch_km_update(const float* objs, uint64_t n, unsigned D, bool box, bool center,
                            const uint32_t* cid, uint32_t K, uint64_t c0, uint64_t nc,
                            double* part, uint32_t* cnt, double* sum, uint64_t* total,
                            hipStream_t s) {
  if (nc == 0 || K == 0) return hipSuccess;
  if (nc > 0x7FFFFFFFull) return hipErrorInvalidValue;
#define QCP_KM_PART(DD, BB, CC) \
  k_km_partial<DD, BB, CC><<<(unsigned)nc, kKmThreads, 0, s>>>(objs, n, cid, K, c0, part, cnt)
  QCP_KM_DISPATCH(QCP_KM_PART);
#undef QCP_KM_PART
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const unsigned g = (K + 255) / 256;
  if (D == 2) k_km_combine<2><<<g, 256, 0, s>>>(part, cnt, nc, K, sum, total);
  else k_km_combine<3><<<g, 256, 0, s>>>(part, cnt, nc, K, sum, total);
  return hipGetLastError();
}

hipError_t launch_km_finish(const double* sum, const uint64_t* total, uint32_t K, unsigned D,
                            float* centers, hipStream_t s) {
  if (K == 0) return hipSuccess;
  const unsigned g = (K + 255) / 256;
  if (D == 2) k_km_finish<2><<<g, 256, 0, s>>>(sum, total, K, centers);
  else if (D == 3) k_km_finish<3><<<g, 256, 0, s>>>(sum, total, K, centers);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}
#undef QCP_KM_DISPATCH

hipError_t launch_compose(uint64_t* keys, const uint32_t* cid, const uint32_t* rank, uint64_t n,
                          unsigned shift_key, unsigned shift_rank, hipStream_t s) {
  k_compose<<<grid_for(n, 256, 8192), 256, 0, s>>>(keys, cid, rank, n, shift_key, shift_rank);
  return hipGetLastError();
}

#define QCP_MW_SWITCH(MWV, ...)                        \
  switch (MWV) {                                       \
    case 1: { constexpr int MW = 1; __VA_ARGS__; } break; \
    case 2: { constexpr int MW = 2; __VA_ARGS__; } break; \
    case 3: { constexpr int MW = 3; __VA_ARGS__; } break; \
    case 4: { constexpr int MW = 4; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue;              \
  }

hipError_t launch_leaf_boxes(const float* obj, uint64_t ostride, uint64_t n, unsigned D,
                             bool box, uint32_t F, uint64_t nleaves, float* nbox,
                             uint64_t nstride, uint64_t leaf_off, hipStream_t s) {
  if ((F == 16 || F == 32 || F == 64 || F == 128 || F == 256) && (D == 2 || D == 3)) {
    // streamed: one wave a run of four leaves
    const unsigned gs = grid_for((n + 4ull * F - 1) / (4ull * F) * kWave, 256, 16384);
#define QCP_LB(DD, BB)                                                                        \
  switch (F) {                                                                                \
    case 16: k_leaf_boxes_row<DD, BB, 16><<<gs, 256, 0, s>>>(obj, ostride, n, nbox, nstride, leaf_off); break; \
    case 32: k_leaf_boxes_row<DD, BB, 32><<<gs, 256, 0, s>>>(obj, ostride, n, nbox, nstride, leaf_off); break; \
    case 64: k_leaf_boxes_row<DD, BB, 64><<<gs, 256, 0, s>>>(obj, ostride, n, nbox, nstride, leaf_off); break; \
    case 128: k_leaf_boxes_row<DD, BB, 128><<<gs, 256, 0, s>>>(obj, ostride, n, nbox, nstride, leaf_off); break; \
    default: k_leaf_boxes_row<DD, BB, 256><<<gs, 256, 0, s>>>(obj, ostride, n, nbox, nstride, leaf_off); break; \
  }
    if (D == 2) {
      if (box) { QCP_LB(2, true) } else { QCP_LB(2, false) }
    } else {
      if (box) { QCP_LB(3, true) } else { QCP_LB(3, false) }
    }
#undef QCP_LB
    return hipGetLastError();
  }
  unsigned g = (unsigned)((nleaves * kWave + 255) / 256);
  int mwv = (int)((F + 63) / 64);
  if (D == 2) {
    if (box) { QCP_MW_SWITCH(mwv, k_leaf_boxes<2, true, MW><<<g, 256, 0, s>>>(obj, ostride, n, F, nleaves, nbox, nstride, leaf_off)) }
    else { QCP_MW_SWITCH(mwv, k_leaf_boxes<2, false, MW><<<g, 256, 0, s>>>(obj, ostride, n, F, nleaves, nbox, nstride, leaf_off)) }
  } else if (D == 3) {
    if (box) { QCP_MW_SWITCH(mwv, k_leaf_boxes<3, true, MW><<<g, 256, 0, s>>>(obj, ostride, n, F, nleaves, nbox, nstride, leaf_off)) }
    else { QCP_MW_SWITCH(mwv, k_leaf_boxes<3, false, MW><<<g, 256, 0, s>>>(obj, ostride, n, F, nleaves, nbox, nstride, leaf_off)) }
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_level_boxes(float* nbox, uint64_t nstride, unsigned D, uint64_t child_off,
                              uint64_t child_cnt, uint32_t F, uint64_t parent_off,
                              uint64_t parent_cnt, hipStream_t s) {
  unsigned g = (unsigned)((parent_cnt * kWave + 255) / 256);
  int mwv = (int)((F + 63) / 64);
  if (D == 2) {
    QCP_MW_SWITCH(mwv, k_level_boxes<2, MW><<<g, 256, 0, s>>>(nbox, nstride, child_off, child_cnt, F, parent_off, parent_cnt))
  } else if (D == 3) {
    QCP_MW_SWITCH(mwv, k_level_boxes<3, MW><<<g, 256, 0, s>>>(nbox, nstride, child_off, child_cnt, F, parent_off, parent_cnt))
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_tree_boxes(const float* obj, uint64_t ostride, const uint32_t* nfirst,
                             const uint32_t* ncount, unsigned D, bool box, uint32_t F,
                             uint32_t pad, uint64_t lo, uint64_t hi, float* nbox,
                             uint64_t nstride, hipStream_t s) {
  if (hi <= lo) return hipSuccess;
  unsigned g = (unsigned)(((hi - lo) * kWave + 255) / 256);
  int mwv = (int)((F + 63) / 64);
  if (D == 2) {
    if (box) { QCP_MW_SWITCH(mwv, k_tree_boxes<2, true, MW><<<g, 256, 0, s>>>(obj, ostride, nfirst, ncount, pad, lo, hi, nbox, nstride)) }
    else { QCP_MW_SWITCH(mwv, k_tree_boxes<2, false, MW><<<g, 256, 0, s>>>(obj, ostride, nfirst, ncount, pad, lo, hi, nbox, nstride)) }
  } else if (D == 3) {
    if (box) { QCP_MW_SWITCH(mwv, k_tree_boxes<3, true, MW><<<g, 256, 0, s>>>(obj, ostride, nfirst, ncount, pad, lo, hi, nbox, nstride)) }
    else { QCP_MW_SWITCH(mwv, k_tree_boxes<3, false, MW><<<g, 256, 0, s>>>(obj, ostride, nfirst, ncount, pad, lo, hi, nbox, nstride)) }
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_iota_u32(uint32_t* p, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_iota_u32<<<grid_for(n, 256, 16384), 256, 0, s>>>(p, n);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_fill_u32<<<grid_for(n, 256, 4096), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_f32(float* p, uint64_t n, float v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_fill_f32<<<grid_for(n, 256, 4096), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

}  // namespace qcp
