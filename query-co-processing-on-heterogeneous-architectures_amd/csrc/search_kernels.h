// search_kernels.h -- range-search kernel templates over a resident tree on gfx950.
//
// Two forms:
//  * k_mprs: the reference's plain form, MPRS "restart scanning + parent
//    check" (src/tree/mphr.cpp:303-438) with one workgroup per query, launched
//    once per batch of queries by the host (src/tree/mphr.cpp:231-258).  A
//    workgroup is a single wave here: the reference's LDS min-reduction over
//    192 threads (FindMinOnGPU, src/common/macro.h:24-42) becomes a ballot and
//    a find-first-set, and the partially zeroed t_hit[] sum
//    (src/tree/mphr.cpp:327-330,429-437) becomes a popcount of the hit ballot.
//  * k_search: the persistent form.  All queries in one launch; each wave
//    takes work items (a query and a subtree root) from a device-wide ticket
//    queue and walks the canonical top-down traversal with a DFS stack of
//    (node, remaining-child bitmask) frames in LDS.  Children that are leaves
//    are streamed in 64-entry chunks, several chunks in flight per wave, from
//    the rank-ordered SoA object arrays; hits are compacted with ballot +
//    mbcnt.  With subtree donation on, a wave that sees idle waves hands the
//    unvisited siblings of its shallowest frame to the queue (work stealing).
//    Visiting children in ascending order makes ID lists come out sorted.
#pragma once

#include <hip/hip_runtime.h>

#include "qcp_device.h"
#include "qcp_kernels.h"

namespace qcp {

template <int MW>
struct Mask {
  uint64_t w[MW];
  __device__ __forceinline__ bool empty() const {
    uint64_t o = 0;
#pragma unroll
    for (int i = 0; i < MW; ++i) o |= w[i];
    return o == 0;
  }
  // lowest set bit, or -1
  __device__ __forceinline__ int lowest() const {
#pragma unroll
    for (int i = 0; i < MW; ++i)
      if (w[i]) return i * 64 + __builtin_ctzll(w[i]);
    return -1;
  }
  __device__ __forceinline__ bool test(int k) const {
    int i = k >> 6;
    uint64_t v = 0;
#pragma unroll
    for (int t = 0; t < MW; ++t)
      if (t == i) v = w[t];
    return (v >> (k & 63)) & 1ull;
  }
  __device__ __forceinline__ void clear(int k) {
#pragma unroll
    for (int t = 0; t < MW; ++t)
      if (t == (k >> 6)) w[t] &= ~(1ull << (k & 63));
  }
  // end (exclusive) of the run of consecutive set bits of `this` starting at k
  __device__ __forceinline__ int run_end(int k) const {
    int e = k;
#pragma unroll
    for (int t = 0; t < MW; ++t) {
      if (e >= t * 64 && e < (t + 1) * 64) {
        uint64_t x = ~(w[t] >> (e & 63));
        int r = x ? __builtin_ctzll(x) : 64 - (e & 63);
        r = r > 64 - (e & 63) ? 64 - (e & 63) : r;
        e += r;
      }
    }
    return e;
  }
  __device__ __forceinline__ void clear_range(int k, int e) {
#pragma unroll
    for (int t = 0; t < MW; ++t) {
      int lo = k > t * 64 ? k - t * 64 : 0;
      int hi = e < (t + 1) * 64 ? e - t * 64 : 64;
      if (hi > lo) {
        uint64_t m = (hi - lo == 64) ? ~0ull : (((1ull << (hi - lo)) - 1ull) << lo);
        w[t] &= ~m;
      }
    }
  }
  __device__ __forceinline__ Mask operator&(const Mask& o) const {
    Mask r;
#pragma unroll
    for (int i = 0; i < MW; ++i) r.w[i] = w[i] & o.w[i];
    return r;
  }
  __device__ __forceinline__ Mask andnot(const Mask& o) const {
    Mask r;
#pragma unroll
    for (int i = 0; i < MW; ++i) r.w[i] = w[i] & ~o.w[i];
    return r;
  }
  __device__ __forceinline__ static Mask none() {
    Mask r;
#pragma unroll
    for (int i = 0; i < MW; ++i) r.w[i] = 0;
    return r;
  }
};

template <int MW>
struct Frame {
  uint32_t node;
  uint32_t level;
  Mask<MW> m;   // children still to visit
  Mask<MW> lm;  // children that are leaves
  Mask<MW> im;  // children whose box lies inside the query
};

// Per-wave traversal context (all members wave-uniform except the *_lane sums).
template <int D>
struct Ctx {
  float ql[D], qh[D];
  // the point test's canonical bounds: NaN bounds as -inf / +inf (a NaN bound
  // rejects nothing); pempty: some dimension has pl > ph (no point can hit)
  float pl[D], ph[D];
  bool pempty;
  uint32_t q;
  uint32_t cnt;           // hits of this item (wave-uniform part)
  uint32_t cnt_lane;      // per-lane hits of the 4-entries-a-lane leaf path
  uint64_t out_base;      // ID write pass: next output slot
  // ID write pass: IDs are staged in a per-wave LDS window of 2 x 64 slots
  // holding output slots [wbase, wbase + 128), wbase a multiple of 64; this
  // query owns the slots from wlo on.  Whole aligned 64-slot blocks are
  // written with one aligned 256-byte store each, so only the blocks at a
  // query's two ends are partial.
  uint32_t* idb;
  uint64_t wbase, wlo;
  uint64_t s1_lane;       // per-lane partial sum of hit IDs
  uint64_t s2_lane;       // per-lane partial sum of squared hit IDs
  uint64_t s1c, s2c;      // power sums of the IDs of contained subtrees
  uint64_t ie, le;        // entries visited (the canonical traversal unless contain is on)
  uint32_t cr;            // contained subtree runs
};

// The leaf test of one entry (lo, hi: its lower and upper corners; the same
// for a point): the box overlap test, or for points the med3 form over the
// canonical bounds (scan_runs skips the leaf runs of a query that no point
// can hit, pempty, after counting their entries).
template <int D, bool BOX>
__device__ __forceinline__ bool leaf_hit(const Ctx<D>& c, const float* lo, const float* hi) {
  if (BOX) return overlap_test<D>(c.ql, c.qh, lo, hi);
  return point_test<D>(c.pl, c.ph, lo);
}

// Output modes of the persistent form.
enum { kOutCount = 0, kOutHash = 1, kOutIds = 2 };

// Shortest leaf run that counts and hashes scan in 256-entry groups (four
// entries a lane) rather than in 64-entry chunks.
constexpr uint32_t kGroupMin = 128;

// FP32 in the global address space (a pointer rebuilt from an integer is
// otherwise generic, which would turn its loads into flat loads)
typedef __attribute__((address_space(1))) float GlobalF32;
typedef float F32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) F32x4 GlobalF32x4;

// Evaluate the children [cbase, cbase+cc) (node ids) of an internal node
// against the query; for explicit layouts also their leaf flags.
template <int D, int MW, bool EXPL>
__device__ __forceinline__ void eval_children(const SearchArgs& a, const Ctx<D>& c,
                                              uint32_t cbase, uint32_t cc, Mask<MW>& m,
                                              Mask<MW>& lm, Mask<MW>& im) {
  const uint32_t lane = lane_id();
  float lo[MW][D], hi[MW][D];
  uint32_t cw[MW];
#pragma unroll
  for (int w = 0; w < MW; ++w) {
    const uint64_t g = (uint64_t)cbase + w * kWave + lane;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      lo[w][d] = a.nbox[(uint64_t)d * a.node_stride + g];
      hi[w][d] = a.nbox[(uint64_t)(D + d) * a.node_stride + g];
    }
    if (EXPL) cw[w] = a.ncount[g];
  }
#pragma unroll
  for (int w = 0; w < MW; ++w) {
    const uint32_t k = w * kWave + lane;
    bool hit = (k < cc) && overlaps<D>(c.ql, c.qh, lo[w], hi[w]);
    m.w[w] = ballot(hit);
    lm.w[w] = EXPL ? ballot((cw[w] & kLeafBit) != 0) : ~0ull;
    im.w[w] = a.contain ? ballot(hit && inside<D>(c.ql, c.qh, lo[w], hi[w])) : 0ull;
  }
}

// ID write pass: write the staged slots [wlo, out_base) of the window's first
// block (a query's last block, or the block before a directly written run).
template <int D>
__device__ __forceinline__ void flush_partial(const SearchArgs& a, Ctx<D>& c) {
  const uint32_t lane = lane_id();
  __builtin_amdgcn_wave_barrier();
  const uint32_t v = c.idb[lane];
  const uint64_t pos = c.wbase + lane;
  if (pos >= c.wlo && pos < c.out_base) __builtin_nontemporal_store(v, a.ids + pos);
  __builtin_amdgcn_wave_barrier();
}

// ID write pass: the window's first block is complete; write it (from wlo on)
// and slide the window by one block.
template <int D>
__device__ __forceinline__ void flush_block(const SearchArgs& a, Ctx<D>& c) {
  const uint32_t lane = lane_id();
  __builtin_amdgcn_wave_barrier();
  const uint32_t v = c.idb[lane];
  const uint32_t v2 = c.idb[kWave + lane];
  if (c.wbase + lane >= c.wlo) __builtin_nontemporal_store(v, a.ids + c.wbase + lane);
  __builtin_amdgcn_wave_barrier();
  c.idb[lane] = v2;
  c.wbase += kWave;
  c.wlo = c.wbase;
}

// Report the ranks [s, e) as hits without reading them (a run of children
// whose boxes lie inside the query).
template <int D, int OUT>
__device__ __forceinline__ void emit_range(const SearchArgs& a, Ctx<D>& c, uint64_t s,
                                           uint64_t e) {
  const uint64_t len = e - s;
  c.cnt += (uint32_t)len;
  c.cr += 1;
  if (OUT == kOutHash) {
    c.s1c += psum1((uint32_t)e) - psum1((uint32_t)s);
    c.s2c += psum2((uint32_t)e) - psum2((uint32_t)s);
  } else if (OUT == kOutIds) {
    const uint32_t lane = lane_id();
    flush_partial<D>(a, c);  // staged IDs before the run, then the run directly
    uint32_t* dst = a.ids + c.out_base;
    // dword stores up to a 16-byte boundary, then 16-byte stores, then the tail
    const uint64_t head = ((4 - (c.out_base & 3)) & 3) < len ? ((4 - (c.out_base & 3)) & 3) : len;
    if (lane < head) dst[lane] = (uint32_t)(s + 1 + lane);
    const uint64_t body = (len - head) & ~3ull;
    uint4* dst4 = reinterpret_cast<uint4*>(dst + head);
    const uint32_t b0 = (uint32_t)(s + 1 + head);
    for (uint64_t i = lane; i < body / 4; i += kWave) {
      const uint32_t v = b0 + (uint32_t)(4 * i);
      __builtin_nontemporal_store(v, &dst4[i].x);
      __builtin_nontemporal_store(v + 1, &dst4[i].y);
      __builtin_nontemporal_store(v + 2, &dst4[i].z);
      __builtin_nontemporal_store(v + 3, &dst4[i].w);
    }
    const uint64_t t0 = head + body;
    if (t0 + lane < len) dst[t0 + lane] = (uint32_t)(s + 1 + t0 + lane);
    c.out_base += len;
    c.wbase = c.out_base & ~(uint64_t)(kWave - 1);
    c.wlo = c.out_base;
  }
}

// Results of one 64-entry chunk whose first rank is st: `hit` is this lane's
// entry overlapping the query.
template <int D, int OUT>
__device__ __forceinline__ void chunk_out(const SearchArgs& a, Ctx<D>& c, uint32_t st, bool hit) {
  const uint32_t lane = lane_id();
  const uint64_t b = ballot(hit);
  const uint32_t nb = (uint32_t)__popcll(b);
  if (OUT == kOutIds) {
    // staged in LDS, written in aligned blocks (nontemporal: the ID stream is
    // written once and read back only by the host, so it should not displace
    // the tree in L2)
    if (hit) c.idb[(uint32_t)(c.out_base - c.wbase) + rank_below(b)] = st + lane + 1u;
    c.out_base += nb;
    if (c.out_base - c.wbase >= kWave) flush_block<D>(a, c);
  } else if (OUT == kOutHash) {
    // branch-free: a select, a 64-bit add and one v_mad_u64_u32 per chunk
    const uint32_t id = st + lane + 1u;
    const uint32_t x = hit ? id : 0u;
    c.s1_lane += x;
    c.s2_lane += (uint64_t)x * x;
  }
  c.cnt += nb;
}

// Stream the leaf runs at the bottom of mask m (bits that are also set in lm),
// U 64-entry chunks in flight, stopping at the first set bit that is neither a
// leaf nor inside the query.  Runs of children inside the query (im) are
// reported whole without a read.  run_range(k, k2, &s, &e) maps a run of leaf
// children [k, k2) to its rank range, sub_range a run of any children.
template <int D, bool BOX, int MW, int OUT, int U, typename RunRange, typename SubRange>
__device__ __forceinline__ void scan_runs(const SearchArgs& a, Ctx<D>& c, Mask<MW>& m,
                                          const Mask<MW>& lm, const Mask<MW>& im,
                                          RunRange run_range, SubRange sub_range) {
  constexpr int E = BOX ? 2 * D : D;
  const uint32_t lane = lane_id();
  // Ranks fit in 32 bits (IDs are u32), so the run bookkeeping stays in SGPRs
  // with scalar compares, and each chunk load is a uniform base pointer plus
  // the lane offset (no per-lane 64-bit address arithmetic).
  const float* col[E];
#pragma unroll
  for (int e = 0; e < E; ++e) col[e] = a.obj + (uint64_t)e * a.obj_stride;
  uint32_t cs = 0, ce = 0;
  while (true) {
    if constexpr (OUT != kOutIds && U >= 4) {
      // counts and hashes: whole groups of 256 entries, four consecutive
      // entries a lane from one 16-byte load per column, U / 4 groups in
      // flight (a quarter of the chunk path's load instructions for the same
      // bytes; the order of the hits does not matter for a count or a sum).
      // Hits are counted per lane and summed when the item ends (a scalar
      // popcount of each entry's hit mask loads the scalar unit, which the
      // traversal also needs).  The run must start on a 16-byte boundary.
      // Hashes take the first block below, counts alone the second.  A run
      // reaches them whole (the refill below leaves aligned runs of at least
      // kGroupMin entries to them): groups of G, then single groups, the
      // last one partial.  A lane whose entries lie past the run's end still
      // loads them (the object arrays hold 1024 slack entries past N) and
      // does not count them.
      constexpr int G = U / 4;
      constexpr uint32_t kPk[4] = {1u, 1u | (1u << 9) | (1u << 18), 1u | (2u << 9) | (4u << 18),
                                   1u | (3u << 9) | (9u << 18)};
      // the leaf test of lane entry j of a group (w: its four entries a column)
      auto hit_of = [&](const F32x4* w, int j) {
        float v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = w[e][j];
        return leaf_hit<D, BOX>(c, v, BOX ? v + D : v);
      };
      auto load_group = [&](F32x4 (&w)[E], uint32_t at) {
#pragma unroll
        for (int e = 0; e < E; ++e)
          w[e] = reinterpret_cast<const GlobalF32x4*>(
              uni64(reinterpret_cast<uint64_t>(col[e] + at)))[lane];
      };
      if (OUT == kOutHash && (cs & 3u) == 0 && ce - cs >= kGroupMin) {
        // ID sums from per-group hit moments.  Lane l's entries j = 0..3 of
        // group k (256 entries from the segment start) have IDs
        // A_k + 4l + j, A_k = R + 256k.  Per group the lane forms, in one
        // register, n (hits), J1 = sum of j and J2 = sum of j^2 over its hits
        // (9, 9 and 14-bit fields), adds that word to cp, and adds k*n, k*J1
        // and k^2*n to g1, gj1, g2; flush() turns these exact small integers
        // into the 64-bit sums of the IDs and of their squares (mod 2^64, as
        // the per-entry sums are), every 64 groups and at the segment's end.
        // That is 12 VALU per four entries where a select, a 64-bit add, a
        // 64-bit multiply-add and the ID's add per entry took 16 and more.
        const uint32_t lane4 = 4u * lane;
        uint32_t r0 = cs + 1u;
        uint32_t cp = 0, g1 = 0, gj1 = 0, g2 = 0, gk = 0;
        auto flush = [&]() {
          const uint64_t N = cp & 0x1FFu, J1 = (cp >> 9) & 0x1FFu, J2 = cp >> 18;
          const uint64_t R = r0, L = lane4;
          const uint64_t an = R * N + 256ull * g1;  // sum over groups of n_k A_k
          c.cnt_lane += (uint32_t)N;
          c.s1_lane += an + L * N + J1;
          c.s2_lane += R * R * N + 512ull * R * g1 + 65536ull * g2 +
                       2ull * (L * an + R * J1 + 256ull * gj1) + L * L * N + 2ull * L * J1 + J2;
          cp = g1 = gj1 = g2 = 0;
        };
        auto add_group = [&](uint32_t pw, uint32_t k) {  // k < 66: 24-bit products
          cp += pw;
          const uint32_t n = pw & 0x1FFu, j1 = (pw >> 9) & 0x1FFu;
          g1 += __umul24(k, n);
          gj1 += __umul24(k, j1);
          g2 += __umul24(uni(k * k), n);
        };
        auto next_group = [&](uint32_t step, uint32_t groups) {
          cs += step;
          gk += groups;
          if (gk >= 64u) {
            flush();
            r0 = cs + 1u;
            gk = 0;
          }
        };
        while (ce - cs >= (uint32_t)(G * 256)) {
          F32x4 w[G][E];
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const GlobalF32x4* base = reinterpret_cast<const GlobalF32x4*>(
                uni64(reinterpret_cast<uint64_t>(col[e] + cs)));
#pragma unroll
            for (int g = 0; g < G; ++g) w[g][e] = base[g * kWave + (int)lane];
          }
#pragma unroll
          for (int g = 0; g < G; ++g) {
            uint32_t pw = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) pw += hit_of(w[g], j) ? kPk[j] : 0u;
            add_group(pw, uni(gk + (uint32_t)g));
          }
          next_group((uint32_t)(G * 256), (uint32_t)G);
        }
        while (ce != cs) {
          const uint32_t rem = ce - cs;
          F32x4 w[E];
          load_group(w, cs);
          uint32_t pw = 0;
#pragma unroll
          for (int j = 0; j < 4; ++j)
            pw += (lane4 + (uint32_t)j < rem && hit_of(w, j)) ? kPk[j] : 0u;
          add_group(pw, uni(gk));
          next_group(rem < 256u ? rem : 256u, 1u);
        }
        flush();
      }
      if (OUT == kOutCount && (cs & 3u) == 0 && ce - cs >= kGroupMin) {
        while (ce - cs >= (uint32_t)(G * 256)) {
          F32x4 w[G][E];
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const GlobalF32x4* base = reinterpret_cast<const GlobalF32x4*>(
                uni64(reinterpret_cast<uint64_t>(col[e] + cs)));
#pragma unroll
            for (int g = 0; g < G; ++g) w[g][e] = base[g * kWave + (int)lane];
          }
#pragma unroll
          for (int g = 0; g < G; ++g) {
#pragma unroll
            for (int j = 0; j < 4; ++j) c.cnt_lane += hit_of(w[g], j) ? 1u : 0u;
          }
          cs += (uint32_t)(G * 256);
        }
        const uint32_t lane4 = 4u * lane;
        while (ce != cs) {
          const uint32_t rem = ce - cs;
          F32x4 w[E];
          load_group(w, cs);
#pragma unroll
          for (int j = 0; j < 4; ++j)
            c.cnt_lane += (lane4 + (uint32_t)j < rem && hit_of(w, j)) ? 1u : 0u;
          cs += rem < 256u ? rem : 256u;
        }
      }
    }
    // fast path: whole groups of U chunks inside the current run, with no
    // per-chunk bounds or refill logic and one base pointer per column per
    // group (the chunks are immediate offsets from it)
    while (ce - cs >= (uint32_t)(U * kWave)) {
      float w[U][E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const GlobalF32* base = reinterpret_cast<const GlobalF32*>(
            uni64(reinterpret_cast<uint64_t>(col[e] + cs)));
        // a signed lane index lets the chunk offsets fold into the loads'
        // immediate offset field (one lane-offset register for all chunks)
#pragma unroll
        for (int u = 0; u < U; ++u) w[u][e] = base[u * kWave + (int)lane];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        chunk_out<D, OUT>(a, c, cs + (uint32_t)(u * kWave),
                          leaf_hit<D, BOX>(c, w[u], BOX ? w[u] + D : w[u]));
      cs += (uint32_t)(U * kWave);
    }
    uint32_t st[U];
    uint32_t len[U];
    bool blocked = false;
    bool defer = false;  // a run for the group path above: no chunks from it
#pragma unroll
    for (int u = 0; u < U; ++u) {
      while (cs == ce && !blocked) {
        const int k = m.lowest();
        if (k < 0) break;
        if (im.test(k)) {
          // IDs are written in rank order: chunks already in flight go first
          if (OUT == kOutIds && u > 0) {
            blocked = true;
            break;
          }
          const int k2 = (m & im).run_end(k);
          m.clear_range(k, k2);
          uint64_t s, e;
          sub_range(k, k2, s, e);
          emit_range<D, OUT>(a, c, uni64(s), uni64(e));
          continue;
        }
        if (lm.test(k)) {
          const int k2 = (m & lm).andnot(im).run_end(k);
          m.clear_range(k, k2);
          uint64_t s, e;
          run_range(k, k2, s, e);
          cs = uni((uint32_t)s);
          ce = uni((uint32_t)e);
          c.le += ce - cs;
          if (!BOX && c.pempty) cs = ce;  // visited, and no point can hit
          if (OUT != kOutIds && U >= 4 && (cs & 3u) == 0 && ce - cs >= kGroupMin) defer = true;
          // a run left empty takes the next one here: this loop may end
          // only on a run with entries to scan, an empty mask, or a child
          // that is neither a leaf nor inside the query (which the caller
          // then descends into), never on a leaf run still in the mask
          continue;
        }
        break;
      }
      // cs <= ce: chunks never pass the run's end
      const uint32_t rem = defer ? 0u : ce - cs;
      len[u] = rem < (uint32_t)kWave ? rem : (uint32_t)kWave;
      st[u] = cs;
      cs += len[u];
    }
    if (len[0] == 0) {
      if (defer) continue;  // the group path takes the new run whole
      break;
    }
    float v[U][E];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (len[u]) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          // the chunk's base pointer is made explicitly wave-uniform so the
          // load takes the SGPR-base + 32-bit lane-offset form
          const GlobalF32* base = reinterpret_cast<const GlobalF32*>(
              uni64(reinterpret_cast<uint64_t>(col[e] + st[u])));
          v[u][e] = base[lane];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (len[u]) {
        const float* lo = v[u];
        const float* hi = BOX ? v[u] + D : v[u];
        chunk_out<D, OUT>(a, c, st[u], (lane < len[u]) & leaf_hit<D, BOX>(c, lo, hi));
      }
    }
  }
}

template <int D>
__device__ __forceinline__ void load_query(const SearchArgs& a, uint32_t q, Ctx<D>& c) {
  const float* qp = a.queries + (uint64_t)q * 2 * D;
  c.pempty = false;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    c.ql[d] = qp[d];
    c.qh[d] = qp[D + d];
    c.pl[d] = c.ql[d] != c.ql[d] ? -__builtin_huge_valf() : c.ql[d];
    c.ph[d] = c.qh[d] != c.qh[d] ? __builtin_huge_valf() : c.qh[d];
    c.pempty |= c.pl[d] > c.ph[d];
  }
}


__device__ __forceinline__ int level_of(const SearchArgs& a, uint32_t node) {
  int l = 0;
  while (l + 1 < a.h && node >= a.lvl_off[l + 1]) ++l;
  return l;
}

__device__ __forceinline__ uint32_t ld_acq(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_acq64(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr unsigned long long kEmptySlot = ~0ull;
constexpr int kXcds = 8;             // workgroup i runs on XCD i % 8

// Donate up to `want` of the unvisited children of one frame to the device
// queue, taking the highest-numbered ones (the donor keeps the next siblings):
// each becomes an item (query, child node).  Children that do not fit in the
// slot array stay in `m`.  pending is raised (and the raise performed) before
// the slots are written, and the donor's own item keeps pending above zero
// meanwhile.
template <int MW, bool EXPL>
__device__ __forceinline__ void donate(const SearchArgs& a, uint32_t q, uint32_t node, int level,
                                       Mask<MW>& m, uint32_t want) {
  const uint32_t lane = lane_id();
  uint32_t avail = 0;
#pragma unroll
  for (int w = 0; w < MW; ++w) avail += (uint32_t)__popcll(m.w[w]);
  uint32_t n = avail < want ? avail : want;
  if (n == 0) return;
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(&a.ctl->tail, n);
  base = uni(base);
  const uint32_t ok = base >= a.slot_cap ? 0u : (a.slot_cap - base < n ? a.slot_cap - base : n);
  if (ok == 0) return;
  if (lane == 0) {
    atomicAdd(&a.ctl->pending, ok);
    atomicAdd(&a.ctl->donated, (unsigned long long)ok);
  }
  // The increment must be performed before any of the items can be seen: a
  // taker that finishes one decrements `pending`, and if that decrement came
  // first while the donor's own item was the last one outstanding, pending
  // would reach zero, `done` would end the search and donated items would be
  // lost.  The release fence waits for lane 0's atomics before the slot stores.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  uint32_t cbase;
  if (EXPL) {
    cbase = a.nfirst[node];
  } else {
    const uint32_t j = node - a.lvl_off[level];
    cbase = a.lvl_off[level + 1] + j * a.F;
  }
  const uint64_t below_me = (2ull << lane) - 1ull;  // bits 0..lane (all bits for lane 63)
  uint32_t higher = 0;                             // set bits in words above w
#pragma unroll
  for (int w = MW - 1; w >= 0; --w) {
    const uint64_t bits = m.w[w];
    const bool mine = (bits >> lane) & 1ull;
    const uint32_t r = higher + (uint32_t)__popcll(bits & ~below_me);  // rank from the top
    const bool give = mine && r < ok;
    if (give) {
      const unsigned long long item =
          ((unsigned long long)q << 32) | (unsigned long long)(cbase + w * kWave + lane);
      __hip_atomic_store(&a.slots[base + r], item, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    m.w[w] = bits & ~ballot(give);
    higher += (uint32_t)__popcll(bits);
  }
}

// Canonical top-down traversal of the subtree rooted at `node`.
template <int D, bool BOX, int MW, bool EXPL, int OUT, int U>
__device__ void walk(const SearchArgs& a, Ctx<D>& c, uint32_t node, Frame<MW>* stk) {
  const uint32_t F = a.F;
  int level = EXPL ? 0 : level_of(a, node);
  auto packed_range = [&](uint32_t j, int k, int k2, uint64_t& s, uint64_t& e) {
    // children of packed node j at the leaf-parent level are leaves jF+k ..;
    // 32-bit (scalar) arithmetic: the leaves' ranks end below n + F < 2^32
    const uint32_t c0 = j * F + (uint32_t)k;
    const uint32_t c1 = j * F + (uint32_t)k2;
    const uint32_t n32 = (uint32_t)a.n;
    s = c0 * F;
    e = c1 * F < n32 ? c1 * F : n32;
  };
  bool is_leaf;
  uint32_t first, cnt;
  if (EXPL) {
    uint32_t cw = a.ncount[node];
    first = a.nfirst[node];
    cnt = cw & ~kLeafBit;
    is_leaf = (cw & kLeafBit) != 0;
  } else {
    is_leaf = (level == a.h - 1);
    uint32_t j = node - a.lvl_off[level];
    first = j * F;
    uint64_t below = is_leaf ? a.n : a.lvl_cnt[level + 1];
    cnt = (uint32_t)((below - first) < F ? (below - first) : F);
  }
  // rank range of the children [k, k2) of node `nd` at `lvl` (contained runs)
  auto sub_range = [&](uint32_t nd, int lvl, int k, int k2, uint64_t& s, uint64_t& e) {
    if (EXPL) {
      const uint32_t cf = a.nfirst[nd];
      s = a.nrlo[cf + k];
      e = a.nrhi[cf + k2 - 1];
    } else {
      const uint64_t j = nd - a.lvl_off[lvl];
      const uint64_t sp = a.span[lvl + 1];
      s = (j * F + (uint64_t)k) * sp;
      const uint64_t e1 = (j * F + (uint64_t)k2) * sp;
      e = e1 < a.n ? e1 : a.n;
    }
  };
  if (is_leaf) {
    Mask<MW> m, lm;
#pragma unroll
    for (int i = 0; i < MW; ++i) { m.w[i] = 0; lm.w[i] = ~0ull; }
    m.w[0] = 1;  // a single pseudo-run covering this leaf's entries
    uint64_t s0 = first, e0 = (uint64_t)first + cnt;
    scan_runs<D, BOX, MW, OUT, U>(a, c, m, lm, Mask<MW>::none(),
                                  [&](int, int, uint64_t& s, uint64_t& e) { s = s0; e = e0; },
                                  [&](int, int, uint64_t& s, uint64_t& e) { s = e = 0; });
    return;
  }
  Mask<MW> m, lm, im;
  uint32_t cbase = EXPL ? first : a.lvl_off[level + 1] + first;
  eval_children<D, MW, EXPL>(a, c, cbase, cnt, m, lm, im);
  c.ie += cnt;
  uint32_t cur = node;
  int sp = 0;
  int since = 0;
  while (true) {
    // leaf children and children inside the query at the bottom of the mask
    // are streamed (or reported whole) right away
    auto subs = [&](int k, int k2, uint64_t& s, uint64_t& e) { sub_range(cur, level, k, k2, s, e); };
    if (EXPL) {
      const uint32_t cf = a.nfirst[cur];
      scan_runs<D, BOX, MW, OUT, U>(a, c, m, lm, im, [&](int k, int k2, uint64_t& s, uint64_t& e) {
        s = a.nfirst[cf + k];
        e = (uint64_t)a.nfirst[cf + k2 - 1] + (a.ncount[cf + k2 - 1] & ~kLeafBit);
      }, subs);
    } else if (level + 1 == a.h - 1) {
      const uint32_t j = cur - a.lvl_off[level];
      scan_runs<D, BOX, MW, OUT, U>(a, c, m, lm, im, [&](int k, int k2, uint64_t& s, uint64_t& e) {
        packed_range(j, k, k2, s, e);
      }, subs);
    } else if (a.contain) {
      scan_runs<D, BOX, MW, OUT, U>(a, c, m, Mask<MW>::none(), im,
                                    [&](int, int, uint64_t& s, uint64_t& e) { s = e = 0; }, subs);
    }
    if (m.empty()) {
      if (sp == 0) break;
      --sp;
      // frames are wave-uniform: readfirstlane keeps them in SGPRs
      cur = uni(stk[sp].node);
      level = (int)uni(stk[sp].level);
#pragma unroll
      for (int i = 0; i < MW; ++i) {
        m.w[i] = uni64(stk[sp].m.w[i]);
        lm.w[i] = uni64(stk[sp].lm.w[i]);
        im.w[i] = uni64(stk[sp].im.w[i]);
      }
      continue;
    }
    const int k = m.lowest();
    m.clear(k);
    // work stealing: when waves wait for items, hand over the unvisited
    // siblings of the shallowest frame that has any, if those subtrees are at
    // least donate_min_height levels above the leaf parents (packed layouts)
    if (OUT != kOutIds && a.donate && ++since >= (int)a.hunger_period) {
      since = 0;
      const int max_level = EXPL ? kMaxDepth : a.h - 2 - (int)a.donate_min_height;
      int f = 0;
      Mask<MW> fm;
      for (; f < sp; ++f) {
#pragma unroll
        for (int i = 0; i < MW; ++i) fm.w[i] = uni64(stk[f].m.w[i]);
        if (!fm.empty()) break;
      }
      const uint32_t f_node = f < sp ? uni(stk[f].node) : cur;
      const int f_level = f < sp ? (int)uni(stk[f].level) : level;
      const bool have = (f < sp) ? (f_level <= max_level) : (!m.empty() && level <= max_level);
      if (have) {
        const uint32_t head = ld_acq(&a.ctl->head);
        const uint32_t tail = ld_acq(&a.ctl->tail);
        if (head > tail && tail < a.slot_cap) {
          const uint32_t waiting = head - tail;
          if (f < sp) {
            donate<MW, EXPL>(a, c.q, f_node, f_level, fm, waiting);
            stk[f].m = fm;
          } else {
            donate<MW, EXPL>(a, c.q, cur, level, m, waiting);
          }
        }
      }
    }
    uint32_t child, cfirst, ccnt;
    if (EXPL) {
      child = a.nfirst[cur] + k;
      cfirst = a.nfirst[child];
      ccnt = a.ncount[child] & ~kLeafBit;
    } else {
      const uint32_t j = cur - a.lvl_off[level];
      const uint32_t cj = j * F + k;
      child = a.lvl_off[level + 1] + cj;
      cfirst = cj * F;
      uint64_t below = a.lvl_cnt[level + 2];
      ccnt = (uint32_t)((below - cfirst) < F ? (below - cfirst) : F);
      cfirst += a.lvl_off[level + 2];
    }
    Mask<MW> cm, clm, cim;
    eval_children<D, MW, EXPL>(a, c, cfirst, ccnt, cm, clm, cim);
    c.ie += ccnt;
    if (cm.empty()) continue;
    if (!m.empty()) {
      stk[sp].node = cur;
      stk[sp].level = (uint32_t)level;
      stk[sp].m = m;
      stk[sp].lm = lm;
      stk[sp].im = im;
      ++sp;
    }
    cur = child;
    level += 1;
    m = cm;
    lm = clm;
    im = cim;
  }
}

// Persistent search.  Root items (query t, from the root) come from eight
// per-XCD ticket ranges over the (spatially ordered) query list: a wave takes
// tickets of its own XCD's range first, so each XCD's L2 sees neighbouring
// queries, and then the other ranges.  With donation on, a wave that finds all
// root tickets gone takes a slot ticket from `head` and waits for that slot to
// be written by a donor, or for `done`; at most max_waiters waves wait, the
// others exit.
template <int D, bool BOX, int MW, bool EXPL, int OUT, int U>
__global__ __launch_bounds__(512) void k_search(SearchArgs a) {
  // one DFS frame stack per wave in dynamic LDS (blocks of 1 to 8 waves)
  constexpr int kDepth = EXPL ? kMaxDepth : kMaxLevels;
  extern __shared__ unsigned char qcp_lds[];
  Frame<MW>* stk = reinterpret_cast<Frame<MW>*>(qcp_lds) + (threadIdx.x / kWave) * kDepth;
  // ID pass: each wave's 2 x 64-slot output window follows all frame stacks
  uint32_t* idb = reinterpret_cast<uint32_t*>(reinterpret_cast<Frame<MW>*>(qcp_lds) +
                                              (blockDim.x / kWave) * kDepth) +
                  (threadIdx.x / kWave) * (2 * kWave);
  const uint32_t lane = lane_id();
  __shared__ unsigned long long wg_stats[4];
  if (threadIdx.x < 4) wg_stats[threadIdx.x] = 0;
  __syncthreads();
  uint64_t ie = 0, le = 0, hits = 0, cruns = 0;
  const bool donating = OUT != kOutIds && a.donate;
  // root items: the queries (from the root), or a caller's (query, subtree) list
  const uint32_t nroot = a.items ? a.n_items : a.nq;
  // several items of one query (donated subtrees, or an item list) add up
  const bool accumulate = donating || a.items != nullptr;
  const uint32_t xcd = blockIdx.x % kXcds;
  bool roots_left = true;
  // root items finished by this wave and not yet taken off `pending`: they
  // are reported in one atomic when the wave runs out of root tickets (one
  // same-address atomic per item would serialise)
  uint32_t roots_done = 0;
  while (true) {
    uint32_t q = 0, node = 0;
    bool got = false;
    if (roots_left) {
      for (int k = 0; k < kXcds && !got; ++k) {
        const uint32_t x = (xcd + k) % kXcds;
        const uint32_t lo = (uint32_t)(((uint64_t)nroot * x) / kXcds);
        const uint32_t hi = (uint32_t)(((uint64_t)nroot * (x + 1)) / kXcds);
        if (lo >= hi) continue;
        if (k > 0 && ld_acq(&a.ctl->xhead[x]) >= hi - lo) continue;
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&a.ctl->xhead[x], 1u);
        t = uni(t);
        if (lo + t < hi) {
          if (a.items) {
            const unsigned long long it = a.items[lo + t];
            q = uni((uint32_t)(it >> 32));
            node = uni((uint32_t)it);
          } else {
            q = a.qperm ? a.qperm[lo + t] : lo + t;
          }
          got = true;
        }
      }
      roots_left = got;
    }
    if (!got) {
      if (!donating) break;
      if (roots_done) {
        uint32_t prev = 0;
        if (lane == 0) prev = atomicSub(&a.ctl->pending, roots_done);
        if (uni(prev) == roots_done && lane == 0)
          __hip_atomic_store(&a.ctl->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        roots_done = 0;
      }
      // Enough waves already wait: leave.  (Signed: donors may have claimed
      // more slots than there are tickets, and those slots need takers.)
      const int32_t waiting = (int32_t)(ld_acq(&a.ctl->head) - ld_acq(&a.ctl->tail));
      if (waiting >= (int32_t)a.max_waiters && !ld_acq(&a.ctl->done)) break;
      uint32_t s = 0;
      if (lane == 0) s = atomicAdd(&a.ctl->head, 1u);
      s = uni(s);
      unsigned long long item = kEmptySlot;
      while (true) {
        if (s < a.slot_cap) {
          item = ld_acq64(&a.slots[s]);
          if (item != kEmptySlot) break;
        }
        if (ld_acq(&a.ctl->done)) break;
        for (int z = 0; z < 8; ++z) __builtin_amdgcn_s_sleep(32);
      }
      item = uni64(item);
      if (item == kEmptySlot) break;
      q = (uint32_t)(item >> 32);
      node = (uint32_t)item;
    }
    Ctx<D> c;
    c.q = q;
    c.cnt = 0;
    c.cnt_lane = 0;
    c.s1_lane = c.s2_lane = 0;
    c.s1c = c.s2c = 0;
    c.cr = 0;
    c.ie = 0;
    c.le = 0;
    c.out_base = (OUT == kOutIds) ? a.offsets[q] : 0;
    c.idb = idb;
    c.wbase = c.out_base & ~(uint64_t)(kWave - 1);
    c.wlo = c.out_base;
    load_query<D>(a, q, c);
    if (a.h > 0) {
      if (EXPL && a.parts > 1 && got && !a.items) {
        // a query over a partitioned MPHR: every partition root, in rank order
        for (uint32_t r = 0; r < a.parts; ++r) walk<D, BOX, MW, EXPL, OUT, U>(a, c, r, stk);
      } else {
        walk<D, BOX, MW, EXPL, OUT, U>(a, c, node, stk);
      }
    }
    if (OUT == kOutIds) flush_partial<D>(a, c);
    if (OUT != kOutIds) c.cnt += (uint32_t)wave_sum_u64(c.cnt_lane);
    if (OUT != kOutIds) {
      // the hash is additive over disjoint ID sets, so donated parts add up
      uint64_t hsum = 0;
      if (OUT == kOutHash)
        hsum = set_hash(c.s1c + wave_sum_u64(c.s1_lane), c.s2c + wave_sum_u64(c.s2_lane));
      if (lane == 0) {
        if (accumulate) {
          if (c.cnt) atomicAdd(&a.counts[q], c.cnt);
          if (OUT == kOutHash && hsum) atomicAdd(&a.hashes[q], (unsigned long long)hsum);
        } else {
          a.counts[q] = c.cnt;
          if (OUT == kOutHash) a.hashes[q] = hsum;
        }
      }
    }
    if (donating) {
      if (got) {
        ++roots_done;
      } else if (lane == 0) {
        if (atomicSub(&a.ctl->pending, 1u) == 1u)
          __hip_atomic_store(&a.ctl->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    ie += c.ie;
    le += c.le;
    hits += c.cnt;
    cruns += c.cr;
  }
  // statistics: waves -> workgroup (LDS) -> this XCD's slot
  if (lane == 0) {
    atomicAdd(&wg_stats[0], (unsigned long long)ie);
    atomicAdd(&wg_stats[1], (unsigned long long)le);
    atomicAdd(&wg_stats[2], (unsigned long long)hits);
    atomicAdd(&wg_stats[3], (unsigned long long)cruns);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&a.ctl->internal_entries[xcd], wg_stats[0]);
    atomicAdd(&a.ctl->leaf_entries[xcd], wg_stats[1]);
    atomicAdd(&a.ctl->total_hits[xcd], wg_stats[2]);
    if (wg_stats[3]) atomicAdd(&a.ctl->contained[xcd], wg_stats[3]);
  }
}

// ---------------------------------------------------------------------------
// Plain MPRS, one workgroup of NW waves per query (packed layouts only).  With
// NW = 1 entry k of a node is lane k % 64 of the one wave and the leftmost
// qualifying child is a ballot + find-first-set; with NW = 3 (the reference's
// 192-thread block, src/tree/mphr.cpp:241-243) entry k belongs to wave
// (k / 64) % 3, and the waves' candidates meet in LDS behind barriers (the
// reference's FindMinOnGPU reduction, src/common/macro.h:24-42).
template <int D, bool BOX, int MW, int NW>
__global__ __launch_bounds__(64 * NW) void k_mprs(SearchArgs a) {
  constexpr int E = BOX ? 2 * D : D;
  constexpr int kNone = 1 << 30;
  const uint32_t lane = lane_id();
  const int wv = NW > 1 ? (int)(threadIdx.x / kWave) : 0;
  const uint32_t q = a.q_begin + blockIdx.x;
  if (q >= a.nq) return;  // uniform over the workgroup
  __shared__ int s_red[NW];
  __shared__ unsigned long long s_sum[3][NW];
  // workgroup-wide minimum of a wave-uniform value (every thread gets it)
  auto group_min = [&](int v) -> int {
    if constexpr (NW == 1) {
      return v;
    } else {
      if (lane == 0) s_red[wv] = v;
      __syncthreads();
      int m = s_red[0];
#pragma unroll
      for (int i = 1; i < NW; ++i) m = s_red[i] < m ? s_red[i] : m;
      __syncthreads();
      return m;
    }
  };
  Ctx<D> c;
  load_query<D>(a, q, c);
  const uint32_t F = a.F;
  const int h = a.h;
  const uint64_t n = a.n;
  uint32_t cnt = 0, rootv = 0, nodev = 0;
  uint64_t s1 = 0, s2 = 0;
  if (h > 0) {
    auto last_index = [&](int l, uint64_t j) -> uint64_t {
      uint64_t e = (j + 1) * a.span[l];
      return e < n ? e : n;
    };
    uint64_t visited = 0;
    const uint64_t last = n;
    int lvl = 0;
    uint64_t j = 0;
    rootv = 1;
    while (visited < last) {
      while (lvl < h - 1) {
        const uint64_t first = j * F;
        const uint64_t below = a.lvl_cnt[lvl + 1];
        const uint32_t cc = (uint32_t)((below - first) < F ? (below - first) : F);
        const uint64_t gbase = a.lvl_off[lvl + 1] + first;
        int pick = kNone;
#pragma unroll
        for (int t = 0; t < (MW + NW - 1) / NW; ++t) {
          const int w = wv + t * NW;  // this wave's 64-entry words of the node
          if (w >= MW) break;
          const uint32_t k = w * kWave + lane;
          float lo[D], hi[D];
#pragma unroll
          for (int d = 0; d < D; ++d) {
            lo[d] = a.nbox[(uint64_t)d * a.node_stride + gbase + k];
            hi[d] = a.nbox[(uint64_t)(D + d) * a.node_stride + gbase + k];
          }
          bool ok = (k < cc) && last_index(lvl + 1, first + k) > visited &&
                    overlaps<D>(c.ql, c.qh, lo, hi);
          uint64_t b = ballot(ok);
          if (pick == kNone && b) pick = w * kWave + __builtin_ctzll(b);
        }
        pick = group_min(pick);
        if (pick == kNone) {
          visited = last_index(lvl, j);
          lvl = 0;
          j = 0;
          ++rootv;
          break;
        }
        ++lvl;
        j = first + (uint64_t)pick;
        ++nodev;
      }
      while (lvl == h - 1) {
        const uint64_t first = j * F;
        const uint32_t cc = (uint32_t)((n - first) < F ? (n - first) : F);
        int none = 1;
#pragma unroll
        for (int t = 0; t < (MW + NW - 1) / NW; ++t) {
          const int w = wv + t * NW;
          if (w >= MW) break;
          const uint32_t k = w * kWave + lane;
          float v[E];
#pragma unroll
          for (int e = 0; e < E; ++e) v[e] = a.obj[(uint64_t)e * a.obj_stride + first + k];
          bool hit = (k < cc) && overlaps<D>(c.ql, c.qh, v, BOX ? v + D : v);
          uint64_t b = ballot(hit);
          cnt += (uint32_t)__popcll(b);
          if (hit) {
            const uint64_t id = first + k + 1;
            s1 += id;
            s2 += id * id;
          }
          if (b) none = 0;
        }
        const bool any = group_min(none) == 0;
        visited = last_index(lvl, j);
        if (visited == last) break;
        if (any) {
          ++j;
          ++nodev;
        } else {
          --lvl;
          j /= F;
          if (lvl == 0) ++rootv; else ++nodev;
        }
      }
    }
  }
  uint64_t t1 = wave_sum_u64(s1), t2 = wave_sum_u64(s2);
  uint64_t tc = cnt;
  if constexpr (NW > 1) {
    if (lane == 0) {
      s_sum[0][wv] = tc;
      s_sum[1][wv] = t1;
      s_sum[2][wv] = t2;
    }
    __syncthreads();
    tc = t1 = t2 = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      tc += s_sum[0][i];
      t1 += s_sum[1][i];
      t2 += s_sum[2][i];
    }
  }
  if (threadIdx.x == 0) {
    a.counts[q] = (uint32_t)tc;
    a.hashes[q] = set_hash(t1, t2);
    if (a.visits) {
      a.visits[2 * (uint64_t)q] = rootv;
      a.visits[2 * (uint64_t)q + 1] = nodev;
    }
  }
}

// The plain form over a partitioned MPHR (explicit node arrays): workgroup p
// runs MPRS over partition p's subtree from its root, node p, for query
// q_begin, as the reference's block bid does from g_root_offset[bid]
// (src/tree/mphr.cpp:314-333).  A node's "last index" is the end of its rank
// range (nrhi), a leaf's parent comes from nparent (the reference's leaf
// back-pointer) and the next leaf is the next node.  Results per workgroup:
// counts[p], hashes[p], visits[2p] (root visits), visits[2p+1] (node visits).
template <int D, bool BOX, int MW, int NW>
__global__ __launch_bounds__(64 * NW) void k_mprs_part(SearchArgs a) {
  constexpr int E = BOX ? 2 * D : D;
  constexpr int kNone = 1 << 30;
  const uint32_t lane = lane_id();
  const int wv = NW > 1 ? (int)(threadIdx.x / kWave) : 0;
  const uint32_t q = a.q_begin;
  const uint32_t root = blockIdx.x;
  if (q >= a.nq || root >= a.parts) return;  // uniform over the workgroup
  __shared__ int s_red[NW];
  __shared__ unsigned long long s_sum[3][NW];
  auto group_min = [&](int v) -> int {
    if constexpr (NW == 1) {
      return v;
    } else {
      if (lane == 0) s_red[wv] = v;
      __syncthreads();
      int m = s_red[0];
#pragma unroll
      for (int i = 1; i < NW; ++i) m = s_red[i] < m ? s_red[i] : m;
      __syncthreads();
      return m;
    }
  };
  Ctx<D> c;
  load_query<D>(a, q, c);
  uint32_t cnt = 0, rootv = 1, nodev = 0;
  uint64_t s1 = 0, s2 = 0;
  uint32_t visited = a.nrlo[root];  // ranks below are done
  const uint32_t last = a.nrhi[root];
  uint32_t nd = root;
  while (visited < last) {
    while (!(a.ncount[nd] & kLeafBit)) {
      const uint32_t cf = a.nfirst[nd];
      const uint32_t cc = a.ncount[nd];
      int pick = kNone;
#pragma unroll
      for (int t = 0; t < (MW + NW - 1) / NW; ++t) {
        const int w = wv + t * NW;
        if (w >= MW) break;
        const uint32_t k = w * kWave + lane;
        const uint64_t ch = (uint64_t)cf + k;  // within the node arrays' slack
        float lo[D], hi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          lo[d] = a.nbox[(uint64_t)d * a.node_stride + ch];
          hi[d] = a.nbox[(uint64_t)(D + d) * a.node_stride + ch];
        }
        const bool ok = (k < cc) && a.nrhi[ch] > visited && overlaps<D>(c.ql, c.qh, lo, hi);
        const uint64_t b = ballot(ok);
        if (pick == kNone && b) pick = w * kWave + __builtin_ctzll(b);
      }
      pick = group_min(pick);
      if (pick == kNone) {
        visited = a.nrhi[nd];
        nd = root;
        ++rootv;
        break;
      }
      nd = cf + (uint32_t)pick;
      ++nodev;
    }
    while (a.ncount[nd] & kLeafBit) {
      const uint32_t f = a.nfirst[nd];
      const uint32_t cc = a.ncount[nd] & ~kLeafBit;
      int none = 1;
#pragma unroll
      for (int t = 0; t < (MW + NW - 1) / NW; ++t) {
        const int w = wv + t * NW;
        if (w >= MW) break;
        const uint32_t k = w * kWave + lane;
        float v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = a.obj[(uint64_t)e * a.obj_stride + f + k];
        const bool hit = (k < cc) && overlaps<D>(c.ql, c.qh, v, BOX ? v + D : v);
        const uint64_t b = ballot(hit);
        cnt += (uint32_t)__popcll(b);
        if (hit) {
          const uint64_t id = (uint64_t)f + k + 1;
          s1 += id;
          s2 += id * id;
        }
        if (b) none = 0;
      }
      const bool any = group_min(none) == 0;
      visited = a.nrhi[nd];
      if (visited == last) break;
      if (any) {
        ++nd;
        ++nodev;
      } else {
        nd = a.nparent[nd];
        if (nd == root) ++rootv; else ++nodev;
      }
    }
  }
  uint64_t t1 = wave_sum_u64(s1), t2 = wave_sum_u64(s2);
  uint64_t tc = cnt;
  if constexpr (NW > 1) {
    if (lane == 0) {
      s_sum[0][wv] = tc;
      s_sum[1][wv] = t1;
      s_sum[2][wv] = t2;
    }
    __syncthreads();
    tc = t1 = t2 = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      tc += s_sum[0][i];
      t1 += s_sum[1][i];
      t2 += s_sum[2][i];
    }
  }
  if (threadIdx.x == 0) {
    a.counts[root] = (uint32_t)tc;
    a.hashes[root] = set_hash(t1, t2);
    a.visits[2 * root] = rootv;
    a.visits[2 * root + 1] = nodev;
  }
}

// ---------------------------------------------------------------------------
// Launchers, instantiated per (D, object kind, layout family) in their own
// translation units (search_inst.hip compiled once per combination), so the
// many kernel variants compile in parallel.
#define QCP_MW_SWITCH(MWV, ...)                        \
  switch (MWV) {                                       \
    case 1: { constexpr int MW = 1; __VA_ARGS__; } break; \
    case 2: { constexpr int MW = 2; __VA_ARGS__; } break; \
    case 3: { constexpr int MW = 3; __VA_ARGS__; } break; \
    case 4: { constexpr int MW = 4; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue;              \
  }

// Chunks in flight: 2/4/8 for the packed counting modes, 4 elsewhere.
#define QCP_U_SWITCH(UV, ...)                          \
  switch (UV) {                                        \
    case 2: { constexpr int U = 2; __VA_ARGS__; } break; \
    case 8: { constexpr int U = 8; __VA_ARGS__; } break; \
    default: { constexpr int U = 4; __VA_ARGS__; } break; \
  }

// Dynamic LDS of a persistent-search workgroup: one frame stack per wave,
// and for the ID pass one 2 x 64-slot output window per wave after them.
template <int MW, bool EXPL, int OUT>
inline size_t lds_for(unsigned block) {
  return (size_t)(block / kWave) * ((EXPL ? kMaxDepth : kMaxLevels) * sizeof(Frame<MW>) +
                                    (OUT == kOutIds ? 2 * kWave * sizeof(uint32_t) : 0));
}

template <int D, bool BOX, bool EXPL, int OUT>
inline hipError_t launch_search_t(const SearchArgs& a, int mw, int u, unsigned grid,
                                  unsigned block, hipStream_t s) {
  if (block % kWave || block < kWave || block > 512) return hipErrorInvalidValue;
  if (!EXPL && OUT != kOutIds) {
    QCP_U_SWITCH(u, QCP_MW_SWITCH(mw, k_search<D, BOX, MW, EXPL, OUT, U>
                                  <<<grid, block, (lds_for<MW, EXPL, OUT>(block)), s>>>(a)));
  } else if (!EXPL && u >= 8) {  // ID pass, packed layouts: 4 or 8 chunks in flight
    constexpr int U = 8;
    QCP_MW_SWITCH(mw, k_search<D, BOX, MW, EXPL, OUT, U>
                      <<<grid, block, (lds_for<MW, EXPL, OUT>(block)), s>>>(a));
  } else {
    constexpr int U = 4;
    QCP_MW_SWITCH(mw, k_search<D, BOX, MW, EXPL, OUT, U>
                      <<<grid, block, (lds_for<MW, EXPL, OUT>(block)), s>>>(a));
  }
  return hipGetLastError();
}

// The persistent search of one (D, object kind, layout family), every output
// mode; mw = 64-entry words per node (F / 64 rounded up).
template <int D, bool BOX, bool EXPL>
hipError_t launch_search_dbe(const SearchArgs& a, int out, int u, unsigned grid, unsigned block,
                             hipStream_t s) {
  const int mw = (int)((a.F + 63) / 64);
  switch (out) {
    case kOutCount: return launch_search_t<D, BOX, EXPL, kOutCount>(a, mw, u, grid, block, s);
    case kOutHash: return launch_search_t<D, BOX, EXPL, kOutHash>(a, mw, u, grid, block, s);
    default: return launch_search_t<D, BOX, EXPL, kOutIds>(a, mw, u, grid, block, s);
  }
}

// The U = 4 instantiation of k_search, for occupancy queries.
template <int D, bool BOX, bool EXPL>
const void* search_ptr_dbe(int out, int mw) {
#define QCP_PTR(OO)                                                          \
  switch (mw) {                                                              \
    case 1: return (const void*)&k_search<D, BOX, 1, EXPL, OO, 4>;           \
    case 2: return (const void*)&k_search<D, BOX, 2, EXPL, OO, 4>;           \
    case 3: return (const void*)&k_search<D, BOX, 3, EXPL, OO, 4>;           \
    case 4: return (const void*)&k_search<D, BOX, 4, EXPL, OO, 4>;           \
    default: return nullptr;                                                 \
  }
  switch (out) {
    case kOutCount: { QCP_PTR(kOutCount) }
    case kOutHash: { QCP_PTR(kOutHash) }
    default: { QCP_PTR(kOutIds) }
  }
#undef QCP_PTR
}

// The plain (MPRS) form of one (D, object kind), packed layouts, with 1 or 3
// waves per workgroup.
template <int D, bool BOX>
hipError_t launch_mprs_db(const SearchArgs& a, unsigned nblocks, unsigned waves, hipStream_t s) {
  const int mw = (int)((a.F + 63) / 64);
  if (waves == 3) {
    QCP_MW_SWITCH(mw, k_mprs<D, BOX, MW, 3><<<nblocks, 192, 0, s>>>(a));
  } else if (waves == 1) {
    QCP_MW_SWITCH(mw, k_mprs<D, BOX, MW, 1><<<nblocks, 64, 0, s>>>(a));
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// The plain form over a partitioned MPHR: one workgroup of 1 or 3 waves per
// partition, for query a.q_begin.
template <int D, bool BOX>
hipError_t launch_mprs_part_db(const SearchArgs& a, unsigned parts, unsigned waves,
                               hipStream_t s) {
  const int mw = (int)((a.F + 63) / 64);
  if (waves == 3) {
    QCP_MW_SWITCH(mw, k_mprs_part<D, BOX, MW, 3><<<parts, 192, 0, s>>>(a));
  } else if (waves == 1) {
    QCP_MW_SWITCH(mw, k_mprs_part<D, BOX, MW, 1><<<parts, 64, 0, s>>>(a));
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Every combination is instantiated in exactly one translation unit.
#define QCP_SEARCH_EXTERN(DD, BB, EE)                                                        \
  extern template hipError_t launch_search_dbe<DD, BB, EE>(const SearchArgs&, int, int,      \
                                                           unsigned, unsigned, hipStream_t); \
  extern template const void* search_ptr_dbe<DD, BB, EE>(int, int);
QCP_SEARCH_EXTERN(2, false, false)
QCP_SEARCH_EXTERN(2, false, true)
QCP_SEARCH_EXTERN(2, true, false)
QCP_SEARCH_EXTERN(2, true, true)
QCP_SEARCH_EXTERN(3, false, false)
QCP_SEARCH_EXTERN(3, false, true)
QCP_SEARCH_EXTERN(3, true, false)
QCP_SEARCH_EXTERN(3, true, true)
#undef QCP_SEARCH_EXTERN
extern template hipError_t launch_mprs_db<2, false>(const SearchArgs&, unsigned, unsigned, hipStream_t);
extern template hipError_t launch_mprs_db<2, true>(const SearchArgs&, unsigned, unsigned, hipStream_t);
extern template hipError_t launch_mprs_db<3, false>(const SearchArgs&, unsigned, unsigned, hipStream_t);
extern template hipError_t launch_mprs_db<3, true>(const SearchArgs&, unsigned, unsigned, hipStream_t);
extern template hipError_t launch_mprs_part_db<2, false>(const SearchArgs&, unsigned, unsigned,
                                                         hipStream_t);
extern template hipError_t launch_mprs_part_db<2, true>(const SearchArgs&, unsigned, unsigned,
                                                        hipStream_t);
extern template hipError_t launch_mprs_part_db<3, false>(const SearchArgs&, unsigned, unsigned,
                                                         hipStream_t);
extern template hipError_t launch_mprs_part_db<3, true>(const SearchArgs&, unsigned, unsigned,
                                                        hipStream_t);

}  // namespace qcp
