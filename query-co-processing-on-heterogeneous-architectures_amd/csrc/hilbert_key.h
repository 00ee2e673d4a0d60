// hilbert_key.h -- Hilbert key of one integer point, written as a per-level
// state machine so that it runs identically on the host and on gfx950.
//
// It computes the same function as HilbertMapper::MappingIntoSingle
// (src/mapper/hilbert_mapper.cpp:25-64): bitTranspose of the packed
// coordinates equals interleaving bit b of every coordinate into digit b with
// dimension d at bit d, so the transpose is done digit by digit here instead
// of with the reference's shift/mask cascade.  The Gray/rotation walk over the
// digits from the top, the final XOR with nthbits>>1 and the prefix XOR are the
// reference's steps.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define QCP_HD __host__ __device__ __forceinline__
#else
#define QCP_HD inline
#endif
// loop unrolling hint for the device compiler (host compilers other than
// clang do not know the pragma)
#if defined(__clang__)
#define QCP_UNROLL _Pragma("unroll")
#else
#define QCP_UNROLL
#endif

namespace qcp {

// One level of the walk.  The reference keeps `flip` (1 << the rotation of
// the level above, 0 at the top) and the digit of the level above; both enter
// only as flip ^ above ^ t, so the walk's whole state is the pair
// s = flip ^ above and the rotation.  Takes this level's digit t (bit d =
// coordinate d), returns the D bits it appends to the index, advances (s, rot).
template <int D>
QCP_HD constexpr uint32_t hilbert_level(uint32_t& s, unsigned& rot, uint32_t t) {
  const uint32_t nd_ones = (1u << D) - 1u;
  const uint32_t x = s ^ t;
  const uint32_t bits = ((x >> rot) | (x << (D - rot))) & nd_ones;
  s = (1u << rot) ^ t;
  // adjust_rotation: rotation += 1 + (position of the lowest set bit of
  // bits, counted from 1, if that bit lies below the top dimension)
  uint32_t low = bits & (0u - bits) & (nd_ones >> 1);
  while (low) {
    low >>= 1;
    ++rot;
  }
  if (++rot >= (unsigned)D) rot -= D;
  return bits;
}

// Final steps: XOR with nthbits >> 1 (nthbits = one 1 per level, at the low
// bit of each D-bit digit) and the prefix XOR (inverse Gray code).
template <int D>
QCP_HD uint64_t hilbert_finish(uint64_t index, unsigned nb) {
  const uint64_t nthbits = ((1ull << (D * nb)) - 1ull) / ((1ull << D) - 1ull);
  index ^= nthbits >> 1;
  for (unsigned s = 1; s < (unsigned)D * nb; s *= 2) index ^= index >> s;
  return index;
}

// Hilbert index of c[0..D) with nb bits per coordinate (D*nb <= 63, nb >= 2).
template <int D>
QCP_HD uint64_t hilbert_key(const uint32_t* c, unsigned nb) {
  uint64_t index = 0;
  uint32_t s = 0;
  unsigned rot = 0;
  for (int b = (int)nb - 1; b >= 0; --b) {
    uint32_t t = 0;
QCP_UNROLL
    for (int d = 0; d < D; ++d) t |= ((c[d] >> b) & 1u) << d;
    index = (index << D) | hilbert_level<D>(s, rot, t);
  }
  return hilbert_finish<D>(index, nb);
}

// The same walk, L levels per step, from a transition table: the entry for
// state (s, rot) and the L digits of L consecutive levels (top level in the
// high bits) holds the D*L index bits those levels append (low 8 bits) and the
// state after them (s | rot << D, bits 8..15).  L = 4 for D = 2 (8 states x
// 256 inputs), L = 2 for D = 3 (24 states x 64 inputs): 2048 and 1536 entries.
template <int D>
struct HilbertTable {
  static constexpr int L = D == 2 ? 4 : 2;
  static constexpr int IN = D * L;                  // input bits per step
  static constexpr int STATES = (1 << D) * D;
  static constexpr uint32_t SIZE = (uint32_t)STATES << IN;
  static constexpr uint32_t PSIZE = 2 * SIZE;       // with the parity bit (below)
};

template <int D>
QCP_HD constexpr uint16_t hilbert_table_entry(uint32_t e) {
  constexpr int L = HilbertTable<D>::L, IN = HilbertTable<D>::IN;
  const uint32_t in = e & ((1u << IN) - 1u);
  const uint32_t st = e >> IN;
  uint32_t s = st & ((1u << D) - 1u);
  unsigned rot = st >> D;
  uint32_t out = 0;
  for (int l = L - 1; l >= 0; --l)
    out = (out << D) | hilbert_level<D>(s, rot, (in >> (D * l)) & ((1u << D) - 1u));
  return (uint16_t)(out | ((s | (rot << D)) << 8));
}

// Bit b of coordinate d moved to bit D*b + d (the digits of every level,
// concatenated from the top): 31-bit coordinates for D = 2, 21-bit for D = 3.
QCP_HD uint64_t spread2(uint32_t v) {
  uint64_t x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}
QCP_HD uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0x1FFFFFu;
  x = (x | (x << 32)) & 0x001F00000000FFFFull;
  x = (x | (x << 16)) & 0x001F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

template <int D>
QCP_HD uint64_t interleave(const uint32_t* c) {
  if (D == 2) return spread2(c[0]) | (spread2(c[1]) << 1);
  return spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
}

// hilbert_key through the table: the top nb % L levels one by one, then L at
// a time.  Equal to hilbert_key<D>(c, nb) for every input (D in {2, 3}).
template <int D>
QCP_HD uint64_t hilbert_key_table(const uint32_t* c, unsigned nb, const uint16_t* tab) {
  constexpr int L = HilbertTable<D>::L, IN = HilbertTable<D>::IN;
  const uint64_t m = interleave<D>(c);
  uint64_t index = 0;
  uint32_t s = 0;
  unsigned rot = 0;
  int b = (int)nb - 1;
  for (const int top = (int)nb - (int)(nb % L); b >= top; --b)
    index = (index << D) | hilbert_level<D>(s, rot, (uint32_t)(m >> (D * b)) & ((1u << D) - 1u));
  for (; b >= 0; b -= L) {
    const uint32_t in = (uint32_t)(m >> (D * (b - L + 1))) & ((1u << IN) - 1u);
    const uint32_t e = tab[((s | (rot << D)) << IN) | in];
    index = (index << IN) | (e & 0xFFu);
    s = (e >> 8) & ((1u << D) - 1u);
    rot = e >> (8 + D);
  }
  return hilbert_finish<D>(index, nb);
}

// The same transition table indexed coordinate-major: entry (s | rot << D)
// << IN | g_{D-1} << (D-1)L | ... | g_0, where g_d holds coordinate d's L bits
// of the L levels (top level in the high bit), so a key never needs the
// coordinates' bits interleaved.  The start table covers the top r = nb % L
// levels from the walk's initial state: entry g_{D-1} << (D-1)r | ... | g_0
// (r bits each) holds the D*r index bits (low 8) and the state after them.
template <int D>
QCP_HD constexpr uint16_t hilbert_table_entry_cm(uint32_t e) {
  constexpr int L = HilbertTable<D>::L, IN = HilbertTable<D>::IN;
  const uint32_t st = e >> IN;
  uint32_t in = 0;  // level-major digits, top level first
  for (int l = L - 1; l >= 0; --l) {
    uint32_t t = 0;
    for (int d = 0; d < D; ++d) t |= ((e >> (d * L + l)) & 1u) << d;
    in = (in << D) | t;
  }
  return hilbert_table_entry<D>((st << IN) | in);
}

template <int D>
QCP_HD constexpr uint16_t hilbert_start_entry(int r, uint32_t e) {
  uint32_t s = 0;
  unsigned rot = 0;
  uint32_t out = 0;
  for (int l = r - 1; l >= 0; --l) {
    uint32_t t = 0;
    for (int d = 0; d < D; ++d) t |= ((e >> (d * r + l)) & 1u) << d;
    out = (out << D) | hilbert_level<D>(s, rot, t);
  }
  return (uint16_t)(out | ((s | (rot << D)) << 8));
}

// Inverse Gray code of the low w bits of x alone: bit i becomes the XOR of
// bits i .. w-1.
QCP_HD constexpr uint32_t prefix_xor_bits(uint32_t x, int w) {
  for (int s = 1; s < w; s *= 2) x ^= x >> s;
  return x & ((1u << w) - 1u);
}

// The walk's tables with the final inverse Gray code folded in.  Bit i of
// hilbert_finish's prefix XOR is the XOR of the index bits at and above i:
// within a step's IN bits that is prefix_xor_bits of the step's output, and
// from the steps above it, the parity P of all their output bits.  So the
// state gains P (entries at state + P * STATES), an entry's output is
// prefix_xor_bits(out) with every bit flipped when P is set, and its next
// state carries P ^ parity(out), which is bit 0 of prefix_xor_bits(out).
// After the last step only the constant hilbert_finish_const remains.
template <int D>
QCP_HD constexpr uint16_t hilbert_ptable_entry_cm(uint32_t e) {
  constexpr int IN = HilbertTable<D>::IN;
  constexpr uint32_t ST = (uint32_t)HilbertTable<D>::STATES;
  const uint32_t in = e & ((1u << IN) - 1u);
  const uint32_t sp = e >> IN;
  const uint32_t p = sp >= ST ? 1u : 0u;
  const uint32_t base = hilbert_table_entry_cm<D>(((sp - p * ST) << IN) | in);
  const uint32_t px = prefix_xor_bits(base & 0xFFu, IN);
  const uint32_t out = px ^ (p ? (1u << IN) - 1u : 0u);
  return (uint16_t)(out | (((base >> 8) + (p ^ (px & 1u)) * ST) << 8));
}

// The start table in the same form (the parity above the top levels is 0).
template <int D>
QCP_HD constexpr uint16_t hilbert_pstart_entry(int r, uint32_t e) {
  constexpr uint32_t ST = (uint32_t)HilbertTable<D>::STATES;
  const uint32_t base = hilbert_start_entry<D>(r, e);
  const uint32_t px = prefix_xor_bits(base & 0xFFu, D * r);
  return (uint16_t)(px | (((base >> 8) + (px & 1u) * ST) << 8));
}

// What hilbert_finish adds to the prefix XOR of the walk's index: the prefix
// XOR of nthbits >> 1 (the prefix XOR is linear).
template <int D, int NB>
QCP_HD constexpr uint64_t hilbert_finish_const() {
  const uint64_t nthbits = ((1ull << (D * NB)) - 1ull) / ((1ull << D) - 1ull);
  uint64_t x = nthbits >> 1;
  for (unsigned s = 1; s < (unsigned)D * NB; s *= 2) x ^= x >> s;
  return x;
}

// R keys of NB bits a coordinate at once, from the coordinate-major table
// `tab` and the start table `start` (entries for the top NB % L levels; not
// read when NB is a multiple of L): the R walks run side by side, so that the
// R chains of dependent table lookups overlap, and every shift and field is
// a constant.  out[r] equals hilbert_key<D>(c[r], NB).
template <int D, int NB, int R>
QCP_HD void hilbert_keys_cm(const uint32_t (&c)[R][D], const uint16_t* tab, const uint16_t* start,
                            uint64_t (&out)[R]) {
  constexpr int L = HilbertTable<D>::L, IN = HilbertTable<D>::IN;
  constexpr int TOP = NB % L;
  uint32_t st[R];  // s | rot << D
  QCP_UNROLL
  for (int r = 0; r < R; ++r) {
    out[r] = 0;
    st[r] = 0;
    if (TOP > 0) {
      uint32_t g = 0;
      QCP_UNROLL
      for (int d = 0; d < D; ++d) g |= ((c[r][d] >> (NB - TOP)) & ((1u << TOP) - 1u)) << (d * TOP);
      const uint32_t e = start[g];
      out[r] = e & 0xFFu;
      st[r] = e >> 8;
    }
  }
  QCP_UNROLL
  for (int b = NB - TOP - L; b >= 0; b -= L) {
    QCP_UNROLL
    for (int r = 0; r < R; ++r) {
      uint32_t g = st[r] << IN;
      QCP_UNROLL
      for (int d = 0; d < D; ++d) g |= ((c[r][d] >> b) & ((1u << L) - 1u)) << (d * L);
      const uint32_t e = tab[g];
      out[r] = (out[r] << IN) | (e & 0xFFu);
      st[r] = e >> 8;
    }
  }
  QCP_UNROLL
  for (int r = 0; r < R; ++r) out[r] = hilbert_finish<D>(out[r], NB);
}

// hilbert_keys_cm from the parity tables (hilbert_ptable_entry_cm,
// hilbert_pstart_entry): the steps already produce the inverse Gray code, so
// each key ends with one constant XOR instead of hilbert_finish's six
// shift-and-XOR rounds.  out[r] equals hilbert_key<D>(c[r], NB).
template <int D, int NB, int R>
QCP_HD void hilbert_keys_pcm(const uint32_t (&c)[R][D], const uint16_t* ptab,
                             const uint16_t* pstart, uint64_t (&out)[R]) {
  constexpr int L = HilbertTable<D>::L, IN = HilbertTable<D>::IN;
  constexpr int TOP = NB % L;
  uint32_t st[R];  // s | rot << D, + STATES when the bits above have odd parity
  QCP_UNROLL
  for (int r = 0; r < R; ++r) {
    out[r] = 0;
    st[r] = 0;
    if (TOP > 0) {
      uint32_t g = 0;
      QCP_UNROLL
      for (int d = 0; d < D; ++d) g |= ((c[r][d] >> (NB - TOP)) & ((1u << TOP) - 1u)) << (d * TOP);
      const uint32_t e = pstart[g];
      out[r] = e & 0xFFu;
      st[r] = e >> 8;
    }
  }
  QCP_UNROLL
  for (int b = NB - TOP - L; b >= 0; b -= L) {
    QCP_UNROLL
    for (int r = 0; r < R; ++r) {
      uint32_t g = st[r] << IN;
      QCP_UNROLL
      for (int d = 0; d < D; ++d) g |= ((c[r][d] >> b) & ((1u << L) - 1u)) << (d * L);
      const uint32_t e = ptab[g];
      out[r] = (out[r] << IN) | (e & 0xFFu);
      st[r] = e >> 8;
    }
  }
  QCP_UNROLL
  for (int r = 0; r < R; ++r) out[r] ^= hilbert_finish_const<D, NB>();
}

// Quantiser of src/mapper/hilbert_mapper.cpp:22 -- (ll)(1000000*p), one FP32
// multiply then truncation.  Returns false outside [0, 2^nb) or for NaN.
QCP_HD bool quant_ref(float p, unsigned nb, uint32_t* out) {
  float x = 1000000.0f * p;
  if (!(x >= 0.0f) || x >= (float)(1ull << nb)) {
    *out = 0;
    return false;
  }
  // 0 <= x < 2^nb <= 2^31: the 32-bit conversion truncates exactly as the
  // reference's (ll) cast does
  *out = (uint32_t)x;
  return true;
}

// 16-bit quantiser of the b16 key mode: clamp(trunc(p*65536), 0, 65535).
QCP_HD uint32_t quant_b16(float p) {
  float x = p * 65536.0f;
  if (x >= 65535.0f) return 65535u;
  if (x > 0.0f) return (uint32_t)x;
  return 0u;
}

}  // namespace qcp
