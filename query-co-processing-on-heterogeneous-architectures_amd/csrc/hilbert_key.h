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

namespace qcp {

// Hilbert index of c[0..D) with nb bits per coordinate (D*nb <= 63, nb >= 2).
template <int D>
QCP_HD uint64_t hilbert_key(const uint32_t* c, unsigned nb) {
  const uint32_t nd_ones = (1u << D) - 1u;
  const uint32_t nd1_ones = nd_ones >> 1;
  uint64_t index = 0;
  unsigned rotation = 0;
  uint32_t flip = 0;
  uint32_t above = 0;  // digit of the level above (0 above the top)
  uint64_t nthbits = 0;
  for (int b = (int)nb - 1; b >= 0; --b) {
    uint32_t t = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) t |= ((c[d] >> b) & 1u) << d;
    uint32_t g = t ^ above;  // coords ^= coords >> nDims
    above = t;
    uint32_t x = flip ^ g;
    uint32_t bits = ((x >> rotation) | (x << (D - rotation))) & nd_ones;
    index = (index << D) | bits;
    flip = 1u << rotation;
    // adjust_rotation: rotation += 1 + (position of the lowest set bit of
    // bits, counted from 1, if that bit lies below the top dimension)
    uint32_t low = bits & (0u - bits) & nd1_ones;
    while (low) {
      low >>= 1;
      ++rotation;
    }
    if (++rotation >= (unsigned)D) rotation -= D;
    nthbits = (nthbits << D) | 1u;
  }
  index ^= nthbits >> 1;
  for (unsigned s = 1; s < (unsigned)D * nb; s *= 2) index ^= index >> s;
  return index;
}

// Quantiser of src/mapper/hilbert_mapper.cpp:22 -- (ll)(1000000*p), one FP32
// multiply then truncation.  Returns false outside [0, 2^nb) or for NaN.
QCP_HD bool quant_ref(float p, unsigned nb, uint32_t* out) {
  float x = 1000000.0f * p;
  if (!(x >= 0.0f) || x >= (float)(1ull << nb)) {
    *out = 0;
    return false;
  }
  long long v = (long long)x;
  if (v >= (1ll << nb)) {
    *out = 0;
    return false;
  }
  *out = (uint32_t)v;
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
