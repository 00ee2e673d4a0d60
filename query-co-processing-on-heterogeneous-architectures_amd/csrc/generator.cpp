// generator.cpp -- seeded, counter-based data generator (replaces the cuRAND
// XORWOW generator of generator/cuRAND.cpp:17-75).
//
// Philox4x32-10 (J. Salmon, M. Moraes, R. Dror, D. Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11): 10 rounds of the 4x32 Philox S-box
// with multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments
// 0x9E3779B9 / 0xBB67AE85.  Word i of a (seed, stream) sequence is word i%4 of
// the block with counter {i/4 low, i/4 high, seed high, 0} and key
// {seed low, stream}, so any slice is generated independently, in parallel.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/qcp.h"

namespace {

inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k[0];
  const uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

inline void philox10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k[0] += 0x9E3779B9u;
      k[1] += 0xBB67AE85u;
    }
    philox_round(c, k);
  }
  out[0] = c[0];
  out[1] = c[1];
  out[2] = c[2];
  out[3] = c[3];
}

template <typename Fn>
void parallel_for(uint64_t n, uint64_t grain, Fn fn) {
  unsigned hw = std::thread::hardware_concurrency();
  if (hw == 0) hw = 1;
  if (hw > 64) hw = 64;
  uint64_t chunks = (n + grain - 1) / grain;
  unsigned nt = (unsigned)std::min<uint64_t>(hw, chunks);
  if (nt <= 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  uint64_t per = (n + nt - 1) / nt;
  per = (per + 3) & ~3ull;
  for (unsigned t = 0; t < nt; ++t) {
    uint64_t b = per * t, e = std::min<uint64_t>(n, b + per);
    if (b >= e) break;
    th.emplace_back([=] { fn(b, e); });
  }
  for (auto& x : th) x.join();
}

void words_range(uint32_t* out, uint64_t b, uint64_t e, uint64_t seed, uint32_t stream,
                 uint64_t offset) {
  const uint32_t key[2] = {(uint32_t)seed, stream};
  uint64_t i = b;
  while (i < e) {
    const uint64_t g = i + offset;
    const uint64_t blk = g >> 2;
    const uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)(seed >> 32), 0u};
    uint32_t w[4];
    philox10(ctr, key, w);
    for (uint32_t j = (uint32_t)(g & 3); j < 4 && i < e; ++j, ++i) out[i - b] = w[j];
  }
}

}  // namespace

extern "C" {

void qcp_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox10(ctr, key, out);
}

void qcp_gen_words(uint32_t* out, uint64_t n, uint64_t seed, uint32_t stream, uint64_t offset) {
  parallel_for(n, 1 << 20, [&](uint64_t b, uint64_t e) {
    words_range(out + b, b, e, seed, stream, offset);
  });
}

void qcp_gen_uniform(float* out, uint64_t n, uint64_t seed, uint32_t stream, uint64_t offset) {
  parallel_for(n, 1 << 20, [&](uint64_t b, uint64_t e) {
    std::vector<uint32_t> w(e - b);
    words_range(w.data(), b, e, seed, stream, offset);
    for (uint64_t i = b; i < e; ++i) out[i] = (float)(w[i - b] >> 8) * (1.0f / 16777216.0f);
  });
}

void qcp_gen_normal(double* out, uint64_t n, uint64_t seed, uint32_t stream, uint64_t offset) {
  // pairs (2i, 2i+1) of words -> (z[2i], z[2i+1]); offset counts normals and
  // must be even
  const uint64_t pairs = (n + 1) / 2;
  const double two_pi = 6.283185307179586476925286766559;
  parallel_for(pairs, 1 << 19, [&](uint64_t b, uint64_t e) {
    std::vector<uint32_t> w(2 * (e - b));
    words_range(w.data(), 2 * b, 2 * e, seed, stream, offset);
    for (uint64_t p = b; p < e; ++p) {
      const double u1 = ((double)(w[2 * (p - b)] >> 8) + 1.0) * (1.0 / 16777216.0);
      const double u2 = (double)(w[2 * (p - b) + 1] >> 8) * (1.0 / 16777216.0);
      const double r = sqrt(-2.0 * log(u1));
      const double z0 = r * cos(two_pi * u2), z1 = r * sin(two_pi * u2);
      out[2 * p] = z0;
      if (2 * p + 1 < n) out[2 * p + 1] = z1;
    }
  });
}

}  // extern "C"
