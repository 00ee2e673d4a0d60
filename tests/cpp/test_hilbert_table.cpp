// CPU check of the library's table-driven Hilbert key (csrc/hilbert_key.h,
// hilbert_key_table, the form the GPU key kernel runs) against the oracle's
// restatement of HilbertMapper::MappingIntoSingle (orc_hilbert_c2i) and the
// library's per-level walk (hilbert_key): every point of small grids, then
// random and edge coordinates at every bit width (D = 2: nb 2..31; D = 3:
// nb 2..21).  Also checks the interleave against bit-by-bit placement.
// Prints "ok <cases>" and exits 0, or names the first mismatch and exits 1.
#include <stdint.h>
#include <stdio.h>

#include <random>
#include <vector>

#include "../../oracle/qcp_oracle.h"
#include "../../query-co-processing-on-heterogeneous-architectures_amd/csrc/hilbert_key.h"

template <int D>
static std::vector<uint16_t> make_table() {
  std::vector<uint16_t> t(qcp::HilbertTable<D>::SIZE);
  for (uint32_t e = 0; e < t.size(); ++e) t[e] = qcp::hilbert_table_entry<D>(e);
  return t;
}

static uint64_t cases = 0;

template <int D>
static bool check(const uint32_t* c, unsigned nb, const std::vector<uint16_t>& tab) {
  uint64_t c64[3] = {0, 0, 0};
  uint64_t m = 0;
  for (int d = 0; d < D; ++d) {
    c64[d] = c[d];
    for (int b = 0; b < 32; ++b) m |= (uint64_t)((c[d] >> b) & 1u) << (D * b + d);
  }
  const uint64_t want = orc_hilbert_c2i(D, nb, c64);
  const uint64_t walk = qcp::hilbert_key<D>(c, nb);
  const uint64_t got = qcp::hilbert_key_table<D>(c, nb, tab.data());
  const uint64_t il = qcp::interleave<D>(c);
  ++cases;
  if (got != want || walk != want || il != m) {
    printf("mismatch D=%d nb=%u c=(%u,%u,%u): oracle %llx walk %llx table %llx "
           "interleave %llx want %llx\n",
           D, nb, c[0], c[1], D > 2 ? c[2] : 0u, (unsigned long long)want,
           (unsigned long long)walk, (unsigned long long)got, (unsigned long long)il,
           (unsigned long long)m);
    return false;
  }
  return true;
}

template <int D>
static bool run(unsigned max_nb, unsigned exhaustive_nb) {
  const std::vector<uint16_t> tab = make_table<D>();
  std::mt19937_64 rng(1234 + D);
  for (unsigned nb = 2; nb <= max_nb; ++nb) {
    const uint32_t lim = nb >= 32 ? 0xFFFFFFFFu : (uint32_t)((1ull << nb) - 1);
    if (nb <= exhaustive_nb) {
      const uint64_t side = 1ull << nb;
      uint64_t total = 1;
      for (int d = 0; d < D; ++d) total *= side;
      for (uint64_t i = 0; i < total; ++i) {
        uint32_t c[3] = {0, 0, 0};
        uint64_t r = i;
        for (int d = 0; d < D; ++d) {
          c[d] = (uint32_t)(r % side);
          r /= side;
        }
        if (!check<D>(c, nb, tab)) return false;
      }
    }
    const uint32_t edge[] = {0u, 1u, 2u, lim >> 1, (lim >> 1) + 1u, lim - 1u, lim,
                             999999u & lim, 524288u & lim};
    for (uint32_t a : edge)
      for (uint32_t b : edge)
        for (uint32_t z : edge) {
          uint32_t c[3] = {a, b, z};
          if (!check<D>(c, nb, tab)) return false;
        }
    for (int i = 0; i < 20000; ++i) {
      uint32_t c[3];
      for (int d = 0; d < 3; ++d) c[d] = (uint32_t)(rng() & lim);
      // also coordinates below the quantiser's 1e6 (the leading levels all zero)
      if (i & 1)
        for (int d = 0; d < D; ++d) c[d] &= (uint32_t)((1u << 20) - 1u);
      if (!check<D>(c, nb, tab)) return false;
    }
  }
  return true;
}

int main() {
  if (!run<2>(31, 6)) return 1;
  if (!run<3>(21, 4)) return 1;
  printf("ok %llu\n", (unsigned long long)cases);
  return 0;
}
