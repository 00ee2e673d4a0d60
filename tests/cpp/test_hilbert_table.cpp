// CPU check of the library's table-driven Hilbert keys (csrc/hilbert_key.h)
// against the oracle's restatement of HilbertMapper::MappingIntoSingle
// (orc_hilbert_c2i) and the library's per-level walk (hilbert_key): the
// level-major table form (hilbert_key_table) on every point of small grids,
// then random and edge coordinates at every bit width (D = 2: nb 2..31;
// D = 3: nb 2..21); the interleave against bit-by-bit placement; and the form
// the GPU key kernels run (hilbert_keys_pcm: coordinate-major table with the
// parity of the bits above folded in, start table for the top nb % L levels,
// four keys side by side), and its form without the parity (hilbert_keys_cm),
// at the key widths they use and at widths that leave every remainder of
// nb % L.
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

template <int D>
static std::vector<uint16_t> make_table_cm() {
  std::vector<uint16_t> t(qcp::HilbertTable<D>::SIZE);
  for (uint32_t e = 0; e < t.size(); ++e) t[e] = qcp::hilbert_table_entry_cm<D>(e);
  return t;
}

template <int D>
static std::vector<uint16_t> make_ptable_cm() {
  std::vector<uint16_t> t(qcp::HilbertTable<D>::PSIZE);
  for (uint32_t e = 0; e < t.size(); ++e) t[e] = qcp::hilbert_ptable_entry_cm<D>(e);
  return t;
}

// hilbert_keys_cm<D, NB, 4> (P = false) or hilbert_keys_pcm<D, NB, 4> (the
// parity tables the GPU kernels use, P = true) on random and edge points.
template <int D, int NB, bool P>
static bool run_cm() {
  constexpr int L = qcp::HilbertTable<D>::L;
  constexpr int TOP = NB % L;
  const std::vector<uint16_t> tab = P ? make_ptable_cm<D>() : make_table_cm<D>();
  std::vector<uint16_t> start(1u << (D * (TOP ? TOP : 1)));
  for (uint32_t e = 0; e < start.size(); ++e)
    start[e] = P ? qcp::hilbert_pstart_entry<D>(TOP, e) : qcp::hilbert_start_entry<D>(TOP, e);
  const uint32_t lim = (uint32_t)((1ull << NB) - 1);
  std::mt19937_64 rng(99 + D * 100 + NB);
  std::vector<uint32_t> pts;  // D coordinates per point
  const uint32_t edge[] = {0u, 1u, 2u, lim >> 1, (lim >> 1) + 1u, lim - 1u, lim, 999999u & lim,
                           524288u & lim};
  for (uint32_t a : edge)
    for (uint32_t b : edge)
      for (uint32_t z : edge) {
        const uint32_t c[3] = {a, b, z};
        for (int d = 0; d < D; ++d) pts.push_back(c[d]);
      }
  for (int i = 0; i < 40000; ++i)
    for (int d = 0; d < D; ++d) pts.push_back((uint32_t)(rng() & lim) & (i & 1 ? 0xFFFFFu : ~0u));
  const size_t n = pts.size() / D;
  for (size_t i = 0; i < n; i += 4) {
    uint32_t c[4][D];
    for (int r = 0; r < 4; ++r)
      for (int d = 0; d < D; ++d) c[r][d] = pts[((i + r) % n) * D + d];
    uint64_t got[4];
    if (P)
      qcp::hilbert_keys_pcm<D, NB, 4>(c, tab.data(), start.data(), got);
    else
      qcp::hilbert_keys_cm<D, NB, 4>(c, tab.data(), start.data(), got);
    for (int r = 0; r < 4; ++r) {
      uint64_t c64[3] = {0, 0, 0};
      for (int d = 0; d < D; ++d) c64[d] = c[r][d];
      const uint64_t want = orc_hilbert_c2i(D, NB, c64);
      ++cases;
      if (got[r] != want) {
        printf("mismatch (coordinate-major%s) D=%d nb=%d c=(%u,%u,%u): %llx want %llx\n",
               P ? ", parity tables" : "", D, NB,
               c[r][0], c[r][1], D > 2 ? c[r][D - 1] : 0u, (unsigned long long)got[r],
               (unsigned long long)want);
        return false;
      }
    }
  }
  return true;
}

int main() {
  if (!run<2>(31, 6)) return 1;
  if (!run<3>(21, 4)) return 1;
  // the GPU kernels' widths (REF 31 / 20, b16 16) and every remainder of nb % L
  if (!run_cm<2, 31, false>() || !run_cm<2, 16, false>() || !run_cm<2, 29, false>() ||
      !run_cm<2, 30, false>() || !run_cm<2, 5, false>() || !run_cm<3, 20, false>() ||
      !run_cm<3, 16, false>() || !run_cm<3, 21, false>() || !run_cm<3, 3, false>())
    return 1;
  if (!run_cm<2, 31, true>() || !run_cm<2, 16, true>() || !run_cm<2, 29, true>() ||
      !run_cm<2, 30, true>() || !run_cm<2, 5, true>() || !run_cm<3, 20, true>() ||
      !run_cm<3, 16, true>() || !run_cm<3, 21, true>() || !run_cm<3, 3, true>())
    return 1;
  printf("ok %llu\n", (unsigned long long)cases);
  return 0;
}
