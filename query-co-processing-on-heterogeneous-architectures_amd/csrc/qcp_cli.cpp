// qcp_cli.cpp -- command-line driver over libqcp.so, flag-compatible with the
// reference's evaluator (Evaluator::ParseArgs, src/evaluator/evaluator.cpp:
// 328-405; index types of AddTrees, 411-444).  The reference reads data and
// queries from fixed paths on its author's machine (GetDataPath/GetQueryPath,
// 480-505); here they are given with --data / --queries (raw little-endian
// FP32: N x D points, or N x 2D boxes with --boxes; Q x 2D query boxes laid out
// [lo(D), hi(D)], the io::DataSet format of src/io/dataset.cpp:36-39,82-88),
// and can be generated with --gen-data / --gen-queries (Philox4x32-10, in
// place of generator/cuRAND.cpp).
//
// Reference flags (upper- and lower-case alike, as in the reference):
//   -i TYPE   m|mphr, r|rtree (the reference RTree: Guttman insertion), b|bvh,
//             h|hybrid (MPHR, co-processing search), hr (Guttman R-tree,
//             co-processing), hb (BVH, co-processing), rls|rtree_ls (RTree_LS,
//             src/tree/rtree_ls.cpp: a Guttman tree over the objects in input
//             order -- RTree_LS::Build never sorts -- whose level-1 nodes are
//             its large leaves, searched by host threads down to the leaves and
//             GPU leaf scans); pr|packed_rtree selects this build's
//             Hilbert-packed R-tree
//   -d N      number of data objects (a trailing "m" multiplies by 10^6)
//   -q Q      number of queries
//   -b B      queries per launch of the plain form (number_of_cuda_blocks)
//   -p P      CPU threads (co-processing search, CPU baseline)
//   -l L      co-processing hand-off level (scan_level)
//   -c C      co-processing pipeline block, in queries (chunk_size)
//   -r R      repeats
//   -e E      evaluation mode: nonzero runs the autotuner (the reference's sweep);
//             with the hybrid form and no -l / -c, also the co-processing tuner
//   -s S      selectivity of --gen-queries (volume fraction)
//   -u U      cluster type (GetClusterType, 431-446): h|hilbert (default),
//             k|kmeans (k-means clusters ranked by their centres' Hilbert keys,
//             --clusters K, default 20000; --kmeans-iters I, default 1001),
//             o|original (the input order is the rank order)
//   -t T      MPHR data partitions (SetNumberOfPartition): T subtrees over
//             contiguous rank ranges; MPHR indexes only (-i m)
//   -y Y, -f  accepted; the data type has no counterpart (data from --data);
//             -f rebuilds even when --index exists
// Options of this driver:
//   --data FILE --queries FILE --dims D --fanout F (default: the reference's
//   degree, 192 for MPHR, 128 otherwise) --boxes --center
//   --key ref|b16 --mode persistent|plain|hybrid --index FILE (load/save)
//   --out FILE (per-query count and hash, text) --contain --json
//   --gen-data FILE --gen-queries FILE --seed S
//   --ref-dump FILE [--ref-degree F]: index the objects of a reference MPHR
//             dump (MPHR::DumpToFile) in its rank order, F = its leaf degree
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/qcp.h"

namespace {

struct Opts {
  std::string index_type = "m";
  uint64_t n = 0, q = 0;
  uint32_t blocks = 128, threads = 0, scan_level = 0, chunk = 0, repeat = 1, eval = 0;
  double selectivity = 1e-4;
  bool force = false;
  std::string data, queries, index_file, out, gen_data, gen_queries;
  uint32_t dims = 3, fanout = 0;  // 0: the reference's degree of the index type
  bool boxes = false, center = false, b16 = false, contain = false, json = false;
  std::string mode;
  uint64_t seed = 1234;
  std::string ref_dump;
  uint32_t ref_degree = 192;
  std::string cluster = "h";
  uint32_t clusters = 0, kmeans_iters = 0;
  uint32_t partitions = 1;
};

uint64_t parse_count(const char* s) {
  std::string v(s);
  uint64_t mult = 1;
  if (!v.empty() && (v.back() == 'm' || v.back() == 'M')) {
    mult = 1000000;
    v.pop_back();
  }
  return std::stoull(v) * mult;
}

bool read_floats(const std::string& path, uint64_t count, std::vector<float>& out) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  out.resize(count);
  const size_t got = count ? fread(out.data(), sizeof(float), count, f) : 0;
  fclose(f);
  return got == count;
}

uint64_t file_floats(const std::string& path) {
  struct stat st;
  if (stat(path.c_str(), &st) != 0) return 0;
  return (uint64_t)st.st_size / sizeof(float);
}

bool write_floats(const std::string& path, const std::vector<float>& v) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  const size_t put = v.empty() ? 0 : fwrite(v.data(), sizeof(float), v.size(), f);
  fclose(f);
  return put == v.size();
}

int fail(const char* what, qcp_status s) {
  fprintf(stderr, "qcp_cli: %s: %s\n", what, qcp_strerror(s));
  return 1;
}

void usage() {
  fprintf(stderr,
          "usage: qcp_cli -i m|r|pr|b|h|hr|hb|rls -d N -q Q [-b B -p P -l L -c C -r R -e E -s S\n"
          "               -u h|k|o -t T] --data FILE --queries FILE [--dims D --fanout F\n"
          "               --boxes --center --key ref|b16 --mode persistent|plain|hybrid\n"
          "               --index FILE --out FILE --contain --json --clusters K\n"
          "               --kmeans-iters I]\n"
          "       qcp_cli --gen-data FILE -d N [--dims D --seed S]\n"
          "       qcp_cli --gen-queries FILE -q Q -s SELECTIVITY [--dims D --seed S]\n");
}

}  // namespace

int main(int argc, char** argv) {
  Opts o;
  static const struct option longopts[] = {
      {"data", required_argument, nullptr, 1000},     {"queries", required_argument, nullptr, 1001},
      {"dims", required_argument, nullptr, 1002},     {"fanout", required_argument, nullptr, 1003},
      {"boxes", no_argument, nullptr, 1004},          {"center", no_argument, nullptr, 1005},
      {"key", required_argument, nullptr, 1006},      {"mode", required_argument, nullptr, 1007},
      {"index", required_argument, nullptr, 1008},    {"out", required_argument, nullptr, 1009},
      {"contain", no_argument, nullptr, 1010},        {"json", no_argument, nullptr, 1011},
      {"gen-data", required_argument, nullptr, 1012}, {"gen-queries", required_argument, nullptr, 1013},
      {"seed", required_argument, nullptr, 1014},     {"help", no_argument, nullptr, 'h'},
      {"ref-dump", required_argument, nullptr, 1015}, {"ref-degree", required_argument, nullptr, 1016},
      {"clusters", required_argument, nullptr, 1017},
      {"kmeans-iters", required_argument, nullptr, 1018},
      {nullptr, 0, nullptr, 0}};
  // the reference's option string: every letter in both cases
  static const char* shortopts = "c:C:i:I:d:D:q:Q:b:B:p:P:s:S:l:L:r:R:e:E:t:T:y:Y:u:U:fFh";
  int c;
  try {
    while ((c = getopt_long(argc, argv, shortopts, longopts, nullptr)) != -1) {
      switch (c) {
        case 'i': case 'I': o.index_type = optarg; break;
        case 'c': case 'C': o.chunk = (uint32_t)atoi(optarg); break;
        case 'd': case 'D': o.n = parse_count(optarg); break;
        case 'q': case 'Q': o.q = parse_count(optarg); break;
        case 'b': case 'B': o.blocks = (uint32_t)atoi(optarg); break;
        case 'p': case 'P': o.threads = (uint32_t)atoi(optarg); break;
        case 's': case 'S': o.selectivity = atof(optarg); break;
        case 'l': case 'L': o.scan_level = (uint32_t)atoi(optarg); break;
        case 'r': case 'R': o.repeat = (uint32_t)atoi(optarg); break;
        case 'e': case 'E': o.eval = (uint32_t)atoi(optarg); break;
        case 'u': case 'U': o.cluster = optarg; break;
        case 't': case 'T': o.partitions = (uint32_t)atoi(optarg); break;
        case 'y': case 'Y': break;
        case 'f': case 'F': o.force = true; break;
        case 1000: o.data = optarg; break;
        case 1001: o.queries = optarg; break;
        case 1002: o.dims = (uint32_t)atoi(optarg); break;
        case 1003: o.fanout = (uint32_t)atoi(optarg); break;
        case 1004: o.boxes = true; break;
        case 1005: o.center = true; break;
        case 1006: o.b16 = strcasecmp(optarg, "b16") == 0; break;
        case 1007: o.mode = optarg; break;
        case 1008: o.index_file = optarg; break;
        case 1009: o.out = optarg; break;
        case 1010: o.contain = true; break;
        case 1011: o.json = true; break;
        case 1012: o.gen_data = optarg; break;
        case 1013: o.gen_queries = optarg; break;
        case 1014: o.seed = std::stoull(optarg); break;
        case 1015: o.ref_dump = optarg; break;
        case 1016: o.ref_degree = (uint32_t)atoi(optarg); break;
        case 1017: o.clusters = (uint32_t)atoi(optarg); break;
        case 1018: o.kmeans_iters = (uint32_t)atoi(optarg); break;
        default: usage(); return 2;
      }
    }
  } catch (...) {
    usage();
    return 2;
  }
  const uint32_t D = o.dims;
  if (D != 2 && D != 3) {
    fprintf(stderr, "qcp_cli: --dims must be 2 or 3\n");
    return 2;
  }
  // ---- generators --------------------------------------------------------
  if (!o.gen_data.empty()) {
    std::vector<float> v((size_t)o.n * D);
    qcp_gen_uniform(v.data(), v.size(), o.seed, 0, 0);
    if (!write_floats(o.gen_data, v)) return fail("write", QCP_EIO);
    printf("wrote %llu %u-D points to %s\n", (unsigned long long)o.n, D, o.gen_data.c_str());
    return 0;
  }
  if (!o.gen_queries.empty()) {
    // square boxes of the given volume fraction, centres uniform (streams 1, 2)
    std::vector<float> cen((size_t)o.q * D), v((size_t)o.q * 2 * D);
    qcp_gen_uniform(cen.data(), cen.size(), o.seed, 1, 0);
    const float side = (float)std::pow(o.selectivity, 1.0 / D);
    const float h = 0.5f * side;
    for (uint64_t i = 0; i < o.q; ++i)
      for (uint32_t d = 0; d < D; ++d) {
        v[i * 2 * D + d] = cen[i * D + d] - h;
        v[i * 2 * D + D + d] = cen[i * D + d] + h;
      }
    if (!write_floats(o.gen_queries, v)) return fail("write", QCP_EIO);
    printf("wrote %llu queries (selectivity %g) to %s\n", (unsigned long long)o.q, o.selectivity,
           o.gen_queries.c_str());
    return 0;
  }
  // ---- index type ----------------------------------------------------------
  std::string t = o.index_type;
  for (auto& ch : t) ch = (char)tolower(ch);
  uint32_t layout = QCP_LAYOUT_MPHR;
  bool input_order = false;  // index the objects in the given order (no keys, no sort)
  std::string mode = o.mode;
  if (t == "m" || t == "mphr") layout = QCP_LAYOUT_MPHR;
  else if (t == "r" || t == "rtree") layout = QCP_LAYOUT_GUTTMAN;
  else if (t == "pr" || t == "packed_rtree") layout = QCP_LAYOUT_RTREE;
  else if (t == "b" || t == "bvh") layout = QCP_LAYOUT_BVH;
  else if (t == "h" || t == "hybrid") { layout = QCP_LAYOUT_MPHR; if (mode.empty()) mode = "hybrid"; }
  else if (t == "hr") { layout = QCP_LAYOUT_GUTTMAN; if (mode.empty()) mode = "hybrid"; }
  else if (t == "hb") { layout = QCP_LAYOUT_BVH; if (mode.empty()) mode = "hybrid"; }
  else if (t == "rls" || t == "rtree_ls") {
    layout = QCP_LAYOUT_GUTTMAN;
    input_order = true;
    if (mode.empty()) mode = "hybrid";
  }
  else {
    fprintf(stderr, "qcp_cli: index type %s has no counterpart here\n", o.index_type.c_str());
    return 2;
  }
  // the reference's compile-time degrees (src/common/config.h:6-8): MPHR leaves
  // 192, R-tree / BVH nodes 128
  if (!o.fanout) o.fanout = layout == QCP_LAYOUT_MPHR ? 192u : 128u;
  if (mode.empty()) mode = "persistent";
  uint32_t smode = QCP_SEARCH_PERSISTENT;
  if (mode == "plain") smode = QCP_SEARCH_PLAIN;
  else if (mode == "hybrid") smode = QCP_SEARCH_HYBRID;
  else if (mode != "persistent") { usage(); return 2; }
  if (qcp_device_count() < 1) return fail("device", QCP_ENODEV);

  // ---- data and index ------------------------------------------------------
  const uint32_t E = o.boxes ? 2 * D : D;
  qcp_search_cfg tuned;  // the autotuner's choice (-e)
  memset(&tuned, 0, sizeof(tuned));
  bool have_tuned = false;
  qcp_index* ix = nullptr;
  qcp_status st;
  struct stat sb;
  const bool have_index = !o.index_file.empty() && stat(o.index_file.c_str(), &sb) == 0;
  if (have_index && !o.force) {
    if ((st = qcp_load(o.index_file.c_str(), 0, &ix)) != QCP_OK) return fail("load", st);
  } else if (!o.ref_dump.empty()) {
    uint64_t n = 0;
    const uint32_t kind = o.boxes ? (uint32_t)QCP_OBJ_BOXES : (uint32_t)QCP_OBJ_POINTS;
    if ((st = qcp_read_reference_mphr(o.ref_dump.c_str(), D, o.ref_degree, kind, nullptr, 0, &n)) != QCP_OK)
      return fail("reference dump", st);
    std::vector<float> objs((size_t)n * E);
    if ((st = qcp_read_reference_mphr(o.ref_dump.c_str(), D, o.ref_degree, kind, objs.data(), n,
                                      &n)) != QCP_OK)
      return fail("reference dump", st);
    qcp_build_cfg cfg = {D, o.fanout, layout, QCP_KEY_REF, QCP_SRC_LO, kind, 0, 0,
                         QCP_ORDER_HILBERT, 0, 0, o.partitions};
    if ((st = qcp_build_ranked(&cfg, objs.data(), n, &ix)) != QCP_OK) return fail("build", st);
    if (!o.index_file.empty() && (st = qcp_save(ix, o.index_file.c_str())) != QCP_OK)
      return fail("save", st);
  } else {
    if (o.data.empty()) { usage(); return 2; }
    if (!o.n) o.n = file_floats(o.data) / E;
    std::vector<float> objs;
    if (!read_floats(o.data, o.n * E, objs)) return fail("read data", QCP_EIO);
    qcp_build_cfg cfg = {D, o.fanout, layout, o.b16 ? (uint32_t)QCP_KEY_B16 : (uint32_t)QCP_KEY_REF,
                         o.center ? (uint32_t)QCP_SRC_CENTER : (uint32_t)QCP_SRC_LO,
                         o.boxes ? (uint32_t)QCP_OBJ_BOXES : (uint32_t)QCP_OBJ_POINTS, 0, 0,
                         QCP_ORDER_HILBERT, o.clusters, o.kmeans_iters, o.partitions};
    std::string ct = o.cluster;
    for (auto& ch : ct) ch = (char)tolower(ch);
    const bool original = input_order || ct == "o" || ct == "original";
    if (ct == "k" || ct == "kmeans") cfg.order = QCP_ORDER_KMEANS_HILBERT;
    else if (!original && ct != "h" && ct != "hilbert") { usage(); return 2; }
    if (o.eval) {
      // the reference's parameter sweep: the autotuner on the first queries
      have_tuned = true;
      std::vector<float> qs;
      const uint64_t nq = std::min<uint64_t>(o.q ? o.q : file_floats(o.queries) / (2 * D), 20000);
      if (!read_floats(o.queries, nq * 2 * D, qs)) return fail("read queries", QCP_EIO);
      void *d_objs = nullptr, *d_q = nullptr;
      if ((st = qcp_device_alloc(0, 4ull * objs.size(), &d_objs)) != QCP_OK ||
          (st = qcp_copy_to_device(d_objs, objs.data(), 4ull * objs.size())) != QCP_OK ||
          (st = qcp_device_alloc(0, 4ull * qs.size(), &d_q)) != QCP_OK ||
          (st = qcp_copy_to_device(d_q, qs.data(), 4ull * qs.size())) != QCP_OK)
        return fail("device copy", st);
      const uint32_t fans[3] = {32, 64, 128};
      std::vector<qcp_tune_entry> table(1024);
      uint32_t tl = 0;
      st = qcp_autotune(&cfg, (const float*)d_objs, o.n, (const float*)d_q, nq, fans, 3,
                        QCP_OUT_COUNTS | QCP_OUT_HASH, &ix, &tuned, table.data(),
                        (uint32_t)table.size(), &tl);
      qcp_device_free(d_objs);
      if (st != QCP_OK) {
        qcp_device_free(d_q);
        return fail("autotune", st);
      }
      qcp_index_info inf;
      qcp_index_info_get(ix, &inf);
      printf("autotune: fanout %u, %u waves per workgroup, %u leaf chunks, donation %s "
             "(%u configurations)\n", inf.fanout, tuned.block_waves, tuned.leaf_chunks,
             tuned.donate ? "on" : "off", tl);
      if (smode == QCP_SEARCH_HYBRID && !o.scan_level && !o.chunk) {
        // the co-processing split (-l, -c not given): the tuner picks both
        qcp_search_cfg base = tuned;
        base.cpu_threads = o.threads;
        base.contain = o.contain ? 1 : 0;
        std::vector<qcp_hybrid_tune_entry> ht(256);
        uint32_t hl = 0;
        st = qcp_autotune_hybrid(ix, (const float*)d_q, nq, &base, &tuned, ht.data(),
                                 (uint32_t)ht.size(), &hl);
        if (st != QCP_OK) {
          qcp_device_free(d_q);
          return fail("hybrid autotune", st);
        }
        o.scan_level = tuned.handoff_level;
        o.chunk = tuned.hybrid_batch;
        printf("hybrid autotune: hand-off level %u%s, %u queries a block (%u configurations)\n",
               tuned.handoff_level, tuned.handoff_level ? "" : " (the leaves)",
               tuned.hybrid_batch, hl);
      }
      qcp_device_free(d_q);
    } else if ((st = original ? qcp_build_ranked(&cfg, objs.data(), o.n, &ix)
                              : qcp_build(&cfg, objs.data(), o.n, &ix)) != QCP_OK) {
      return fail("build", st);
    }
    if (!o.index_file.empty() && (st = qcp_save(ix, o.index_file.c_str())) != QCP_OK)
      return fail("save", st);
  }
  qcp_index_info info;
  qcp_index_info_get(ix, &info);
  printf("index: %llu objects, %llu nodes, %u levels, fanout %u, build %.3f ms\n",
         (unsigned long long)info.n, (unsigned long long)info.n_nodes, info.height, info.fanout,
         info.build_ms);

  // ---- queries and search ----------------------------------------------------
  if (o.queries.empty()) {
    qcp_index_destroy(ix);
    return 0;
  }
  if (!o.q) o.q = file_floats(o.queries) / (2 * D);
  std::vector<float> qs;
  if (!read_floats(o.queries, o.q * 2 * D, qs)) return fail("read queries", QCP_EIO);
  qcp_batch* b = nullptr;
  if ((st = qcp_batch_create(ix, qs.data(), o.q, &b)) != QCP_OK) return fail("batch", st);
  qcp_search_cfg sc;
  memset(&sc, 0, sizeof(sc));
  sc.mode = smode;
  sc.outputs = QCP_OUT_COUNTS | QCP_OUT_HASH;
  sc.plain_batch = o.blocks;
  sc.contain = o.contain ? 1 : 0;
  sc.cpu_threads = o.threads;
  sc.handoff_level = o.scan_level;
  sc.hybrid_batch = o.chunk;
  if (have_tuned) {
    sc.block_waves = tuned.block_waves;
    sc.leaf_chunks = tuned.leaf_chunks;
    sc.donate = tuned.donate;
  }
  qcp_search_stats ss;
  double best_ms = 1e30;
  for (uint32_t r = 0; r < std::max(1u, o.repeat); ++r) {
    const auto t0 = std::chrono::steady_clock::now();
    if ((st = qcp_batch_run(b, &sc, &ss)) != QCP_OK) return fail("search", st);
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    best_ms = std::min(best_ms, ms);
    // the reference's log lines (src/tree/mphr.cpp:262-265)
    printf("#%u) Hit : %llu\n", r + 1, (unsigned long long)ss.total_hits);
    printf("Avg. Search Time on the GPU (ms)\n%.6f\n", ss.total_ms / (double)std::max<uint64_t>(o.q, 1));
    printf("Total Search Time (ms) %.6f, device %.6f\n", ms, ss.kernel_ms);
  }
  std::vector<uint32_t> counts(o.q);
  std::vector<uint64_t> hashes(o.q);
  if ((st = qcp_batch_results(b, counts.data(), hashes.data(), nullptr, nullptr, 0, nullptr)) != QCP_OK)
    return fail("results", st);
  if (!o.out.empty()) {
    FILE* f = fopen(o.out.c_str(), "w");
    if (!f) return fail("write", QCP_EIO);
    for (uint64_t i = 0; i < o.q; ++i)
      fprintf(f, "%u %llu\n", counts[i], (unsigned long long)hashes[i]);
    fclose(f);
  }
  if (o.json)
    printf("{\"queries\": %llu, \"hits\": %llu, \"best_ms\": %.6f, \"queries_per_s\": %.6e, "
           "\"mode\": \"%s\", \"layout\": %u}\n",
           (unsigned long long)o.q, (unsigned long long)ss.total_hits, best_ms,
           o.q / (best_ms * 1e-3), mode.c_str(), layout);
  qcp_batch_destroy(b);
  qcp_index_destroy(ix);
  return 0;
}
