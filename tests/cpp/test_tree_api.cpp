// Drives the tree::Tree mirror (include/qcp_tree.hpp) the way the reference's
// Evaluator drives its trees: AddTrees-style construction by index type,
// Build over a data set read from a raw file, Search with repeats, dump and
// reload, and the error paths.  Writes "count hash" per query for each type to
// OUTDIR/<type>.txt; tests/test_gpu_tree_api.py compares them with the oracle.
//
// usage: test_tree_api DATA QUERIES DIMS N Q OUTDIR
#include <stdio.h>
#include <stdlib.h>

#include <string>

#include "../../include/qcp_tree.hpp"

using namespace qcp;

static int failures = 0;
#define CHECK(cond, what)                                  \
  do {                                                     \
    if (!(cond)) {                                         \
      fprintf(stderr, "FAILED: %s (%s)\n", what, #cond);   \
      ++failures;                                          \
    }                                                      \
  } while (0)

static bool write_results(const std::string& path, const tree::Tree& t) {
  FILE* f = fopen(path.c_str(), "w");
  if (!f) return false;
  for (size_t i = 0; i < t.GetCounts().size(); ++i)
    fprintf(f, "%u %llu\n", t.GetCounts()[i], (unsigned long long)t.GetHashes()[i]);
  fclose(f);
  return true;
}

int main(int argc, char** argv) {
  if (argc != 7) {
    fprintf(stderr, "usage: test_tree_api DATA QUERIES DIMS N Q OUTDIR\n");
    return 2;
  }
  const std::string data = argv[1], queries = argv[2], out = argv[6];
  const ui dims = (ui)atoi(argv[3]), n = (ui)atoi(argv[4]), q = (ui)atoi(argv[5]);
  auto input = io::DataSet::Read(data, dims, n);
  auto qs = io::DataSet::Read(queries, dims, q, true);
  CHECK(input && qs, "read data and queries");
  if (!input || !qs) return 1;
  CHECK(!io::DataSet::Read(data + ".missing", dims, n), "a missing file gives nullptr");
  CHECK(!io::DataSet::Read(data, dims, n + 1), "a short file gives nullptr");

  const char* types[] = {"m", "r", "b", "h", "hr", "hb", "rls", "pr"};
  for (const char* ty : types) {
    auto t = tree::MakeTree(ty);
    CHECK(t != nullptr, ty);
    if (!t) continue;
    CHECK(t->Search(qs, q, 1) == 0 && t->GetLastStatus() == QCP_EINVAL, "search before build");
    bool ok = t->Build(input);
    CHECK(ok, ty);
    if (!ok) {
      fprintf(stderr, "%s: build: %s\n", ty, qcp_strerror(t->GetLastStatus()));
      continue;
    }
    CHECK(t->Search(qs, q, 2) == 1, ty);
    printf("%s: Hit : %llu\n", ty, (unsigned long long)t->GetTotalHit());
    unsigned long long sum = 0;
    for (uint32_t c : t->GetCounts()) sum += c;
    CHECK(sum == t->GetTotalHit(), "total equals the sum of the counts");
    CHECK(write_results(out + "/" + ty + ".txt", *t), "write results");
    CHECK(t->Search(qs, q + 1, 1) == 0 && t->GetLastStatus() == QCP_EINVAL,
          "more queries than the set holds");
  }

  // dump and reload (DumpToFile / DumpFromFile), for a packed and an inserted tree
  for (const char* ty : {"m", "r"}) {
    auto a = tree::MakeTree(ty);
    auto b = tree::MakeTree(ty);
    const std::string path = out + "/ix_" + ty + ".qcp";
    CHECK(a->Build(input) && a->DumpToFile(path), "dump");
    CHECK(b->DumpFromFile(path), "reload");
    CHECK(a->Search(qs, q, 1) == 1 && b->Search(qs, q, 1) == 1, "search after reload");
    CHECK(a->GetCounts() == b->GetCounts() && a->GetHashes() == b->GetHashes(),
          "reloaded tree gives the same results");
    CHECK(!b->DumpFromFile(path + ".missing") && b->GetLastStatus() == QCP_EIO,
          "reload of a missing file");
    CHECK(b->Search(qs, q, 1) == 1, "a failed reload keeps the loaded tree");
  }

  // the reference's setters
  tree::MPHR m;
  // data partitions (the reference's -t): 7 subtrees, searched in the
  // persistent form and in the reference's partition mode (plain)
  m.SetNumberOfPartition(7);
  CHECK(m.GetTreeType() == tree::TREE_TYPE_MPHR_PARTITION, "partitioned tree type");
  CHECK(m.Build(input) && m.Search(qs, q, 1) == 1, "MPHR with 7 partitions");
  CHECK(write_results(out + "/m_part7.txt", m), "write partition results");
  m.SetSearchMode(QCP_SEARCH_PLAIN);
  CHECK(m.Search(qs, q, 1) == 1, "MPHR with 7 partitions, plain form");
  CHECK(write_results(out + "/m_part7_plain.txt", m), "write partition plain results");
  m.SetSearchMode(QCP_SEARCH_PERSISTENT);
  m.SetNumberOfPartition(n + 1);
  CHECK(!m.Build(input) && m.GetLastStatus() == QCP_EINVAL, "more partitions than objects");
  m.SetNumberOfPartition(1);
  m.SetNumberOfCUDABlocks(64);
  m.SetSearchMode(QCP_SEARCH_PLAIN);
  CHECK(m.Build(input) && m.Search(qs, q, 1) == 1, "MPHR plain form");
  CHECK(write_results(out + "/m_plain.txt", m), "write plain results");
  tree::Hybrid h;
  h.SetScanLevel(1);
  h.SetChunkSize(500);
  h.SetNumberOfCPUThreads(3);
  h.SetUpperTreeType(tree::TREE_TYPE_BVH);
  CHECK(h.Build(input) && h.Search(qs, q, 1) == 1, "Hybrid over a BVH, level 1");
  CHECK(write_results(out + "/hb1.txt", h), "write hybrid results");
  CHECK(!m.Build(nullptr) && m.GetLastStatus() == QCP_EINVAL, "build without data");
  auto wrong = std::make_shared<io::DataSet>(dims == 3 ? 2 : 3, std::vector<float>(12, 0.5f), true);
  CHECK(m.Search(wrong, 1, 1) == 0 && m.GetLastStatus() == QCP_EINVAL, "queries of another dimension");

  printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
  return failures ? 1 : 0;
}
