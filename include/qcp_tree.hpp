/*
 * qcp_tree.hpp -- the reference's tree::Tree interface in C++, over the C ABI
 * of qcp.h (header only; link with -lqcp).
 *
 * The reference's driver (Evaluator::Build / Search, src/evaluator/evaluator.cpp:
 * 121-280) holds std::shared_ptr<tree::Tree> objects made by AddTrees (411-444)
 * and calls Build(std::shared_ptr<io::DataSet>), Search(queries,
 * number_of_search, number_of_repeat), DumpFromFile / DumpToFile and the
 * per-type setters (src/tree/tree.h:24-48; MPHR src/tree/mphr.h:36-38; BVH
 * src/tree/bvh.h:39; RTree src/tree/rtree.h:1848; Hybrid src/tree/hybrid.h:
 * 55-67; RTree_LS src/tree/rtree_ls.h:50-54).  The classes below keep those
 * names, argument meanings and bool / int returns, and run every build and
 * search on the GPU through libqcp.so.  Differences, by design:
 *  - errors are returned (false / 0) with the status kept in GetLastStatus();
 *    nothing calls exit() or asserts, and no exception leaves these classes;
 *  - Search keeps what the reference only logged: the total hit count, the
 *    per-query counts and ID-set hashes of the last repeat, and its time;
 *  - one tree per object, no device globals, so several trees can coexist;
 *  - the degree is a runtime value (SetFanout) with the reference's
 *    compile-time degrees as defaults (src/common/config.h:6-8);
 *  - MPHR partitions (SetNumberOfPartition) are built and searched as the
 *    reference does (one subtree per contiguous rank range; the plain search
 *    launches one workgroup per partition per query), without the
 *    reference's requirement that the block count equal the partition count.
 */
#ifndef QCP_TREE_HPP
#define QCP_TREE_HPP

#include <stdint.h>
#include <stdio.h>

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "qcp.h"

namespace qcp {

typedef unsigned int ui;

namespace io {

/* Cluster types of the data set (src/common/types.h:58-63): the object order
 * a tree is built in. */
enum ClusterType {
  CLUSTER_TYPE_INVALID = -1,
  CLUSTER_TYPE_NONE = 1,           /* the given order (no keys, no sort)        */
  CLUSTER_TYPE_HILBERT = 2,        /* Hilbert order of the keys (default)       */
  CLUSTER_TYPE_KMEANSHILBERT = 3   /* k-means clusters, then Hilbert order      */
};

/* io::DataSet (src/io/dataset.h:17-47): rows of little-endian FP32, points
 * (D floats) or boxes (2D floats, [lo(D), hi(D)]); a query set is boxes. */
class DataSet {
 public:
  DataSet(ui number_of_dimensions, std::vector<float> values, bool boxes = false,
          ClusterType cluster_type = CLUSTER_TYPE_HILBERT)
      : dims_(number_of_dimensions), boxes_(boxes), cluster_(cluster_type),
        values_(std::move(values)) {}

  /* Reads number_of_data rows of a raw file (the reference's DataSet
   * constructor, src/io/dataset.cpp:10-40).  Returns nullptr when the file is
   * missing or shorter than asked (where the reference exits). */
  static std::shared_ptr<DataSet> Read(const std::string& path, ui number_of_dimensions,
                                       ui number_of_data, bool boxes = false,
                                       ClusterType cluster_type = CLUSTER_TYPE_HILBERT) {
    const size_t row = (size_t)number_of_dimensions * (boxes ? 2 : 1);
    std::vector<float> v(row * number_of_data);
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return nullptr;
    const size_t got = v.empty() ? 0 : fread(v.data(), sizeof(float), v.size(), f);
    fclose(f);
    if (got != v.size()) return nullptr;
    return std::make_shared<DataSet>(number_of_dimensions, std::move(v), boxes, cluster_type);
  }

  ui GetNumberOfDims() const { return dims_; }
  ui GetNumberOfData() const {
    const size_t row = (size_t)dims_ * (boxes_ ? 2 : 1);
    return row ? (ui)(values_.size() / row) : 0;
  }
  ClusterType GetClusterType() const { return cluster_; }
  bool IsBoxes() const { return boxes_; }
  /* row-major values (the reference's GetPoints returns a copy) */
  const std::vector<float>& GetPoints() const { return values_; }

 private:
  ui dims_;
  bool boxes_;
  ClusterType cluster_;
  std::vector<float> values_;
};

}  // namespace io

namespace tree {

/* src/common/types.h:45-53, plus this build's Hilbert-packed R-tree. */
enum TreeType {
  TREE_TYPE_INVALID = -1,
  TREE_TYPE_HYBRID = 1,
  TREE_TYPE_MPHR = 2,
  TREE_TYPE_MPHR_PARTITION = 3,
  TREE_TYPE_BVH = 4,
  TREE_TYPE_RTREE = 5,
  TREE_TYPE_RTREE_LS = 6,
  TREE_TYPE_PACKED_RTREE = 7
};

class Tree {
 public:
  virtual ~Tree() { qcp_index_destroy(index_); }
  Tree(const Tree&) = delete;
  Tree& operator=(const Tree&) = delete;

  /* Builds the index over the data set's objects in its cluster order. */
  virtual bool Build(std::shared_ptr<io::DataSet> input_data_set) {
    if (!input_data_set || input_data_set->GetNumberOfDims() < 2 ||
        input_data_set->GetNumberOfDims() > 3)
      return Fail(QCP_EINVAL);
    const io::ClusterType ct = input_data_set->GetClusterType();
    if (ct == io::CLUSTER_TYPE_INVALID) return Fail(QCP_EINVAL);
    qcp_build_cfg cfg = build_cfg_;
    cfg.dims = input_data_set->GetNumberOfDims();
    cfg.obj_kind = input_data_set->IsBoxes() ? QCP_OBJ_BOXES : QCP_OBJ_POINTS;
    cfg.order = ct == io::CLUSTER_TYPE_KMEANSHILBERT ? QCP_ORDER_KMEANS_HILBERT : QCP_ORDER_HILBERT;
    const std::vector<float>& v = input_data_set->GetPoints();
    const uint64_t n = input_data_set->GetNumberOfData();
    qcp_index* ix = nullptr;
    const bool given_order = input_order_ || ct == io::CLUSTER_TYPE_NONE;
    const qcp_status s = given_order ? qcp_build_ranked(&cfg, n ? v.data() : nullptr, n, &ix)
                                     : qcp_build(&cfg, n ? v.data() : nullptr, n, &ix);
    if (s != QCP_OK) return Fail(s);
    qcp_index_destroy(index_);
    index_ = ix;
    return Ok();
  }

  /* Loads an index written by DumpToFile (the reference's own dump format is
   * read by qcp_read_reference_mphr). */
  virtual bool DumpFromFile(std::string index_name) {
    qcp_index* ix = nullptr;
    const qcp_status s = qcp_load(index_name.c_str(), build_cfg_.device, &ix);
    if (s != QCP_OK) return Fail(s);
    qcp_index_destroy(index_);
    index_ = ix;
    return Ok();
  }

  virtual bool DumpToFile(std::string index_name) {
    if (!index_) return Fail(QCP_EINVAL);
    const qcp_status s = qcp_save(index_, index_name.c_str());
    return s == QCP_OK ? Ok() : Fail(s);
  }

  /* Searches the first number_of_search queries number_of_repeat times;
   * returns 1, or 0 on error.  The results of the last repeat are kept. */
  virtual int Search(std::shared_ptr<io::DataSet> query_data_set, ui number_of_search,
                     ui number_of_repeat) {
    if (!index_ || !query_data_set || !query_data_set->IsBoxes() ||
        number_of_search > query_data_set->GetNumberOfData() || number_of_repeat == 0)
      return Fail(QCP_EINVAL);
    qcp_index_info info;
    qcp_status s = qcp_index_info_get(index_, &info);
    if (s != QCP_OK) return Fail(s);
    if (info.dims != query_data_set->GetNumberOfDims()) return Fail(QCP_EINVAL);
    qcp_batch* b = nullptr;
    const std::vector<float>& q = query_data_set->GetPoints();
    s = qcp_batch_create(index_, number_of_search ? q.data() : nullptr, number_of_search, &b);
    if (s != QCP_OK) return Fail(s);
    qcp_search_stats st{};
    for (ui r = 0; r < number_of_repeat && s == QCP_OK; ++r) s = qcp_batch_run(b, &search_cfg_, &st);
    if (s == QCP_OK) {
      counts_.assign(number_of_search, 0);
      hashes_.assign(number_of_search, 0);
      s = qcp_batch_results(b, counts_.data(), hashes_.data(), nullptr, nullptr, 0, nullptr);
    }
    qcp_batch_destroy(b);
    if (s != QCP_OK) return Fail(s);
    total_hit_ = st.total_hits;
    search_ms_ = st.total_ms;
    return Ok();
  }

  TreeType GetTreeType() const { return tree_type; }

  /* ---- beyond the reference ------------------------------------------- */
  qcp_status GetLastStatus() const { return status_; }
  uint64_t GetTotalHit() const { return total_hit_; }
  const std::vector<uint32_t>& GetCounts() const { return counts_; }
  const std::vector<uint64_t>& GetHashes() const { return hashes_; }
  double GetSearchTimeMs() const { return search_ms_; }  /* device ms, last repeat */
  qcp_index* GetIndex() const { return index_; }
  void SetFanout(ui fanout) { build_cfg_.fanout = fanout; }
  void SetKeyMode(uint32_t key_mode) { build_cfg_.key_mode = key_mode; }
  void SetKeySource(uint32_t key_src) { build_cfg_.key_src = key_src; }
  void SetDevice(ui device) { build_cfg_.device = device; }
  /* QCP_SEARCH_PERSISTENT (default), QCP_SEARCH_PLAIN (MPHR), QCP_SEARCH_HYBRID */
  void SetSearchMode(uint32_t mode) { search_cfg_.mode = mode; }

 protected:
  Tree(TreeType type, uint32_t layout, ui fanout) : tree_type(type) {
    build_cfg_.dims = 3;
    build_cfg_.fanout = fanout;
    build_cfg_.layout = layout;
    build_cfg_.key_mode = QCP_KEY_REF;
    build_cfg_.key_src = QCP_SRC_LO;
    search_cfg_.mode = QCP_SEARCH_PERSISTENT;
    search_cfg_.outputs = QCP_OUT_COUNTS | QCP_OUT_HASH;
    search_cfg_.plain_batch = 128;
  }
  bool Fail(qcp_status s) {
    status_ = s;
    return false;
  }
  bool Ok() {
    status_ = QCP_OK;
    return true;
  }

  TreeType tree_type;
  qcp_build_cfg build_cfg_{};
  qcp_search_cfg search_cfg_{};
  bool input_order_ = false;

 private:
  qcp_index* index_ = nullptr;
  qcp_status status_ = QCP_OK;
  uint64_t total_hit_ = 0;
  double search_ms_ = 0.0;
  std::vector<uint32_t> counts_;
  std::vector<uint64_t> hashes_;
};

/* MPHR (src/tree/mphr.h): Hilbert-packed, degree 192. */
class MPHR : public Tree {
 public:
  MPHR() : Tree(TREE_TYPE_MPHR, QCP_LAYOUT_MPHR, 192) {}
  /* queries per launch of the plain (MPRS) form */
  void SetNumberOfCUDABlocks(ui number_of_cuda_blocks) {
    search_cfg_.plain_batch = number_of_cuda_blocks;
  }
  /* MPHR::SetNumberOfPartition (src/tree/mphr.cpp:279-285): P subtrees over
   * contiguous rank ranges (0 or 1 = one tree); Build refuses P > N. */
  void SetNumberOfPartition(ui number_of_partition) {
    build_cfg_.partitions = number_of_partition;
    if (number_of_partition > 1) tree_type = TREE_TYPE_MPHR_PARTITION;
  }
};

/* BVH (src/tree/bvh.h): rank-prefix splits, degree 128.  The reference
 * searches it on CPU threads; here the threads serve the co-processing form. */
class BVH : public Tree {
 public:
  BVH() : Tree(TREE_TYPE_BVH, QCP_LAYOUT_BVH, 128) {}
  void SetNumberOfCPUThreads(ui number_of_cpu_threads) {
    search_cfg_.cpu_threads = number_of_cpu_threads;
  }
};

/* RTree (src/tree/rtree.cpp): Guttman insertion of the sorted objects,
 * degree 128, depth-first entry IDs. */
class RTree : public Tree {
 public:
  RTree() : Tree(TREE_TYPE_RTREE, QCP_LAYOUT_GUTTMAN, 128) {}
  void SetNumberOfCPUThreads(ui number_of_cpu_threads) {
    search_cfg_.cpu_threads = number_of_cpu_threads;
  }
};

/* Hybrid (src/tree/hybrid.h): host threads walk the upper tree, the GPU
 * searches what they hand over (MPHR layout unless SetUpperTreeType). */
class Hybrid : public Tree {
 public:
  Hybrid() : Tree(TREE_TYPE_HYBRID, QCP_LAYOUT_MPHR, 192) { search_cfg_.mode = QCP_SEARCH_HYBRID; }
  void SetScanLevel(ui scan_level) { search_cfg_.handoff_level = scan_level; }
  void SetChunkSize(ui chunk_size) { search_cfg_.hybrid_batch = chunk_size; }
  void SetNumberOfCPUThreads(ui number_of_cpu_threads) {
    search_cfg_.cpu_threads = number_of_cpu_threads;
  }
  void SetNumberOfCUDABlocks(ui) {}  /* the GPU side is one persistent launch per block */
  void SetUpperTreeType(TreeType upper) {
    build_cfg_.layout = upper == TREE_TYPE_BVH     ? (uint32_t)QCP_LAYOUT_BVH
                        : upper == TREE_TYPE_RTREE ? (uint32_t)QCP_LAYOUT_GUTTMAN
                                                   : (uint32_t)QCP_LAYOUT_MPHR;
    build_cfg_.fanout = upper == TREE_TYPE_MPHR ? 192 : 128;
  }
};

/* RTree_LS (src/tree/rtree_ls.h): Guttman insertion in the given order, host
 * traversal down to the leaves, GPU leaf scans. */
class RTree_LS : public Tree {
 public:
  RTree_LS() : Tree(TREE_TYPE_RTREE_LS, QCP_LAYOUT_GUTTMAN, 128) {
    search_cfg_.mode = QCP_SEARCH_HYBRID;
    input_order_ = true;
  }
  void SetChunkSize(ui chunk_size) { search_cfg_.hybrid_batch = chunk_size; }
  void SetNumberOfCPUThreads(ui number_of_cpu_threads) {
    search_cfg_.cpu_threads = number_of_cpu_threads;
  }
  void SetNumberOfCUDABlocks(ui) {}
  void SetUpperTreeType(TreeType) {}  /* the upper tree is the Guttman tree's */
};

/* The Hilbert-packed R-tree (no counterpart in the reference). */
class PackedRTree : public Tree {
 public:
  PackedRTree() : Tree(TREE_TYPE_PACKED_RTREE, QCP_LAYOUT_RTREE, 128) {}
};

/* Evaluator::AddTrees (src/evaluator/evaluator.cpp:411-444) by index-type
 * name: m|mphr, r|rtree, b|bvh, h|hybrid, hr, hb, rls|rtree_ls, and
 * pr|packed_rtree; nullptr for anything else. */
inline std::shared_ptr<Tree> MakeTree(std::string index_type) {
  for (auto& c : index_type) c = (char)((c >= 'A' && c <= 'Z') ? c - 'A' + 'a' : c);
  if (index_type == "m" || index_type == "mphr") return std::make_shared<MPHR>();
  if (index_type == "r" || index_type == "rtree") return std::make_shared<RTree>();
  if (index_type == "b" || index_type == "bvh") return std::make_shared<BVH>();
  if (index_type == "pr" || index_type == "packed_rtree") return std::make_shared<PackedRTree>();
  if (index_type == "rls" || index_type == "rtree_ls") return std::make_shared<RTree_LS>();
  if (index_type == "h" || index_type == "hybrid" || index_type == "hr" || index_type == "hb") {
    auto t = std::make_shared<Hybrid>();
    if (index_type == "hr") t->SetUpperTreeType(TREE_TYPE_RTREE);
    if (index_type == "hb") t->SetUpperTreeType(TREE_TYPE_BVH);
    return t;
  }
  return nullptr;
}

}  // namespace tree
}  // namespace qcp

#endif  /* QCP_TREE_HPP */
