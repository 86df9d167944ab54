// pac_graph.hpp -- C++ drop-in facade over the C ABI (include/pac.h) with the
// reference's class interface for the intersection path: Graph (src/Graph.hpp:38-103)
// and Partitioned_Graph (src/Graph.hpp:105-155) of Byy-0024/PartitionAndConquer.
//
// Semantics kept from the reference: get_edge_num() returns the directed entry count
// 2m; *_galloping triangle counts return 6T; LCC is double with 0.0 for d=0 and NaN
// for d=1; Jaccard query results are n*n doubles.  The num_threads arguments are
// accepted and ignored (the work runs on the GPU).  Errors throw pac_ref::Error.
#pragma once
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pac.h"

namespace pac_ref {

typedef int VID;                 // Graph.hpp:23
typedef unsigned long long EID;  // Graph.hpp:24

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Process-wide context on device 0 (created on first use).
pac_ctx *default_context();

class Graph {
 public:
  explicit Graph(const std::string graphfile);                 // text edge list, Graph.cpp:27-90
  Graph(const std::string csrdir, bool is_weighted);           // binary CSR, Graph.cpp:92-137
  explicit Graph(const std::vector<std::vector<VID>> &adj_list);  // Graph.cpp:139-149
  Graph(pac_ctx *ctx, pac_graph *g);                           // adopt a device graph
  Graph(const Graph &) = delete;
  Graph &operator=(const Graph &) = delete;
  Graph(Graph &&o) noexcept;
  ~Graph();

  const EID *get_csr_vlist() const;
  const VID *get_csr_elist() const;
  EID get_edge_num() const;
  VID get_vertex_num() const;
  size_t get_deg(VID u) const;
  size_t get_max_deg() const;
  bool has_edge(VID u, VID v) const;
  size_t cnt_common_neis_merge(VID u, VID v) const;
  size_t cnt_common_neis_galloping(VID u, VID v) const;
  // neighbour-list queries (Graph.cpp:163-184, 241-253, 272-285, 405-411, 468-508)
  bool has_directed_edge(VID u, VID v) const;
  bool is_weighted() const { return false; }
  VID get_nei(VID u, EID offset) const;
  std::pair<VID, VID> get_edge(EID offset) const;
  unsigned long long size_in_bytes() const;
  void get_degs(std::vector<int> &degs) const;
  void get_neis(VID u, std::vector<VID> &neis) const;
  size_t cnt_common_neis_merge(VID u, const std::vector<VID> &nodes) const;
  size_t cnt_common_neis_less_than_merge(VID u, VID v, VID threshold) const;
  void get_common_neis(VID u, VID v, std::vector<VID> &res) const;  // appends, as the reference
  void get_common_neis(VID u, const std::vector<VID> &nodes, std::vector<VID> &res) const;
  size_t solve_triangle_counting_galloping(const int num_threads) const;
  size_t solve_triangle_counting_merge(const int num_threads) const;
  void solve_lcc_galloping(std::vector<double> &res, int num_threads) const;
  void solve_lcc_merge(std::vector<double> &res, int num_threads) const;
  void solve_jsim_galloping(const std::vector<VID> &queries, std::vector<double> &res,
                            int num_threads) const;
  void solve_jsim_merge(const std::vector<VID> &queries, std::vector<double> &res,
                        int num_threads) const;
  void batch_cnt_common_neis_merge(const std::vector<std::pair<VID, VID>> &queries,
                                   std::vector<size_t> &res, int num_threads) const;
  void batch_cnt_common_neis_galloping(const std::vector<std::pair<VID, VID>> &queries,
                                       std::vector<size_t> &res, int num_threads) const;
  // k-way intersections and friend triangles (Graph.cpp:200-225, 255-270, 309-338, 423-447)
  size_t cnt_common_neis_merge(const std::vector<VID> &q) const;
  size_t cnt_common_neis_galloping(const std::vector<VID> &q) const;
  void batch_cnt_common_neis_merge(const std::vector<std::vector<VID>> &queries,
                                   std::vector<size_t> &res, int num_threads) const;
  void batch_cnt_common_neis_galloping(const std::vector<std::vector<VID>> &queries,
                                       std::vector<size_t> &res, int num_threads) const;
  size_t solve_friend_triangle_merge(const VID root, const std::vector<VID> &nodes,
                                     const int num_threads) const;
  size_t solve_friend_triangle_galloping(const VID root, const std::vector<VID> &nodes,
                                         const int num_threads) const;
  void greedy_partition(const size_t partition_number, std::vector<int> &vertex2partition);
  void greedy_partition_2(const size_t partition_num, std::vector<int> &vertex2partition);
  void random_partition(const size_t partition_num, std::vector<int> &vertex2partition);
  void save_as_csr_binfile(const std::string csr_vlist_binfile,
                           const std::string csr_elist_binfile);

  // GPU-native extras
  uint64_t triangle_count(int dispatch, int launch) const;
  void per_edge_jaccard(std::vector<int32_t> &src, std::vector<int32_t> &dst,
                        std::vector<uint32_t> &inter, std::vector<uint32_t> &uni,
                        std::vector<float> &ratio, int dispatch, int launch) const;
  pac_graph *handle() const { return g_; }
  pac_ctx *context() const { return ctx_; }

 private:
  void ensure_host() const;
  void ensure_prepared() const;
  pac_ctx *ctx_ = nullptr;
  pac_graph *g_ = nullptr;
  mutable bool prepared_ = false;
  mutable std::vector<EID> vlist_;
  mutable std::vector<VID> elist_;
};

class Partitioned_Graph {
 public:
  // Graph.cpp:1017-1069: records the vertex->partition map (P = max+1) and builds
  // the device work queue for it on its own copy of the graph.
  Partitioned_Graph(const Graph &g, const std::vector<int> &vertex2partition,
                    const int thread_num);
  ~Partitioned_Graph();
  Partitioned_Graph(const Partitioned_Graph &) = delete;
  Partitioned_Graph &operator=(const Partitioned_Graph &) = delete;

  size_t get_vertex_num() const;
  size_t get_edge_num() const;
  size_t get_partition_num() const;
  size_t solve_triangle_counting_galloping() const;  // 6T, Graph.cpp:1326-1344
  size_t solve_triangle_counting_merge() const;      // T, Graph.cpp:1346-1362
  void solve_lcc_galloping(std::vector<double> &res) const;
  void solve_lcc_merge(std::vector<double> &res) const;
  void evaluate_partition() const;                   // Graph.cpp:1460-1473
  size_t solve_friend_triangle_merge(const VID root, const std::vector<VID> &nodes) const;
  size_t solve_friend_triangle_galloping(const VID root, const std::vector<VID> &nodes) const;
  size_t solve_square_counting() const;              // Graph.cpp:1291-1324
  // Per-partition intersections: only common neighbours in partition_id count
  // (Graph.cpp:1071-1195).
  size_t cnt_common_neis_galloping(const VID u, const VID v, const int partition_id) const;
  size_t cnt_common_neis_galloping(const std::vector<VID> &q, const int partition_id) const;
  size_t cnt_common_neis_less_than_merge(const VID u, const VID v, const VID threshold,
                                         const int partition_id) const;
  size_t cnt_common_neis_merge(const VID u, const VID v, const int partition_id) const;
  size_t cnt_common_neis_merge(const std::vector<VID> &q, const int partition_id) const;
  size_t cnt_common_neis_merge(const VID u, const std::vector<VID> &nodes,
                               const int partition_id) const;
  void get_common_neis(const VID u, const std::vector<VID> &nodes, const int partition_id,
                       std::vector<VID> &res) const;          // Graph.cpp:1475-1493
  void get_neis(const VID u, const int partition_id, std::vector<VID> &res) const;
  // Batched queries summed over all partitions (Graph.cpp:1382-1458).  As in the
  // reference, res is resized to the query count and the counts are added to it.
  void batch_cnt_common_neis_galloping(const std::vector<std::pair<VID, VID>> &queries,
                                       std::vector<size_t> &res) const;
  void batch_cnt_common_neis_galloping(const std::vector<std::vector<VID>> &queries,
                                       std::vector<size_t> &res) const;
  void batch_cnt_common_neis_galloping_fine_grained(
      const std::vector<std::pair<VID, VID>> &queries, std::vector<size_t> &res) const;
  void batch_cnt_common_neis_merge(const std::vector<std::pair<VID, VID>> &queries,
                                   std::vector<size_t> &res) const;
  void batch_cnt_common_neis_merge(const std::vector<std::vector<VID>> &queries,
                                   std::vector<size_t> &res) const;
  void batch_cnt_common_neis_merge_fine_grained(const std::vector<std::pair<VID, VID>> &queries,
                                                std::vector<size_t> &res) const;
  // Jaccard query matrix with the Graph:: formula (Graph.cpp:798-812); the partitioned
  // reference's mis-indexed union (Graph.cpp:2157, 2192) is not reproduced.
  void solve_jsim_galloping(const std::vector<VID> &queries, std::vector<double> &res) const;
  void solve_jsim_merge(const std::vector<VID> &queries, std::vector<double> &res) const;
  uint64_t triangle_count(int dispatch, int launch) const;
  void per_edge_jaccard(std::vector<int32_t> &src, std::vector<int32_t> &dst,
                        std::vector<uint32_t> &inter, std::vector<uint32_t> &uni,
                        std::vector<float> &ratio, int dispatch, int launch) const;
  pac_graph *handle() const { return g_; }

 private:
  pac_ctx *ctx_ = nullptr;
  pac_graph *g_ = nullptr;
  std::vector<int> v2p_;
  size_t partition_num_ = 0;
};

}  // namespace pac_ref
