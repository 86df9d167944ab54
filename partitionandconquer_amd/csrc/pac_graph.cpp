// pac_graph.cpp -- the reference-compatible C++ facade (see pac_graph.hpp).
#include "pac_graph.hpp"

#include <algorithm>
#include <cstdio>
#include <mutex>

namespace pac_ref {

namespace {

void check(int rc) {
  if (rc != PAC_OK) throw Error(rc, pac_last_error());
}

pac_ctx *g_ctx = nullptr;
std::once_flag g_once;

}  // namespace

pac_ctx *default_context() {
  std::call_once(g_once, [] { check(pac_ctx_create(0, &g_ctx)); });
  return g_ctx;
}

Graph::Graph(const std::string graphfile) : ctx_(default_context()) {
  check(pac_graph_load_text(ctx_, graphfile.c_str(), &g_));
  pac_graph_info i;
  check(pac_graph_get_info(g_, &i));
  printf("Vertex num: %llu.\n", (unsigned long long)i.num_vertices);
}

Graph::Graph(const std::string csrdir, bool is_weighted) : ctx_(default_context()) {
  (void)is_weighted;  // edge weights are only used by SSSP, outside this path
  check(pac_graph_load_csr_dir(ctx_, csrdir.c_str(), &g_));
  printf("Has read graph! #Vertices: %d, #Edges: %llu, #Weights: 0.\n", get_vertex_num(),
         get_edge_num());
}

Graph::Graph(const std::vector<std::vector<VID>> &adj_list) : ctx_(default_context()) {
  std::vector<EID> vl(1, 0);
  std::vector<VID> el;
  for (const auto &row : adj_list) {
    el.insert(el.end(), row.begin(), row.end());
    vl.push_back(el.size());
  }
  check(pac_graph_upload(ctx_, reinterpret_cast<const uint64_t *>(vl.data()), adj_list.size(),
                         el.data(), el.size(), &g_));
}

Graph::Graph(pac_ctx *ctx, pac_graph *g) : ctx_(ctx), g_(g) {}

Graph::Graph(Graph &&o) noexcept
    : ctx_(o.ctx_), g_(o.g_), prepared_(o.prepared_), vlist_(std::move(o.vlist_)),
      elist_(std::move(o.elist_)) {
  o.g_ = nullptr;
}

Graph::~Graph() {
  if (g_) pac_graph_free(g_);
}

void Graph::ensure_host() const {
  if (!vlist_.empty()) return;
  pac_graph_info i;
  check(pac_graph_get_info(g_, &i));
  vlist_.resize(i.num_vertices + 1);
  elist_.resize(i.num_entries);
  check(pac_graph_download_csr(g_, reinterpret_cast<uint64_t *>(vlist_.data()),
                               elist_.data()));
}

void Graph::ensure_prepared() const {
  if (prepared_) return;
  check(pac_graph_prepare(g_, 1, nullptr));
  prepared_ = true;
}

const EID *Graph::get_csr_vlist() const {
  ensure_host();
  return vlist_.data();
}

const VID *Graph::get_csr_elist() const {
  ensure_host();
  return elist_.data();
}

EID Graph::get_edge_num() const {
  pac_graph_info i;
  check(pac_graph_get_info(g_, &i));
  return i.num_entries;
}

VID Graph::get_vertex_num() const {
  pac_graph_info i;
  check(pac_graph_get_info(g_, &i));
  return (VID)i.num_vertices;
}

size_t Graph::get_deg(VID u) const {
  ensure_host();
  return vlist_[u + 1] - vlist_[u];
}

size_t Graph::get_max_deg() const {
  pac_graph_info i;
  check(pac_graph_get_info(g_, &i));
  return i.max_degree;
}

bool Graph::has_edge(VID u, VID v) const {
  ensure_host();
  const VID *b = elist_.data() + vlist_[u], *e = elist_.data() + vlist_[u + 1];
  return std::binary_search(b, e, v);
}

size_t Graph::cnt_common_neis_merge(VID u, VID v) const {
  uint64_t r = 0;
  check(pac_cnt_common_pairs(g_, &u, &v, 1, &r));
  return r;
}

size_t Graph::cnt_common_neis_galloping(VID u, VID v) const { return cnt_common_neis_merge(u, v); }

bool Graph::has_directed_edge(VID u, VID v) const {
  ensure_host();
  const VID *b = elist_.data() + vlist_[u], *e = elist_.data() + vlist_[u + 1];
  return std::binary_search(b, e, v);
}

VID Graph::get_nei(VID u, EID offset) const {
  ensure_host();
  return elist_[vlist_[u] + offset];
}

std::pair<VID, VID> Graph::get_edge(EID offset) const {
  ensure_host();
  const VID u = (VID)(std::upper_bound(vlist_.begin(), vlist_.end(), offset) - vlist_.begin()) - 1;
  return {u, elist_[offset]};
}

unsigned long long Graph::size_in_bytes() const {
  // Graph.cpp:409-411 counts the offsets at sizeof(int)
  return (unsigned long long)(get_vertex_num() + 1) * sizeof(int) +
         (unsigned long long)get_edge_num() * sizeof(VID);
}

void Graph::get_degs(std::vector<int> &degs) const {
  ensure_host();
  degs.resize(vlist_.size() - 1);
  for (size_t u = 0; u + 1 < vlist_.size(); ++u) degs[u] = (int)(vlist_[u + 1] - vlist_[u]);
}

void Graph::get_neis(VID u, std::vector<VID> &neis) const {
  ensure_host();
  neis.assign(elist_.begin() + vlist_[u], elist_.begin() + vlist_[u + 1]);
}

namespace {
// pac_common_neis into a vector (count first, then the list)
std::vector<VID> common_list(pac_graph *g, VID u, VID v, const std::vector<VID> *nodes,
                             int partition) {
  const int32_t *nd = nodes ? nodes->data() : nullptr;
  const uint64_t n = nodes ? nodes->size() : 0;
  uint64_t c = 0;
  check(pac_common_neis(g, u, v, nd, n, partition, nullptr, 0, &c));
  std::vector<VID> out(c);
  if (c) check(pac_common_neis(g, u, v, nd, n, partition, out.data(), c, &c));
  return out;
}
}  // namespace

size_t Graph::cnt_common_neis_merge(VID u, const std::vector<VID> &nodes) const {
  uint64_t c = 0;
  check(pac_common_neis(g_, u, -1, nodes.data(), nodes.size(), -1, nullptr, 0, &c));
  return c;
}

size_t Graph::cnt_common_neis_less_than_merge(VID u, VID v, VID threshold) const {
  uint64_t r = 0;
  check(pac_cnt_common_pairs_filtered(g_, &u, &v, &threshold, 1, -1, &r));
  return r;
}

void Graph::get_common_neis(VID u, VID v, std::vector<VID> &res) const {
  const std::vector<VID> c = common_list(g_, u, v, nullptr, -1);
  res.insert(res.end(), c.begin(), c.end());
}

void Graph::get_common_neis(VID u, const std::vector<VID> &nodes, std::vector<VID> &res) const {
  res = common_list(g_, u, -1, &nodes, -1);
}

uint64_t Graph::triangle_count(int dispatch, int launch) const {
  ensure_prepared();
  uint64_t t = 0;
  check(pac_tc(g_, dispatch, launch, &t, nullptr));
  return t;
}

size_t Graph::solve_triangle_counting_galloping(const int) const {
  return 6 * triangle_count(PAC_FORCE_GALLOP, PAC_LAUNCH_PERSISTENT);
}

size_t Graph::solve_triangle_counting_merge(const int) const {
  return triangle_count(PAC_DISPATCH_FIXED, PAC_LAUNCH_PERSISTENT);
}

void Graph::solve_lcc_galloping(std::vector<double> &res, int) const {
  ensure_prepared();
  res.assign(get_vertex_num(), 0.0);
  check(pac_lcc(g_, PAC_FORCE_GALLOP, PAC_LAUNCH_PERSISTENT, nullptr, res.data(), nullptr));
}

void Graph::solve_lcc_merge(std::vector<double> &res, int) const {
  ensure_prepared();
  res.assign(get_vertex_num(), 0.0);
  check(pac_lcc(g_, PAC_DISPATCH_FIXED, PAC_LAUNCH_PERSISTENT, nullptr, res.data(), nullptr));
}

void Graph::solve_jsim_merge(const std::vector<VID> &q, std::vector<double> &res, int) const {
  res.assign(q.size() * q.size(), 0.0);
  check(pac_jaccard_queries(g_, q.data(), (uint32_t)q.size(), res.data()));
}

void Graph::solve_jsim_galloping(const std::vector<VID> &q, std::vector<double> &res,
                                 int nt) const {
  solve_jsim_merge(q, res, nt);
}

void Graph::batch_cnt_common_neis_merge(const std::vector<std::pair<VID, VID>> &queries,
                                        std::vector<size_t> &res, int) const {
  std::vector<int32_t> u(queries.size()), v(queries.size());
  for (size_t i = 0; i < queries.size(); ++i) {
    u[i] = queries[i].first;
    v[i] = queries[i].second;
  }
  std::vector<uint64_t> r(queries.size());
  check(pac_cnt_common_pairs(g_, u.data(), v.data(), u.size(), r.data()));
  res.assign(r.begin(), r.end());
}

void Graph::batch_cnt_common_neis_galloping(const std::vector<std::pair<VID, VID>> &queries,
                                            std::vector<size_t> &res, int nt) const {
  batch_cnt_common_neis_merge(queries, res, nt);
}

void Graph::batch_cnt_common_neis_merge(const std::vector<std::vector<VID>> &queries,
                                        std::vector<size_t> &res, int) const {
  std::vector<uint64_t> qoff(queries.size() + 1, 0);
  std::vector<int32_t> qv;
  for (size_t i = 0; i < queries.size(); ++i) {
    qv.insert(qv.end(), queries[i].begin(), queries[i].end());
    qoff[i + 1] = qv.size();
  }
  std::vector<uint64_t> r(queries.size());
  check(pac_cnt_common_kway(g_, qoff.data(), qv.data(), queries.size(), r.data()));
  res.assign(r.begin(), r.end());
}

void Graph::batch_cnt_common_neis_galloping(const std::vector<std::vector<VID>> &queries,
                                            std::vector<size_t> &res, int nt) const {
  batch_cnt_common_neis_merge(queries, res, nt);
}

size_t Graph::cnt_common_neis_merge(const std::vector<VID> &q) const {
  std::vector<size_t> r;
  batch_cnt_common_neis_merge(std::vector<std::vector<VID>>{q}, r, 1);
  return r[0];
}

size_t Graph::cnt_common_neis_galloping(const std::vector<VID> &q) const {
  return cnt_common_neis_merge(q);
}

size_t Graph::solve_friend_triangle_merge(const VID root, const std::vector<VID> &nodes,
                                          const int) const {
  uint64_t r = 0;
  check(pac_friend_triangle(g_, root, nodes.data(), nodes.size(), &r));
  return r;
}

size_t Graph::solve_friend_triangle_galloping(const VID root, const std::vector<VID> &nodes,
                                              const int nt) const {
  return solve_friend_triangle_merge(root, nodes, nt);
}

void Graph::greedy_partition(const size_t P, std::vector<int> &v2p) {
  v2p.assign(get_vertex_num(), 0);
  check(pac_greedy_partition(g_, (uint32_t)P, 1, v2p.data()));
}

// Graph::greedy_partition_2, Graph.cpp:969-1015 (seeded; see pac_greedy_partition_2).
void Graph::greedy_partition_2(const size_t P, std::vector<int> &v2p) {
  v2p.assign(get_vertex_num(), -1);
  check(pac_greedy_partition_2(g_, (uint32_t)P, 1, v2p.data()));
}

void Graph::random_partition(const size_t P, std::vector<int> &v2p) {
  v2p.assign(get_vertex_num(), 0);
  check(pac_random_partition(get_vertex_num(), (uint32_t)P, 1, v2p.data()));
}

void Graph::save_as_csr_binfile(const std::string vfile, const std::string efile) {
  ensure_host();
  FILE *f = fopen(vfile.c_str(), "wb");
  if (!f) throw Error(PAC_ERR_IO, "cannot create " + vfile);
  fwrite(vlist_.data(), sizeof(EID), vlist_.size(), f);
  fclose(f);
  printf("Size of csr_vlist: %zu.\n", vlist_.size());
  f = fopen(efile.c_str(), "wb");
  if (!f) throw Error(PAC_ERR_IO, "cannot create " + efile);
  fwrite(elist_.data(), sizeof(VID), elist_.size(), f);
  fclose(f);
  printf("Size of csr_elist: %zu.\n", elist_.size());
}

namespace {
void edge_jaccard(pac_graph *g_, std::vector<int32_t> &src, std::vector<int32_t> &dst,
                  std::vector<uint32_t> &inter, std::vector<uint32_t> &uni,
                  std::vector<float> &ratio, int dispatch, int launch) {
  pac_graph_info i;
  check(pac_graph_get_info(g_, &i));
  src.resize(i.num_edges);
  dst.resize(i.num_edges);
  inter.resize(i.num_edges);
  uni.resize(i.num_edges);
  ratio.resize(i.num_edges);
  check(pac_jaccard_edges(g_, dispatch, launch, src.data(), dst.data(), inter.data(), uni.data(),
                          ratio.data()));
}
}  // namespace

void Graph::per_edge_jaccard(std::vector<int32_t> &src, std::vector<int32_t> &dst,
                             std::vector<uint32_t> &inter, std::vector<uint32_t> &uni,
                             std::vector<float> &ratio, int dispatch, int launch) const {
  ensure_prepared();
  edge_jaccard(g_, src, dst, inter, uni, ratio, dispatch, launch);
}

void Partitioned_Graph::per_edge_jaccard(std::vector<int32_t> &src, std::vector<int32_t> &dst,
                                         std::vector<uint32_t> &inter, std::vector<uint32_t> &uni,
                                         std::vector<float> &ratio, int dispatch,
                                         int launch) const {
  edge_jaccard(g_, src, dst, inter, uni, ratio, dispatch, launch);
}

// ---- Partitioned_Graph -------------------------------------------------------------
Partitioned_Graph::Partitioned_Graph(const Graph &g, const std::vector<int> &v2p, const int)
    : ctx_(g.context()), v2p_(v2p) {
  const EID *vl = g.get_csr_vlist();
  const VID *el = g.get_csr_elist();
  check(pac_graph_upload(ctx_, reinterpret_cast<const uint64_t *>(vl), g.get_vertex_num(), el,
                         g.get_edge_num(), &g_));
  int pmax = 0;
  for (int p : v2p_) pmax = std::max(pmax, p);
  partition_num_ = (size_t)pmax + 1;
  check(pac_graph_prepare(g_, (uint32_t)partition_num_, v2p_.empty() ? nullptr : v2p_.data()));
  printf("Partitioned Graph Constructed!\n");
}

Partitioned_Graph::~Partitioned_Graph() {
  if (g_) pac_graph_free(g_);
}

size_t Partitioned_Graph::get_vertex_num() const {
  pac_graph_info i;
  check(pac_graph_get_info(g_, &i));
  return i.num_vertices;
}

size_t Partitioned_Graph::get_edge_num() const {
  pac_graph_info i;
  check(pac_graph_get_info(g_, &i));
  return i.num_entries;
}

size_t Partitioned_Graph::get_partition_num() const { return partition_num_; }

uint64_t Partitioned_Graph::triangle_count(int dispatch, int launch) const {
  uint64_t t = 0;
  check(pac_tc(g_, dispatch, launch, &t, nullptr));
  return t;
}

size_t Partitioned_Graph::solve_triangle_counting_galloping() const {
  return 6 * triangle_count(PAC_FORCE_GALLOP, PAC_LAUNCH_PERSISTENT);
}

size_t Partitioned_Graph::solve_triangle_counting_merge() const {
  return triangle_count(PAC_DISPATCH_FIXED, PAC_LAUNCH_PERSISTENT);
}

void Partitioned_Graph::solve_lcc_galloping(std::vector<double> &res) const {
  res.assign(get_vertex_num(), 0.0);
  check(pac_lcc(g_, PAC_FORCE_GALLOP, PAC_LAUNCH_PERSISTENT, nullptr, res.data(), nullptr));
}

void Partitioned_Graph::solve_lcc_merge(std::vector<double> &res) const {
  res.assign(get_vertex_num(), 0.0);
  check(pac_lcc(g_, PAC_DISPATCH_FIXED, PAC_LAUNCH_PERSISTENT, nullptr, res.data(), nullptr));
}

void Partitioned_Graph::evaluate_partition() const {
  double ratio = 0;
  check(pac_evaluate_partition(g_, v2p_.data(), (uint32_t)partition_num_, &ratio));
  printf("balance ratio: %f\n", ratio);
}

// The per-partition partials of the reference sum to the unpartitioned quantity.
size_t Partitioned_Graph::solve_friend_triangle_merge(const VID root,
                                                      const std::vector<VID> &nodes) const {
  uint64_t r = 0;
  check(pac_friend_triangle(g_, root, nodes.data(), nodes.size(), &r));
  return r;
}

size_t Partitioned_Graph::solve_friend_triangle_galloping(const VID root,
                                                          const std::vector<VID> &nodes) const {
  return solve_friend_triangle_merge(root, nodes);
}

size_t Partitioned_Graph::solve_square_counting() const {
  uint64_t r = 0;
  check(pac_square_count(g_, &r));
  return r;
}

size_t Partitioned_Graph::cnt_common_neis_merge(const VID u, const VID v,
                                                const int partition_id) const {
  uint64_t r = 0;
  check(pac_cnt_common_pairs_filtered(g_, &u, &v, nullptr, 1, partition_id, &r));
  return r;
}

size_t Partitioned_Graph::cnt_common_neis_galloping(const VID u, const VID v,
                                                    const int partition_id) const {
  return cnt_common_neis_merge(u, v, partition_id);
}

size_t Partitioned_Graph::cnt_common_neis_less_than_merge(const VID u, const VID v,
                                                          const VID threshold,
                                                          const int partition_id) const {
  uint64_t r = 0;
  check(pac_cnt_common_pairs_filtered(g_, &u, &v, &threshold, 1, partition_id, &r));
  return r;
}

size_t Partitioned_Graph::cnt_common_neis_merge(const std::vector<VID> &q,
                                                const int partition_id) const {
  const uint64_t qoff[2] = {0, q.size()};
  uint64_t r = 0;
  check(pac_cnt_common_kway_filtered(g_, qoff, q.data(), 1, partition_id, &r));
  return r;
}

size_t Partitioned_Graph::cnt_common_neis_galloping(const std::vector<VID> &q,
                                                    const int partition_id) const {
  return cnt_common_neis_merge(q, partition_id);
}

size_t Partitioned_Graph::cnt_common_neis_merge(const VID u, const std::vector<VID> &nodes,
                                                const int partition_id) const {
  uint64_t c = 0;
  check(pac_common_neis(g_, u, -1, nodes.data(), nodes.size(), partition_id, nullptr, 0, &c));
  return c;
}

void Partitioned_Graph::get_common_neis(const VID u, const std::vector<VID> &nodes,
                                        const int partition_id, std::vector<VID> &res) const {
  res = common_list(g_, u, -1, &nodes, partition_id);
}

void Partitioned_Graph::get_neis(const VID u, const int partition_id,
                                 std::vector<VID> &res) const {
  res = common_list(g_, u, u, nullptr, partition_id);  // N(u) & N(u) = N(u)
}

namespace {
void add_counts(std::vector<size_t> &res, const std::vector<uint64_t> &r) {
  res.resize(r.size(), 0);
  for (size_t i = 0; i < r.size(); ++i) res[i] += r[i];
}
}  // namespace

void Partitioned_Graph::batch_cnt_common_neis_merge(
    const std::vector<std::pair<VID, VID>> &queries, std::vector<size_t> &res) const {
  std::vector<int32_t> u(queries.size()), v(queries.size());
  for (size_t i = 0; i < queries.size(); ++i) {
    u[i] = queries[i].first;
    v[i] = queries[i].second;
  }
  std::vector<uint64_t> r(queries.size());
  check(pac_cnt_common_pairs(g_, u.data(), v.data(), u.size(), r.data()));
  add_counts(res, r);
}

void Partitioned_Graph::batch_cnt_common_neis_merge(const std::vector<std::vector<VID>> &queries,
                                                    std::vector<size_t> &res) const {
  std::vector<uint64_t> qoff(queries.size() + 1, 0);
  std::vector<int32_t> qv;
  for (size_t i = 0; i < queries.size(); ++i) {
    qv.insert(qv.end(), queries[i].begin(), queries[i].end());
    qoff[i + 1] = qv.size();
  }
  std::vector<uint64_t> r(queries.size());
  check(pac_cnt_common_kway(g_, qoff.data(), qv.data(), queries.size(), r.data()));
  add_counts(res, r);
}

void Partitioned_Graph::batch_cnt_common_neis_galloping(
    const std::vector<std::pair<VID, VID>> &queries, std::vector<size_t> &res) const {
  batch_cnt_common_neis_merge(queries, res);
}

void Partitioned_Graph::batch_cnt_common_neis_galloping(
    const std::vector<std::vector<VID>> &queries, std::vector<size_t> &res) const {
  batch_cnt_common_neis_merge(queries, res);
}

void Partitioned_Graph::batch_cnt_common_neis_merge_fine_grained(
    const std::vector<std::pair<VID, VID>> &queries, std::vector<size_t> &res) const {
  batch_cnt_common_neis_merge(queries, res);
}

void Partitioned_Graph::batch_cnt_common_neis_galloping_fine_grained(
    const std::vector<std::pair<VID, VID>> &queries, std::vector<size_t> &res) const {
  batch_cnt_common_neis_merge(queries, res);
}

void Partitioned_Graph::solve_jsim_merge(const std::vector<VID> &q,
                                         std::vector<double> &res) const {
  res.assign(q.size() * q.size(), 0.0);
  check(pac_jaccard_queries(g_, q.data(), (uint32_t)q.size(), res.data()));
}

void Partitioned_Graph::solve_jsim_galloping(const std::vector<VID> &q,
                                             std::vector<double> &res) const {
  solve_jsim_merge(q, res);
}

}  // namespace pac_ref
