// pac_oracle.cpp -- CPU/OpenMP restatement of the reference's intersection path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product
// library never does.  Each function names the reference function (file:line,
// Byy-0024/PartitionAndConquer src/) whose loop structure it follows.  The
// loops are written fresh; where the reference has undefined behaviour (the
// out-of-bounds read of Partitioned_Graph::cnt_common_neis_less_than_merge,
// Graph.cpp:1136-1149) the restatement checks bounds first, which leaves every
// result unchanged.
#include "pac_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

namespace {

inline uint64_t deg(const uint64_t *vl, int64_t u) { return vl[u + 1] - vl[u]; }

// Graph::cnt_common_neis_merge, Graph.cpp:227-239: two-pointer merge.
inline uint64_t merge_count(const int32_t *a, uint64_t na, const int32_t *b, uint64_t nb) {
  uint64_t i = 0, j = 0, res = 0;
  while (i < na && j < nb) {
    if (a[i] == b[j]) {
      ++res;
      ++i;
      ++j;
    } else if (a[i] < b[j]) {
      ++i;
    } else {
      ++j;
    }
  }
  return res;
}

// Graph::cnt_common_neis_less_than_merge, Graph.cpp:272-285: merge that stops as
// soon as either head reaches the threshold.
inline uint64_t merge_count_below(const int32_t *a, uint64_t na, const int32_t *b,
                                  uint64_t nb, int32_t thr) {
  uint64_t i = 0, j = 0, res = 0;
  while (i < na && j < nb && a[i] < thr && b[j] < thr) {
    if (a[i] == b[j]) {
      ++res;
      ++i;
      ++j;
    } else if (a[i] < b[j]) {
      ++i;
    } else {
      ++j;
    }
  }
  return res;
}

// Graph::cnt_common_neis_galloping, Graph.cpp:186-198: every element of the
// smaller list is looked up with lower_bound over the whole larger list.
inline uint64_t gallop_count(const int32_t *a, uint64_t na, const int32_t *b, uint64_t nb) {
  if (na > nb) {
    std::swap(a, b);
    std::swap(na, nb);
  }
  uint64_t res = 0;
  for (uint64_t i = 0; i < na; ++i) {
    const int32_t t = a[i];
    const int32_t *p = std::lower_bound(b, b + nb, t);
    if (p < b + nb && *p == t) ++res;
  }
  return res;
}

// SplitMix64 (Steele, Lea, Flood 2014): the finaliser and the golden gamma.
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;
inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int csr_from_keys(uint64_t V, std::vector<uint64_t> &keys, uint64_t **vlist,
                  int32_t **elist, uint64_t *nnz) {
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  uint64_t *vl = (uint64_t *)malloc(sizeof(uint64_t) * (V + 1));
  int32_t *el = (int32_t *)malloc(sizeof(int32_t) * (keys.empty() ? 1 : keys.size()));
  if (!vl || !el) {
    free(vl);
    free(el);
    return -2;
  }
  std::fill(vl, vl + V + 1, 0);
  for (uint64_t k : keys) vl[(k >> 32) + 1]++;
  for (uint64_t u = 0; u < V; ++u) vl[u + 1] += vl[u];
  for (size_t i = 0; i < keys.size(); ++i) el[i] = (int32_t)(uint32_t)(keys[i] & 0xffffffffu);
  *vlist = vl;
  *elist = el;
  *nnz = keys.size();
  return 0;
}

}  // namespace

extern "C" {

void orc_free(void *p) { free(p); }

// Graph::Graph(const string graphfile), Graph.cpp:27-90.  Same parse: lines
// starting with '#' or '%' are skipped, the first number ends at the first
// non-digit, exactly one character is skipped, the second number follows; a 0
// anywhere (including in self-loops and in lines with no second number) marks the
// file as 0-based; self-loops are dropped before V is updated; without a 0 every
// ID is shifted down by one; rows are sorted and deduplicated.
int orc_text_to_csr(const char *path, uint64_t *V, uint64_t **vlist, int32_t **elist,
                    uint64_t *nnz) {
  FILE *fp = fopen(path, "r");
  if (!fp) return -1;
  bool has_zero = false;
  uint64_t vertex_num = 0;
  std::vector<std::pair<int32_t, int32_t>> edges;
  char line[512];
  while (fgets(line, 512, fp) != NULL) {
    if (line[0] == '#' || line[0] == '%') continue;
    int32_t u = 0, v = 0;
    const char *c = line;
    while (isdigit((unsigned char)*c)) u = u * 10 + (*c++ - '0');
    c++;
    while (isdigit((unsigned char)*c)) v = v * 10 + (*c++ - '0');
    if (u == 0 || v == 0) has_zero = true;
    if (u == v) continue;
    if ((uint64_t)u > vertex_num) vertex_num = (uint64_t)u;
    if ((uint64_t)v > vertex_num) vertex_num = (uint64_t)v;
    edges.push_back({u, v});
  }
  fclose(fp);
  if (has_zero) vertex_num++;
  std::vector<std::vector<int32_t>> adj(vertex_num);
  for (auto &e : edges) {
    int32_t a = has_zero ? e.first : e.first - 1;
    int32_t b = has_zero ? e.second : e.second - 1;
    adj[a].push_back(b);
    adj[b].push_back(a);
  }
  uint64_t total = 0;
  for (auto &row : adj) {
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    total += row.size();
  }
  uint64_t *vl = (uint64_t *)malloc(sizeof(uint64_t) * (vertex_num + 1));
  int32_t *el = (int32_t *)malloc(sizeof(int32_t) * (total ? total : 1));
  uint64_t pos = 0;
  for (uint64_t u = 0; u < vertex_num; ++u) {
    vl[u] = pos;
    for (int32_t x : adj[u]) el[pos++] = x;
  }
  vl[vertex_num] = pos;
  *V = vertex_num;
  *vlist = vl;
  *elist = el;
  *nnz = total;
  return 0;
}

int orc_edges_to_csr(uint64_t V, const uint32_t *src, const uint32_t *dst, uint64_t ne,
                     uint64_t **vlist, int32_t **elist, uint64_t *nnz) {
  std::vector<uint64_t> keys;
  keys.reserve(2 * ne);
  for (uint64_t i = 0; i < ne; ++i) {
    if (src[i] == dst[i]) continue;
    if (src[i] >= V || dst[i] >= V) return -1;
    keys.push_back(((uint64_t)src[i] << 32) | dst[i]);
    keys.push_back(((uint64_t)dst[i] << 32) | src[i]);
  }
  return csr_from_keys(V, keys, vlist, elist, nnz);
}

uint64_t orc_cnt_merge(const uint64_t *vl, const int32_t *el, int32_t u, int32_t v) {
  return merge_count(el + vl[u], deg(vl, u), el + vl[v], deg(vl, v));
}

uint64_t orc_cnt_less_than_merge(const uint64_t *vl, const int32_t *el, int32_t u,
                                 int32_t v, int32_t thr) {
  return merge_count_below(el + vl[u], deg(vl, u), el + vl[v], deg(vl, v), thr);
}

uint64_t orc_cnt_galloping(const uint64_t *vl, const int32_t *el, int32_t u, int32_t v) {
  return gallop_count(el + vl[u], deg(vl, u), el + vl[v], deg(vl, v));
}

// Graph::solve_triangle_counting_merge, Graph.cpp:355-368: for every u and every
// neighbour v < u, count common neighbours below v.  Each triangle w<v<u is
// counted once.
uint64_t orc_tc_merge(const uint64_t *vl, const int32_t *el, uint64_t V, int nt) {
  uint64_t res = 0;
#pragma omp parallel for num_threads(nt) reduction(+ : res)
  for (int64_t u = 0; u < (int64_t)V; ++u) {
    for (uint64_t j = vl[u]; j < vl[u + 1]; ++j) {
      const int32_t v = el[j];
      if (u <= v) continue;
      res += orc_cnt_less_than_merge(vl, el, (int32_t)u, v, v);
    }
  }
  return res;
}

// Graph::solve_triangle_counting_galloping, Graph.cpp:342-353: every directed
// edge, full intersection, no orientation -> 6T.
uint64_t orc_tc_galloping(const uint64_t *vl, const int32_t *el, uint64_t V, int nt) {
  uint64_t res = 0;
#pragma omp parallel for num_threads(nt) reduction(+ : res)
  for (int64_t u = 0; u < (int64_t)V; ++u) {
    for (uint64_t j = vl[u]; j < vl[u + 1]; ++j) res += orc_cnt_galloping(vl, el, (int32_t)u, el[j]);
  }
  return res;
}

uint64_t orc_tc_merge_rows(const uint64_t *vl, const int32_t *el, const int32_t *rows,
                           uint64_t nrows, int nt) {
  uint64_t res = 0;
#pragma omp parallel for num_threads(nt) reduction(+ : res)
  for (int64_t i = 0; i < (int64_t)nrows; ++i) {
    const int32_t u = rows[i];
    for (uint64_t j = vl[u]; j < vl[u + 1]; ++j) {
      const int32_t v = el[j];
      if (u <= v) continue;
      res += orc_cnt_less_than_merge(vl, el, u, v, v);
    }
  }
  return res;
}

// Graph::solve_lcc_merge, Graph.cpp:828-840.  d=0 keeps 0.0 (the `continue`),
// d=1 gives 0.0/0.0 = NaN, exactly as the reference.
void orc_lcc_merge(const uint64_t *vl, const int32_t *el, uint64_t V, double *res, int nt) {
#pragma omp parallel for num_threads(nt)
  for (int64_t u = 0; u < (int64_t)V; ++u) {
    res[u] = 0.0;
    const uint64_t d = deg(vl, u);
    if (d < 1) continue;
    for (uint64_t i = vl[u]; i < vl[u + 1]; ++i) res[u] += (double)orc_cnt_merge(vl, el, (int32_t)u, el[i]);
    res[u] /= (double)(d * (d - 1));
  }
}

// Graph::solve_lcc_galloping, Graph.cpp:814-826.
void orc_lcc_galloping(const uint64_t *vl, const int32_t *el, uint64_t V, double *res,
                       int nt) {
#pragma omp parallel for num_threads(nt)
  for (int64_t u = 0; u < (int64_t)V; ++u) {
    res[u] = 0.0;
    const uint64_t d = deg(vl, u);
    if (d < 1) continue;
    for (uint64_t i = vl[u]; i < vl[u + 1]; ++i)
      res[u] += (double)orc_cnt_galloping(vl, el, (int32_t)u, el[i]);
    res[u] /= (double)(d * (d - 1));
  }
}

void orc_lcc_numer(const uint64_t *vl, const int32_t *el, uint64_t V, uint64_t *numer,
                   int nt) {
#pragma omp parallel for num_threads(nt) schedule(dynamic, 64)
  for (int64_t u = 0; u < (int64_t)V; ++u) {
    uint64_t s = 0;
    for (uint64_t i = vl[u]; i < vl[u + 1]; ++i) s += orc_cnt_merge(vl, el, (int32_t)u, el[i]);
    numer[u] = s;
  }
}

uint64_t orc_lcc_merge_rows(const uint64_t *vl, const int32_t *el, const int32_t *rows,
                            uint64_t nrows, double *res_rows, int nt) {
  uint64_t total = 0;
#pragma omp parallel for num_threads(nt) reduction(+ : total)
  for (int64_t k = 0; k < (int64_t)nrows; ++k) {
    const int32_t u = rows[k];
    double r = 0.0;
    const uint64_t d = deg(vl, u);
    if (d >= 1) {
      for (uint64_t i = vl[u]; i < vl[u + 1]; ++i) {
        const uint64_t c = orc_cnt_merge(vl, el, u, el[i]);
        r += (double)c;
        total += c;
      }
      r /= (double)(d * (d - 1));
    }
    res_rows[k] = r;
  }
  return total;
}

// Graph::solve_jsim_merge, Graph.cpp:798-812.  The denominator is
// (double)(d_u + d_v) - c evaluated in double, as in the reference.
void orc_jsim_merge(const uint64_t *vl, const int32_t *el, const int32_t *q, uint64_t n,
                    double *res, int nt) {
#pragma omp parallel for num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const int32_t u = q[i];
    const uint64_t du = deg(vl, u);
    for (uint64_t j = 0; j < n; ++j) {
      const int32_t v = q[j];
      const uint64_t dv = deg(vl, v);
      double c = (double)orc_cnt_merge(vl, el, u, v);
      res[i * n + j] = c / ((double)(du + dv) - c);
    }
  }
}

// Graph::solve_jsim_galloping, Graph.cpp:782-796.
void orc_jsim_galloping(const uint64_t *vl, const int32_t *el, const int32_t *q,
                        uint64_t n, double *res, int nt) {
#pragma omp parallel for num_threads(nt)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const int32_t u = q[i];
    const uint64_t du = deg(vl, u);
    for (uint64_t j = 0; j < n; ++j) {
      const int32_t v = q[j];
      const uint64_t dv = deg(vl, v);
      double c = (double)orc_cnt_galloping(vl, el, u, v);
      res[i * n + j] = c / ((double)(du + dv) - c);
    }
  }
}

void orc_jaccard_edges(const uint64_t *vl, const int32_t *el, const int32_t *src,
                       const int32_t *dst, uint64_t ne, uint32_t *inter, uint32_t *uni,
                       float *ratio, int nt) {
#pragma omp parallel for num_threads(nt) schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)ne; ++i) {
    const int32_t u = src[i], v = dst[i];
    const uint64_t c = orc_cnt_merge(vl, el, u, v);
    const uint64_t un = deg(vl, u) + deg(vl, v) - c;
    inter[i] = (uint32_t)c;
    uni[i] = (uint32_t)un;
    ratio[i] = (float)((double)c / (double)un);
  }
}

// ---- Partitioned_Graph ----------------------------------------------------------
struct orc_pgraph {
  uint64_t V = 0, P = 0;
  std::vector<uint64_t> vl;  // partition-major offsets: vl[k*V + u], V*P + 1 entries
  std::vector<int32_t> el;
};

// Partitioned_Graph::Partitioned_Graph, Graph.cpp:1017-1069: P = max(v2p)+1, every
// row N(u) is split into P sub-rows by the partition of the neighbour, stored
// partition-major so that sub-row (k,u) is el[vl[k*V+u] .. vl[k*V+u+1]).
orc_pgraph *orc_pgraph_build(const uint64_t *vl, const int32_t *el, uint64_t V,
                             const int32_t *v2p, int nt) {
  orc_pgraph *g = new orc_pgraph;
  g->V = V;
  int32_t pmax = 0;
  for (uint64_t u = 0; u < V; ++u) pmax = std::max(pmax, v2p[u]);
  g->P = (uint64_t)pmax + 1;
  const uint64_t P = g->P;
  // count entries per (k,u), then scatter in row order (neighbours stay sorted)
  std::vector<uint64_t> cnt(V * P + 1, 0);
#pragma omp parallel for num_threads(nt)
  for (int64_t u = 0; u < (int64_t)V; ++u)
    for (uint64_t i = vl[u]; i < vl[u + 1]; ++i) cnt[(uint64_t)v2p[el[i]] * V + u]++;
  g->vl.assign(V * P + 1, 0);
  uint64_t run = 0;
  for (uint64_t x = 0; x < V * P; ++x) {
    g->vl[x] = run;
    run += cnt[x];
  }
  g->vl[V * P] = run;
  g->el.assign(run, 0);
#pragma omp parallel for num_threads(nt)
  for (int64_t u = 0; u < (int64_t)V; ++u) {
    std::vector<uint64_t> cur(P);
    for (uint64_t k = 0; k < P; ++k) cur[k] = g->vl[k * V + u];
    for (uint64_t i = vl[u]; i < vl[u + 1]; ++i) {
      const int32_t v = el[i];
      g->el[cur[v2p[v]]++] = v;
    }
  }
  return g;
}

uint64_t orc_pgraph_num_partitions(const orc_pgraph *g) { return g->P; }

// Partitioned_Graph::cnt_common_neis_less_than_merge, Graph.cpp:1132-1153 (with the
// bounds checked before the element reads).
static inline uint64_t pg_below(const orc_pgraph *g, int32_t u, int32_t v, int32_t thr,
                                uint64_t k) {
  const uint64_t V = g->V;
  const uint64_t ub = g->vl[k * V + u], ue = g->vl[k * V + u + 1];
  const uint64_t vb = g->vl[k * V + v], ve = g->vl[k * V + v + 1];
  return merge_count_below(g->el.data() + ub, ue - ub, g->el.data() + vb, ve - vb, thr);
}

// Partitioned_Graph::cnt_common_neis_galloping, Graph.cpp:1071-1085.
static inline uint64_t pg_gallop(const orc_pgraph *g, int32_t u, int32_t v, uint64_t k) {
  const uint64_t V = g->V;
  const uint64_t ub = g->vl[k * V + u], ue = g->vl[k * V + u + 1];
  const uint64_t vb = g->vl[k * V + v], ve = g->vl[k * V + v + 1];
  return gallop_count(g->el.data() + ub, ue - ub, g->el.data() + vb, ve - vb);
}

// Partitioned_Graph::cnt_common_neis_merge, Graph.cpp:1155-1173.
static inline uint64_t pg_merge(const orc_pgraph *g, int32_t u, int32_t v, uint64_t k) {
  const uint64_t V = g->V;
  const uint64_t ub = g->vl[k * V + u], ue = g->vl[k * V + u + 1];
  const uint64_t vb = g->vl[k * V + v], ve = g->vl[k * V + v + 1];
  return merge_count(g->el.data() + ub, ue - ub, g->el.data() + vb, ve - vb);
}

// Partitioned_Graph::evaluate_partition, Graph.cpp:1460-1473: the largest per-partition
// degree |N_i(u)| and the largest degree (summed over the partitions), balance ratio
// P * max_part_deg / max_deg in double (0/0 = NaN for a graph without edges).
double orc_pgraph_evaluate_partition(const orc_pgraph *g, uint64_t *max_part_deg,
                                     uint64_t *max_deg) {
  const uint64_t V = g->V, P = g->P;
  uint64_t mp = 0, md = 0;
  for (uint64_t u = 0; u < V; ++u) {
    uint64_t du = 0;
    for (uint64_t i = 0; i < P; ++i) {
      const uint64_t pd = g->vl[i * V + u + 1] - g->vl[i * V + u];
      du += pd;
      mp = std::max(mp, pd);
    }
    md = std::max(md, du);
  }
  *max_part_deg = mp;
  *max_deg = md;
  return (double)(P * mp) / (double)md;
}

// Per-partition pair counts, Graph.cpp:1071-1085 (galloping), 1155-1173 (merge).
uint64_t orc_pgraph_cnt(const orc_pgraph *g, int32_t u, int32_t v, uint64_t k, int galloping) {
  return galloping ? pg_gallop(g, u, v, k) : pg_merge(g, u, v, k);
}

// Partitioned_Graph::cnt_common_neis_less_than_merge, Graph.cpp:1132-1153.
uint64_t orc_pgraph_cnt_less_than(const orc_pgraph *g, int32_t u, int32_t v, int32_t thr,
                                  uint64_t k) {
  return pg_below(g, u, v, thr, k);
}

// Partitioned_Graph::cnt_common_neis_merge(u, nodes, k), Graph.cpp:1175-1195: N_k(u)
// merged against a sorted node list.
uint64_t orc_pgraph_cnt_nodes(const orc_pgraph *g, int32_t u, const int32_t *nodes, uint64_t n,
                              uint64_t k) {
  const uint64_t V = g->V;
  const uint64_t ub = g->vl[k * V + u], ue = g->vl[k * V + u + 1];
  return merge_count(g->el.data() + ub, ue - ub, nodes, n);
}

// Partitioned_Graph::get_common_neis(u, nodes, k, res), Graph.cpp:1475-1493: the merge
// of N_k(u) and a sorted node list, emitting each match.  Returns the count.
uint64_t orc_pgraph_common_neis(const orc_pgraph *g, int32_t u, const int32_t *nodes,
                                uint64_t n, uint64_t k, int32_t *out) {
  const uint64_t V = g->V;
  uint64_t i = g->vl[k * V + u], j = 0, c = 0;
  const uint64_t ie = g->vl[k * V + u + 1];
  while (i < ie && j < n) {
    const int32_t a = g->el[i];
    if (a == nodes[j]) {
      out[c++] = nodes[j];
      ++i;
      ++j;
    } else if (a < nodes[j]) {
      ++i;
    } else {
      ++j;
    }
  }
  return c;
}

// Partitioned_Graph::get_neis(u, k, res), Graph.cpp:1495-1498.  Returns the count.
uint64_t orc_pgraph_get_neis(const orc_pgraph *g, int32_t u, uint64_t k, int32_t *out) {
  const uint64_t V = g->V;
  const uint64_t b = g->vl[k * V + u], e = g->vl[k * V + u + 1];
  for (uint64_t i = b; i < e; ++i) out[i - b] = g->el[i];
  return e - b;
}

// Partitioned_Graph::cnt_common_neis_merge(q, k), Graph.cpp:1115-1130: N_k(q[0]) is
// narrowed by N_k(q[1]) .. N_k(q[n-2]) and the last sub-list is counted against it.
// Partitioned_Graph::cnt_common_neis_galloping(q, k), Graph.cpp:1087-1113: the
// sub-list of smallest size drives and each element is looked up in the others.
uint64_t orc_pgraph_kway(const orc_pgraph *g, const int32_t *q, uint64_t n, uint64_t k,
                         int galloping) {
  const uint64_t V = g->V;
  const int32_t *el = g->el.data();
  auto lo = [&](int32_t x) { return g->vl[k * V + x]; };
  auto hi = [&](int32_t x) { return g->vl[k * V + x + 1]; };
  if (galloping) {
    uint64_t min_deg = V;
    int32_t u = q[0];
    for (uint64_t i = 0; i < n; ++i)
      if (hi(q[i]) - lo(q[i]) < min_deg) {
        min_deg = hi(q[i]) - lo(q[i]);
        u = q[i];
      }
    uint64_t res = 0;
    for (uint64_t i = lo(u); i < hi(u); ++i) {
      const int32_t t = el[i];
      bool found = true;
      for (uint64_t j = 0; j < n && found; ++j) {
        if (q[j] == u) continue;
        const int32_t *b = el + lo(q[j]), *e = el + hi(q[j]);
        const int32_t *p = std::lower_bound(b, e, t);
        found = p < e && *p == t;
      }
      if (found) ++res;
    }
    return res;
  }
  std::vector<int32_t> cur(el + lo(q[0]), el + hi(q[0])), nxt;
  for (uint64_t i = 1; i + 1 < n; ++i) {
    nxt.resize(cur.size());
    nxt.resize(orc_pgraph_common_neis(g, q[i], cur.data(), cur.size(), k, nxt.data()));
    cur.swap(nxt);
  }
  return merge_count(el + lo(q[n - 1]), hi(q[n - 1]) - lo(q[n - 1]), cur.data(), cur.size());
}

uint64_t orc_pgraph_tc_merge_partial(const orc_pgraph *g, uint64_t i) {
  const uint64_t V = g->V, P = g->P;
  uint64_t res = 0;
  for (uint64_t j = 0; j < P; ++j)
    for (uint64_t u = 0; u < V; ++u)
      for (uint64_t k = g->vl[j * V + u]; k < g->vl[j * V + u + 1]; ++k) {
        const int32_t v = g->el[k];
        if ((uint64_t)v >= u) break;
        res += pg_below(g, (int32_t)u, v, v, i);
      }
  return res;
}

// Partitioned_Graph::solve_triangle_counting_merge, Graph.cpp:1346-1362: one
// OpenMP iteration per intersection partition i; every (j,u,v<u) is enumerated and
// intersected inside partition i only.
uint64_t orc_pgraph_tc_merge(const orc_pgraph *g) {
  uint64_t res = 0;
#pragma omp parallel for reduction(+ : res) schedule(dynamic, 1)
  for (int64_t i = 0; i < (int64_t)g->P; ++i) res += orc_pgraph_tc_merge_partial(g, (uint64_t)i);
  return res;
}

// Partitioned_Graph::solve_triangle_counting_galloping, Graph.cpp:1326-1344 -> 6T.
uint64_t orc_pgraph_tc_galloping(const orc_pgraph *g) {
  const uint64_t V = g->V, P = g->P;
  uint64_t res = 0;
#pragma omp parallel for reduction(+ : res) schedule(dynamic, 1)
  for (int64_t i = 0; i < (int64_t)P; ++i)
    for (uint64_t j = 0; j < P; ++j)
      for (uint64_t u = 0; u < V; ++u)
        for (uint64_t k = g->vl[j * V + u]; k < g->vl[j * V + u + 1]; ++k)
          res += pg_gallop(g, (int32_t)u, g->el[k], (uint64_t)i);
  return res;
}

// Partitioned_Graph::solve_lcc_merge, Graph.cpp:2230-2261: per-partition partial
// numerators summed, degree recovered as the sum of sub-row lengths, then
// res[u] /= d*(d-1) in double unless d < 0.5.
void orc_pgraph_lcc_merge(const orc_pgraph *g, double *res) {
  const uint64_t V = g->V, P = g->P;
  std::vector<double> acc(V, 0.0);
#pragma omp parallel
  {
    std::vector<double> mine(V, 0.0);
#pragma omp for schedule(dynamic, 1)
    for (int64_t i = 0; i < (int64_t)P; ++i)
      for (uint64_t u = 0; u < V; ++u)
        for (uint64_t j = 0; j < P; ++j)
          for (uint64_t k = g->vl[j * V + u]; k < g->vl[j * V + u + 1]; ++k)
            mine[u] += (double)pg_merge(g, (int32_t)u, g->el[k], (uint64_t)i);
#pragma omp critical
    for (uint64_t u = 0; u < V; ++u) acc[u] += mine[u];
  }
  for (uint64_t u = 0; u < V; ++u) {
    double d = 0.0;
    for (uint64_t i = 0; i < P; ++i) d += (double)(g->vl[i * V + u + 1] - g->vl[i * V + u]);
    res[u] = acc[u];
    if (d < 0.5) continue;
    res[u] /= (d * (d - 1));
  }
}

void orc_pgraph_free(orc_pgraph *g) { delete g; }

// Graph::random_partition, Graph.cpp:564-570 (rand() replaced by a seeded stream).
void orc_random_partition(uint64_t V, uint64_t P, uint64_t seed, int32_t *v2p) {
  const uint64_t key = mix64(seed + kGamma);
  for (uint64_t u = 0; u < V; ++u) v2p[u] = (int32_t)(mix64(key + u * kGamma) % P);
}

// Graph::greedy_partition_2, Graph.cpp:969-1015, following its loops: the degree list
// sorted descending (stable, so equal degrees keep ID order -- the reference's std::sort
// leaves that order open), per vertex with neighbours the first k minimising the largest
// partition_neis_counts[k][v] over v in N(u) (starting from vertex_num), then the counts of
// its neighbours in that partition raised; an isolated vertex u takes
// mix64(key + u * gamma) % P (the random_partition stream) in place of rand() % P.
void orc_greedy_partition_2(const uint64_t *vl, const int32_t *el, uint64_t V, uint64_t P,
                            uint64_t seed, int32_t *v2p) {
  std::vector<std::vector<uint64_t>> counts(P, std::vector<uint64_t>(V, 0));
  std::vector<std::pair<int64_t, uint64_t>> degree_list(V);
  for (uint64_t i = 0; i < V; ++i) {
    v2p[i] = -1;
    degree_list[i] = {(int64_t)i, deg(vl, (int64_t)i)};
  }
  std::stable_sort(degree_list.begin(), degree_list.end(),
                   [](const auto &a, const auto &b) { return a.second > b.second; });
  const uint64_t key = mix64(seed + kGamma);
  for (const auto &item : degree_list) {
    const int64_t u = item.first;
    int64_t u_part = -1;
    if (deg(vl, u) > 0) {
      uint64_t minimax = V;
      for (uint64_t k = 0; k < P; ++k) {
        uint64_t max_curr = 0;
        for (uint64_t j = vl[u]; j < vl[u + 1]; ++j)
          if (counts[k][el[j]] > max_curr) max_curr = counts[k][el[j]];
        if (max_curr < minimax) {
          minimax = max_curr;
          u_part = (int64_t)k;
        }
      }
    }
    if (u_part == -1) u_part = (int64_t)(mix64(key + (uint64_t)u * kGamma) % P);
    v2p[u] = (int32_t)u_part;
    for (uint64_t j = vl[u]; j < vl[u + 1]; ++j) counts[u_part][el[j]]++;
  }
}

void orc_contiguous_partition(uint64_t V, uint64_t P, int32_t *v2p) {
  const uint64_t w = (V + P - 1) / P;
  for (uint64_t u = 0; u < V; ++u) v2p[u] = (int32_t)(u / (w ? w : 1));
}

// ---- generators ------------------------------------------------------------------
// R-MAT sample i draws one SplitMix64 stream seeded with mix64(key + i*gamma); level
// l takes the top 53 bits of the (l+1)-th output and picks quadrant 0..3 against
// integer thresholds floor(a*2^53), floor((a+b)*2^53), floor((a+b+c)*2^53).  Bits
// are produced most-significant first: quadrant q contributes (q>>1) to the source
// and (q&1) to the destination.
int orc_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                 uint64_t seed, uint64_t *V, uint64_t **vlist, int32_t **elist,
                 uint64_t *nnz) {
  if (scale > 30) return -1;
  const uint64_t n = 1ull << scale;
  const uint64_t ns = (uint64_t)edge_factor << scale;
  const double two53 = 9007199254740992.0;
  const uint64_t t1 = (uint64_t)(a * two53), t2 = (uint64_t)((a + b) * two53),
                 t3 = (uint64_t)((a + b + c) * two53);
  const uint64_t key = mix64(seed + kGamma);
  std::vector<uint64_t> keys;
  keys.resize(2 * ns);
  std::vector<uint8_t> keep(ns, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)ns; ++i) {
    uint64_t st = mix64(key + (uint64_t)i * kGamma);
    uint32_t u = 0, v = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      st += kGamma;
      const uint64_t r = mix64(st) >> 11;
      const uint32_t q = r < t1 ? 0u : (r < t2 ? 1u : (r < t3 ? 2u : 3u));
      u = (u << 1) | (q >> 1);
      v = (v << 1) | (q & 1u);
    }
    keys[2 * i] = ((uint64_t)u << 32) | v;
    keys[2 * i + 1] = ((uint64_t)v << 32) | u;
    keep[i] = (u != v);
  }
  size_t w = 0;
  for (uint64_t i = 0; i < ns; ++i)
    if (keep[i]) {
      keys[w++] = keys[2 * i];
      keys[w++] = keys[2 * i + 1];
    }
  keys.resize(w);
  *V = n;
  return csr_from_keys(n, keys, vlist, elist, nnz);
}

// G(n,p) by Batagelj-Brandes geometric skipping, one stream per row v (seeded with
// mix64(key + v*gamma)) so rows are independent: w jumps by 1 + floor(log(r)/log(1-p))
// with r in (0,1]; every w < v reached is an edge (v,w).
int orc_gen_er(uint64_t n, double avg_deg, uint64_t seed, uint64_t *V, uint64_t **vlist,
               int32_t **elist, uint64_t *nnz) {
  if (n < 2 || n > (1ull << 31)) return -1;
  double p = avg_deg / (double)(n - 1);
  if (p > 1.0) p = 1.0;
  const uint64_t key = mix64(seed + kGamma);
  std::vector<std::vector<uint64_t>> per(omp_get_max_threads());
  const double lq = std::log1p(-p);
#pragma omp parallel
  {
    std::vector<uint64_t> &out = per[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 1024)
    for (int64_t v = 1; v < (int64_t)n; ++v) {
      uint64_t st = mix64(key + (uint64_t)v * kGamma);
      int64_t w = -1;
      while (true) {
        st += kGamma;
        const double r = (double)((mix64(st) >> 11) + 1) * (1.0 / 9007199254740992.0);
        const double skip = (p >= 1.0) ? 0.0 : std::floor(std::log(r) / lq);
        if (skip >= (double)n) break;
        w += 1 + (int64_t)skip;
        if (w >= v) break;
        out.push_back(((uint64_t)v << 32) | (uint64_t)w);
        out.push_back(((uint64_t)w << 32) | (uint64_t)v);
      }
    }
  }
  std::vector<uint64_t> keys;
  for (auto &x : per) keys.insert(keys.end(), x.begin(), x.end());
  *V = n;
  return csr_from_keys(n, keys, vlist, elist, nnz);
}

// Degree-ordered orientation u->v iff (d(u),u) < (d(v),v); W is the total input
// length of all oriented merges, the bench's algorithmic work.
void orc_oriented_stats(const uint64_t *vl, const int32_t *el, uint64_t V, orc_stats *s) {
  std::vector<uint64_t> dplus(V, 0);
  uint64_t maxd = 0, sdsq = 0;
#pragma omp parallel for reduction(max : maxd) reduction(+ : sdsq)
  for (int64_t u = 0; u < (int64_t)V; ++u) {
    const uint64_t du = deg(vl, u);
    maxd = std::max(maxd, du);
    sdsq += du * du;
    uint64_t c = 0;
    for (uint64_t i = vl[u]; i < vl[u + 1]; ++i) {
      const int64_t v = el[i];
      const uint64_t dv = deg(vl, v);
      if (du < dv || (du == dv && u < v)) ++c;
    }
    dplus[u] = c;
  }
  uint64_t W = 0, m = 0, maxo = 0, sosq = 0;
#pragma omp parallel for reduction(+ : W, m, sosq) reduction(max : maxo)
  for (int64_t u = 0; u < (int64_t)V; ++u) {
    const uint64_t du = deg(vl, u);
    maxo = std::max(maxo, dplus[u]);
    sosq += dplus[u] * dplus[u];
    for (uint64_t i = vl[u]; i < vl[u + 1]; ++i) {
      const int64_t v = el[i];
      const uint64_t dv = deg(vl, v);
      if (du < dv || (du == dv && u < v)) {
        ++m;
        W += dplus[u] + dplus[v];
      }
    }
  }
  s->V = V;
  s->nnz = vl[V];
  s->m = m;
  s->max_deg = maxd;
  s->max_outdeg = maxo;
  s->W = W;
  s->sum_outdeg_sq = sosq;
  s->sum_deg_sq = sdsq;
}

namespace {
// d+(u) under the orientation u->v iff (d(u),u) < (d(v),v).
std::vector<uint64_t> out_degrees(const uint64_t *vl, const int32_t *el, uint64_t V) {
  std::vector<uint64_t> dplus(V, 0);
#pragma omp parallel for
  for (int64_t u = 0; u < (int64_t)V; ++u) {
    const uint64_t du = deg(vl, u);
    uint64_t c = 0;
    for (uint64_t i = vl[u]; i < vl[u + 1]; ++i) {
      const int64_t v = el[i];
      const uint64_t dv = deg(vl, v);
      if (du < dv || (du == dv && u < v)) ++c;
    }
    dplus[u] = c;
  }
  return dplus;
}

inline bool oriented(const uint64_t *vl, int64_t u, int64_t v) {
  const uint64_t du = deg(vl, u), dv = deg(vl, v);
  return du < dv || (du == dv && u < v);
}
}  // namespace

void orc_pair_stats(const uint64_t *vl, const int32_t *el, uint64_t V, orc_pairs *s) {
  const std::vector<uint64_t> dplus = out_degrees(vl, el, V);
  std::vector<uint8_t> owner(V, 0);
  uint64_t pairs = 0, streamed = 0;
#pragma omp parallel reduction(+ : pairs, streamed)
  {
    std::vector<std::pair<uint64_t, int64_t>> out;  // N+(t) as (degree, id), sorted
#pragma omp for schedule(dynamic, 256)
    for (int64_t t = 0; t < (int64_t)V; ++t) {
      out.clear();
      for (uint64_t i = vl[t]; i < vl[t + 1]; ++i)
        if (oriented(vl, t, el[i])) out.emplace_back(deg(vl, el[i]), el[i]);
      std::sort(out.begin(), out.end());
      const uint64_t n = out.size();
      for (uint64_t i = 0; i < n; ++i) {
        const int64_t h = out[i].second;
        const uint64_t suffix = n - 1 - i;  // the out-neighbours of t ranked above h
        if (suffix == 0 || dplus[h] == 0) continue;
        ++pairs;
        streamed += suffix;
        // the head owns the pair (a byte store: races only ever write 1)
        __atomic_store_n(&owner[h], (uint8_t)1, __ATOMIC_RELAXED);
      }
    }
  }
  uint64_t rows = 0, staged = 0;
#pragma omp parallel for reduction(+ : rows, staged)
  for (int64_t x = 0; x < (int64_t)V; ++x)
    if (owner[x]) {
      ++rows;
      staged += dplus[x];
    }
  s->pairs = pairs;
  s->rows = rows;
  s->streamed = streamed;
  s->staged = staged;
}

uint64_t orc_tc_oriented(const uint64_t *vl, const int32_t *el, uint64_t V, int nt) {
  // oriented CSR: each row keeps its out-neighbours in ID order
  const std::vector<uint64_t> dplus = out_degrees(vl, el, V);
  std::vector<uint64_t> ooff(V + 1, 0);
  for (uint64_t u = 0; u < V; ++u) ooff[u + 1] = ooff[u] + dplus[u];
  std::vector<int32_t> oadj(ooff[V]);
#pragma omp parallel for
  for (int64_t u = 0; u < (int64_t)V; ++u) {
    uint64_t k = ooff[u];
    for (uint64_t i = vl[u]; i < vl[u + 1]; ++i)
      if (oriented(vl, u, el[i])) oadj[k++] = el[i];
  }
  // Per thread, N+(u) is marked in a byte array over the vertices; each out-neighbour
  // v's list is then scanned against the marks.
  uint64_t res = 0;
#pragma omp parallel num_threads(nt) reduction(+ : res)
  {
    std::vector<uint8_t> mark(V, 0);
#pragma omp for schedule(dynamic, 256)
    for (int64_t u = 0; u < (int64_t)V; ++u) {
      for (uint64_t i = ooff[u]; i < ooff[u + 1]; ++i) mark[oadj[i]] = 1;
      for (uint64_t i = ooff[u]; i < ooff[u + 1]; ++i) {
        const int32_t v = oadj[i];
        for (uint64_t j = ooff[v]; j < ooff[v + 1]; ++j) res += mark[oadj[j]];
      }
      for (uint64_t i = ooff[u]; i < ooff[u + 1]; ++i) mark[oadj[i]] = 0;
    }
  }
  return res;
}

// ---- MLP ---------------------------------------------------------------------------
// 16 -> 64 -> 64 -> 3 with ReLU; parameter layout W1[64][16] b1[64] W2[64][64] b2[64]
// W3[3][64] b3[3]; dot products accumulate in double.
namespace {
constexpr int kIn = 16, kH = 64, kOut = 3;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kOut * kH, kNParams = oB3 + kOut;

void mlp_fwd_one(const float *P, const float *x, float *h1, float *h2, float *z) {
  for (int j = 0; j < kH; ++j) {
    double s = P[oB1 + j];
    for (int i = 0; i < kIn; ++i) s += (double)P[oW1 + j * kIn + i] * x[i];
    h1[j] = s > 0 ? (float)s : 0.f;
  }
  for (int j = 0; j < kH; ++j) {
    double s = P[oB2 + j];
    for (int i = 0; i < kH; ++i) s += (double)P[oW2 + j * kH + i] * h1[i];
    h2[j] = s > 0 ? (float)s : 0.f;
  }
  for (int j = 0; j < kOut; ++j) {
    double s = P[oB3 + j];
    for (int i = 0; i < kH; ++i) s += (double)P[oW3 + j * kH + i] * h2[i];
    z[j] = (float)s;
  }
}
}  // namespace

void orc_mlp_forward(const float *params, const float *x, uint64_t n, float *logits) {
  for (uint64_t s = 0; s < n; ++s) {
    float h1[kH], h2[kH];
    mlp_fwd_one(params, x + s * kIn, h1, h2, logits + s * kOut);
  }
}

float orc_mlp_backward(const float *params, const float *x, const int32_t *labels,
                       const float *weights, uint64_t n, float *grads) {
  std::vector<double> g(kNParams, 0.0);
  double loss = 0.0;
  for (uint64_t s = 0; s < n; ++s) {
    float h1[kH], h2[kH], z[kOut];
    const float *xs = x + s * kIn;
    mlp_fwd_one(params, xs, h1, h2, z);
    const double wgt = (weights ? (double)weights[s] : 1.0) / (double)n;
    double mx = std::max(z[0], std::max(z[1], z[2]));
    double e[kOut], se = 0;
    for (int j = 0; j < kOut; ++j) {
      e[j] = std::exp((double)z[j] - mx);
      se += e[j];
    }
    const int y = labels[s];
    loss += wgt * (std::log(se) + mx - (double)z[y]);
    double dz[kOut];
    for (int j = 0; j < kOut; ++j) dz[j] = wgt * (e[j] / se - (j == y ? 1.0 : 0.0));
    double dh2[kH];
    for (int i = 0; i < kH; ++i) dh2[i] = 0;
    for (int j = 0; j < kOut; ++j) {
      g[oB3 + j] += dz[j];
      for (int i = 0; i < kH; ++i) {
        g[oW3 + j * kH + i] += dz[j] * h2[i];
        dh2[i] += dz[j] * params[oW3 + j * kH + i];
      }
    }
    double dh1[kH];
    for (int i = 0; i < kH; ++i) dh1[i] = 0;
    for (int j = 0; j < kH; ++j) {
      if (h2[j] <= 0.f) continue;
      g[oB2 + j] += dh2[j];
      for (int i = 0; i < kH; ++i) {
        g[oW2 + j * kH + i] += dh2[j] * h1[i];
        dh1[i] += dh2[j] * params[oW2 + j * kH + i];
      }
    }
    for (int j = 0; j < kH; ++j) {
      if (h1[j] <= 0.f) continue;
      g[oB1 + j] += dh1[j];
      for (int i = 0; i < kIn; ++i) g[oW1 + j * kIn + i] += dh1[j] * xs[i];
    }
  }
  for (int k = 0; k < kNParams; ++k) grads[k] = (float)g[k];
  return (float)loss;
}


// ---- k-way intersections, friend triangles, squares -------------------------------

// Graph::cnt_common_neis_merge(const vector<VID>&), Graph.cpp:255-270: the running
// intersection starts as N(q[0]), is merged with N(q[1]) .. N(q[k-2]), and the last
// list is only counted against it.
uint64_t orc_kway_merge(const uint64_t *vl, const int32_t *el, const int32_t *q, uint64_t k) {
  std::vector<int32_t> cur(el + vl[q[0]], el + vl[q[0] + 1]), nxt;
  for (uint64_t i = 1; i + 1 < k; ++i) {
    nxt.clear();
    const int32_t *b = el + vl[q[i]];
    const uint64_t nb = deg(vl, q[i]);
    uint64_t x = 0, y = 0;
    while (x < cur.size() && y < nb) {
      if (cur[x] == b[y]) {
        nxt.push_back(cur[x]);
        ++x;
        ++y;
      } else if (cur[x] < b[y]) {
        ++x;
      } else {
        ++y;
      }
    }
    cur.swap(nxt);
  }
  return merge_count(cur.data(), cur.size(), el + vl[q[k - 1]], deg(vl, q[k - 1]));
}

// Graph::cnt_common_neis_merge(u, nodes), Graph.cpp:241-253: N(u) merged against a
// sorted node list.
uint64_t orc_cnt_merge_nodes(const uint64_t *vl, const int32_t *el, int32_t u,
                             const int32_t *nodes, uint64_t n) {
  return merge_count(el + vl[u], deg(vl, u), nodes, n);
}

// Graph::get_common_neis(u, v, res), Graph.cpp:468-479 (with v's list as `nodes`) and
// Graph::get_common_neis(u, nodes, res), Graph.cpp:481-493: the merge emits each match.
// Returns the count.
uint64_t orc_common_neis_nodes(const uint64_t *vl, const int32_t *el, int32_t u,
                               const int32_t *nodes, uint64_t n, int32_t *out) {
  uint64_t i = vl[u], j = 0, c = 0;
  while (i < vl[u + 1] && j < n) {
    if (el[i] == nodes[j]) {
      out[c++] = nodes[j];
      ++i;
      ++j;
    } else if (el[i] < nodes[j]) {
      ++i;
    } else {
      ++j;
    }
  }
  return c;
}

// Graph::cnt_common_neis_galloping(const vector<VID>&), Graph.cpp:200-225: the
// minimum-degree list drives; each of its elements is looked up with lower_bound in
// every other list.
uint64_t orc_kway_galloping(const uint64_t *vl, const int32_t *el, const int32_t *q, uint64_t k) {
  int32_t u = q[0];
  for (uint64_t i = 0; i < k; ++i)
    if (deg(vl, q[i]) < deg(vl, u)) u = q[i];
  uint64_t res = 0;
  for (uint64_t i = vl[u]; i < vl[u + 1]; ++i) {
    const int32_t t = el[i];
    bool found = true;
    for (uint64_t j = 0; j < k && found; ++j) {
      const int32_t v = q[j];
      if (v == u) continue;
      const int32_t *b = el + vl[v], *e = el + vl[v + 1];
      const int32_t *p = std::lower_bound(b, e, t);
      found = p < e && *p == t;
    }
    if (found) ++res;
  }
  return res;
}

// Graph::solve_friend_triangle_merge, Graph.cpp:323-338: over all index pairs i < j of
// `nodes` with neither equal to root, the 3-way common-neighbour count of
// {root, nodes[i], nodes[j]}.
uint64_t orc_friend_triangle_merge(const uint64_t *vl, const int32_t *el, int32_t root,
                                   const int32_t *nodes, uint64_t n, int nt) {
  uint64_t res = 0;
#pragma omp parallel for num_threads(nt) reduction(+ : res) schedule(dynamic, 4)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    if (nodes[i] == root) continue;
    for (uint64_t j = (uint64_t)i + 1; j < n; ++j) {
      if (nodes[j] == root) continue;
      const int32_t q[3] = {root, nodes[i], nodes[j]};
      res += orc_kway_merge(vl, el, q, 3);
    }
  }
  return res;
}

// Partitioned_Graph::solve_square_counting, Graph.cpp:1291-1324: work items (v, w) for
// every u and neighbours w < v < u (sub-rows scanned partition by partition, each
// stopping at the first neighbour >= the bound), then for every partition k the
// common neighbours x < w of v and w in N_k.
uint64_t orc_pgraph_square_counting(const orc_pgraph *g) {
  const uint64_t V = g->V, P = g->P;
  std::vector<std::pair<int32_t, int32_t>> items;
  for (uint64_t u = 0; u < V; ++u)
    for (uint64_t i = 0; i < P; ++i)
      for (uint64_t vp = g->vl[i * V + u]; vp < g->vl[i * V + u + 1]; ++vp) {
        const int32_t v = g->el[vp];
        if ((int64_t)u <= v) break;
        for (uint64_t j = 0; j < P; ++j)
          for (uint64_t wp = g->vl[j * V + u]; wp < g->vl[j * V + u + 1]; ++wp) {
            const int32_t w = g->el[wp];
            if (v <= w) break;
            items.emplace_back(v, w);
          }
      }
  uint64_t res = 0;
#pragma omp parallel for reduction(+ : res) schedule(dynamic, 1)
  for (int64_t k = 0; k < (int64_t)P; ++k)
    for (const auto &e : items) res += pg_below(g, e.first, e.second, e.second, (uint64_t)k);
  return res;
}

}  // extern "C"
