/*
 * pac_oracle.h -- CPU/OpenMP restatement of the PartitionAndConquer neighbour-set
 * intersection path (triangle counting, LCC, Jaccard), used ONLY as test
 * infrastructure: by tests/, by __graft_entry__.smoke() as the checker, and by
 * bench.py's cpu_baseline leg.  The product library (libpac.so) never links,
 * loads or calls anything here.
 *
 * Every function cites the reference function it restates (file:line in
 * Byy-0024/PartitionAndConquer).  Types follow the reference: vertex IDs are
 * int32 (Graph.hpp:23 VID), CSR offsets are uint64 (Graph.hpp:24 EID), counts are
 * size_t/uint64 and ratios double.
 *
 * Pinning (what checks each function independently of its own formula):
 *   orc_text_to_csr      hand-written text fixtures (tests/golden/text_*.txt)
 *   orc_cnt_*            std::set_intersection / Python set on random sorted lists
 *   orc_tc_*             closed forms (K_n, wheel, cycle, K_ab, friendship),
 *                        trace(A^3)/6 dense brute force, networkx.triangles
 *   orc_lcc_*            diag(A^3), networkx.clustering (differs only at d=1)
 *   orc_jsim_*           (A^2)_uv dense, networkx.jaccard_coefficient, closed forms
 *   orc_jaccard_edges    (A^2)_uv dense on the edge set
 *   orc_pgraph_*         partition-sum invariance against orc_tc_merge and
 *                        brute force, for random and contiguous vertex2partition;
 *                        evaluate_partition: K_n closed forms, numpy per-partition
 *                        neighbour histograms
 *   orc_kway_*           Python set intersections of k neighbourhoods
 *   orc_friend_*         brute-force pairs over Python sets; closed form on K_n
 *   orc_pgraph_square_*  brute-force enumeration of 4-cycles (u,v,x,w) with
 *                        u > v > w > x on tiny graphs; K_n -> C(n,4); K_{a,b}
 *                        labelled side by side -> 0
 *   orc_greedy_part*_2   hand-derived placements (path, star), the minimax property of
 *                        every placement replayed in Python, seeded isolated vertices
 *   orc_pair_stats       closed forms on K_n and stars; a per-vertex brute-force walk
 *                        of the ownership rule on R-MAT and Erdos-Renyi graphs
 *   orc_tc_oriented      golden fixtures; equality with orc_tc_merge on R-MAT / ER
 *   orc_gen_*            structural invariants (symmetric, sorted, unique, no
 *                        self-loops), edge-count expectations; the GPU generator
 *                        must reproduce orc_gen_rmat bit-exactly
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- loaders / builders ---------------------------------------------------- */

/* Graph::Graph(const std::string graphfile), Graph.cpp:27-90.
 * Returns 0 on success, -1 if the file cannot be opened.  *vlist (V+1 uint64) and
 * *elist (nnz int32) are malloc'd; free with orc_free. */
int orc_text_to_csr(const char *path, uint64_t *V, uint64_t **vlist, int32_t **elist,
                    uint64_t *nnz);

/* Build a symmetric, sorted, deduplicated, self-loop-free CSR from an undirected
 * edge list (the row-building half of Graph.cpp:70-85). */
int orc_edges_to_csr(uint64_t V, const uint32_t *src, const uint32_t *dst, uint64_t ne,
                     uint64_t **vlist, int32_t **elist, uint64_t *nnz);

void orc_free(void *p);

/* ---- pairwise intersections (Graph.cpp:186-198, 227-239, 272-285) ---------- */
uint64_t orc_cnt_merge(const uint64_t *vl, const int32_t *el, int32_t u, int32_t v);
uint64_t orc_cnt_less_than_merge(const uint64_t *vl, const int32_t *el, int32_t u,
                                 int32_t v, int32_t thr);
uint64_t orc_cnt_galloping(const uint64_t *vl, const int32_t *el, int32_t u, int32_t v);

/* ---- triangle counting ---------------------------------------------------- */
/* Graph::solve_triangle_counting_merge, Graph.cpp:355-368 -> T */
uint64_t orc_tc_merge(const uint64_t *vl, const int32_t *el, uint64_t V, int nt);
/* Graph::solve_triangle_counting_galloping, Graph.cpp:342-353 -> 6T */
uint64_t orc_tc_galloping(const uint64_t *vl, const int32_t *el, uint64_t V, int nt);
/* The loop of Graph.cpp:355-368 restricted to the listed rows u (CPU baseline
 * sampling); returns the triangles counted at those rows. */
uint64_t orc_tc_merge_rows(const uint64_t *vl, const int32_t *el, const int32_t *rows,
                           uint64_t nrows, int nt);

/* ---- LCC ------------------------------------------------------------------ */
/* Graph::solve_lcc_merge, Graph.cpp:828-840 (res has V doubles) */
void orc_lcc_merge(const uint64_t *vl, const int32_t *el, uint64_t V, double *res, int nt);
/* Graph::solve_lcc_galloping, Graph.cpp:814-826 */
void orc_lcc_galloping(const uint64_t *vl, const int32_t *el, uint64_t V, double *res,
                       int nt);
/* The integer numerator of Graph.cpp:834-837: sum over v in N(u) of
 * cnt_common_neis_merge(u,v), which equals 2*t(u). */
void orc_lcc_numer(const uint64_t *vl, const int32_t *el, uint64_t V, uint64_t *numer,
                   int nt);
/* Graph.cpp:828-840 restricted to listed rows (CPU baseline sampling). Returns the
 * sum of the numerators (to keep the work observable). */
uint64_t orc_lcc_merge_rows(const uint64_t *vl, const int32_t *el, const int32_t *rows,
                            uint64_t nrows, double *res_rows, int nt);

/* ---- Jaccard ---------------------------------------------------------------- */
/* Graph::solve_jsim_merge, Graph.cpp:798-812 (res has n*n doubles) */
void orc_jsim_merge(const uint64_t *vl, const int32_t *el, const int32_t *q, uint64_t n,
                    double *res, int nt);
/* Graph::solve_jsim_galloping, Graph.cpp:782-796 */
void orc_jsim_galloping(const uint64_t *vl, const int32_t *el, const int32_t *q,
                        uint64_t n, double *res, int nt);
/* Per-edge form of the pair formula of Graph.cpp:808-809: for each listed edge
 * (src[i],dst[i]): inter = cnt_common_neis_merge, uni = d(u)+d(v)-inter,
 * ratio = (float)((double)inter/(double)uni). */
void orc_jaccard_edges(const uint64_t *vl, const int32_t *el, const int32_t *src,
                       const int32_t *dst, uint64_t ne, uint32_t *inter, uint32_t *uni,
                       float *ratio, int nt);

/* ---- Partitioned_Graph (Graph.cpp:1017-1069 and its solvers) ---------------- */
typedef struct orc_pgraph orc_pgraph;
orc_pgraph *orc_pgraph_build(const uint64_t *vl, const int32_t *el, uint64_t V,
                             const int32_t *v2p, int nt);
uint64_t orc_pgraph_num_partitions(const orc_pgraph *g);
/* Partitioned_Graph::solve_triangle_counting_merge, Graph.cpp:1346-1362 -> T */
uint64_t orc_pgraph_tc_merge(const orc_pgraph *g);
/* Partitioned_Graph::solve_triangle_counting_galloping, Graph.cpp:1326-1344 -> 6T */
uint64_t orc_pgraph_tc_galloping(const orc_pgraph *g);
/* Partitioned_Graph::solve_lcc_merge, Graph.cpp:2230-2261 */
void orc_pgraph_lcc_merge(const orc_pgraph *g, double *res);
/* Per-partition partial of Graph.cpp:1346-1362 for intersection partition i. */
uint64_t orc_pgraph_tc_merge_partial(const orc_pgraph *g, uint64_t i);
/* Partitioned_Graph::evaluate_partition, Graph.cpp:1460-1473: returns the balance ratio
 * P * max_part_deg / max_deg (NaN without edges) and the two maxima. */
double orc_pgraph_evaluate_partition(const orc_pgraph *g, uint64_t *max_part_deg,
                                     uint64_t *max_deg);
/* Per-partition pair counts |N_k(u) & N_k(v)|: Graph.cpp:1155-1173 (merge) and
 * 1071-1085 (galloping = nonzero). */
uint64_t orc_pgraph_cnt(const orc_pgraph *g, int32_t u, int32_t v, uint64_t k, int galloping);
/* Partitioned_Graph::cnt_common_neis_less_than_merge, Graph.cpp:1132-1153. */
uint64_t orc_pgraph_cnt_less_than(const orc_pgraph *g, int32_t u, int32_t v, int32_t thr,
                                  uint64_t k);
/* Partitioned_Graph::cnt_common_neis_merge(u, nodes, k), Graph.cpp:1175-1195. */
uint64_t orc_pgraph_cnt_nodes(const orc_pgraph *g, int32_t u, const int32_t *nodes, uint64_t n,
                              uint64_t k);
/* Partitioned_Graph::get_common_neis(u, nodes, k), Graph.cpp:1475-1493 -> count; out holds
 * at least min(n, |N_k(u)|) entries. */
uint64_t orc_pgraph_common_neis(const orc_pgraph *g, int32_t u, const int32_t *nodes,
                                uint64_t n, uint64_t k, int32_t *out);
/* Partitioned_Graph::get_neis(u, k), Graph.cpp:1495-1498 -> count. */
uint64_t orc_pgraph_get_neis(const orc_pgraph *g, int32_t u, uint64_t k, int32_t *out);
/* Partitioned_Graph::cnt_common_neis_{merge,galloping}(q, k), Graph.cpp:1087-1130 (n >= 1). */
uint64_t orc_pgraph_kway(const orc_pgraph *g, const int32_t *q, uint64_t n, uint64_t k,
                         int galloping);
void orc_pgraph_free(orc_pgraph *g);

/* ---- random_partition (Graph.cpp:564-570) with a seeded stream instead of rand() */
/* Graph::greedy_partition_2, Graph.cpp:969-1015 (ties in degree by ID, seeded draws for
 * isolated vertices: see the .cpp). */
void orc_greedy_partition_2(const uint64_t *vl, const int32_t *el, uint64_t V, uint64_t P,
                            uint64_t seed, int32_t *v2p);
void orc_random_partition(uint64_t V, uint64_t P, uint64_t seed, int32_t *v2p);
/* Contiguous-ID partition: part(v) = v / ceil(V/P). */
void orc_contiguous_partition(uint64_t V, uint64_t P, int32_t *v2p);

/* ---- seeded generators (no reference counterpart; test infrastructure) ----- */
/* R-MAT, ef * 2^scale samples, quadrant probabilities (a,b,c,1-a-b-c), integer
 * arithmetic only, so the GPU generator reproduces it bit for bit. */
int orc_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                 uint64_t seed, uint64_t *V, uint64_t **vlist, int32_t **elist,
                 uint64_t *nnz);
/* Erdos-Renyi G(n,p), p = avg_deg/(n-1), row-parallel Batagelj-Brandes skipping. */
int orc_gen_er(uint64_t n, double avg_deg, uint64_t seed, uint64_t *V, uint64_t **vlist,
               int32_t **elist, uint64_t *nnz);

/* ---- degree-ordered orientation statistics (the bench's work model) -------- */
typedef struct {
  uint64_t V, nnz, m;          /* vertices, directed entries (2m), undirected edges */
  uint64_t max_deg, max_outdeg;
  uint64_t W;                  /* sum over oriented u->v of (d+(u) + d+(v)) */
  uint64_t sum_outdeg_sq;      /* sum_u d+(u)^2 */
  uint64_t sum_deg_sq;         /* sum_u d(u)^2 */
} orc_stats;
void orc_oriented_stats(const uint64_t *vl, const int32_t *el, uint64_t V, orc_stats *s);

/* Intersection pairs under the same orientation.  Each oriented edge t->h is one pair,
 * owned by the head h, whose out-list a GPU row stages, while the suffix of N+(t) above
 * h -- S(t,h) = #{w in N+(t) : (d(h),h) < (d(w),w)}, the only part of N+(t) that can meet
 * N+(h) -- is streamed; a pair with S(t,h) = 0 or d+(h) = 0 holds no triangle and is
 * dropped.  These counts are the work model of the hash path (its algorithmic bytes),
 * not a reference function. */
typedef struct {
  uint64_t pairs;     /* #{oriented t->h : S(t,h) > 0 and d+(h) > 0}              */
  uint64_t rows;      /* vertices owning at least one pair                         */
  uint64_t streamed;  /* sum of S(t,h) over the pairs                              */
  uint64_t staged;    /* sum of d+(x) over the vertices x owning at least one pair */
} orc_pairs;
void orc_pair_stats(const uint64_t *vl, const int32_t *el, uint64_t V, orc_pairs *s);

/* ---- a fast exact checker (not a restatement) ---------------------------------- */
/* T by the degree-ordered forward algorithm (Schank and Wagner, 2005): the sum over
 * oriented u->v of |N+(u) & N+(v)|, N+(u) marked in a per-thread byte array and each
 * N+(v) scanned against the marks.  A triangle a < b < c in (degree, id) order is counted once, at the edge
 * a->b, with c in both out-lists.  Used where orc_tc_merge's ID-ordered
 * prefix merges take too long (hub rows at low IDs); the CPU tests pin it against
 * orc_tc_merge. */
uint64_t orc_tc_oriented(const uint64_t *vl, const int32_t *el, uint64_t V, int nt);

/* ---- MLP oracle (16 -> 64 -> 64 -> 3, ReLU, softmax cross-entropy) ---------- */
/* Weights are given as float (already rounded to bf16 by the caller when the GPU
 * path stores bf16); all arithmetic is float with double accumulation where noted.
 * Layout: W1[64][16], b1[64], W2[64][64], b2[64], W3[3][64], b3[3]. */
void orc_mlp_forward(const float *params, const float *x, uint64_t n, float *logits);
/* Mean softmax cross-entropy over n samples with integer labels; writes dL/dparams
 * (same layout as params) and returns the loss. Sample weights may be NULL. */
float orc_mlp_backward(const float *params, const float *x, const int32_t *labels,
                       const float *weights, uint64_t n, float *grads);

/* ---- k-way intersections, friend triangles, squares ------------------------- */

/* Graph::cnt_common_neis_merge(const vector<VID>&), Graph.cpp:255-270 (k >= 1). */
uint64_t orc_kway_merge(const uint64_t *vl, const int32_t *el, const int32_t *q, uint64_t k);
/* Graph::cnt_common_neis_merge(u, nodes), Graph.cpp:241-253 (nodes sorted). */
uint64_t orc_cnt_merge_nodes(const uint64_t *vl, const int32_t *el, int32_t u,
                             const int32_t *nodes, uint64_t n);
/* Graph::get_common_neis(u, nodes), Graph.cpp:468-493 -> count; out holds at least
 * min(n, d(u)) entries. */
uint64_t orc_common_neis_nodes(const uint64_t *vl, const int32_t *el, int32_t u,
                               const int32_t *nodes, uint64_t n, int32_t *out);
/* Graph::cnt_common_neis_galloping(const vector<VID>&), Graph.cpp:200-225. */
uint64_t orc_kway_galloping(const uint64_t *vl, const int32_t *el, const int32_t *q, uint64_t k);
/* Graph::solve_friend_triangle_merge, Graph.cpp:323-338. */
uint64_t orc_friend_triangle_merge(const uint64_t *vl, const int32_t *el, int32_t root,
                                   const int32_t *nodes, uint64_t n, int nt);
/* Partitioned_Graph::solve_square_counting, Graph.cpp:1291-1324. */
uint64_t orc_pgraph_square_counting(const orc_pgraph *g);

#ifdef __cplusplus
}
#endif
