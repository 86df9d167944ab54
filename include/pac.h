/*
 * pac.h -- C ABI of partitionandconquer_amd: the MI355X (gfx950) implementation of
 * the partitioned neighbour-set intersection path of Byy-0024/PartitionAndConquer
 * (triangle counting, local clustering coefficient, Jaccard similarity).
 *
 * The reference exposes this path as the C++ classes Graph / Partitioned_Graph
 * (src/Graph.hpp:38-155) plus the argv protocol of bin/main (src/test.cpp:301-389).
 * Every entry point below names the reference interface it replaces.  Plain
 * pointers and sizes only; no C++ types or exceptions cross this boundary.
 *
 * Conventions
 *   - Every int-returning function returns PAC_OK (0) or a negative PAC_ERR_*;
 *     pac_last_error() then describes the failure (thread-local string).
 *   - The caller owns every host buffer; the library owns device memory behind the
 *     opaque handles.  Host output buffers must have the documented length.
 *   - A pac_ctx owns one HIP stream on one device.  Handles are not re-entrant:
 *     use one host thread per context.
 *   - Vertex IDs are int32 on the boundary (reference VID, Graph.hpp:23) and uint32
 *     inside; CSR offsets are uint64 on the boundary (reference EID, Graph.hpp:24)
 *     and uint32 on the device, which limits a graph to < 2^32 directed entries.
 *   - Counts are exact integers.  Ratios are computed from the exact integers as
 *     (double)num/(double)den, rounded once to float where a float is returned.
 */
#ifndef PAC_H
#define PAC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PAC_OK 0
#define PAC_ERR_ARG (-1)         /* bad argument (NULL, size, range)                 */
#define PAC_ERR_HIP (-2)         /* HIP runtime error                                */
#define PAC_ERR_OOM (-3)         /* device or host allocation failed / budget hit    */
#define PAC_ERR_FORMAT (-4)      /* malformed CSR: offsets, ordering, IDs, loops     */
#define PAC_ERR_IO (-5)          /* file missing, unreadable or truncated            */
#define PAC_ERR_STATE (-6)       /* call order (e.g. solve before prepare)           */
#define PAC_ERR_UNSUPPORTED (-7) /* graph exceeds a documented limit                 */
#define PAC_ERR_TIMEOUT (-8)     /* a solve did not complete within the context's    *
                                  * timeout (pac_ctx_set_timeout)                     */

typedef struct pac_ctx pac_ctx;
typedef struct pac_graph pac_graph;

/* Which intersection strategy runs on each partition. */
enum {
  PAC_DISPATCH_FIXED = 0,   /* fixed degree-threshold heuristic                      */
  PAC_DISPATCH_LEARNED = 1, /* MLP over per-partition features (falls back to FIXED  *
                             * when no model is loaded)                              */
  PAC_FORCE_MERGE = 2,      /* wave64 merge-path sorted merge everywhere             */
  PAC_FORCE_GALLOP = 3,     /* binary search of the shorter list in the longer one   */
  PAC_FORCE_BITMAP = 4      /* LDS hash of N+(u): two-choice cuckoo table, exact      */
};
#define PAC_NUM_STRATEGIES 3 /* strategy ids: 0 merge, 1 gallop, 2 bitmap */

/* How the per-partition work is launched. */
enum {
  PAC_LAUNCH_PLAIN = 0,      /* one launch per partition group (a run of consecutive  *
                              * partitions with the same strategy), one workgroup/row */
  PAC_LAUNCH_PERSISTENT = 1, /* one persistent kernel over an atomic work queue       */
  PAC_LAUNCH_GRAPH = 2       /* PERSISTENT, captured once in a hipGraph and replayed  */
};

typedef struct {
  uint64_t num_vertices;   /* V                                                     */
  uint64_t num_entries;    /* directed CSR entries = 2m (Graph::get_edge_num)        */
  uint64_t num_edges;      /* undirected edges m = oriented edges                    */
  uint64_t max_degree;
  uint64_t max_out_degree; /* under degree orientation (d(u),u) < (d(v),v)           */
  uint64_t work_W;         /* sum over oriented u->v of d+(u)+d+(v)                  */
  uint64_t num_partitions; /* 0 before pac_graph_prepare                             */
  uint64_t num_items;      /* persistent work items after pac_graph_prepare          */
  double prepare_ms;       /* GPU time of the last pac_graph_prepare                 */
  /* The intersection pairs (after pac_graph_prepare; 0 before).  Each oriented edge
   * t->h is one pair, owned by the head h, which stages N+(h) while the suffix of N+(t)
   * above h (the out-lists are in (degree, id) order; every element of N+(h) ranks above
   * h) is streamed; pairs whose suffix or N+(h) is empty are dropped.  These are the
   * work model of the hash path:
   * its algorithmic bytes per launch are S*streamed + 4*staged_item_elems +
   * 8*num_pairs (each pair's list bounds) + 20*num_rows + 32*num_items (TC), plus
   * 4*num_pairs (partner IDs) and 16*num_vertices (per-vertex counts read and written
   * once) with per-vertex counts, or 16*num_pairs (partner, edge and hit-word indices) +
   * streamed/8 (one hit bit per streamed element) + 8*num_edges (supports read and written
   * once) for per-edge supports; S = stream_elem_bytes, 4 for per-edge supports. */
  uint64_t num_pairs;
  uint64_t num_rows;          /* vertices owning at least one pair                     */
  uint64_t streamed_elems;    /* sum over pairs t->h of #{w in N+(t) ranked above h}   */
  uint64_t staged_elems;      /* sum of d+(x) over the owning vertices                 */
  uint64_t staged_item_elems; /* the same, once per work item holding the row (a row  *
                               * split over several items is staged by each)          */
  uint64_t stream_elem_bytes; /* bytes per streamed element on the hash path of TC and *
                               * per-vertex counts when this graph is solved alone:   *
                               * 3 or 4 as pac_ctx_set_packed_lists decides; per-edge *
                               * supports always stream 4-byte lists                  */
} pac_graph_info;

typedef struct {
  double total_ms;      /* device time of the last solve (zeroing+intersection+epilogue) */
  double isect_ms;      /* device time of the intersection launches alone               */
  double epilogue_ms;   /* device time of the epilogue kernel                           */
  double kernel_ms;     /* device time of the dominant intersection kernel alone (events  *
                         * around its launch; 0 when replayed from a hipGraph)            */
  uint64_t launches;    /* kernel launches in the intersection phase                    */
  uint64_t strategy_partitions[PAC_NUM_STRATEGIES]; /* partitions per strategy          */
} pac_timing;

/* ---- context ------------------------------------------------------------------- */
/* No reference counterpart (the reference is single-process OpenMP): binds a HIP
 * device and creates the context's stream. */
int pac_ctx_create(int device, pac_ctx **out);
/* Destroys the context and every graph still alive on it (their handles become
 * invalid; do not pass them to pac_graph_free afterwards). */
void pac_ctx_destroy(pac_ctx *ctx);
/* Thread-local description of the last failure on this thread. */
const char *pac_last_error(void);
/* Device timing of the last solve on this context (hipEvents on the ctx stream): TC, LCC,
 * Jaccard and batches as described above; the pair, k-way, friend-triangle and square
 * queries set total_ms = isect_ms to their device work (inputs already uploaded). */
int pac_ctx_last_timing(const pac_ctx *ctx, pac_timing *out);
/* Number of compute units of the context's device. */
int pac_ctx_num_cus(const pac_ctx *ctx);
/* Failure detection (no reference counterpart: the reference waits without bound).  With
 * ms > 0 every solve that waits for its results (TC, LCC, Jaccard, batches, the queries)
 * polls its completion for at most ms of wall time and then returns PAC_ERR_TIMEOUT; the
 * kernels keep running on the context's stream (a launch cannot be withdrawn), so later
 * calls on the context wait behind them.  ms = 0 (the default) waits without bound. */
int pac_ctx_set_timeout(pac_ctx *ctx, double ms);
/* No reference counterpart (the reference's lists are int arrays).  pac_graph_prepare gives
 * every graph whose IDs are all below 2^24 a copy of its oriented lists with 3-byte
 * elements (about 0.75 bytes per undirected edge and 8 bytes per intersection pair more
 * device memory).  The hash path of TC and per-vertex counts (LCC) streams the partner
 * lists from them -- a quarter fewer streamed bytes, the same results -- when every graph
 * of the launch has them and the mode allows: PAC_PACKED_AUTO (the default) when the
 * launch's largest graph has at least 2^24 undirected edges (smaller graphs' lists stay
 * in the caches, where unpacking costs more than the bytes save), PAC_PACKED_ALWAYS
 * whatever the size, PAC_PACKED_OFF never.  Per-edge supports (Jaccard) always stream the
 * 4-byte lists.  Applies from the next launch, hipGraph replays included. */
enum { PAC_PACKED_OFF = 0, PAC_PACKED_AUTO = 1, PAC_PACKED_ALWAYS = 2 };
int pac_ctx_set_packed_lists(pac_ctx *ctx, int mode);
/* Device-to-device copy of `bytes` (multiple of 16) by a plain streaming kernel,
 * `reps` times; *ms receives the median kernel time.  Used to measure the achievable
 * HBM bandwidth and to calibrate hardware byte counters.  No reference counterpart. */
int pac_stream_copy(pac_ctx *ctx, uint64_t bytes, int reps, double *ms);

/* ---- graph construction (L1/L2 of the reference) ---------------------------------- */
/* Replaces Graph::Graph(const std::vector<...>&) / the CSR members (Graph.hpp:99-100):
 * uploads a symmetric CSR with sorted, duplicate-free, self-loop-free rows.
 * vlist has V+1 entries, elist has nnz = vlist[V] entries.  Offsets and IDs are
 * validated (PAC_ERR_FORMAT). */
int pac_graph_upload(pac_ctx *ctx, const uint64_t *vlist, uint64_t num_vertices,
                     const int32_t *elist, uint64_t nnz, pac_graph **out);
/* Replaces Graph::Graph(const std::string csrdir, bool), Graph.cpp:92-137: reads
 * <dir>/csr_vlist.bin (uint64 [V+1]) and <dir>/csr_elist.bin (int32 [nnz]). */
int pac_graph_load_csr_dir(pac_ctx *ctx, const char *csrdir, pac_graph **out);
/* Replaces Graph::Graph(const std::string graphfile), Graph.cpp:27-90 (text edge
 * list: '#'/'%' comments, one-character separator, 1-based IDs shifted when no 0
 * appears, self-loops dropped, symmetrized, rows sorted and deduplicated).  Parsed on
 * the device, one thread per line; a file with a line over 511 characters or a line
 * ending in its first number is parsed on the host with the reference's 512-byte
 * read loop, whose buffer those lines depend on. */
int pac_graph_load_text(pac_ctx *ctx, const char *path, pac_graph **out);
/* Builds the graph from an undirected edge list on the device: both directions,
 * self-loops dropped, duplicates removed, rows sorted (the row building of
 * Graph.cpp:70-85, done by a device radix sort + unique + compaction).  num_vertices
 * = 0 takes max ID + 1.  The text loader above parses on the host and uses this. */
int pac_graph_from_edges(pac_ctx *ctx, const int32_t *src, const int32_t *dst, uint64_t ne,
                         uint64_t num_vertices, pac_graph **out);
/* Replaces Graph::save_as_csr_binfile, Graph.cpp:595-605 (same two files). */
int pac_graph_save_csr_dir(const pac_graph *g, const char *csrdir);
/* Seeded R-MAT generator on the GPU: ef*2^scale samples with quadrant probabilities
 * (a,b,c,1-a-b-c), no noise, no label permutation, self-loops and duplicates
 * removed, symmetrized.  No reference counterpart (the reference reads files). */
int pac_graph_generate_rmat(pac_ctx *ctx, uint32_t scale, uint32_t edge_factor,
                            double a, double b, double c, uint64_t seed,
                            pac_graph **out);
/* Seeded Erdos-Renyi G(n,p), p = avg_degree/(n-1). No reference counterpart. */
int pac_graph_generate_er(pac_ctx *ctx, uint64_t n, double avg_degree, uint64_t seed,
                          pac_graph **out);
/* Copies the symmetric CSR back in the reference's types (Graph::get_csr_vlist /
 * get_csr_elist, Graph.cpp:151-157).  vlist: V+1, elist: nnz. */
int pac_graph_download_csr(const pac_graph *g, uint64_t *vlist, int32_t *elist);
int pac_graph_get_info(const pac_graph *g, pac_graph_info *out);
/* Structural check of the uploaded graph on the GPU; level 1 adds the symmetry
 * check (every (u,v) has (v,u)). */
int pac_graph_validate(const pac_graph *g, int level);
void pac_graph_free(pac_graph *g);

/* ---- partitioning (L3) ---------------------------------------------------------- */
/* Replaces Partitioned_Graph::Partitioned_Graph(g, vertex2partition, threads),
 * Graph.cpp:1017-1069: builds the degree-ordered oriented CSR on the GPU
 * (histogram / scan / scatter into uint32 CSR), groups rows by partition and builds
 * the persistent work queue.  v2p (V int32) assigns each vertex to a partition, as
 * the reference's P<n>.bin does (test.cpp:242-243); NULL selects contiguous-ID
 * partitions part(v) = v / ceil(V/num_partitions).  P = max(num_partitions,
 * max(v2p)+1) (the reference takes max(v2p)+1, Graph.cpp:1021). */
int pac_graph_prepare(pac_graph *g, uint32_t num_partitions, const int32_t *v2p);
/* Per-partition oriented edge counts (P entries) after prepare. */
int pac_graph_partition_sizes(const pac_graph *g, uint64_t *edges_per_partition);
/* Graph::random_partition, Graph.cpp:564-570, with a seeded stream (host). */
int pac_random_partition(uint64_t V, uint32_t P, uint64_t seed, int32_t *v2p);
/* Graph::greedy_partition, Graph.cpp:898-967 (degree-descending,
 * deficit-proportional; seeded instead of random_device) (host). */
int pac_greedy_partition(const pac_graph *g, uint32_t P, uint64_t seed, int32_t *v2p);
/* Graph::greedy_partition_2, Graph.cpp:969-1015 (minimax of the per-partition neighbour
 * counts, vertices by degree descending, ties by ID; isolated vertices take the seeded
 * draws of pac_random_partition instead of rand()) (host, P x V counters). */
int pac_greedy_partition_2(const pac_graph *g, uint32_t P, uint64_t seed, int32_t *v2p);
/* Partitioned_Graph::evaluate_partition, Graph.cpp:1460-1473: returns the balance
 * ratio P*max_part_deg/max_deg for a vertex2partition. */
int pac_evaluate_partition(const pac_graph *g, const int32_t *v2p, uint32_t P,
                           double *balance_ratio);

/* ---- solvers (L4/L5) -------------------------------------------------------------- */
/* Replaces Graph::solve_triangle_counting_merge (Graph.cpp:355-368) and
 * Partitioned_Graph::solve_triangle_counting_merge (Graph.cpp:1346-1362): *total = T.
 * per_vertex (NULL or V uint64) receives t(u), the triangles through u.  The
 * reference's *_galloping variants return 6T (Graph.cpp:342-353, 1326-1344); the
 * C++ facade multiplies. */
int pac_tc(pac_graph *g, int dispatch, int launch, uint64_t *total, uint64_t *per_vertex);
/* Replaces Graph::solve_lcc_merge / _galloping (Graph.cpp:814-840) and the
 * Partitioned_Graph versions (Graph.cpp:2197-2261): lcc64[u] =
 * 2t(u)/(d(d-1)) in double (0.0 for d=0, NaN for d=1, as the reference), lcc32 =
 * (float)lcc64.  Either output may be NULL.  total (may be NULL) receives T. */
int pac_lcc(pac_graph *g, int dispatch, int launch, float *lcc32, double *lcc64,
            uint64_t *total);
/* Per-edge Jaccard over every undirected edge, in oriented-edge order (m entries: rows in
 * ID order, each row's edges in (degree, id) order of their heads):
 * src/dst (original IDs, src has the smaller (degree, id)), inter = |N(u)&N(v)|,
 * uni = d(u)+d(v)-inter, ratio = (float)((double)inter/(double)uni) -- the pair
 * formula of Graph::solve_jsim_merge (Graph.cpp:808-809) applied to each edge.  Any
 * output pointer may be NULL; when all are NULL the intersection counts (support),
 * unions and ratios stay in device memory. */
int pac_jaccard_edges(pac_graph *g, int dispatch, int launch, int32_t *src, int32_t *dst,
                      uint32_t *inter, uint32_t *uni, float *ratio);
/* Replaces Graph::solve_jsim_merge / _galloping (Graph.cpp:782-812): the n x n matrix
 * res[i*n+j] = c/((double)(d_i+d_j) - c) over the query vertices q[0..n). */
int pac_jaccard_queries(pac_graph *g, const int32_t *q, uint32_t n, double *res);
/* Replaces Graph::batch_cnt_common_neis_merge(pairs) (Graph.cpp:413-447):
 * res[i] = |N(u[i]) & N(v[i])| over the full symmetric lists. */
int pac_cnt_common_pairs(pac_graph *g, const int32_t *u, const int32_t *v, uint64_t n,
                         uint64_t *res);
/* Filtered pair counts: res[i] = #{w in N(u[i]) & N(v[i]) : w < thr[i], and part(w) ==
 * partition}.  thr may be NULL (no bound); partition = -1 means every partition, else the
 * graph must be prepared and 0 <= partition < P, part() being the map of the last
 * pac_graph_prepare.  Replaces Graph::cnt_common_neis_less_than_merge (Graph.cpp:272-285),
 * Partitioned_Graph::cnt_common_neis_{merge,galloping}(u, v, k) (Graph.cpp:1155-1173,
 * 1071-1085), Partitioned_Graph::cnt_common_neis_less_than_merge(u, v, thr, k)
 * (Graph.cpp:1132-1153) and, summed over k, Partitioned_Graph::batch_cnt_common_neis_*
 * (Graph.cpp:1382-1458). */
int pac_cnt_common_pairs_filtered(pac_graph *g, const int32_t *u, const int32_t *v,
                                  const int32_t *thr, uint64_t n, int32_t partition,
                                  uint64_t *res);
/* Common-neighbour lists, ascending, each value once: N(u) & N(v) when nodes is NULL,
 * else N(u) & nodes[0..n) (nodes sorted ascending, else PAC_ERR_ARG; values outside
 * [0, V) never match), restricted to `partition` as in pac_cnt_common_pairs_filtered.
 * *count receives the number found; out (capacity cap entries, may be NULL with cap 0
 * to ask for the count only) receives the list, PAC_ERR_ARG when it does not fit.
 * Replaces Graph::get_common_neis(u, v) and (u, nodes) (Graph.cpp:468-493),
 * Graph::cnt_common_neis_merge(u, nodes) (Graph.cpp:241-253), and the per-partition
 * Partitioned_Graph::get_common_neis, cnt_common_neis_merge(u, nodes, k) and
 * get_neis(u, k) (= u with v = u) (Graph.cpp:1175-1195, 1475-1498). */
int pac_common_neis(pac_graph *g, int32_t u, int32_t v, const int32_t *nodes, uint64_t n,
                    int32_t partition, int32_t *out, uint64_t cap, uint64_t *count);

/* ---- multi-way intersections ----------------------------------------------------- */
/* |N(q_0) & N(q_1) & ... & N(q_{k-1})| for nq queries in CSR form: query i is
 * qv[qoff[i] .. qoff[i+1]) (qoff[0] = 0, every query non-empty), res[i] its count.
 * Replaces Graph::batch_cnt_common_neis_{merge,galloping}(const vector<vector<VID>>&,
 * vector<size_t>&, int) (Graph.cpp:423-429, 441-447) and the single-query
 * Graph::cnt_common_neis_{merge,galloping}(const vector<VID>&) (Graph.cpp:200-225,
 * 255-270); both reference variants return the same count. */
int pac_cnt_common_kway(pac_graph *g, const uint64_t *qoff, const int32_t *qv, uint64_t nq,
                        uint64_t *res);
/* pac_cnt_common_kway restricted to the common neighbours in `partition` (-1: all), as in
 * pac_cnt_common_pairs_filtered.  Replaces Partitioned_Graph::cnt_common_neis_{merge,
 * galloping}(q, k) (Graph.cpp:1087-1130) and, summed over k, the k-way
 * Partitioned_Graph::batch_cnt_common_neis_* (Graph.cpp:1395-1406, 1434-1445). */
int pac_cnt_common_kway_filtered(pac_graph *g, const uint64_t *qoff, const int32_t *qv,
                                 uint64_t nq, int32_t partition, uint64_t *res);
/* Sum over index pairs i < j of nodes (neither equal to root) of
 * |N(root) & N(nodes[i]) & N(nodes[j])|.  Replaces Graph::solve_friend_triangle_{merge,
 * galloping} (Graph.cpp:309-338) and Partitioned_Graph::solve_friend_triangle_*
 * (Graph.cpp:1222-1264), whose per-partition partials sum to the same value. */
int pac_friend_triangle(pac_graph *g, int32_t root, const int32_t *nodes, uint64_t n,
                        uint64_t *res);
/* Partitioned_Graph::solve_square_counting (Graph.cpp:1291-1324): the sum over every u
 * and neighbours w < v < u of u of #{x < w : x ~ v, x ~ w} -- each 4-cycle whose
 * largest and smallest vertex IDs are opposite corners, counted once.  Independent of
 * the partitioning; needs no pac_graph_prepare. */
int pac_square_count(pac_graph *g, uint64_t *res);

/* ---- solve without allocation ---------------------------------------------------- */
/* Size of the static device array (one per GPU, held until the process ends) that
 * pac_tc_noalloc uses for its temporaries. */
#define PAC_NOALLOC_SCRATCH_BYTES (2u << 20)
/* No counterpart in the reference.  For callers that cannot allocate device memory while
 * they solve (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): pac_tc on the persistent kernel (the context's tile) with its temporaries
 * -- work-queue heads, graph descriptor, deferred-row queue -- in a static device array
 * of PAC_NOALLOC_SCRATCH_BYTES per GPU instead of library allocations; the call makes no
 * allocation and no host-device copy.  The graph must be prepared (pac_graph_prepare
 * allocates).  d_total (device memory, 1 uint64) receives T and d_per_vertex (device
 * memory, V uint64, or NULL) t(u): both are zeroed and written by kernels on the
 * context's stream, and the call returns without waiting for them.  dispatch: FIXED or a
 * FORCE_* mode; LEARNED only once a solve with PAC_DISPATCH_LEARNED has computed the
 * choice for the loaded model (PAC_ERR_STATE before; with no model loaded LEARNED means
 * FIXED).  PAC_ERR_UNSUPPORTED when the graph's deferred-row queue (work items plus rows
 * above 1024 out-neighbours, when its longest out-list exceeds the tile) would exceed the
 * array.  Every call on one GPU uses that GPU's one array, which keeps the last call's
 * contents: calls that overlap on one GPU (other contexts or threads) race. */
int pac_tc_noalloc(pac_graph *g, int dispatch, uint64_t *d_total, uint64_t *d_per_vertex);

/* ---- batching ------------------------------------------------------------------- */
/* Triangle counts of G prepared graphs in one persistent launch (or one hipGraph
 * replay when launch == PAC_LAUNCH_GRAPH): totals[i] = T of gs[i]; a graph may appear
 * only once (PAC_ERR_ARG otherwise).  No reference counterpart (the reference handles
 * one graph per process). */
int pac_tc_batch(pac_ctx *ctx, pac_graph *const *gs, uint32_t G, int dispatch, int launch,
                 uint64_t *totals);

/* ---- autotuning ------------------------------------------------------------------ */
/* Tile configurations (waves per row, LDS list capacity). */
#define PAC_NUM_TILES 4
/* What a solve accumulates: the total only (pac_tc without per-vertex counts, batches),
 * per-vertex counts (pac_tc with per_vertex, pac_lcc) or per-edge supports
 * (pac_jaccard_edges).  The intersection kernel's registers and LDS differ by mode. */
enum { PAC_MODE_TC = 0, PAC_MODE_LCC = 1, PAC_MODE_JACCARD = 2 };
/* No reference counterpart (the reference's knobs are the partition count and the
 * thread count, scripts/run.py:36, test.cpp:371-379).  Times the intersection launch of
 * the persistent kernel (HIP events around it, median of 3 after a warm-up) for every
 * tile configuration on the given graphs, in `mode` with `dispatch`, and keeps the
 * fastest for this context.  ms_out (PAC_NUM_TILES, may be NULL) receives the times.
 * cache_path (may be NULL) is read first: a line for the same graphs' shape, mode and
 * dispatch gives the tile and the times without timing anything; otherwise the result
 * is appended to it.  The context's tile is left unchanged when an error is returned.
 * Returns the chosen tile id in *tile. */
int pac_autotune(pac_ctx *ctx, pac_graph *const *gs, uint32_t G, int mode, int dispatch,
                 const char *cache_path, int *tile, double *ms_out);
int pac_set_tile(pac_ctx *ctx, int tile);
int pac_get_tile(const pac_ctx *ctx, int *tile);

/* ---- learned dispatch ------------------------------------------------------------ */
#define PAC_MLP_IN 16
#define PAC_MLP_HIDDEN 64
#define PAC_MLP_OUT 3
#define PAC_MLP_NPARAMS (PAC_MLP_HIDDEN * PAC_MLP_IN + PAC_MLP_HIDDEN + \
                         PAC_MLP_HIDDEN * PAC_MLP_HIDDEN + PAC_MLP_HIDDEN + \
                         PAC_MLP_OUT * PAC_MLP_HIDDEN + PAC_MLP_OUT)
/* Per-partition features (P x 16 floats) of a prepared graph. */
int pac_partition_features(pac_graph *g, float *features);
/* Times each strategy on each partition inside the persistent launch (LCC mode): in
 * launch r of a round partition k runs strategy (k + r) mod 3, and the kernel's per-item
 * clock stamps give each partition's time under each strategy; ms (P x 3) holds the
 * median over `reps` rounds, 0 for partitions owning no pairs.  These are the learned
 * dispatch's training labels. */
int pac_profile_partitions(pac_graph *g, int reps, double *ms);
/* The strategy the fixed heuristic picks per partition (P bytes). */
int pac_fixed_choice(pac_graph *g, uint8_t *choice);
/* The strategy the loaded MLP picks per partition (P bytes). */
int pac_mlp_choice(pac_graph *g, uint8_t *choice);
/* Set / get the MLP parameters (PAC_MLP_NPARAMS floats; stored on the device as
 * bf16, so get returns the bf16-rounded values). */
int pac_mlp_set_params(pac_ctx *ctx, const float *params);
int pac_mlp_get_params(const pac_ctx *ctx, float *params);
int pac_mlp_init(pac_ctx *ctx, uint64_t seed);
/* Forward pass on the GPU: logits (n x 3) for features x (n x 16). */
int pac_mlp_forward(pac_ctx *ctx, const float *x, uint64_t n, float *logits);
/* Mean (weighted) softmax cross-entropy and its gradient w.r.t. the parameters,
 * computed by the GPU backward kernels (weights may be NULL). */
int pac_mlp_loss_grad(pac_ctx *ctx, const float *x, const int32_t *labels,
                      const float *weights, uint64_t n, float *loss, float *grads);
/* Trains the MLP on the GPU with Adam for `epochs` full-batch steps; loss_out
 * (may be NULL) receives the final loss. */
int pac_mlp_train(pac_ctx *ctx, const float *x, const int32_t *labels,
                  const float *weights, uint64_t n, uint32_t epochs, float lr,
                  float *loss_out);
int pac_mlp_save(const pac_ctx *ctx, const char *path);
int pac_mlp_load(pac_ctx *ctx, const char *path);

#ifdef __cplusplus
}
#endif
#endif /* PAC_H */
