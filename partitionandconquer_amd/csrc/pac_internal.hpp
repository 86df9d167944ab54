// pac_internal.hpp -- shared declarations of the gfx950 library behind include/pac.h.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pac.h"

namespace pac {

void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

#define PAC_CHECK_HIP(call)                                                       \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ::pac::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,                 \
                       hipGetErrorString(e_));                                    \
      return e_ == hipErrorOutOfMemory ? PAC_ERR_OOM : PAC_ERR_HIP;               \
    }                                                                             \
  } while (0)

#define PAC_TRY(expr)         \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != PAC_OK) return rc_; \
  } while (0)

// Strategy ids (the enum order of the dispatch table and of the MLP outputs).
enum Strategy : uint8_t { kMerge = 0, kGallop = 1, kBitmap = 2 };
// What the intersection kernels accumulate.
enum Mode : int { kModeTC = 0, kModeLCC = 1, kModeSUP = 2 };
static_assert(kModeTC == PAC_MODE_TC && kModeLCC == PAC_MODE_LCC &&
                  kModeSUP == PAC_MODE_JACCARD,
              "the ABI's solve modes are the kernels' modes");

// Owning device buffer (hipMalloc / hipFree).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  // Grow-only (re)allocation; contents are not preserved.
  int reserve(size_t n) {
    if (n <= bytes && p) return PAC_OK;
    release();
    if (n == 0) n = 8;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc(%zu bytes) failed: %s", n, hipGetErrorString(e));
      return PAC_ERR_OOM;
    }
    bytes = n;
    return PAC_OK;
  }
  template <class T>
  T *as() const { return static_cast<T *>(p); }
};

// One graph as the persistent kernel sees it (device-side descriptor).
struct GraphDev {
  const uint32_t *ooff;  // V+1 oriented row offsets
  const uint32_t *oadj;  // m oriented neighbours, each row in (degree, id) order
  const uint32_t *okey;  // per oriented edge, its head's (degree, id) rank: each row's keys
                         // ascend, so the sorted strategies compare these, not the IDs
  const uint32_t *poff;  // V+1 offsets of the intersection pairs owned by each vertex
  const uint32_t *ppar;  // partner of each pair (its out-list is the one streamed)
  const uint32_t *pedge; // oriented-edge index of each pair (position in oadj)
  // per pair: the part of the partner's out-list it streams, as (first position in oadj,
  // length) -- read with the pair, so its bounds need no gather
  const uint2 *pinfo;
  // The lists the hash path streams and their bounds per pair: oadj with pinfo
  // (selem = 4), or, for a graph whose IDs are all below 2^24, its 3-byte lists opk with
  // pkinfo (selem = 3).  Pair p's stream starts at byte 4 * sinfo[p].x of sadj (element i
  // at + selem * i) and has sinfo[p].y elements.
  const uint8_t *sadj;
  const uint2 *sinfo;
  uint32_t selem;
  const uint32_t *rows;  // rows owning at least one pair, grouped by partition
  unsigned long long *total;  // T accumulator
  unsigned long long *tvert;  // LCC mode: per-vertex triangle counts, every credit added
                              // atomically by the intersection kernel
  // LCC mode (may be null: then these credits go to tvert): the pairs' credits to their
  // partners, i.e. per vertex the triangles in which it ranks lowest, as 32-bit atomics;
  // folded into tvert after the launch.  A vertex t is lowest in at most C(d+(t), 2) < m
  // triangles (d+(t)^2 <= the sum of its out-neighbours' degrees <= 2m), so they fit.
  uint32_t *tlow;
  uint32_t *sup;              // per-oriented-edge support (SUP mode)
  // SUP mode: the hit words of the streamed lists.  The pair with rank r among the pairs
  // streaming N+(y) owns the 64-bit words hitw[wbase[y] + r*nw ..+ nw), nw = ceil(d+(y)/64);
  // bit l of word c is set when element 64c + l of N+(y) was found in the owner's list.
  // The ranks follow the pairs' positions in N+(y), so the rows whose streams reach
  // column c (start at or before it) are the first ones.
  const unsigned long long *pword;  // per pair: its first hit word, wbase[y] + r*nw
  unsigned long long *hitw;
  uint32_t V;
  uint32_t m;                 // oriented edges; oadj[m] holds the sentinel kListEnd
  uint32_t part_base;         // offset of this graph's partitions in the dispatch table
};

// A unit of the persistent work queue: rows rows[lo..hi) of graph g, partition part;
// when hi == lo+1 only the pairs at positions [c_lo, c_hi) of the row.
struct WorkItem {
  uint32_t g, part, lo, hi, c_lo, c_hi, cost_hi, cost_lo;
};

// A row (pairs [c_lo, c_hi)) whose out-list exceeds the LDS capacity of the launch's
// tile, queued by the persistent kernel for a second pass with the widest tile.
struct DeferItem {
  uint32_t g, r, c_lo, c_hi, st, pad;
};
// The out-list length above which a row can exceed the smallest tile's capacity.
constexpr uint32_t kMinTileCap = 1024;

// Tile configurations: waves per row, LDS list capacity (entries).
struct TileCfg {
  int waves, cap;
};
extern const TileCfg kTiles[PAC_NUM_TILES];

struct Ctx;

struct Graph {
  Ctx *ctx = nullptr;
  uint32_t V = 0;
  uint64_t nnz = 0;  // 2m
  uint64_t m = 0;    // oriented edges
  uint64_t max_deg = 0, max_outdeg = 0, W = 0;
  // the pair work model (pac_graph_info), set by pac_graph_prepare
  uint64_t n_pairs = 0, n_rows = 0, streamed = 0, staged = 0, staged_items = 0;
  // symmetric CSR
  DevBuf off, adj, deg;
  // Oriented CSR.  Rows in ID order; each row's out-neighbours in (degree, id) order, with
  // okey[e] = the rank of oadj[e] in that order (vertex rank[v] = its position among all
  // vertices sorted by (degree, id)).
  DevBuf odeg, ooff, oadj, okey, rank;
  // When every ID is below 2^24, built with the pairs: the oriented lists with 3-byte
  // elements (little-endian), the list of vertex y from byte 4 ceil(3 (ooff[y] + y) / 4) --
  // every list starts on a 4-byte boundary and ends before the next one starts, since each
  // is given 3 bytes more than its length; and per pair, pkinfo = (4-byte word where its
  // stream starts, length): the suffix start rounded down to a multiple of 4 elements (a
  // word boundary of the list), the length grown by the up to 3 elements before it, which
  // rank at most h and so are never found in N+(h).  Empty otherwise.
  DevBuf opk, pkinfo;
  // Intersection pairs.  Each oriented edge t->h stands for N+(t) & N+(h), whose elements
  // all rank above h: the pair is owned by the head h, which stages N+(h) in LDS, and
  // streams only the suffix of N+(t) after h (N+(t) is in rank order, and h is in it).
  // Pairs with an empty suffix or an empty N+(h) are dropped (they hold no triangle).
  // pedge = the position of t->h in oadj; pinfo = (pedge + 1, suffix length).
  DevBuf poff, ppar, pedge, pinfo;
  // SUP mode streams from a 64-element boundary of N+(t) (so a block's hit word is a
  // column of the list): per pair (first position, length) with the start rounded down
  // to a multiple of 64 within the list.  Built with the hit words.
  DevBuf pinfo64;
  // Construction state, kept for the later passes: crow[c] = the row holding entry 256c;
  // oflags / pflags = one bit per symmetric entry (kept by the orientation / owned as a
  // pair), owpre / ppre = exclusive scans of their 64-bit words' popcounts.
  DevBuf crow, oflags, owpre, pflags, ppre;
  bool pairs_built = false;
  // Per-edge supports (SUP mode) from hit words instead of one atomic per triangle:
  // pstreams[y] = number of pairs streaming N+(y); wbase[y] = first word of y's streams;
  // prank[e] = the number of pairs' edges before position e of oadj (the rank of the pair
  // of edge e among its partner's pairs is prank[e] - prank[ooff[y]]); pword, hitw as in
  // GraphDev.
  DevBuf pstreams, pword, wbase, hitw, supbig, prank;
  std::vector<uint32_t> h_supbig;  // (y, first stream) units of kSupRun streams of the
                                   // vertices streamed by more than 64 pairs
  bool hitw_built = false;
  DevBuf dsum;  // per oriented edge t->h, d(t) + d(h) (Jaccard unions), built on first use
  // partitioning and work items
  bool prepared = false;
  uint32_t P = 0;
  std::vector<uint32_t> h_pstart;      // P+1: rows of partition k are rows[pstart[k]..pstart[k+1])
  std::vector<uint64_t> h_pedges;      // P: oriented edges per partition
  std::vector<float> h_features;       // P x 16
  std::vector<uint8_t> h_fixed;        // P: fixed-heuristic strategy
  std::vector<uint8_t> h_learned;      // P: MLP choice cached for mlp_version
  uint64_t learned_version = 0;
  std::vector<WorkItem> h_items;
  DevBuf rows, items;
  uint64_t n_long_rows = 0;  // rows owning pairs whose out-list exceeds kMinTileCap
  // Partition map of the last prepare: the caller's vertex2partition, or (empty)
  // contiguous ranges of part_width IDs.  Its device copy is built by partition_map().
  std::vector<int32_t> h_v2p;
  uint64_t part_width = 1;
  DevBuf part;
  bool part_ready = false;
  DevBuf rowwork;  // V: streamed length of row u = sum of d+(partner) over u's pairs
  // Strategy tables on the device, P bytes each: all merge, all gallop, all hash, the fixed
  // heuristic (uploaded by pac_graph_prepare), and the MLP's choice for dlearned_version
  // (uploaded when a learned solve computes it).  Read by pac_tc_noalloc.
  DevBuf dstrat, dlearned;
  uint64_t dlearned_version = 0;
  // outputs
  DevBuf total, tvert, tlow, sup, lcc32, lcc64, juni, jratio;
  double prepare_ms = 0;
};

struct Ctx {
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[6] = {};
  hipEvent_t kev[2] = {};     // around the dominant kernel when time_kernel is set
  bool time_kernel = false;
  // when set (pac_profile_partitions, never inside a captured graph), the persistent
  // kernel writes each item's start and end clock here: 2 x items entries
  unsigned long long *item_stamps = nullptr;
  pac_timing last{};
  double timeout_ms = 0;  // bound on waiting for a solve's completion (0: none)
  int packed_lists = PAC_PACKED_AUTO;  // when launches stream the 3-byte lists
  int tile = 1;
  void *noalloc_scratch = nullptr;  // this device's static scratch of pac_tc_noalloc
  DevBuf scratch;     // generic temporary storage (library-owned)
  DevBuf scratch2;
  DevBuf queue;       // persistent work-queue heads: [0] items, [1] deferred rows, [2] their queue
  DevBuf defer;       // DeferItem queue of the wide second pass
  DevBuf dispatch;    // per-partition strategy table
  DevBuf gdesc;       // GraphDev array
  DevBuf batch_items;
  std::vector<const void *> batch_key;  // graph set whose items are in batch_items
  // MLP
  DevBuf mlp_params;  // bf16 parameters on device
  bool mlp_loaded = false;
  uint64_t mlp_version = 0;   // bumped whenever the parameters change
  std::vector<float> mlp_host;  // float copy (bf16-rounded)
  std::vector<void *> graphs;  // live pac_graph handles, freed with the context
  // hipGraph cache for the batched persistent launch
  hipGraph_t hgraph = nullptr;
  hipGraphExec_t hexec = nullptr;
  std::vector<const void *> hgraph_key;
  int hgraph_mode = -1, hgraph_tile = -1;
};

// ---- prepare.hip -----------------------------------------------------------------
int build_symmetric_from_host(Graph *g, const uint64_t *vlist, const int32_t *elist);
int orient(Graph *g);
// Build the pair lists and each row's streamed length (rowwork) of an oriented graph
// (once; later calls return at once).
int build_pairs(Graph *g);
// What pac_graph_prepare needs of the rows that own pairs, from the device: the rows in ID
// order (also left in g->rows), per row (d+(r), streamed length, pairs, d(r)), W and the
// largest out-degree; with part_width > 0 also the oriented edges of each of the P
// contiguous partitions of part_width IDs.
struct RowSummary {
  std::vector<uint32_t> rows;
  std::vector<uint4> stats;
  std::vector<uint64_t> part_edges;
  uint64_t W = 0, max_outdeg = 0;
};
int row_summary(Graph *g, uint64_t part_width, uint32_t P, RowSummary &out);
// Value of the entry past the last oriented edge (oadj[m]): neither a vertex ID (IDs
// are < 2^31) nor the empty hash-table slot, so a lookup of it never hits.
constexpr uint32_t kListEnd = 0xFFFFFFFEu;
constexpr uint32_t kPackMaxV = 1u << 24;  // 3-byte lists: every ID below this
// PAC_PACKED_AUTO streams them when the launch's largest graph has this many edges: at
// R-MAT scale 20 (m = 1.6e7, 62 MB of 4-byte lists) packed and 4-byte launches took the
// same time, at scale 22 (m = 6.4e7) the packed TC was 9 % faster, and a batch of 256
// graphs of at most 4e6 edges ran 2.4 % slower packed.
constexpr uint64_t kPackMinEdges = 1ull << 24;
// The element width of the lists one launch streams, the same for all its graphs (its
// kernels are compiled for one): 3 for TC or LCC when every graph has its 3-byte lists and
// the context's mode allows (PAC_PACKED_AUTO: the largest graph has kPackMinEdges edges),
// else 4.
inline uint32_t launch_elem_bytes(const Ctx *ctx, Graph *const *gs, uint32_t G, int mode) {
  if (mode == kModeSUP || ctx->packed_lists == PAC_PACKED_OFF || G == 0) return 4u;
  uint64_t max_m = 0;
  for (uint32_t i = 0; i < G; ++i) {
    if (gs[i]->V > kPackMaxV || gs[i]->opk.p == nullptr || gs[i]->pkinfo.p == nullptr)
      return 4u;
    max_m = gs[i]->m > max_m ? gs[i]->m : max_m;
  }
  return ctx->packed_lists == PAC_PACKED_ALWAYS || max_m >= kPackMinEdges ? 3u : 4u;
}
// Vertices streamed by at most 64 pairs have their hit words summed by one wave (one
// load per word column); the others are split into units of kSupRun streams.
constexpr uint32_t kSupRun = 256;
// Build the pair ranks, word offsets and (zeroed) hit words of SUP mode (once).
int build_hit_words(Graph *g);
// sup[y->w] += number of pair streams of N+(y) that found w (after a SUP-mode launch).
int sup_from_hit_words(Ctx *ctx, Graph *g);
// tvert += tlow (after an LCC-mode launch).
int fold_low_credits(Ctx *ctx, Graph *g);
int graph_validate(const Graph *g, int level);

// ---- gen.hip ---------------------------------------------------------------------
int generate_rmat(Graph *g, uint32_t scale, uint32_t ef, double a, double b, double c,
                  uint64_t seed);
// Symmetric, deduplicated, self-loop-free CSR from an undirected edge list on the device.
// Parse a text edge list (n bytes, host memory) on the device with the reference's
// rules; on success d_src/d_dst hold the *ne kept edges with 0-based IDs and *V the vertex
// count.  *host_needed is set instead when a line can only be parsed with the reference's
// buffer discipline on the host (a line over 511 characters, or one ending in its first
// number).
int parse_text_device(Ctx *ctx, const char *text, uint64_t n, DevBuf &d_src, DevBuf &d_dst,
                      uint64_t *ne, uint64_t *V, bool *host_needed);
int graph_from_device_edges(Graph *g, const uint32_t *d_src, const uint32_t *d_dst, uint64_t ne,
                            uint32_t V);

// ---- isect.hip -------------------------------------------------------------------
// Allocate / zero the outputs of `mode` for each graph.
int alloc_outputs(Graph *const *gs, uint32_t G, int mode);
int zero_outputs(Ctx *ctx, Graph *const *gs, uint32_t G, int mode);
// Plain path: one launch per partition with the partition's strategy.
int launch_plain(Ctx *ctx, Graph *g, int mode, const uint8_t *strategy_per_part,
                 uint64_t *launches);
// Persistent path over the concatenated work items of gs.  Without use_graph the
// caller zeroes the outputs first; with use_graph the zeroing is part of the graph.
int launch_persistent(Ctx *ctx, Graph *const *gs, uint32_t G, int mode,
                      const uint8_t *strategy_per_part_concat, bool use_graph,
                      uint64_t *launches);
// The no-allocation scratch: its device address for ctx->device (hipGetSymbolAddress once
// per device; call with the device current).
int noalloc_scratch_init(Ctx *ctx);
// Persistent launch of a prepared graph in TC or LCC mode with its temporaries in the
// static scratch: zeroes and fills d_total (and d_tvert in LCC mode); allocates nothing.
int launch_persistent_noalloc(Ctx *ctx, Graph *g, int mode, const uint8_t *d_strat,
                              unsigned long long *d_total, unsigned long long *d_tvert);
// LCC mode: per-vertex triangle counts = row totals + per-edge credits (after a launch).
int epilogue_lcc(Ctx *ctx, Graph *g, bool want32, bool want64);
int epilogue_jaccard(Ctx *ctx, Graph *g, int32_t *src, int32_t *dst, uint32_t *inter,
                     uint32_t *uni, float *ratio);
// Common-neighbour counts of (u[i], v[i]); a common neighbour w counts only when
// w < thr[i] (d_thr null: no bound) and, with d_part not null, d_part[w] == pid.
int pair_counts(Ctx *ctx, Graph *g, const int32_t *d_u, const int32_t *d_v, uint64_t n,
                uint64_t *d_res, const int32_t *d_thr = nullptr, const uint32_t *d_part = nullptr,
                uint32_t pid = 0);
int query_matrix(Ctx *ctx, Graph *g, const int32_t *d_q, uint32_t nq, double *d_res);
int fill_u32(hipStream_t s, uint32_t *p, uint64_t n, uint32_t val);

// ---- multi.hip -------------------------------------------------------------------
// |N(q_0) & ... & N(q_k-1)| per query (query i = qv[qoff[i] .. qoff[i+1]), non-empty).
int kway_counts(Ctx *ctx, Graph *g, const uint64_t *d_qoff, const int32_t *d_qv, uint64_t nq,
                uint64_t *d_res, const uint32_t *d_part = nullptr, uint32_t pid = 0);
// The common neighbours of u and the sorted list d_list[0..n) (each value once, in
// ascending order) whose partition is pid when d_part is not null: written to d_out,
// their number to *d_count.
int common_list(Ctx *ctx, Graph *g, uint32_t u, const int32_t *d_list, uint64_t n,
                const uint32_t *d_part, uint32_t pid, int32_t *d_out,
                unsigned long long *d_count);
// The partition of every vertex under the map of the last prepare (built on first use).
int partition_map(Graph *g, const uint32_t **d_part);
// Friend triangles of root over nodes; d_c holds root_deg counters, d_res one u64.
int friend_triangles(Ctx *ctx, Graph *g, uint32_t root, const int32_t *d_nodes, uint64_t n,
                     uint32_t root_deg, uint32_t *d_c, uint64_t *d_res);
// The reference's square count (4-cycles with the extreme IDs at opposite corners).
int square_count(Ctx *ctx, Graph *g, uint64_t *h_res);
int stream_copy(Ctx *ctx, uint64_t bytes, int reps, double *ms);

// ---- mlp.hip ---------------------------------------------------------------------
int mlp_forward_dev(Ctx *ctx, const float *d_x, uint64_t n, float *d_logits);
int mlp_loss_grad_dev(Ctx *ctx, const float *d_x, const int32_t *d_labels,
                      const float *d_w, uint64_t n, float *d_loss, float *d_grads);
int mlp_adam_step(Ctx *ctx, float *d_master, float *d_m, float *d_v, const float *d_g,
                  float lr, int step);
int mlp_store_bf16(Ctx *ctx, const float *d_master);

// ---- host helpers ----------------------------------------------------------------
int fixed_heuristic(const Graph *g, std::vector<uint8_t> &out);
// One batched persistent solve of `mode` over G prepared graphs (launch PERSISTENT or
// GRAPH); totals (may be null) receives each graph's T.
int solve_batch(Ctx *c, pac_graph *const *hs, uint32_t G, int mode, int dispatch, int launch,
                uint64_t *totals);
// Per-partition features and the fixed heuristic's choice, from the per-row summaries
// (d+(r), streamed length, pairs, d(r)) in the order of the grouped row list.
void partition_features(Graph *g, const std::vector<uint4> &stats);

}  // namespace pac

// The opaque handles of include/pac.h.
struct pac_ctx {
  pac::Ctx c;
};
struct pac_graph {
  pac::Graph g;
};
