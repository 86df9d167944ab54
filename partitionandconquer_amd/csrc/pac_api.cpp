// pac_api.cpp -- host side of the C ABI (include/pac.h): contexts, graph lifecycle,
// file formats, partitioning, the work-queue builder, strategy dispatch and the
// solver entry points.  Kernels live in the .hip files.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "pac_internal.hpp"
#include "pac_wait.hpp"

namespace pac {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;
inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Target work (streamed elements plus overheads) of one persistent work item.  Rows
// are packed until an item reaches it -- about sixteen items per resident workgroup (six
// per CU), which keeps the queue's tail short -- and a row above twice the target is split
// by pair range into pieces of about twice it (each piece stages the row again and
// flushes its own counters, so rows are split no finer than at eight items per resident
// workgroup).  Against eight items per workgroup for both: per-edge Jaccard at R-MAT
// scale 25 137-139 -> 122-123 ms, TC at scale 24 25.0 -> 24.7 ms, LCC at scales 22 and 24
// unchanged; splitting at the finer target as well cost LCC at scale 22 8.3 -> 9.3 ms.
uint64_t item_target(uint64_t total_cost, int num_cus) {
  const char *e = getenv("PAC_ITEM_WORK");
  if (e) {
    const long long v = atoll(e);
    if (v >= 256) return (uint64_t)v;
  }
  const uint64_t slots = (uint64_t)(num_cus > 0 ? num_cus : 1) * 6 * 16;
  return std::max<uint64_t>(1ull << 15, total_cost / slots);
}

int ensure_ctx(pac_ctx *ctx) {
  if (!ctx) {
    set_error("null context");
    return PAC_ERR_ARG;
  }
  PAC_CHECK_HIP(hipSetDevice(ctx->c.device));
  return PAC_OK;
}

int new_graph(pac_ctx *ctx, pac_graph **out) {
  if (!out) {
    set_error("null output handle");
    return PAC_ERR_ARG;
  }
  pac_graph *h = new (std::nothrow) pac_graph;
  if (!h) {
    set_error("host allocation failed");
    return PAC_ERR_OOM;
  }
  h->g.ctx = &ctx->c;
  ctx->c.graphs.push_back(h);
  *out = h;
  return PAC_OK;
}

// Unregister and delete a graph (device buffers are released by its destructors).
void delete_graph(pac_graph *h) {
  Ctx *ctx = h->g.ctx;
  auto &v = ctx->graphs;
  v.erase(std::remove(v.begin(), v.end(), (void *)h), v.end());
  delete h;
}

// Host validation of the boundary offsets, then upload, device validation and the
// degree-ordered orientation.
int upload_csr(pac_ctx *ctx, const uint64_t *vlist, uint64_t V, const int32_t *elist,
               uint64_t nnz, pac_graph **out) {
  PAC_TRY(ensure_ctx(ctx));
  if (!vlist || (nnz && !elist)) {
    set_error("null CSR array");
    return PAC_ERR_ARG;
  }
  if (V >= (1ull << 31)) {
    set_error("%llu vertices exceed the int32 vertex ID range", (unsigned long long)V);
    return PAC_ERR_UNSUPPORTED;
  }
  if (nnz >= (1ull << 32)) {
    set_error("%llu entries exceed the uint32 CSR limit", (unsigned long long)nnz);
    return PAC_ERR_UNSUPPORTED;
  }
  if (vlist[0] != 0 || vlist[V] != nnz) {
    set_error("offsets must start at 0 and end at nnz (%llu != %llu)",
              (unsigned long long)vlist[V], (unsigned long long)nnz);
    return PAC_ERR_FORMAT;
  }
  for (uint64_t u = 0; u < V; ++u)
    if (vlist[u + 1] < vlist[u]) {
      set_error("offsets decrease at vertex %llu", (unsigned long long)u);
      return PAC_ERR_FORMAT;
    }
  if (nnz & 1ull) {
    set_error("odd number of directed entries: the CSR is not symmetric");
    return PAC_ERR_FORMAT;
  }
  pac_graph *h = nullptr;
  PAC_TRY(new_graph(ctx, &h));
  Graph *g = &h->g;
  g->V = (uint32_t)V;
  g->nnz = nnz;
  int rc = build_symmetric_from_host(g, vlist, elist);
  if (rc == PAC_OK) rc = graph_validate(g, 0);
  if (rc == PAC_OK) rc = orient(g);
  if (rc != PAC_OK) {
    delete_graph(h);
    return rc;
  }
  *out = h;
  return PAC_OK;
}

// Upload an undirected edge list and build the symmetric CSR on the device.
int edges_to_graph(pac_ctx *ctx, const uint32_t *src, const uint32_t *dst, uint64_t ne,
                   uint64_t V, pac_graph **out) {
  if (2 * ne >= (1ull << 32)) {
    set_error("%llu edges exceed the uint32 CSR limit", (unsigned long long)ne);
    return PAC_ERR_UNSUPPORTED;
  }
  pac_graph *h = nullptr;
  PAC_TRY(new_graph(ctx, &h));
  DevBuf ds, dd;
  int rc = ds.reserve(4 * (ne ? ne : 1));
  if (rc == PAC_OK) rc = dd.reserve(4 * (ne ? ne : 1));
  if (rc == PAC_OK && ne) {
    if (hipMemcpy(ds.p, src, 4 * ne, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dd.p, dst, 4 * ne, hipMemcpyHostToDevice) != hipSuccess) {
      set_error("edge upload failed");
      rc = PAC_ERR_HIP;
    }
  }
  if (rc == PAC_OK)
    rc = graph_from_device_edges(&h->g, ds.as<uint32_t>(), dd.as<uint32_t>(), ne, (uint32_t)V);
  if (rc == PAC_OK) rc = orient(&h->g);
  if (rc != PAC_OK) {
    delete_graph(h);
    return rc;
  }
  *out = h;
  return PAC_OK;
}

int finish_generated(pac_ctx *ctx, pac_graph *h, pac_graph **out) {
  int rc = orient(&h->g);
  if (rc != PAC_OK) {
    delete_graph(h);
    return rc;
  }
  (void)ctx;
  *out = h;
  return PAC_OK;
}

template <class T>
int read_file(const std::string &path, std::vector<T> &v) {
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) {
    set_error("cannot open %s", path.c_str());
    return PAC_ERR_IO;
  }
  in.seekg(0, std::ios::end);
  const std::streamoff size = in.tellg();
  in.seekg(0, std::ios::beg);
  if (size < 0 || size % (std::streamoff)sizeof(T) != 0) {
    set_error("%s: size %lld is not a multiple of %zu", path.c_str(), (long long)size, sizeof(T));
    return PAC_ERR_IO;
  }
  v.resize((size_t)size / sizeof(T));
  if (size > 0 && !in.read(reinterpret_cast<char *>(v.data()), size)) {
    set_error("%s: short read", path.c_str());
    return PAC_ERR_IO;
  }
  return PAC_OK;
}

template <class T>
int write_file(const std::string &path, const T *p, size_t n) {
  std::ofstream out(path, std::ios::binary);
  if (!out.is_open()) {
    set_error("cannot create %s", path.c_str());
    return PAC_ERR_IO;
  }
  out.write(reinterpret_cast<const char *>(p), (std::streamsize)(n * sizeof(T)));
  if (!out) {
    set_error("%s: write failed", path.c_str());
    return PAC_ERR_IO;
  }
  return PAC_OK;
}

int download_csr(const Graph *g, std::vector<uint64_t> &vl, std::vector<int32_t> &el) {
  PAC_CHECK_HIP(hipStreamSynchronize(g->ctx->stream));
  std::vector<uint32_t> off(g->V + 1ull);
  PAC_CHECK_HIP(hipMemcpy(off.data(), g->off.p, sizeof(uint32_t) * off.size(),
                          hipMemcpyDeviceToHost));
  vl.assign(off.begin(), off.end());
  el.resize(g->nnz);
  if (g->nnz)
    PAC_CHECK_HIP(hipMemcpy(el.data(), g->adj.p, sizeof(int32_t) * g->nnz, hipMemcpyDeviceToHost));
  return PAC_OK;
}

template <class T>
int d2h(Ctx *ctx, std::vector<T> &h, const DevBuf &d, size_t n) {
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  h.resize(n);
  if (n) PAC_CHECK_HIP(hipMemcpy(h.data(), d.p, sizeof(T) * n, hipMemcpyDeviceToHost));
  return PAC_OK;
}

// Strategy per partition for a dispatch mode.
int choose_strategies(Graph *g, int dispatch, std::vector<uint8_t> &out) {
  out.assign(g->P, (uint8_t)kBitmap);
  switch (dispatch) {
    case PAC_DISPATCH_FIXED: out = g->h_fixed; return PAC_OK;
    case PAC_FORCE_MERGE: std::fill(out.begin(), out.end(), (uint8_t)kMerge); return PAC_OK;
    case PAC_FORCE_GALLOP: std::fill(out.begin(), out.end(), (uint8_t)kGallop); return PAC_OK;
    case PAC_FORCE_BITMAP: std::fill(out.begin(), out.end(), (uint8_t)kBitmap); return PAC_OK;
    case PAC_DISPATCH_LEARNED: {
      if (!g->ctx->mlp_loaded) {
        out = g->h_fixed;  // documented fallback: no model loaded
        return PAC_OK;
      }
      if (g->learned_version != g->ctx->mlp_version || g->h_learned.size() != g->P) {
        pac_graph *h =
            reinterpret_cast<pac_graph *>(reinterpret_cast<char *>(g) - offsetof(pac_graph, g));
        g->h_learned.assign(g->P, (uint8_t)kBitmap);
        PAC_TRY(pac_mlp_choice(h, g->h_learned.data()));
        g->learned_version = g->ctx->mlp_version;
        PAC_TRY(g->dlearned.reserve(g->P));
        PAC_CHECK_HIP(hipMemcpy(g->dlearned.p, g->h_learned.data(), g->P, hipMemcpyHostToDevice));
        g->dlearned_version = g->learned_version;
      }
      out = g->h_learned;
      return PAC_OK;
    }
    default:
      set_error("unknown dispatch mode %d", dispatch);
      return PAC_ERR_ARG;
  }
}

int check_solve_args(pac_graph *h, int dispatch, int launch) {
  if (!h) {
    set_error("null graph");
    return PAC_ERR_ARG;
  }
  if (!h->g.prepared) {
    set_error("graph not prepared: call pac_graph_prepare first");
    return PAC_ERR_STATE;
  }
  if (dispatch < PAC_DISPATCH_FIXED || dispatch > PAC_FORCE_BITMAP) {
    set_error("unknown dispatch mode %d", dispatch);
    return PAC_ERR_ARG;
  }
  if (launch < PAC_LAUNCH_PLAIN || launch > PAC_LAUNCH_GRAPH) {
    set_error("unknown launch mode %d", launch);
    return PAC_ERR_ARG;
  }
  PAC_CHECK_HIP(hipSetDevice(h->g.ctx->device));
  return PAC_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

// The timed region of every solve: zero outputs, intersection, epilogue.
struct Timed {
  Ctx *ctx;
  bool kernel_timed = false;
  explicit Timed(Ctx *c) : ctx(c) { ctx->last = pac_timing{}; }
  int mark(int i) {
    PAC_CHECK_HIP(hipEventRecord(ctx->ev[i], ctx->stream));
    return PAC_OK;
  }
  int finish(bool kernel_timed = false) {
    if (ctx->timeout_ms > 0) {
      // bounded wait: the solve's last event, polled until the context's deadline
      const int rc = wait_bounded(
          [&]() -> int {
            const hipError_t e = hipEventQuery(ctx->ev[3]);
            if (e == hipSuccess) return 1;
            if (e == hipErrorNotReady) return 0;
            set_error("solve failed: %s", hipGetErrorString(e));
            return PAC_ERR_HIP;
          },
          ctx->timeout_ms, [](unsigned us) { std::this_thread::sleep_for(std::chrono::microseconds(us)); });
      if (rc == PAC_ERR_TIMEOUT)
        set_error("the solve did not complete within %.1f ms; its kernels keep running on the "
                  "context's stream, which later calls wait behind",
                  ctx->timeout_ms);
      if (rc != PAC_OK) return rc;
    }
    PAC_CHECK_HIP(hipEventSynchronize(ctx->ev[3]));
    ctx->last.total_ms = elapsed(ctx->ev[0], ctx->ev[3]);
    ctx->last.isect_ms = elapsed(ctx->ev[1], ctx->ev[2]);
    ctx->last.epilogue_ms = elapsed(ctx->ev[2], ctx->ev[3]);
    ctx->last.kernel_ms = (kernel_timed || this->kernel_timed) ? elapsed(ctx->kev[0], ctx->kev[1]) : 0.0;
    return PAC_OK;
  }
};

// Zeroing + intersection for one graph with the chosen launch path.
int run_intersection(Graph *g, int mode, int dispatch, int launch, Timed &t) {
  Ctx *ctx = g->ctx;
  std::vector<uint8_t> strat;
  PAC_TRY(choose_strategies(g, dispatch, strat));
  for (uint8_t s : strat) ctx->last.strategy_partitions[s < 3 ? s : 2]++;
  Graph *gs[1] = {g};
  if (mode == kModeSUP) PAC_TRY(build_hit_words(g));  // once per graph, outside the timing
  PAC_TRY(t.mark(0));
  if (launch != PAC_LAUNCH_GRAPH) PAC_TRY(zero_outputs(ctx, gs, 1, mode));
  PAC_TRY(t.mark(1));
  uint64_t launches = 0;
  if (launch == PAC_LAUNCH_PLAIN) {
    PAC_TRY(launch_plain(ctx, g, mode, strat.data(), &launches));
  } else {
    ctx->time_kernel = (launch == PAC_LAUNCH_PERSISTENT) && !g->h_items.empty();
    const int rc = launch_persistent(ctx, gs, 1, mode, strat.data(), launch == PAC_LAUNCH_GRAPH,
                                     &launches);
    t.kernel_timed = ctx->time_kernel;
    ctx->time_kernel = false;
    PAC_TRY(rc);
  }
  if (mode == kModeSUP) PAC_TRY(sup_from_hit_words(ctx, g));
  if (mode == kModeLCC) PAC_TRY(fold_low_credits(ctx, g));
  ctx->last.launches = launches;
  PAC_TRY(t.mark(2));
  return PAC_OK;
}

// Graph::greedy_partition, Graph.cpp:898-967, with a seeded stream: vertices in
// degree-descending order; each vertex's still unassigned neighbours are spread over
// the partitions in proportion to how far each is below deg/P of this vertex's
// already-assigned neighbours; leftovers and untouched vertices go to random
// partitions.
void greedy_partition_host(const std::vector<uint64_t> &vl, const std::vector<int32_t> &el,
                           uint64_t V, uint32_t P, uint64_t seed, int32_t *v2p) {
  std::vector<uint32_t> order(V);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return vl[a + 1] - vl[a] > vl[b + 1] - vl[b];
  });
  std::fill(v2p, v2p + V, -1);
  uint64_t st = mix64(seed + kGamma);
  auto next = [&]() {
    st += kGamma;
    return mix64(st);
  };
  std::vector<int64_t> cnt(P), quota(P);
  std::vector<uint32_t> todo;
  for (uint32_t u : order) {
    std::fill(cnt.begin(), cnt.end(), 0);
    todo.clear();
    for (uint64_t i = vl[u]; i < vl[u + 1]; ++i) {
      const int32_t v = el[i];
      if (v2p[v] > -1)
        cnt[v2p[v]]++;
      else
        todo.push_back((uint32_t)v);
    }
    if (todo.empty()) continue;
    for (size_t i = todo.size(); i > 1; --i) std::swap(todo[i - 1], todo[next() % i]);
    const int64_t lb = (int64_t)((vl[u + 1] - vl[u]) / P);
    int64_t sum = 0;
    for (uint32_t k = 0; k < P; ++k) {
      quota[k] = std::max<int64_t>(0, lb - cnt[k]);
      sum += quota[k];
    }
    size_t ptr = 0;
    if (sum > 0)
      for (uint32_t k = 0; k < P; ++k) {
        const size_t len = (size_t)(quota[k] * (int64_t)todo.size() / sum);
        for (size_t j = 0; j < len && ptr < todo.size(); ++j) v2p[todo[ptr++]] = (int32_t)k;
      }
    for (; ptr < todo.size(); ++ptr) v2p[todo[ptr]] = (int32_t)(next() % P);
  }
  for (uint64_t u = 0; u < V; ++u)
    if (v2p[u] == -1) v2p[u] = (int32_t)(next() % P);
}

// Graph::greedy_partition_2, Graph.cpp:969-1015: vertices by degree, descending (ties by
// ID: the reference's std::sort leaves their order to the library); a vertex with
// neighbours joins the first partition k whose largest count[k][v] over its neighbours v is
// the smallest (count[k][v] = how many of v's neighbours k already holds), then adds 1 to
// count[k][v] for each neighbour; an isolated vertex u takes a seeded draw
// (mix64(key + u * gamma) mod P, the stream of pac_random_partition) instead of rand().
// The counts are P x V, as in the reference.
int greedy_partition_2_host(const std::vector<uint64_t> &vl, const std::vector<int32_t> &el,
                            uint64_t V, uint32_t P, uint64_t seed, int32_t *v2p) {
  std::vector<uint32_t> order(V);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return vl[a + 1] - vl[a] > vl[b + 1] - vl[b];
  });
  std::vector<uint32_t> cnt;
  try {
    cnt.assign((size_t)P * V, 0u);
  } catch (const std::bad_alloc &) {
    set_error("greedy_partition_2 needs %llu x %llu counters", (unsigned long long)P,
              (unsigned long long)V);
    return PAC_ERR_OOM;
  }
  const uint64_t key = mix64(seed + kGamma);
  for (uint32_t u : order) {
    int64_t part = -1;
    if (vl[u + 1] > vl[u]) {
      uint64_t best = V;  // the reference starts the minimum at vertex_num
      for (uint32_t k = 0; k < P; ++k) {
        const uint32_t *ck = &cnt[(size_t)k * V];
        uint64_t mx = 0;
        for (uint64_t i = vl[u]; i < vl[u + 1]; ++i) mx = std::max<uint64_t>(mx, ck[el[i]]);
        if (mx < best) {
          best = mx;
          part = k;
        }
      }
    }
    if (part < 0) part = (int64_t)(mix64(key + (uint64_t)u * kGamma) % P);
    v2p[u] = (int32_t)part;
    uint32_t *cp = &cnt[(size_t)part * V];
    for (uint64_t i = vl[u]; i < vl[u + 1]; ++i) cp[el[i]]++;
  }
  return PAC_OK;
}

}  // namespace

// ---- fixed heuristic and features ---------------------------------------------------
// Feature vector of a partition (16 floats, log-scaled then divided by 16).  Per row x
// (stats[q] of the row at position q of the grouped row list): a = d+(x) (the list staged
// in LDS), p = x's pairs, w = the total length streamed for them, d = d(x); mean_b = w/p
// per pair.
void partition_features(Graph *g, const std::vector<uint4> &stats) {
  g->h_features.assign((size_t)g->P * PAC_MLP_IN, 0.f);
  g->h_fixed.assign(g->P, (uint8_t)kBitmap);
  for (uint32_t k = 0; k < g->P; ++k) {
    double n = 0, sa = 0, sa2 = 0, maxa = 0, sd = 0, maxd = 0, sp = 0, sap = 0, sb = 0,
           maxw = 0, big = 0, small = 0, gal = 0;
    for (uint32_t q = g->h_pstart[k]; q < g->h_pstart[k + 1]; ++q) {
      const double a = stats[q].x, p = stats[q].z, w = stats[q].y, d = stats[q].w;
      n += 1;
      sa += a;
      sa2 += a * a;
      maxa = std::max(maxa, a);
      sd += d;
      maxd = std::max(maxd, d);
      sp += p;
      sap += a * p;
      sb += w;
      maxw = std::max(maxw, w);
      big += a > 64 ? 1 : 0;
      small += a <= 4 ? 1 : 0;
      const double mb = p > 0 ? w / p : 0;
      gal += p * std::min(a, mb) * std::log2(2.0 + std::max(a, mb));
    }
    const double mean_a = n ? sa / n : 0, mean_b = sp ? sb / sp : 0;
    const double var_a = n ? std::max(0.0, sa2 / n - mean_a * mean_a) : 0;
    const double wmean_a = sp ? sap / sp : 0;  // staged length seen by the average pair
    auto L = [](double x) { return (float)(std::log2(1.0 + x) / 16.0); };
    float *f = &g->h_features[(size_t)k * PAC_MLP_IN];
    f[0] = L(n);
    f[1] = L(sp);
    f[2] = L(sb);
    f[3] = L(mean_a);
    f[4] = L(std::sqrt(var_a));
    f[5] = L(maxa);
    f[6] = L(n ? sd / n : 0);
    f[7] = L(maxd);
    f[8] = L(mean_b);
    f[9] = (float)((std::log2(1.0 + mean_b) - std::log2(1.0 + wmean_a)) / 16.0);
    f[10] = L(wmean_a);
    f[11] = n ? (float)(big / n) : 0.f;
    f[12] = n ? (float)(small / n) : 0.f;
    f[13] = L(sap + sb);
    f[14] = L(gal);
    f[15] = L(maxw);
    // Fixed degree-threshold heuristic: partitions whose pairs see long staged lists
    // (pair-weighted mean out-degree >= 32) use the LDS hash; strongly skewed ones
    // (staged and streamed lists 8x apart) use galloping; the rest merge.
    uint8_t s = kMerge;
    if (wmean_a >= 32.0)
      s = kBitmap;
    else if (mean_b >= 8.0 * std::max(1.0, wmean_a) || wmean_a >= 8.0 * std::max(1.0, mean_b))
      s = kGallop;
    g->h_fixed[k] = s;
  }
}

int fixed_heuristic(const Graph *g, std::vector<uint8_t> &out) {
  out = g->h_fixed;
  return PAC_OK;
}

// One persistent launch (or hipGraph replay) over the concatenated work items of G
// prepared graphs of this context, accumulating `mode`; totals (may be null) receives
// each graph's T.  The intersection launch is timed with events unless replayed.
int solve_batch(Ctx *c, pac_graph *const *hs, uint32_t G, int mode, int dispatch, int launch,
                uint64_t *totals) {
  if (!hs || G == 0 || launch == PAC_LAUNCH_PLAIN || launch > PAC_LAUNCH_GRAPH ||
      mode < kModeTC || mode > kModeSUP) {
    set_error("need graphs, a batched launch mode and a solve mode");
    return PAC_ERR_ARG;
  }
  std::vector<Graph *> gs(G);
  std::vector<uint8_t> strat;
  {
    // each graph's count accumulates in its own device total: a graph listed twice would
    // receive both counts
    std::vector<const pac_graph *> seen(hs, hs + G);
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end()) {
      set_error("a graph appears more than once in the batch");
      return PAC_ERR_ARG;
    }
  }
  for (uint32_t i = 0; i < G; ++i) {
    if (!hs[i] || !hs[i]->g.prepared || hs[i]->g.ctx != c) {
      set_error("graph %u is null, unprepared or from another context", i);
      return PAC_ERR_STATE;
    }
    gs[i] = &hs[i]->g;
    std::vector<uint8_t> s;
    PAC_TRY(choose_strategies(gs[i], dispatch, s));
    strat.insert(strat.end(), s.begin(), s.end());
    if (mode == kModeSUP) PAC_TRY(build_hit_words(gs[i]));  // once per graph, untimed
  }
  PAC_CHECK_HIP(hipSetDevice(c->device));
  Timed t(c);
  PAC_TRY(t.mark(0));
  if (launch != PAC_LAUNCH_GRAPH) PAC_TRY(zero_outputs(c, gs.data(), G, mode));
  PAC_TRY(t.mark(1));
  uint64_t launches = 0;
  c->time_kernel = (launch == PAC_LAUNCH_PERSISTENT);
  const int rc = launch_persistent(c, gs.data(), G, mode, strat.data(),
                                   launch == PAC_LAUNCH_GRAPH, &launches);
  t.kernel_timed = c->time_kernel;
  c->time_kernel = false;
  PAC_TRY(rc);
  if (mode == kModeSUP)
    for (uint32_t i = 0; i < G; ++i) PAC_TRY(sup_from_hit_words(c, gs[i]));
  if (mode == kModeLCC)
    for (uint32_t i = 0; i < G; ++i) PAC_TRY(fold_low_credits(c, gs[i]));
  PAC_TRY(t.mark(2));
  PAC_TRY(t.mark(3));
  PAC_TRY(t.finish());
  c->last.launches = launches;
  if (totals)
    for (uint32_t i = 0; i < G; ++i)
      PAC_CHECK_HIP(hipMemcpy(&totals[i], gs[i]->total.p, 8, hipMemcpyDeviceToHost));
  return PAC_OK;
}

}  // namespace pac

using namespace pac;

extern "C" {

const char *pac_last_error(void) { return g_err.c_str(); }

int pac_ctx_create(int device, pac_ctx **out) {
  if (!out) {
    set_error("null output handle");
    return PAC_ERR_ARG;
  }
  int n = 0;
  PAC_CHECK_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    set_error("device %d not present (%d devices)", device, n);
    return PAC_ERR_ARG;
  }
  PAC_CHECK_HIP(hipSetDevice(device));
  pac_ctx *c = new (std::nothrow) pac_ctx;
  if (!c) {
    set_error("host allocation failed");
    return PAC_ERR_OOM;
  }
  c->c.device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->c.stream, hipStreamDefault) != hipSuccess) {
    set_error("device %d: cannot query properties or create a stream", device);
    delete c;
    return PAC_ERR_HIP;
  }
  c->c.num_cus = prop.multiProcessorCount;
  for (auto &e : c->c.kev)
    if (hipEventCreate(&e) != hipSuccess) {
      set_error("hipEventCreate failed");
      pac_ctx_destroy(c);
      return PAC_ERR_HIP;
    }
  for (auto &e : c->c.ev)
    if (hipEventCreate(&e) != hipSuccess) {
      set_error("hipEventCreate failed");
      pac_ctx_destroy(c);
      return PAC_ERR_HIP;
    }
  const char *t = getenv("PAC_TILE");
  if (t) c->c.tile = std::max(0, std::min(PAC_NUM_TILES - 1, atoi(t)));
  // the device is current: find its copy of the no-allocation scratch
  const int rc = noalloc_scratch_init(&c->c);
  if (rc != PAC_OK) {
    pac_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return PAC_OK;
}

void pac_ctx_destroy(pac_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
  // graphs still alive belong to this context: free them now (their handles die here)
  std::vector<void *> live;
  live.swap(ctx->c.graphs);
  for (void *p : live) delete static_cast<pac_graph *>(p);
  if (ctx->c.hexec) (void)hipGraphExecDestroy(ctx->c.hexec);
  if (ctx->c.hgraph) (void)hipGraphDestroy(ctx->c.hgraph);
  for (auto &e : ctx->c.ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : ctx->c.kev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
}

int pac_ctx_last_timing(const pac_ctx *ctx, pac_timing *out) {
  if (!ctx || !out) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  *out = ctx->c.last;
  return PAC_OK;
}

int pac_ctx_num_cus(const pac_ctx *ctx) { return ctx ? ctx->c.num_cus : 0; }

int pac_ctx_set_timeout(pac_ctx *ctx, double ms) {
  if (!ctx || !(ms >= 0)) {
    set_error("need a context and a timeout >= 0 ms");
    return PAC_ERR_ARG;
  }
  ctx->c.timeout_ms = ms;
  return PAC_OK;
}

int pac_ctx_set_packed_lists(pac_ctx *ctx, int mode) {
  if (!ctx || mode < PAC_PACKED_OFF || mode > PAC_PACKED_ALWAYS) {
    set_error("need a context and a mode of PAC_PACKED_OFF, _AUTO or _ALWAYS");
    return PAC_ERR_ARG;
  }
  ctx->c.packed_lists = mode;
  return PAC_OK;
}

int pac_stream_copy(pac_ctx *ctx, uint64_t bytes, int reps, double *ms) {
  PAC_TRY(ensure_ctx(ctx));
  if (!ms) {
    set_error("null output");
    return PAC_ERR_ARG;
  }
  return stream_copy(&ctx->c, bytes, reps, ms);
}

int pac_graph_upload(pac_ctx *ctx, const uint64_t *vlist, uint64_t num_vertices,
                     const int32_t *elist, uint64_t nnz, pac_graph **out) {
  return upload_csr(ctx, vlist, num_vertices, elist, nnz, out);
}

int pac_graph_load_csr_dir(pac_ctx *ctx, const char *csrdir, pac_graph **out) {
  if (!csrdir) {
    set_error("null directory");
    return PAC_ERR_ARG;
  }
  std::vector<uint64_t> vl;
  std::vector<int32_t> el;
  const std::string d(csrdir);
  PAC_TRY(read_file(d + "/csr_vlist.bin", vl));
  PAC_TRY(read_file(d + "/csr_elist.bin", el));
  if (vl.empty()) {
    set_error("%s/csr_vlist.bin is empty", csrdir);
    return PAC_ERR_FORMAT;
  }
  return upload_csr(ctx, vl.data(), vl.size() - 1, el.data(), el.size(), out);
}

namespace {
// Graph::Graph(graphfile), Graph.cpp:27-90, on the host: the reference's fgets loop with its
// 512-byte buffer, for files with a line the device parser leaves to it.
int load_text_host(pac_ctx *ctx, const char *path, pac_graph **out) {
  FILE *fp = fopen(path, "r");
  if (!fp) {
    set_error("cannot open %s", path);
    return PAC_ERR_IO;
  }
  bool zero_based = false;
  uint64_t maxid = 0;
  std::vector<std::pair<uint32_t, uint32_t>> edges;
  char line[512];
  while (fgets(line, sizeof line, fp)) {
    if (line[0] == '#' || line[0] == '%') continue;
    const char *c = line;
    int64_t u = 0, v = 0;
    while (isdigit((unsigned char)*c)) u = u * 10 + (*c++ - '0');
    ++c;  // exactly one separator character
    while (isdigit((unsigned char)*c)) v = v * 10 + (*c++ - '0');
    if (u == 0 || v == 0) zero_based = true;
    if (u == v) continue;
    if (u >= (1ll << 31) || v >= (1ll << 31)) {
      fclose(fp);
      set_error("%s: vertex ID exceeds int32", path);
      return PAC_ERR_UNSUPPORTED;
    }
    maxid = std::max<uint64_t>(maxid, (uint64_t)std::max(u, v));
    edges.push_back({(uint32_t)u, (uint32_t)v});
  }
  fclose(fp);
  const uint64_t V = zero_based ? maxid + 1 : maxid;
  const uint32_t shift = zero_based ? 0u : 1u;
  std::vector<uint32_t> src(edges.size()), dst(edges.size());
  for (size_t i = 0; i < edges.size(); ++i) {
    src[i] = edges[i].first - shift;
    dst[i] = edges[i].second - shift;
  }
  std::vector<std::pair<uint32_t, uint32_t>>().swap(edges);
  // symmetrize, sort and deduplicate on the device
  return edges_to_graph(ctx, src.data(), dst.data(), src.size(), V, out);
}
}  // namespace

// The file is read whole and parsed on the device (line starts by a flag compaction, one
// thread per line); the edges then go through the device symmetrize / sort / dedup.
int pac_graph_load_text(pac_ctx *ctx, const char *path, pac_graph **out) {
  PAC_TRY(ensure_ctx(ctx));
  if (!path || !out) {
    set_error("null path or output handle");
    return PAC_ERR_ARG;
  }
  std::vector<char> text;
  {
    FILE *fp = fopen(path, "rb");
    if (!fp) {
      set_error("cannot open %s", path);
      return PAC_ERR_IO;
    }
    char buf[1 << 16];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, fp)) > 0) text.insert(text.end(), buf, buf + got);
    fclose(fp);
  }
  DevBuf d_src, d_dst;
  uint64_t ne = 0, V = 0;
  bool host_needed = false;
  PAC_TRY(parse_text_device(&ctx->c, text.data(), text.size(), d_src, d_dst, &ne, &V,
                            &host_needed));
  std::vector<char>().swap(text);
  if (host_needed) return load_text_host(ctx, path, out);
  if (2 * ne >= (1ull << 32) || V > 0xFFFFFFFFull) {
    set_error("%s: %llu edges / %llu vertices exceed the uint32 CSR limit", path,
              (unsigned long long)ne, (unsigned long long)V);
    return PAC_ERR_UNSUPPORTED;
  }
  pac_graph *h = nullptr;
  PAC_TRY(new_graph(ctx, &h));
  int rc = graph_from_device_edges(&h->g, d_src.as<uint32_t>(), d_dst.as<uint32_t>(), ne,
                                   (uint32_t)V);
  if (rc == PAC_OK) rc = orient(&h->g);
  if (rc != PAC_OK) {
    delete_graph(h);
    return rc;
  }
  *out = h;
  return PAC_OK;
}

int pac_graph_from_edges(pac_ctx *ctx, const int32_t *src, const int32_t *dst, uint64_t ne,
                         uint64_t num_vertices, pac_graph **out) {
  PAC_TRY(ensure_ctx(ctx));
  if (ne && (!src || !dst)) {
    set_error("null edge array");
    return PAC_ERR_ARG;
  }
  uint64_t V = num_vertices;
  int64_t mx = -1;
  for (uint64_t i = 0; i < ne; ++i) {
    if (src[i] < 0 || dst[i] < 0) {
      set_error("negative vertex ID in edge %llu", (unsigned long long)i);
      return PAC_ERR_FORMAT;
    }
    mx = std::max<int64_t>(mx, std::max(src[i], dst[i]));
  }
  if (V == 0) V = (uint64_t)(mx + 1);
  if ((int64_t)V <= mx) {
    set_error("edge endpoint %lld >= num_vertices %llu", (long long)mx, (unsigned long long)V);
    return PAC_ERR_FORMAT;
  }
  if (V >= (1ull << 31)) {
    set_error("%llu vertices exceed the int32 vertex ID range", (unsigned long long)V);
    return PAC_ERR_UNSUPPORTED;
  }
  return edges_to_graph(ctx, reinterpret_cast<const uint32_t *>(src),
                        reinterpret_cast<const uint32_t *>(dst), ne, V, out);
}

int pac_graph_save_csr_dir(const pac_graph *h, const char *csrdir) {
  if (!h || !csrdir) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  std::vector<uint64_t> vl;
  std::vector<int32_t> el;
  PAC_CHECK_HIP(hipSetDevice(h->g.ctx->device));
  PAC_TRY(download_csr(&h->g, vl, el));
  const std::string d(csrdir);
  PAC_TRY(write_file(d + "/csr_vlist.bin", vl.data(), vl.size()));
  PAC_TRY(write_file(d + "/csr_elist.bin", el.data(), el.size()));
  return PAC_OK;
}

int pac_graph_generate_rmat(pac_ctx *ctx, uint32_t scale, uint32_t edge_factor, double a,
                            double b, double c, uint64_t seed, pac_graph **out) {
  PAC_TRY(ensure_ctx(ctx));
  pac_graph *h = nullptr;
  PAC_TRY(new_graph(ctx, &h));
  int rc = generate_rmat(&h->g, scale, edge_factor, a, b, c, seed);
  if (rc != PAC_OK) {
    delete_graph(h);
    return rc;
  }
  return finish_generated(ctx, h, out);
}

// G(n,p) by Batagelj-Brandes geometric skipping with one stream per row v (seeded by
// mix64(key + v*gamma)): w advances by 1 + floor(log(r)/log(1-p)), r in (0,1].  Pairs
// come out in increasing (v, w) order, so the symmetric CSR is built by count / scan /
// fill with every row already sorted: row u holds its lower neighbours (from row u, in
// increasing w) followed by its upper neighbours (from rows v > u, in increasing v).
int pac_graph_generate_er(pac_ctx *ctx, uint64_t n, double avg_degree, uint64_t seed,
                          pac_graph **out) {
  PAC_TRY(ensure_ctx(ctx));
  if (n < 2 || n >= (1ull << 31) || !(avg_degree > 0)) {
    set_error("er: need 2 <= n < 2^31 and avg_degree > 0");
    return PAC_ERR_ARG;
  }
  double p = avg_degree / (double)(n - 1);
  if (p > 1.0) p = 1.0;
  const uint64_t key = mix64(seed + kGamma);
  const double lq = std::log1p(-p);
  std::vector<uint32_t> lo;          // lower endpoints, grouped by upper endpoint v
  std::vector<uint64_t> vstart(n + 1, 0);
  lo.reserve((size_t)(avg_degree * (double)n * 0.55) + 16);
  for (uint64_t v = 1; v < n; ++v) {
    vstart[v] = lo.size();
    uint64_t st = mix64(key + v * kGamma);
    int64_t w = -1;
    while (true) {
      st += kGamma;
      const double r = (double)((mix64(st) >> 11) + 1) * (1.0 / 9007199254740992.0);
      const double skip = (p >= 1.0) ? 0.0 : std::floor(std::log(r) / lq);
      if (skip >= (double)n) break;
      w += 1 + (int64_t)skip;
      if (w >= (int64_t)v) break;
      lo.push_back((uint32_t)w);
    }
  }
  vstart[n] = lo.size();
  vstart[0] = 0;
  std::vector<uint64_t> vl(n + 1, 0);
  for (uint64_t v = 1; v < n; ++v) {
    vl[v + 1] += vstart[v + 1] - vstart[v];  // lower neighbours of v
    for (uint64_t i = vstart[v]; i < vstart[v + 1]; ++i) vl[lo[i] + 1]++;  // v is upper of lo[i]
  }
  for (uint64_t u = 0; u < n; ++u) vl[u + 1] += vl[u];
  std::vector<int32_t> el(vl[n]);
  std::vector<uint64_t> cur(vl.begin(), vl.end() - 1);
  for (uint64_t v = 1; v < n; ++v)
    for (uint64_t i = vstart[v]; i < vstart[v + 1]; ++i) el[cur[v]++] = (int32_t)lo[i];
  for (uint64_t v = 1; v < n; ++v)
    for (uint64_t i = vstart[v]; i < vstart[v + 1]; ++i) el[cur[lo[i]]++] = (int32_t)v;
  return upload_csr(ctx, vl.data(), n, el.data(), el.size(), out);
}

int pac_graph_download_csr(const pac_graph *h, uint64_t *vlist, int32_t *elist) {
  if (!h || !vlist || (h->g.nnz && !elist)) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  PAC_CHECK_HIP(hipSetDevice(h->g.ctx->device));
  std::vector<uint64_t> vl;
  std::vector<int32_t> el;
  PAC_TRY(download_csr(&h->g, vl, el));
  std::copy(vl.begin(), vl.end(), vlist);
  std::copy(el.begin(), el.end(), elist);
  return PAC_OK;
}

int pac_graph_get_info(const pac_graph *h, pac_graph_info *out) {
  if (!h || !out) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  const Graph &g = h->g;
  out->num_vertices = g.V;
  out->num_entries = g.nnz;
  out->num_edges = g.m;
  out->max_degree = g.max_deg;
  out->max_out_degree = g.max_outdeg;
  out->work_W = g.W;
  out->num_partitions = g.prepared ? g.P : 0;
  out->num_items = g.h_items.size();
  out->prepare_ms = g.prepare_ms;
  out->num_pairs = g.prepared ? g.n_pairs : 0;
  out->num_rows = g.prepared ? g.n_rows : 0;
  out->streamed_elems = g.prepared ? g.streamed : 0;
  out->staged_elems = g.prepared ? g.staged : 0;
  out->staged_item_elems = g.prepared ? g.staged_items : 0;
  Graph *gs[1] = {const_cast<Graph *>(&g)};
  out->stream_elem_bytes = launch_elem_bytes(g.ctx, gs, 1, kModeTC);
  return PAC_OK;
}

int pac_graph_validate(const pac_graph *h, int level) {
  if (!h) {
    set_error("null graph");
    return PAC_ERR_ARG;
  }
  PAC_CHECK_HIP(hipSetDevice(h->g.ctx->device));
  return graph_validate(&h->g, level);
}

void pac_graph_free(pac_graph *h) {
  if (!h) return;
  (void)hipSetDevice(h->g.ctx->device);
  (void)hipStreamSynchronize(h->g.ctx->stream);
  Ctx *ctx = h->g.ctx;
  // drop cached batch state that refers to this graph
  for (const void *p : ctx->batch_key)
    if (p == h) {
      ctx->batch_key.clear();
      break;
    }
  for (const void *p : ctx->hgraph_key)
    if (p == h) {
      if (ctx->hexec) (void)hipGraphExecDestroy(ctx->hexec);
      if (ctx->hgraph) (void)hipGraphDestroy(ctx->hgraph);
      ctx->hexec = nullptr;
      ctx->hgraph = nullptr;
      ctx->hgraph_key.clear();
      break;
    }
  delete_graph(h);
}

int pac_graph_prepare(pac_graph *h, uint32_t num_partitions, const int32_t *v2p) {
  if (!h) {
    set_error("null graph");
    return PAC_ERR_ARG;
  }
  Graph *g = &h->g;
  Ctx *ctx = g->ctx;
  PAC_CHECK_HIP(hipSetDevice(ctx->device));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  const uint32_t V = g->V;
  uint32_t P = num_partitions;
  if (v2p) {
    for (uint32_t u = 0; u < V; ++u) {
      if (v2p[u] < 0) {
        set_error("vertex2partition[%u] = %d is negative", u, v2p[u]);
        return PAC_ERR_ARG;
      }
      P = std::max(P, (uint32_t)v2p[u] + 1);
    }
  }
  if (P == 0) {
    set_error("need at least one partition");
    return PAC_ERR_ARG;
  }
  const uint64_t width = ((uint64_t)V + P - 1) / P;
  PAC_CHECK_HIP(hipEventRecord(ctx->ev[4], ctx->stream));
  // pair lists and row work (once per graph), then the rows that own pairs, in ID order
  RowSummary rs;
  PAC_TRY(row_summary(g, v2p ? 0 : (width ? width : 1), P, rs));
  PAC_CHECK_HIP(hipEventRecord(ctx->ev[5], ctx->stream));
  g->W = rs.W;
  g->max_outdeg = rs.max_outdeg;
  // rows grouped by partition, stable in ID (contiguous partitions: ID order already);
  // rowlist[q] and its summary stats[q] = (d+(r), streamed length, pairs, d(r))
  auto part_of = [&](uint32_t u) -> uint32_t {
    return v2p ? (uint32_t)v2p[u] : (uint32_t)(u / (width ? width : 1));
  };
  const size_t nrows = rs.rows.size();
  std::vector<uint32_t> count(P + 1, 0);
  for (uint32_t r : rs.rows) count[part_of(r) + 1]++;
  for (uint32_t k = 0; k < P; ++k) count[k + 1] += count[k];
  std::vector<uint32_t> rowlist;
  std::vector<uint4> stats;
  if (v2p) {
    rowlist.resize(nrows);
    stats.resize(nrows);
    std::vector<uint32_t> cur(count.begin(), count.end() - 1);
    for (size_t i = 0; i < nrows; ++i) {
      const uint32_t q = cur[part_of(rs.rows[i])]++;
      rowlist[q] = rs.rows[i];
      stats[q] = rs.stats[i];
    }
    if (nrows)
      PAC_CHECK_HIP(hipMemcpy(g->rows.p, rowlist.data(), sizeof(uint32_t) * nrows,
                              hipMemcpyHostToDevice));
    // oriented edges per partition, from the out-degrees of every vertex
    std::vector<uint32_t> odeg;
    PAC_TRY(d2h(ctx, odeg, g->odeg, V));
    g->h_pedges.assign(P, 0);
    for (uint32_t u = 0; u < V; ++u) g->h_pedges[part_of(u)] += odeg[u];
  } else {
    rowlist.swap(rs.rows);
    stats.swap(rs.stats);
    g->h_pedges = rs.part_edges;
  }
  g->P = P;
  g->h_pstart = count;
  g->part_width = width ? width : 1;
  if (v2p)
    g->h_v2p.assign(v2p, v2p + V);
  else
    g->h_v2p.clear();
  g->part_ready = false;
  // Work items: pack small rows, split heavy rows by pair range.  Row cost = staged
  // list + streamed lists + a per-pair and a per-row overhead.
  auto row_cost = [&](uint32_t q) {
    return (uint64_t)stats[q].x + stats[q].y + 4ull * stats[q].z + 64;
  };
  uint64_t total_cost = 0;
  for (size_t q = 0; q < nrows; ++q) total_cost += row_cost((uint32_t)q);
  const uint64_t target = item_target(total_cost, ctx->num_cus);
  g->n_long_rows = 0;
  for (size_t q = 0; q < nrows; ++q)
    if (stats[q].x > kMinTileCap) ++g->n_long_rows;
  g->h_items.clear();
  for (uint32_t k = 0; k < P; ++k) {
    uint32_t lo = count[k];
    uint64_t acc = 0;
    for (uint32_t q = count[k]; q < count[k + 1]; ++q) {
      const uint64_t cost = row_cost(q);
      const uint32_t np = stats[q].z;
      if (cost > 2 * target) {
        if (q > lo) g->h_items.push_back({0, k, lo, q, 0, 0xFFFFFFFFu, 0, 0});
        const uint64_t nchunk = std::min<uint64_t>(np, (cost + 2 * target - 1) / (2 * target));
        for (uint64_t c = 0; c < nchunk; ++c) {
          const uint32_t c_lo = (uint32_t)(np * c / nchunk);
          const uint32_t c_hi = (uint32_t)(np * (c + 1) / nchunk);
          if (c_hi > c_lo) g->h_items.push_back({0, k, q, q + 1, c_lo, c_hi, 0, 0});
        }
        lo = q + 1;
        acc = 0;
        continue;
      }
      acc += cost;
      if (acc >= target) {
        g->h_items.push_back({0, k, lo, q + 1, 0, 0xFFFFFFFFu, 0, 0});
        lo = q + 1;
        acc = 0;
      }
    }
    if (count[k + 1] > lo) g->h_items.push_back({0, k, lo, count[k + 1], 0, 0xFFFFFFFFu, 0, 0});
  }
  // Larger items first: the queue hands out the long ones early so the tail is short.
  {
    std::vector<std::pair<uint64_t, uint32_t>> cost(g->h_items.size());
    for (size_t i = 0; i < g->h_items.size(); ++i) {
      const WorkItem &w = g->h_items[i];
      uint64_t c = 0;
      if (w.hi == w.lo + 1 && w.c_hi != 0xFFFFFFFFu) {
        c = row_cost(w.lo) * (w.c_hi - w.c_lo) / std::max(1u, stats[w.lo].z);
      } else {
        for (uint32_t q = w.lo; q < w.hi; ++q) c += row_cost(q);
      }
      cost[i] = {c, (uint32_t)i};
    }
    std::stable_sort(cost.begin(), cost.end(),
                     [](const auto &a, const auto &b) { return a.first > b.first; });
    std::vector<WorkItem> sorted(g->h_items.size());
    for (size_t i = 0; i < cost.size(); ++i) {
      sorted[i] = g->h_items[cost[i].second];
      sorted[i].cost_hi = (uint32_t)(cost[i].first >> 32);
      sorted[i].cost_lo = (uint32_t)cost[i].first;
    }
    g->h_items.swap(sorted);
  }
  // The pair work model: pairs, owning rows, streamed and staged elements (staged once
  // per row, and once per item that holds the row).
  g->n_pairs = 0;
  g->n_rows = nrows;
  g->streamed = g->staged = g->staged_items = 0;
  for (size_t q = 0; q < nrows; ++q) {
    g->n_pairs += stats[q].z;
    g->streamed += stats[q].y;
    g->staged += stats[q].x;
  }
  for (const WorkItem &w : g->h_items)
    for (uint32_t q = w.lo; q < w.hi; ++q) g->staged_items += stats[q].x;
  partition_features(g, stats);
  g->h_learned.clear();
  g->learned_version = 0;
  g->dlearned_version = 0;
  {
    // device strategy tables: merge, gallop and hash everywhere, then the fixed heuristic
    std::vector<uint8_t> tables((size_t)4 * P);
    for (uint32_t k = 0; k < P; ++k) {
      tables[k] = kMerge;
      tables[(size_t)P + k] = kGallop;
      tables[(size_t)2 * P + k] = kBitmap;
      tables[(size_t)3 * P + k] = g->h_fixed[k];
    }
    PAC_TRY(g->dstrat.reserve(tables.size()));
    PAC_CHECK_HIP(hipMemcpy(g->dstrat.p, tables.data(), tables.size(), hipMemcpyHostToDevice));
  }
  PAC_TRY(g->items.reserve(sizeof(WorkItem) * (g->h_items.empty() ? 1 : g->h_items.size())));
  if (!g->h_items.empty())
    PAC_CHECK_HIP(hipMemcpy(g->items.p, g->h_items.data(), sizeof(WorkItem) * g->h_items.size(),
                            hipMemcpyHostToDevice));
  g->prepare_ms = elapsed(ctx->ev[4], ctx->ev[5]);
  g->prepared = true;
  // drop batch caches that may refer to this graph's old buffers
  ctx->batch_key.clear();
  ctx->hgraph_key.clear();
  return PAC_OK;
}

int pac_graph_partition_sizes(const pac_graph *h, uint64_t *edges_per_partition) {
  if (!h || !edges_per_partition) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  if (!h->g.prepared) {
    set_error("graph not prepared");
    return PAC_ERR_STATE;
  }
  std::copy(h->g.h_pedges.begin(), h->g.h_pedges.end(), edges_per_partition);
  return PAC_OK;
}

int pac_random_partition(uint64_t V, uint32_t P, uint64_t seed, int32_t *v2p) {
  if (!v2p || P == 0) {
    set_error("need v2p and P > 0");
    return PAC_ERR_ARG;
  }
  const uint64_t key = mix64(seed + kGamma);
  for (uint64_t u = 0; u < V; ++u) v2p[u] = (int32_t)(mix64(key + u * kGamma) % P);
  return PAC_OK;
}

int pac_greedy_partition(const pac_graph *h, uint32_t P, uint64_t seed, int32_t *v2p) {
  if (!h || !v2p || P == 0) {
    set_error("need a graph, v2p and P > 0");
    return PAC_ERR_ARG;
  }
  std::vector<uint64_t> vl;
  std::vector<int32_t> el;
  PAC_CHECK_HIP(hipSetDevice(h->g.ctx->device));
  PAC_TRY(download_csr(&h->g, vl, el));
  greedy_partition_host(vl, el, h->g.V, P, seed, v2p);
  return PAC_OK;
}

int pac_greedy_partition_2(const pac_graph *h, uint32_t P, uint64_t seed, int32_t *v2p) {
  if (!h || !v2p || P == 0) {
    set_error("need a graph, v2p and P > 0");
    return PAC_ERR_ARG;
  }
  std::vector<uint64_t> vl;
  std::vector<int32_t> el;
  PAC_CHECK_HIP(hipSetDevice(h->g.ctx->device));
  PAC_TRY(download_csr(&h->g, vl, el));
  return greedy_partition_2_host(vl, el, h->g.V, P, seed, v2p);
}

// Partitioned_Graph::evaluate_partition, Graph.cpp:1460-1473.
int pac_evaluate_partition(const pac_graph *h, const int32_t *v2p, uint32_t P,
                           double *balance_ratio) {
  if (!h || !v2p || !balance_ratio || P == 0) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  std::vector<uint64_t> vl;
  std::vector<int32_t> el;
  PAC_CHECK_HIP(hipSetDevice(h->g.ctx->device));
  PAC_TRY(download_csr(&h->g, vl, el));
  uint64_t max_part_deg = 0, max_deg = 0;
  std::vector<uint64_t> per(P);
  for (uint64_t u = 0; u < h->g.V; ++u) {
    std::fill(per.begin(), per.end(), 0);
    for (uint64_t i = vl[u]; i < vl[u + 1]; ++i) {
      const int32_t p = v2p[el[i]];
      if (p < 0 || (uint32_t)p >= P) {
        set_error("vertex2partition out of range");
        return PAC_ERR_ARG;
      }
      per[p]++;
    }
    for (uint64_t x : per) max_part_deg = std::max(max_part_deg, x);
    max_deg = std::max<uint64_t>(max_deg, vl[u + 1] - vl[u]);
  }
  // as Graph.cpp:1471: a graph without edges gives 0/0 = NaN
  *balance_ratio = (double)(P * max_part_deg) / (double)max_deg;
  return PAC_OK;
}

int pac_tc(pac_graph *h, int dispatch, int launch, uint64_t *total, uint64_t *per_vertex) {
  PAC_TRY(check_solve_args(h, dispatch, launch));
  Graph *g = &h->g;
  Ctx *ctx = g->ctx;
  const int mode = per_vertex ? kModeLCC : kModeTC;
  Timed t(ctx);
  PAC_TRY(run_intersection(g, mode, dispatch, launch, t));
  PAC_TRY(t.mark(3));
  PAC_TRY(t.finish());
  if (total)
    PAC_CHECK_HIP(hipMemcpy(total, g->total.p, sizeof(uint64_t), hipMemcpyDeviceToHost));
  if (per_vertex && g->V)
    PAC_CHECK_HIP(hipMemcpy(per_vertex, g->tvert.p, sizeof(uint64_t) * g->V,
                            hipMemcpyDeviceToHost));
  return PAC_OK;
}

int pac_lcc(pac_graph *h, int dispatch, int launch, float *lcc32, double *lcc64,
            uint64_t *total) {
  PAC_TRY(check_solve_args(h, dispatch, launch));
  Graph *g = &h->g;
  Ctx *ctx = g->ctx;
  Timed t(ctx);
  PAC_TRY(run_intersection(g, kModeLCC, dispatch, launch, t));
  PAC_TRY(epilogue_lcc(ctx, g, true, lcc64 != nullptr));
  PAC_TRY(t.mark(3));
  PAC_TRY(t.finish());
  if (total)
    PAC_CHECK_HIP(hipMemcpy(total, g->total.p, sizeof(uint64_t), hipMemcpyDeviceToHost));
  if (lcc32 && g->V)
    PAC_CHECK_HIP(hipMemcpy(lcc32, g->lcc32.p, sizeof(float) * g->V, hipMemcpyDeviceToHost));
  if (lcc64 && g->V)
    PAC_CHECK_HIP(hipMemcpy(lcc64, g->lcc64.p, sizeof(double) * g->V, hipMemcpyDeviceToHost));
  return PAC_OK;
}

int pac_jaccard_edges(pac_graph *h, int dispatch, int launch, int32_t *src, int32_t *dst,
                      uint32_t *inter, uint32_t *uni, float *ratio) {
  PAC_TRY(check_solve_args(h, dispatch, launch));
  Graph *g = &h->g;
  Ctx *ctx = g->ctx;
  const uint64_t m = g->m ? g->m : 1;
  DevBuf d_src, d_dst, d_inter, d_uni, d_ratio;
  const bool device_only = !src && !dst && !inter && !uni && !ratio;
  if (src) PAC_TRY(d_src.reserve(4 * m));
  if (dst) PAC_TRY(d_dst.reserve(4 * m));
  if (inter) PAC_TRY(d_inter.reserve(4 * m));
  if (uni) PAC_TRY(d_uni.reserve(4 * m));
  if (ratio) PAC_TRY(d_ratio.reserve(4 * m));
  if (device_only) {  // results stay in graph-owned device buffers
    PAC_TRY(g->juni.reserve(4 * m));
    PAC_TRY(g->jratio.reserve(4 * m));
  }
  Timed t(ctx);
  PAC_TRY(run_intersection(g, kModeSUP, dispatch, launch, t));
  if (device_only)
    PAC_TRY(epilogue_jaccard(ctx, g, nullptr, nullptr, nullptr, g->juni.as<uint32_t>(),
                             g->jratio.as<float>()));
  else
    PAC_TRY(epilogue_jaccard(ctx, g, d_src.as<int32_t>(), d_dst.as<int32_t>(),
                             d_inter.as<uint32_t>(), d_uni.as<uint32_t>(), d_ratio.as<float>()));
  PAC_TRY(t.mark(3));
  PAC_TRY(t.finish());
  if (g->m) {
    if (src) PAC_CHECK_HIP(hipMemcpy(src, d_src.p, 4 * g->m, hipMemcpyDeviceToHost));
    if (dst) PAC_CHECK_HIP(hipMemcpy(dst, d_dst.p, 4 * g->m, hipMemcpyDeviceToHost));
    if (inter) PAC_CHECK_HIP(hipMemcpy(inter, d_inter.p, 4 * g->m, hipMemcpyDeviceToHost));
    if (uni) PAC_CHECK_HIP(hipMemcpy(uni, d_uni.p, 4 * g->m, hipMemcpyDeviceToHost));
    if (ratio) PAC_CHECK_HIP(hipMemcpy(ratio, d_ratio.p, 4 * g->m, hipMemcpyDeviceToHost));
  }
  return PAC_OK;
}

int pac_jaccard_queries(pac_graph *h, const int32_t *q, uint32_t n, double *res) {
  if (!h || (n && (!q || !res))) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  Graph *g = &h->g;
  PAC_CHECK_HIP(hipSetDevice(g->ctx->device));
  for (uint32_t i = 0; i < n; ++i)
    if (q[i] < 0 || (uint32_t)q[i] >= g->V) {
      set_error("query vertex %d out of range", q[i]);
      return PAC_ERR_ARG;
    }
  if (n == 0) return PAC_OK;
  DevBuf d_q, d_res;
  PAC_TRY(d_q.reserve(4ull * n));
  PAC_TRY(d_res.reserve(8ull * n * n));
  PAC_CHECK_HIP(hipMemcpy(d_q.p, q, 4ull * n, hipMemcpyHostToDevice));
  PAC_TRY(query_matrix(g->ctx, g, d_q.as<int32_t>(), n, d_res.as<double>()));
  PAC_CHECK_HIP(hipStreamSynchronize(g->ctx->stream));
  PAC_CHECK_HIP(hipMemcpy(res, d_res.p, 8ull * n * n, hipMemcpyDeviceToHost));
  return PAC_OK;
}

int pac_cnt_common_pairs(pac_graph *h, const int32_t *u, const int32_t *v, uint64_t n,
                         uint64_t *res) {
  return pac_cnt_common_pairs_filtered(h, u, v, nullptr, n, -1, res);
}

namespace {
// Resolves a partition argument: -1 = no filter (*d_part = null); else the graph must be
// prepared and 0 <= partition < P.
int partition_filter(Graph *g, int32_t partition, const uint32_t **d_part) {
  *d_part = nullptr;
  if (partition < 0) {
    if (partition != -1) {
      set_error("partition %d: use -1 for no partition filter", partition);
      return PAC_ERR_ARG;
    }
    return PAC_OK;
  }
  if (!g->prepared) {
    set_error("a partition filter needs a prepared graph");
    return PAC_ERR_STATE;
  }
  if ((uint32_t)partition >= g->P) {
    set_error("partition %d out of range (P = %u)", partition, g->P);
    return PAC_ERR_ARG;
  }
  return partition_map(g, d_part);
}
}  // namespace

int pac_cnt_common_pairs_filtered(pac_graph *h, const int32_t *u, const int32_t *v,
                                  const int32_t *thr, uint64_t n, int32_t partition,
                                  uint64_t *res) {
  if (!h || (n && (!u || !v || !res))) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  Graph *g = &h->g;
  PAC_CHECK_HIP(hipSetDevice(g->ctx->device));
  for (uint64_t i = 0; i < n; ++i)
    if (u[i] < 0 || v[i] < 0 || (uint32_t)u[i] >= g->V || (uint32_t)v[i] >= g->V) {
      set_error("query pair %llu out of range", (unsigned long long)i);
      return PAC_ERR_ARG;
    }
  const uint32_t *d_part = nullptr;
  PAC_TRY(partition_filter(g, partition, &d_part));
  if (n == 0) return PAC_OK;
  DevBuf du, dv, dt, dr;
  PAC_TRY(du.reserve(4 * n));
  PAC_TRY(dv.reserve(4 * n));
  PAC_TRY(dr.reserve(8 * n));
  PAC_CHECK_HIP(hipMemcpy(du.p, u, 4 * n, hipMemcpyHostToDevice));
  PAC_CHECK_HIP(hipMemcpy(dv.p, v, 4 * n, hipMemcpyHostToDevice));
  if (thr) {
    PAC_TRY(dt.reserve(4 * n));
    PAC_CHECK_HIP(hipMemcpy(dt.p, thr, 4 * n, hipMemcpyHostToDevice));
  }
  Timed t(g->ctx);  // the device work alone, queries already resident
  PAC_TRY(t.mark(0));
  PAC_TRY(t.mark(1));
  PAC_TRY(pair_counts(g->ctx, g, du.as<int32_t>(), dv.as<int32_t>(), n, dr.as<uint64_t>(),
                      thr ? dt.as<int32_t>() : nullptr, d_part, (uint32_t)partition));
  PAC_TRY(t.mark(2));
  PAC_TRY(t.mark(3));
  PAC_TRY(t.finish());
  PAC_CHECK_HIP(hipMemcpy(res, dr.p, 8 * n, hipMemcpyDeviceToHost));
  return PAC_OK;
}

int pac_common_neis(pac_graph *h, int32_t u, int32_t v, const int32_t *nodes, uint64_t n,
                    int32_t partition, int32_t *out, uint64_t cap, uint64_t *count) {
  if (!h || !count || (n && !nodes) || (cap && !out)) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  Graph *g = &h->g;
  PAC_CHECK_HIP(hipSetDevice(g->ctx->device));
  if (u < 0 || (uint32_t)u >= g->V || (!nodes && (v < 0 || (uint32_t)v >= g->V))) {
    set_error("vertex out of range (u = %d, v = %d, V = %u)", u, v, g->V);
    return PAC_ERR_ARG;
  }
  for (uint64_t i = 1; nodes && i < n; ++i)
    if (nodes[i] < nodes[i - 1]) {
      set_error("nodes must be sorted ascending (nodes[%llu] < nodes[%llu])",
                (unsigned long long)i, (unsigned long long)(i - 1));
      return PAC_ERR_ARG;
    }
  const uint32_t *d_part = nullptr;
  PAC_TRY(partition_filter(g, partition, &d_part));
  uint32_t ru[2], rv[2] = {0, 0};
  PAC_CHECK_HIP(hipMemcpy(ru, g->off.as<uint32_t>() + u, 8, hipMemcpyDeviceToHost));
  // the list that is walked: `nodes`, or the shorter of N(u), N(v) (the other is searched)
  uint32_t s = (uint32_t)u;
  uint64_t len = n;
  const int32_t *d_list = nullptr;
  DevBuf dn;
  if (nodes) {
    if (n) {
      PAC_TRY(dn.reserve(4 * n));
      PAC_CHECK_HIP(hipMemcpy(dn.p, nodes, 4 * n, hipMemcpyHostToDevice));
    }
    d_list = dn.as<int32_t>();
  } else {
    PAC_CHECK_HIP(hipMemcpy(rv, g->off.as<uint32_t>() + v, 8, hipMemcpyDeviceToHost));
    const bool walk_u = ru[1] - ru[0] < rv[1] - rv[0];
    s = walk_u ? (uint32_t)v : (uint32_t)u;
    const uint32_t *lo = walk_u ? ru : rv;
    len = lo[1] - lo[0];
    d_list = reinterpret_cast<const int32_t *>(g->adj.as<uint32_t>() + lo[0]);
  }
  const uint64_t most = std::min<uint64_t>(len, ru[1] - ru[0]);
  DevBuf dout, dc;
  PAC_TRY(dout.reserve(4 * (most ? most : 1)));
  PAC_TRY(dc.reserve(8));
  PAC_TRY(common_list(g->ctx, g, s, d_list, len, d_part, partition < 0 ? 0u : (uint32_t)partition,
                      dout.as<int32_t>(), reinterpret_cast<unsigned long long *>(dc.p)));
  PAC_CHECK_HIP(hipStreamSynchronize(g->ctx->stream));
  uint64_t c = 0;
  PAC_CHECK_HIP(hipMemcpy(&c, dc.p, 8, hipMemcpyDeviceToHost));
  *count = c;
  if (out && c > cap) {
    set_error("%llu common neighbours, output holds %llu", (unsigned long long)c,
              (unsigned long long)cap);
    return PAC_ERR_ARG;
  }
  if (out && c) PAC_CHECK_HIP(hipMemcpy(out, dout.p, 4 * c, hipMemcpyDeviceToHost));
  return PAC_OK;
}

int pac_cnt_common_kway(pac_graph *h, const uint64_t *qoff, const int32_t *qv, uint64_t nq,
                        uint64_t *res) {
  return pac_cnt_common_kway_filtered(h, qoff, qv, nq, -1, res);
}

int pac_cnt_common_kway_filtered(pac_graph *h, const uint64_t *qoff, const int32_t *qv,
                                 uint64_t nq, int32_t partition, uint64_t *res) {
  if (!h || (nq && (!qoff || !qv || !res))) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  Graph *g = &h->g;
  PAC_CHECK_HIP(hipSetDevice(g->ctx->device));
  const uint32_t *d_part = nullptr;
  PAC_TRY(partition_filter(g, partition, &d_part));
  if (nq == 0) return PAC_OK;
  if (qoff[0] != 0) {
    set_error("query offsets must start at 0");
    return PAC_ERR_ARG;
  }
  for (uint64_t i = 0; i < nq; ++i)
    if (qoff[i + 1] <= qoff[i]) {
      set_error("query %llu is empty or its offsets decrease", (unsigned long long)i);
      return PAC_ERR_ARG;
    }
  const uint64_t nv = qoff[nq];
  for (uint64_t i = 0; i < nv; ++i)
    if (qv[i] < 0 || (uint32_t)qv[i] >= g->V) {
      set_error("query vertex %llu (= %d) out of range", (unsigned long long)i, qv[i]);
      return PAC_ERR_ARG;
    }
  DevBuf dq, dv, dr;
  PAC_TRY(dq.reserve(8 * (nq + 1)));
  PAC_TRY(dv.reserve(4 * nv));
  PAC_TRY(dr.reserve(8 * nq));
  PAC_CHECK_HIP(hipMemcpy(dq.p, qoff, 8 * (nq + 1), hipMemcpyHostToDevice));
  PAC_CHECK_HIP(hipMemcpy(dv.p, qv, 4 * nv, hipMemcpyHostToDevice));
  Timed t(g->ctx);
  PAC_TRY(t.mark(0));
  PAC_TRY(t.mark(1));
  PAC_TRY(kway_counts(g->ctx, g, dq.as<uint64_t>(), dv.as<int32_t>(), nq, dr.as<uint64_t>(),
                      d_part, partition < 0 ? 0u : (uint32_t)partition));
  PAC_TRY(t.mark(2));
  PAC_TRY(t.mark(3));
  PAC_TRY(t.finish());
  PAC_CHECK_HIP(hipMemcpy(res, dr.p, 8 * nq, hipMemcpyDeviceToHost));
  return PAC_OK;
}

int pac_friend_triangle(pac_graph *h, int32_t root, const int32_t *nodes, uint64_t n,
                        uint64_t *res) {
  if (!h || !res || (n && !nodes)) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  Graph *g = &h->g;
  PAC_CHECK_HIP(hipSetDevice(g->ctx->device));
  if (root < 0 || (uint32_t)root >= g->V) {
    set_error("root %d out of range", root);
    return PAC_ERR_ARG;
  }
  for (uint64_t i = 0; i < n; ++i)
    if (nodes[i] < 0 || (uint32_t)nodes[i] >= g->V) {
      set_error("node %llu (= %d) out of range", (unsigned long long)i, nodes[i]);
      return PAC_ERR_ARG;
    }
  uint32_t rb[2];
  PAC_CHECK_HIP(hipMemcpy(rb, g->off.as<uint32_t>() + root, 8, hipMemcpyDeviceToHost));
  const uint32_t rdeg = rb[1] - rb[0];
  DevBuf dn, dc, dr;
  PAC_TRY(dn.reserve(4 * (n ? n : 1)));
  PAC_TRY(dc.reserve(4 * (uint64_t)(rdeg ? rdeg : 1)));
  PAC_TRY(dr.reserve(8));
  if (n) PAC_CHECK_HIP(hipMemcpy(dn.p, nodes, 4 * n, hipMemcpyHostToDevice));
  Timed t(g->ctx);
  PAC_TRY(t.mark(0));
  PAC_TRY(t.mark(1));
  PAC_TRY(friend_triangles(g->ctx, g, (uint32_t)root, dn.as<int32_t>(), n, rdeg,
                           dc.as<uint32_t>(), dr.as<uint64_t>()));
  PAC_TRY(t.mark(2));
  PAC_TRY(t.mark(3));
  PAC_TRY(t.finish());
  PAC_CHECK_HIP(hipMemcpy(res, dr.p, 8, hipMemcpyDeviceToHost));
  return PAC_OK;
}

int pac_square_count(pac_graph *h, uint64_t *res) {
  if (!h || !res) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  Graph *g = &h->g;
  PAC_CHECK_HIP(hipSetDevice(g->ctx->device));
  Timed t(g->ctx);  // includes the scratch allocation and the 8-byte result copy
  PAC_TRY(t.mark(0));
  PAC_TRY(t.mark(1));
  PAC_TRY(square_count(g->ctx, g, res));
  PAC_TRY(t.mark(2));
  PAC_TRY(t.mark(3));
  return t.finish();
}

int pac_tc_noalloc(pac_graph *h, int dispatch, uint64_t *d_total, uint64_t *d_per_vertex) {
  PAC_TRY(check_solve_args(h, dispatch, PAC_LAUNCH_PERSISTENT));
  if (!d_total) {
    set_error("need a device pointer for the total");
    return PAC_ERR_ARG;
  }
  Graph *g = &h->g;
  const uint8_t *tables = g->dstrat.as<const uint8_t>();
  const uint8_t *strat = nullptr;
  switch (dispatch) {
    case PAC_FORCE_MERGE: strat = tables; break;
    case PAC_FORCE_GALLOP: strat = tables + g->P; break;
    case PAC_FORCE_BITMAP: strat = tables + 2ull * g->P; break;
    case PAC_DISPATCH_FIXED: strat = tables + 3ull * g->P; break;
    default:  // PAC_DISPATCH_LEARNED
      if (!g->ctx->mlp_loaded) {
        strat = tables + 3ull * g->P;  // the documented fallback: no model loaded
      } else if (g->dlearned_version == g->ctx->mlp_version && g->dlearned_version != 0 &&
                 g->dlearned.p) {
        strat = g->dlearned.as<const uint8_t>();
      } else {
        set_error("the learned choice for the loaded model is not computed yet: run one "
                  "solve with PAC_DISPATCH_LEARNED first");
        return PAC_ERR_STATE;
      }
  }
  return launch_persistent_noalloc(g->ctx, g, d_per_vertex ? kModeLCC : kModeTC, strat,
                                   reinterpret_cast<unsigned long long *>(d_total),
                                   reinterpret_cast<unsigned long long *>(d_per_vertex));
}

int pac_tc_batch(pac_ctx *ctx, pac_graph *const *hs, uint32_t G, int dispatch, int launch,
                 uint64_t *totals) {
  PAC_TRY(ensure_ctx(ctx));
  if (!hs || !totals || G == 0) {
    set_error("need graphs and totals");
    return PAC_ERR_ARG;
  }
  if (launch == PAC_LAUNCH_PLAIN) {
    // one plain solve per graph (the unbatched baseline)
    double ms = 0;
    uint64_t launches = 0;
    for (uint32_t i = 0; i < G; ++i) {
      PAC_TRY(pac_tc(hs[i], dispatch, PAC_LAUNCH_PLAIN, &totals[i], nullptr));
      ms += ctx->c.last.total_ms;
      launches += ctx->c.last.launches;
    }
    ctx->c.last.total_ms = ms;
    ctx->c.last.isect_ms = ms;
    ctx->c.last.launches = launches;
    return PAC_OK;
  }
  return pac::solve_batch(&ctx->c, hs, G, kModeTC, dispatch, launch, totals);
}

int pac_set_tile(pac_ctx *ctx, int tile) {
  if (!ctx || tile < 0 || tile >= PAC_NUM_TILES) {
    set_error("tile must be in [0,%d)", PAC_NUM_TILES);
    return PAC_ERR_ARG;
  }
  ctx->c.tile = tile;
  return PAC_OK;
}

int pac_get_tile(const pac_ctx *ctx, int *tile) {
  if (!ctx || !tile) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  *tile = ctx->c.tile;
  return PAC_OK;
}

}  // extern "C"
