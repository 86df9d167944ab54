// pac_tune.cpp -- per-partition strategy timing (the MLP's training labels), the
// per-partition feature/heuristic accessors and the tile autotuner.
//
// No reference counterpart: the reference has one hard-coded loop per strategy and
// no tuning (Graph.cpp:342-368).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "pac_internal.hpp"

using namespace pac;

namespace {

double median(std::vector<double> v) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  return v[v.size() / 2];
}

std::string shape_key(pac_graph *const *gs, uint32_t G) {
  unsigned long long h = 1469598103934665603ull;  // FNV-1a over (V, m) of every graph
  auto mix = [&](uint64_t x) {
    for (int i = 0; i < 8; ++i) {
      h ^= (x >> (8 * i)) & 0xFF;
      h *= 1099511628211ull;
    }
  };
  mix(G);
  for (uint32_t i = 0; i < G; ++i) {
    mix(gs[i]->g.V);
    mix(gs[i]->g.m);
    mix(gs[i]->g.P);
  }
  char buf[32];
  snprintf(buf, sizeof buf, "%016llx", h);
  return buf;
}

}  // namespace

extern "C" {

int pac_partition_features(pac_graph *h, float *features) {
  if (!h || !features) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  if (!h->g.prepared) {
    set_error("graph not prepared");
    return PAC_ERR_STATE;
  }
  std::copy(h->g.h_features.begin(), h->g.h_features.end(), features);
  return PAC_OK;
}

int pac_fixed_choice(pac_graph *h, uint8_t *choice) {
  if (!h || !choice) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  if (!h->g.prepared) {
    set_error("graph not prepared");
    return PAC_ERR_STATE;
  }
  std::copy(h->g.h_fixed.begin(), h->g.h_fixed.end(), choice);
  return PAC_OK;
}

// Every strategy on every partition, timed inside the persistent launch the solves use
// (LCC mode, the headline's).  In launch r of a round, partition k runs strategy
// (k + r) mod 3, so every launch mixes the three strategies and each partition meets each
// of them once per round; a partition's time under a strategy is the sum of its items'
// durations from the kernel's per-item clock stamps.  A warm-up launch precedes the
// rounds; ms[k][s] is the median over `reps` rounds (0 for a partition owning no pairs).
int pac_profile_partitions(pac_graph *h, int reps, double *ms) {
  if (!h || !ms || reps < 1) {
    set_error("need a graph, reps >= 1 and an output buffer");
    return PAC_ERR_ARG;
  }
  Graph *g = &h->g;
  if (!g->prepared) {
    set_error("graph not prepared");
    return PAC_ERR_STATE;
  }
  Ctx *ctx = g->ctx;
  PAC_CHECK_HIP(hipSetDevice(ctx->device));
  const uint32_t P = g->P;
  const size_t nit = g->h_items.size();
  std::fill(ms, ms + (size_t)P * PAC_NUM_STRATEGIES, 0.0);
  if (nit == 0) return PAC_OK;
  int khz = 0;
  PAC_CHECK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
  if (khz <= 0) {
    set_error("device reports no wall-clock rate");
    return PAC_ERR_HIP;
  }
  DevBuf st;
  PAC_TRY(st.reserve(2 * sizeof(unsigned long long) * nit));
  std::vector<unsigned long long> hs(2 * nit);
  std::vector<uint8_t> strat(P);
  std::vector<double> acc((size_t)P * PAC_NUM_STRATEGIES);
  std::vector<std::vector<double>> t((size_t)P * PAC_NUM_STRATEGIES);
  Graph *gs[1] = {g};
  struct StampsOff {  // the context never keeps the stamps pointer past this call
    Ctx *c;
    ~StampsOff() { c->item_stamps = nullptr; }
  } off{ctx};
  for (int round = -1; round < reps; ++round) {  // round -1: warm-up, not recorded
    std::fill(acc.begin(), acc.end(), 0.0);
    for (uint32_t r = 0; r < PAC_NUM_STRATEGIES; ++r) {
      for (uint32_t k = 0; k < P; ++k) strat[k] = (uint8_t)((k + r) % PAC_NUM_STRATEGIES);
      PAC_TRY(zero_outputs(ctx, gs, 1, kModeLCC));
      ctx->item_stamps = round < 0 ? nullptr : st.as<unsigned long long>();
      PAC_TRY(launch_persistent(ctx, gs, 1, kModeLCC, strat.data(), false, nullptr));
      ctx->item_stamps = nullptr;
      if (round < 0) break;
      PAC_CHECK_HIP(hipMemcpyAsync(hs.data(), st.p, 2 * sizeof(unsigned long long) * nit,
                                   hipMemcpyDeviceToHost, ctx->stream));
      PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
      for (size_t i = 0; i < nit; ++i) {
        const uint32_t k = g->h_items[i].part;
        acc[(size_t)k * PAC_NUM_STRATEGIES + strat[k]] += (double)(hs[2 * i + 1] - hs[2 * i]);
      }
    }
    if (round < 0) continue;
    for (uint32_t k = 0; k < P; ++k)
      if (g->h_pstart[k + 1] > g->h_pstart[k])
        for (int s = 0; s < PAC_NUM_STRATEGIES; ++s)
          t[(size_t)k * PAC_NUM_STRATEGIES + s].push_back(
              acc[(size_t)k * PAC_NUM_STRATEGIES + s] / khz);  // ticks / kHz = ms
  }
  for (size_t i = 0; i < t.size(); ++i)
    if (!t[i].empty()) ms[i] = median(t[i]);
  return PAC_OK;
}

// Times the intersection launch of the batched persistent kernel for every tile
// configuration in the caller's mode and dispatch and keeps the fastest.  The cache file
// holds lines "<shape-key> <mode> <dispatch> <tile> <ms per tile...>".
int pac_autotune(pac_ctx *ctx, pac_graph *const *gs, uint32_t G, int mode, int dispatch,
                 const char *cache_path, int *tile, double *ms_out) {
  if (!ctx || !gs || G == 0 || !tile) {
    set_error("need a context, graphs and a tile output");
    return PAC_ERR_ARG;
  }
  if (mode < PAC_MODE_TC || mode > PAC_MODE_JACCARD) {
    set_error("unknown solve mode %d", mode);
    return PAC_ERR_ARG;
  }
  if (dispatch < PAC_DISPATCH_FIXED || dispatch > PAC_FORCE_BITMAP) {
    set_error("unknown dispatch mode %d", dispatch);
    return PAC_ERR_ARG;
  }
  const std::string key = shape_key(gs, G);
  if (cache_path) {
    std::ifstream in(cache_path);
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::string k;
      int md = -1, dp = -1, t = -1;
      double ms[PAC_NUM_TILES];
      if (!(ls >> k >> md >> dp >> t)) continue;
      bool whole = true;
      for (int i = 0; i < PAC_NUM_TILES; ++i) whole = whole && (bool)(ls >> ms[i]);
      if (whole && k == key && md == mode && dp == dispatch && t >= 0 && t < PAC_NUM_TILES) {
        ctx->c.tile = t;
        *tile = t;
        if (ms_out) std::copy(ms, ms + PAC_NUM_TILES, ms_out);
        return PAC_OK;
      }
    }
  }
  const int saved = ctx->c.tile;
  double ms[PAC_NUM_TILES];
  int best_t = saved;
  for (int t = 0; t < PAC_NUM_TILES; ++t) {
    ctx->c.tile = t;
    std::vector<double> v;
    int rc = PAC_OK;
    for (int r = 0; r < 4 && rc == PAC_OK; ++r) {  // a warm-up, then three timed
      rc = solve_batch(&ctx->c, gs, G, mode, dispatch, PAC_LAUNCH_PERSISTENT, nullptr);
      if (r) v.push_back(ctx->c.last.isect_ms);
    }
    if (rc != PAC_OK) {
      ctx->c.tile = saved;
      return rc;
    }
    ms[t] = median(v);
    if (t == 0 || ms[t] < ms[best_t]) best_t = t;
  }
  ctx->c.tile = best_t;
  *tile = best_t;
  if (ms_out) std::copy(ms, ms + PAC_NUM_TILES, ms_out);
  if (cache_path) {
    std::ofstream out(cache_path, std::ios::app);
    if (out.is_open()) {
      out << key << " " << mode << " " << dispatch << " " << best_t;
      char buf[32];
      for (int t = 0; t < PAC_NUM_TILES; ++t) {
        snprintf(buf, sizeof buf, " %.17g", ms[t]);
        out << buf;
      }
      out << "\n";
    }
  }
  return PAC_OK;
}

}  // extern "C"
