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

// Every strategy on every partition, one plain launch each (LCC mode, the mode of the
// headline benchmark), timed with events on the context stream; median of `reps`.
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
  Graph *gs[1] = {g};
  PAC_TRY(zero_outputs(ctx, gs, 1, kModeLCC));
  const int nev = 2 * reps;
  std::vector<hipEvent_t> ev(nev);
  for (auto &e : ev) PAC_CHECK_HIP(hipEventCreate(&e));
  int rc = PAC_OK;
  for (uint32_t k = 0; k < g->P && rc == PAC_OK; ++k) {
    for (int s = 0; s < PAC_NUM_STRATEGIES && rc == PAC_OK; ++s) {
      if (g->h_pstart[k + 1] == g->h_pstart[k]) {
        ms[(size_t)k * PAC_NUM_STRATEGIES + s] = 0.0;
        continue;
      }
      // warm-up launch, then reps timed launches
      rc = launch_partition(ctx, g, kModeLCC, k, s);
      for (int r = 0; r < reps && rc == PAC_OK; ++r) {
        if (hipEventRecord(ev[2 * r], ctx->stream) != hipSuccess) rc = PAC_ERR_HIP;
        if (rc == PAC_OK) rc = launch_partition(ctx, g, kModeLCC, k, s);
        if (rc == PAC_OK && hipEventRecord(ev[2 * r + 1], ctx->stream) != hipSuccess)
          rc = PAC_ERR_HIP;
      }
      if (rc != PAC_OK) break;
      if (hipEventSynchronize(ev[nev - 1]) != hipSuccess) {
        set_error("event synchronize failed");
        rc = PAC_ERR_HIP;
        break;
      }
      std::vector<double> t(reps);
      for (int r = 0; r < reps; ++r) {
        float x = 0.f;
        (void)hipEventElapsedTime(&x, ev[2 * r], ev[2 * r + 1]);
        t[r] = x;
      }
      ms[(size_t)k * PAC_NUM_STRATEGIES + s] = median(t);
    }
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  return rc;
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
