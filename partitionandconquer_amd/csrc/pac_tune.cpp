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

// Times the batched persistent TC kernel for every tile configuration and keeps the
// fastest.  The cache file holds lines "<shape-key> <tile>".
int pac_autotune(pac_ctx *ctx, pac_graph *const *gs, uint32_t G, const char *cache_path,
                 int *tile, double *ms_out) {
  if (!ctx || !gs || G == 0 || !tile) {
    set_error("need a context, graphs and a tile output");
    return PAC_ERR_ARG;
  }
  const std::string key = shape_key(gs, G);
  if (cache_path && !ms_out) {
    std::ifstream in(cache_path);
    std::string k;
    int t;
    while (in >> k >> t)
      if (k == key && t >= 0 && t < PAC_NUM_TILES) {
        ctx->c.tile = t;
        *tile = t;
        return PAC_OK;
      }
  }
  std::vector<uint64_t> totals(G);
  double best = 1e300;
  int best_t = ctx->c.tile;
  const int saved = ctx->c.tile;
  for (int t = 0; t < PAC_NUM_TILES; ++t) {
    ctx->c.tile = t;
    PAC_TRY(pac_tc_batch(ctx, gs, G, PAC_DISPATCH_FIXED, PAC_LAUNCH_PERSISTENT, totals.data()));
    std::vector<double> v;
    for (int r = 0; r < 3; ++r) {
      PAC_TRY(pac_tc_batch(ctx, gs, G, PAC_DISPATCH_FIXED, PAC_LAUNCH_PERSISTENT, totals.data()));
      v.push_back(ctx->c.last.total_ms);
    }
    const double m = median(v);
    if (ms_out) ms_out[t] = m;
    if (m < best) {
      best = m;
      best_t = t;
    }
  }
  (void)saved;
  ctx->c.tile = best_t;
  *tile = best_t;
  if (cache_path) {
    std::ofstream out(cache_path, std::ios::app);
    if (out.is_open()) out << key << " " << best_t << "\n";
  }
  return PAC_OK;
}

}  // extern "C"
