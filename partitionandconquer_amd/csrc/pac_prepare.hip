// pac_prepare.hip -- symmetric CSR upload/validation and the degree-ordered orientation.
//
// Reference counterparts (Byy-0024/PartitionAndConquer):
//   Graph::Graph(csrdir, bool) / Graph(adj_list)  Graph.cpp:92-149 (CSR members)
//   orientation: the reference orients implicitly by vertex ID (the `u <= v` skip of
//   Graph.cpp:361 and the threshold merge of Graph.cpp:272-285).  Here the orientation
//   is by (degree, id), which bounds every out-list by sqrt(2m); triangle counts do
//   not depend on the orientation.
//
// Orientation = histogram / exclusive scan / scatter into a uint32 CSR:
//   1. d+(u) = |{v in N(u) : (d(u),u) < (d(v),v)}|    (one wave per row, ballot+popcount)
//   2. ooff = exclusive_scan(d+)                        (rocPRIM device scan)
//   3. each row's kept neighbours are written in input order (ballot prefix), so
//      rows stay sorted by ID and no re-sort is needed.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "pac_internal.hpp"

namespace pac {
namespace {

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ bool before(uint32_t du, uint32_t u, uint32_t dv, uint32_t v) {
  return du < dv || (du == dv && u < v);
}

uint32_t grid_waves(uint64_t n) {
  uint64_t b = (n + 3) / 4;
  if (b < 1) b = 1;
  if (b > (1u << 20)) b = 1u << 20;
  return (uint32_t)b;
}

__global__ void k_deg(const uint32_t *__restrict__ off, uint32_t V, uint32_t *__restrict__ deg,
                      uint32_t *maxdeg) {
  uint32_t mx = 0;
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < V;
       u += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = off[u + 1] - off[u];
    deg[u] = d;
    mx = max(mx, d);
  }
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, o, 64));
  // one global atomic per block (per-wave atomics on one address serialise)
  __shared__ uint32_t wmax[16];
  if ((threadIdx.x & 63u) == 0) wmax[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) mx = max(mx, wmax[w]);
    if (mx) atomicMax(maxdeg, mx);
  }
}

// Error bits: 1 ID >= V, 2 row not strictly increasing, 4 self-loop, 8 asymmetric.
__global__ void k_check_rows(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
                             uint32_t V, int check_sym, uint32_t *err) {
  const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6;
  uint32_t bits = 0;
  for (uint64_t u = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); u < V;
       u += (uint64_t)gridDim.x * wpb) {
    const uint32_t a0 = off[u], a1 = off[u + 1];
    for (uint32_t i = a0 + lane; i < a1; i += 64) {
      const uint32_t v = adj[i];
      if (v >= V) {
        bits |= 1u;
        continue;
      }
      if (i + 1 < a1 && adj[i + 1] <= v) bits |= 2u;
      if (v == u) bits |= 4u;
      if (check_sym) {
        // binary search u in N(v)
        uint32_t lo = off[v], n = off[v + 1] - lo;
        bool found = false;
        while (n > 0) {
          const uint32_t half = n >> 1;
          const uint32_t x = adj[lo + half];
          if (x == u) {
            found = true;
            break;
          }
          if (x < u) {
            lo += half + 1;
            n -= half + 1;
          } else {
            n = half;
          }
        }
        if (!found) bits |= 8u;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) bits |= (uint32_t)__shfl_xor(bits, o, 64);
  if (lane == 0 && bits) atomicOr(err, bits);
}

// 1. out-degree histogram (one wave per row)
__global__ void k_outdeg(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
                         const uint32_t *__restrict__ deg, uint32_t V,
                         uint32_t *__restrict__ odeg) {
  const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6;
  for (uint64_t u = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); u < V;
       u += (uint64_t)gridDim.x * wpb) {
    const uint32_t a0 = off[u], a1 = off[u + 1], du = a1 - a0;
    uint32_t cnt = 0;
    for (uint32_t i0 = a0; i0 < a1; i0 += 64) {
      const uint32_t i = i0 + lane;
      bool keep = false;
      if (i < a1) {
        const uint32_t v = adj[i];
        keep = before(du, (uint32_t)u, deg[v], v);
      }
      cnt += (uint32_t)__popcll(__ballot(keep));
    }
    if (lane == 0) odeg[u] = cnt;
  }
}

// 3. order-preserving scatter of the kept neighbours
__global__ void k_scatter(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
                          const uint32_t *__restrict__ deg, const uint32_t *__restrict__ ooff,
                          uint32_t V, uint32_t *__restrict__ oadj) {
  const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6;
  for (uint64_t u = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); u < V;
       u += (uint64_t)gridDim.x * wpb) {
    const uint32_t a0 = off[u], a1 = off[u + 1], du = a1 - a0;
    uint32_t out = ooff[u];
    for (uint32_t i0 = a0; i0 < a1; i0 += 64) {
      const uint32_t i = i0 + lane;
      bool keep = false;
      uint32_t v = 0;
      if (i < a1) {
        v = adj[i];
        keep = before(du, (uint32_t)u, deg[v], v);
      }
      const uint64_t mask = __ballot(keep);
      if (keep) oadj[out + lanes_below(mask)] = v;
      out += (uint32_t)__popcll(mask);
    }
  }
}

// ---- intersection pairs --------------------------------------------------------------
// The oriented edge t->h stands for the intersection N+(t) & N+(h).  Vertex x owns the
// pair {x,u} (u in N(x)) when x's out-list is the longer of the two, ties going to the
// tail:  x->u: d+(u) <= d+(x) (and d+(u) > 0, else the pair holds no triangle);
//        u->x: d+(x) >  d+(u).
// Exactly one endpoint owns each edge with d+(h) > 0 or d+(t) < d+(h).
__device__ __forceinline__ bool owns(uint32_t dx, uint32_t x, uint32_t ox, uint32_t du, uint32_t u,
                                     uint32_t ou, bool *fwd) {
  *fwd = before(dx, x, du, u);
  return *fwd ? (ou <= ox && ou > 0u) : (ox > ou);
}

__global__ void k_iota(uint32_t *__restrict__ p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

// pairs owned per vertex (one wave per vertex, over its symmetric row) and in-degree
__global__ void k_pair_count(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
                             const uint32_t *__restrict__ deg, const uint32_t *__restrict__ odeg,
                             uint32_t V, uint32_t *__restrict__ npairs,
                             uint32_t *__restrict__ nin) {
  const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6;
  for (uint64_t x = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); x < V;
       x += (uint64_t)gridDim.x * wpb) {
    const uint32_t a0 = off[x], a1 = off[x + 1], dx = a1 - a0, ox = odeg[x];
    uint32_t cnt = 0;
    for (uint32_t i0 = a0; i0 < a1; i0 += 64) {
      const uint32_t i = i0 + lane;
      bool own = false;
      if (i < a1) {
        const uint32_t u = adj[i];
        bool fwd;
        own = owns(dx, (uint32_t)x, ox, deg[u], u, odeg[u], &fwd);
      }
      cnt += (uint32_t)__popcll(__ballot(own));
    }
    if (lane == 0) {
      npairs[x] = cnt;
      nin[x] = dx - ox;
    }
  }
}

// Order-preserving fill: x's pairs come out sorted by partner ID.  The edge index of a
// pair x->u is ooff[x] + (rank of u among x's out-neighbours); of u->x it is the next
// entry of x's in-edge list, which holds the edges u->x in u order.
__global__ void k_pair_fill(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
                            const uint32_t *__restrict__ deg, const uint32_t *__restrict__ odeg,
                            const uint32_t *__restrict__ ooff, const uint32_t *__restrict__ oadj,
                            const uint32_t *__restrict__ poff, const uint32_t *__restrict__ inoff,
                            const uint32_t *__restrict__ inidx, uint32_t V,
                            uint32_t *__restrict__ ppar, uint32_t *__restrict__ pedge) {
  const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6;
  for (uint64_t x = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); x < V;
       x += (uint64_t)gridDim.x * wpb) {
    const uint32_t a0 = off[x], a1 = off[x + 1], dx = a1 - a0, ox = odeg[x];
    uint32_t out = poff[x], rank = ooff[x], in = inoff[x];
    for (uint32_t i0 = a0; i0 < a1; i0 += 64) {
      const uint32_t i = i0 + lane;
      bool own = false, fwd = false;
      uint32_t u = 0;
      if (i < a1) {
        u = adj[i];
        own = owns(dx, (uint32_t)x, ox, deg[u], u, odeg[u], &fwd);
      }
      const bool back = i < a1 && !fwd;
      const uint64_t fmask = __ballot(fwd);
      const uint64_t omask = __ballot(own);
      const uint64_t bmask = __ballot(back);
      uint32_t e = 0;
      if (fwd) {
        e = rank + lanes_below(fmask);
      } else if (back) {
        e = inidx[in + lanes_below(bmask)];
      }
      if (own) {
        const uint32_t k = out + lanes_below(omask);
        ppar[k] = u;
        pedge[k] = e;
      }
      rank += (uint32_t)__popcll(fmask);
      out += (uint32_t)__popcll(omask);
      in += (uint32_t)__popcll(bmask);
    }
  }
}

}  // namespace

int build_symmetric_from_host(Graph *g, const uint64_t *vlist, const int32_t *elist) {
  Ctx *ctx = g->ctx;
  const uint64_t V = g->V, nnz = g->nnz;
  std::vector<uint32_t> off32(V + 1);
  for (uint64_t u = 0; u <= V; ++u) off32[u] = (uint32_t)vlist[u];
  PAC_TRY(g->off.reserve(sizeof(uint32_t) * (V + 1)));
  PAC_TRY(g->adj.reserve(sizeof(uint32_t) * (nnz ? nnz : 1)));
  PAC_CHECK_HIP(hipMemcpyAsync(g->off.p, off32.data(), sizeof(uint32_t) * (V + 1),
                               hipMemcpyHostToDevice, ctx->stream));
  if (nnz)
    PAC_CHECK_HIP(hipMemcpyAsync(g->adj.p, elist, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice,
                                 ctx->stream));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return PAC_OK;
}

int graph_validate(const Graph *g, int level) {
  Ctx *ctx = g->ctx;
  if (g->V == 0) return PAC_OK;
  PAC_TRY(ctx->scratch2.reserve(64));
  uint32_t *err = ctx->scratch2.as<uint32_t>();
  PAC_TRY(fill_u32(ctx->stream, err, 1, 0u));
  hipLaunchKernelGGL(k_check_rows, dim3(grid_waves(g->V)), dim3(256), 0, ctx->stream,
                     g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(), g->V,
                     level >= 1 ? 1 : 0, err);
  PAC_CHECK_HIP(hipGetLastError());
  uint32_t h = 0;
  PAC_CHECK_HIP(hipMemcpyAsync(&h, err, 4, hipMemcpyDeviceToHost, ctx->stream));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (h) {
    set_error("malformed CSR:%s%s%s%s", (h & 1) ? " neighbour ID out of range" : "",
              (h & 2) ? " row not strictly increasing (unsorted or duplicate)" : "",
              (h & 4) ? " self-loop" : "", (h & 8) ? " asymmetric (u,v) without (v,u)" : "");
    return PAC_ERR_FORMAT;
  }
  return PAC_OK;
}

int orient(Graph *g) {
  Ctx *ctx = g->ctx;
  const uint32_t V = g->V;
  PAC_TRY(g->deg.reserve(sizeof(uint32_t) * (V ? V : 1)));
  PAC_TRY(g->odeg.reserve(sizeof(uint32_t) * ((uint64_t)V + 1)));
  PAC_TRY(g->ooff.reserve(sizeof(uint32_t) * ((uint64_t)V + 1)));
  PAC_TRY(ctx->scratch2.reserve(64));
  uint32_t *maxdeg = ctx->scratch2.as<uint32_t>();
  PAC_TRY(fill_u32(ctx->stream, maxdeg, 1, 0u));
  PAC_TRY(fill_u32(ctx->stream, g->odeg.as<uint32_t>(), (uint64_t)V + 1, 0u));
  if (V) {
    hipLaunchKernelGGL(k_deg, dim3(std::max<uint32_t>(1u, std::min<uint32_t>(4096u, (V + 255) / 256))),
                       dim3(256), 0, ctx->stream,
                       g->off.as<const uint32_t>(), V, g->deg.as<uint32_t>(), maxdeg);
    hipLaunchKernelGGL(k_outdeg, dim3(grid_waves(V)), dim3(256), 0, ctx->stream,
                       g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(),
                       g->deg.as<const uint32_t>(), V, g->odeg.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
  }
  size_t tmp_bytes = 0;
  PAC_CHECK_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, g->odeg.as<uint32_t>(),
                                        g->ooff.as<uint32_t>(), 0u, (size_t)V + 1,
                                        rocprim::plus<uint32_t>(), ctx->stream));
  PAC_TRY(ctx->scratch.reserve(tmp_bytes));
  PAC_CHECK_HIP(rocprim::exclusive_scan(ctx->scratch.p, tmp_bytes, g->odeg.as<uint32_t>(),
                                        g->ooff.as<uint32_t>(), 0u, (size_t)V + 1,
                                        rocprim::plus<uint32_t>(), ctx->stream));
  uint32_t hm[2] = {0, 0};
  PAC_CHECK_HIP(hipMemcpyAsync(&hm[0], g->ooff.as<uint32_t>() + V, 4, hipMemcpyDeviceToHost,
                               ctx->stream));
  PAC_CHECK_HIP(hipMemcpyAsync(&hm[1], maxdeg, 4, hipMemcpyDeviceToHost, ctx->stream));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  g->m = hm[0];
  g->max_deg = hm[1];
  if (2 * g->m != g->nnz) {
    set_error("orientation kept %llu of %llu entries: the CSR is not symmetric",
              (unsigned long long)g->m, (unsigned long long)g->nnz);
    return PAC_ERR_FORMAT;
  }
  // m entries plus one sentinel (kListEnd) that list streams past their end read
  PAC_TRY(g->oadj.reserve(sizeof(uint32_t) * (g->m + 1)));
  if (V) {
    hipLaunchKernelGGL(k_scatter, dim3(grid_waves(V)), dim3(256), 0, ctx->stream,
                       g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(),
                       g->deg.as<const uint32_t>(), g->ooff.as<const uint32_t>(), V,
                       g->oadj.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
  }
  PAC_TRY(fill_u32(ctx->stream, g->oadj.as<uint32_t>() + g->m, 1, kListEnd));
  g->pairs_built = false;
  return PAC_OK;
}

// Pair lists = histogram / exclusive scan / order-preserving fill, like the orientation.
int build_pairs(Graph *g) {
  if (g->pairs_built) return PAC_OK;
  Ctx *ctx = g->ctx;
  const uint32_t V = g->V;
  DevBuf npairs, nin;
  PAC_TRY(npairs.reserve(sizeof(uint32_t) * ((uint64_t)V + 1)));
  PAC_TRY(nin.reserve(sizeof(uint32_t) * ((uint64_t)V + 1)));
  PAC_TRY(g->poff.reserve(sizeof(uint32_t) * ((uint64_t)V + 1)));
  PAC_TRY(g->inoff.reserve(sizeof(uint32_t) * ((uint64_t)V + 1)));
  PAC_TRY(g->inidx.reserve(sizeof(uint32_t) * (g->m ? g->m : 1)));
  PAC_TRY(fill_u32(ctx->stream, npairs.as<uint32_t>(), (uint64_t)V + 1, 0u));
  PAC_TRY(fill_u32(ctx->stream, nin.as<uint32_t>(), (uint64_t)V + 1, 0u));
  if (V) {
    hipLaunchKernelGGL(k_pair_count, dim3(grid_waves(V)), dim3(256), 0, ctx->stream,
                       g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(),
                       g->deg.as<const uint32_t>(), g->odeg.as<const uint32_t>(), V,
                       npairs.as<uint32_t>(), nin.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
  }
  size_t tmp_in = 0;
  PAC_CHECK_HIP(rocprim::exclusive_scan(nullptr, tmp_in, nin.as<uint32_t>(),
                                        g->inoff.as<uint32_t>(), 0u, (size_t)V + 1,
                                        rocprim::plus<uint32_t>(), ctx->stream));
  PAC_TRY(ctx->scratch.reserve(tmp_in));
  PAC_CHECK_HIP(rocprim::exclusive_scan(ctx->scratch.p, tmp_in, nin.as<uint32_t>(),
                                        g->inoff.as<uint32_t>(), 0u, (size_t)V + 1,
                                        rocprim::plus<uint32_t>(), ctx->stream));
  // In-edge lists = the oriented edges sorted by head (a CSR transpose).  The radix sort
  // is stable and the edges enter in tail order, so each in-list is in tail order.
  if (g->m) {
    DevBuf keys_out, vals_in;
    PAC_TRY(keys_out.reserve(sizeof(uint32_t) * g->m));
    PAC_TRY(vals_in.reserve(sizeof(uint32_t) * g->m));
    hipLaunchKernelGGL(k_iota, dim3(std::min<uint64_t>(1u << 16, (g->m + 255) / 256)), dim3(256), 0,
                       ctx->stream, vals_in.as<uint32_t>(), g->m);
    PAC_CHECK_HIP(hipGetLastError());
    uint32_t bits = 1;
    while (bits < 32 && (1ull << bits) < V) ++bits;
    size_t tmp_sort = 0;
    PAC_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, tmp_sort, g->oadj.as<const uint32_t>(),
                                            keys_out.as<uint32_t>(), vals_in.as<const uint32_t>(),
                                            g->inidx.as<uint32_t>(), (size_t)g->m, 0u, bits,
                                            ctx->stream));
    PAC_TRY(ctx->scratch.reserve(tmp_sort));
    PAC_CHECK_HIP(rocprim::radix_sort_pairs(ctx->scratch.p, tmp_sort, g->oadj.as<const uint32_t>(),
                                            keys_out.as<uint32_t>(), vals_in.as<const uint32_t>(),
                                            g->inidx.as<uint32_t>(), (size_t)g->m, 0u, bits,
                                            ctx->stream));
    // keys_out and vals_in are freed on return
    PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  }
  size_t tmp_bytes = 0;
  PAC_CHECK_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, npairs.as<uint32_t>(),
                                        g->poff.as<uint32_t>(), 0u, (size_t)V + 1,
                                        rocprim::plus<uint32_t>(), ctx->stream));
  PAC_TRY(ctx->scratch.reserve(tmp_bytes));
  PAC_CHECK_HIP(rocprim::exclusive_scan(ctx->scratch.p, tmp_bytes, npairs.as<uint32_t>(),
                                        g->poff.as<uint32_t>(), 0u, (size_t)V + 1,
                                        rocprim::plus<uint32_t>(), ctx->stream));
  uint32_t np = 0;
  PAC_CHECK_HIP(hipMemcpyAsync(&np, g->poff.as<uint32_t>() + V, 4, hipMemcpyDeviceToHost,
                               ctx->stream));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (np > g->m) {
    set_error("pair construction produced %u pairs for %llu oriented edges", np,
              (unsigned long long)g->m);
    return PAC_ERR_HIP;
  }
  PAC_TRY(g->ppar.reserve(sizeof(uint32_t) * (np ? np : 1)));
  PAC_TRY(g->pedge.reserve(sizeof(uint32_t) * (np ? np : 1)));
  if (V && g->m) {
    hipLaunchKernelGGL(k_pair_fill, dim3(grid_waves(V)), dim3(256), 0, ctx->stream,
                       g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(),
                       g->deg.as<const uint32_t>(), g->odeg.as<const uint32_t>(),
                       g->ooff.as<const uint32_t>(), g->oadj.as<const uint32_t>(),
                       g->poff.as<const uint32_t>(), g->inoff.as<const uint32_t>(),
                       g->inidx.as<const uint32_t>(), V, g->ppar.as<uint32_t>(),
                       g->pedge.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
  }
  // npairs and nin are freed on return: finish their readers first
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  g->pairs_built = true;
  return PAC_OK;
}

namespace {

// Number of pairs streaming each vertex's out-list.
__global__ void k_partner_count(const uint32_t *__restrict__ ppar, uint64_t np,
                                uint32_t *__restrict__ R) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np;
       p += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&R[ppar[p]], 1u);
}

// Hit words per vertex: pairs streaming it x words per stream; entry V is 0, so the
// exclusive scan ends with the total.
__global__ void k_word_counts(const uint32_t *__restrict__ R, const uint32_t *__restrict__ odeg,
                              uint32_t V, unsigned long long *__restrict__ w) {
  for (uint64_t y = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; y <= V;
       y += (uint64_t)gridDim.x * blockDim.x)
    w[y] = y < V ? (unsigned long long)R[y] * ((odeg[y] + 63u) >> 6) : 0ull;
}

// First hit word of each pair: its partner's word base plus its rank among the pairs
// streaming that partner times the words per stream.  The ranks come in any order (the
// words of all of a vertex's streams are summed, so only their distinctness matters).
__global__ void k_pair_words(const uint32_t *__restrict__ ppar, uint64_t np,
                             const uint32_t *__restrict__ odeg,
                             const unsigned long long *__restrict__ wbase,
                             uint32_t *__restrict__ cursor, unsigned long long *__restrict__ pword) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t y = ppar[p];
    const uint32_t r = atomicAdd(&cursor[y], 1u);
    pword[p] = wbase[y] + (unsigned long long)r * ((odeg[y] + 63u) >> 6);
  }
}

uint32_t grid_elems(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 1u << 16));
}

}  // namespace

int build_hit_words(Graph *g) {
  if (g->hitw_built) return PAC_OK;
  PAC_TRY(build_pairs(g));
  Ctx *ctx = g->ctx;
  const uint32_t V = g->V;
  uint32_t np = 0;
  PAC_CHECK_HIP(hipMemcpy(&np, g->poff.as<uint32_t>() + V, 4, hipMemcpyDeviceToHost));
  PAC_TRY(g->pstreams.reserve(sizeof(uint32_t) * ((uint64_t)V + 1)));
  PAC_TRY(g->pword.reserve(sizeof(unsigned long long) * (np ? np : 1)));
  PAC_TRY(g->wbase.reserve(sizeof(unsigned long long) * ((uint64_t)V + 1)));
  DevBuf words, cursor;
  PAC_TRY(words.reserve(sizeof(unsigned long long) * ((uint64_t)V + 1)));
  PAC_TRY(cursor.reserve(sizeof(uint32_t) * ((uint64_t)V + 1)));
  PAC_TRY(fill_u32(ctx->stream, g->pstreams.as<uint32_t>(), (uint64_t)V + 1, 0u));
  PAC_TRY(fill_u32(ctx->stream, cursor.as<uint32_t>(), (uint64_t)V + 1, 0u));
  if (np) {
    hipLaunchKernelGGL(k_partner_count, dim3(grid_elems(np)), dim3(256), 0, ctx->stream,
                       g->ppar.as<const uint32_t>(), (uint64_t)np, g->pstreams.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_word_counts, dim3(grid_elems((uint64_t)V + 1)), dim3(256), 0, ctx->stream,
                     g->pstreams.as<const uint32_t>(), g->odeg.as<const uint32_t>(), V,
                     words.as<unsigned long long>());
  PAC_CHECK_HIP(hipGetLastError());
  size_t tmp = 0;
  PAC_CHECK_HIP(rocprim::exclusive_scan(nullptr, tmp, words.as<unsigned long long>(),
                                        g->wbase.as<unsigned long long>(), 0ull, (size_t)V + 1,
                                        rocprim::plus<unsigned long long>(), ctx->stream));
  PAC_TRY(ctx->scratch.reserve(tmp));
  PAC_CHECK_HIP(rocprim::exclusive_scan(ctx->scratch.p, tmp, words.as<unsigned long long>(),
                                        g->wbase.as<unsigned long long>(), 0ull, (size_t)V + 1,
                                        rocprim::plus<unsigned long long>(), ctx->stream));
  if (np) {
    hipLaunchKernelGGL(k_pair_words, dim3(grid_elems(np)), dim3(256), 0, ctx->stream,
                       g->ppar.as<const uint32_t>(), (uint64_t)np, g->odeg.as<const uint32_t>(),
                       g->wbase.as<const unsigned long long>(), cursor.as<uint32_t>(),
                       g->pword.as<unsigned long long>());
    PAC_CHECK_HIP(hipGetLastError());
  }
  unsigned long long nwords = 0;
  PAC_CHECK_HIP(hipMemcpyAsync(&nwords, g->wbase.as<unsigned long long>() + V, 8,
                               hipMemcpyDeviceToHost, ctx->stream));
  std::vector<uint32_t> R(V);
  if (V)
    PAC_CHECK_HIP(hipMemcpyAsync(R.data(), g->pstreams.p, 4ull * V, hipMemcpyDeviceToHost,
                                 ctx->stream));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // words and cursor are freed on return
  // Every pair writes all of its words in each SUP launch; the zero fill only makes the
  // words of a graph that has never been solved well defined.
  // 8 words of padding: the column sums load 8 words per stream from any word position
  PAC_TRY(g->hitw.reserve(sizeof(unsigned long long) * (nwords + 8)));
  PAC_TRY(fill_u32(ctx->stream, g->hitw.as<uint32_t>(), 2ull * (nwords + 8), 0u));
  // units (y, first stream) of kSupRun streams of the vertices streamed more than 64 times
  g->h_supbig.clear();
  for (uint32_t y = 0; y < V; ++y)
    if (R[y] > 64u)
      for (uint32_t r0 = 0; r0 < R[y]; r0 += kSupRun) {
        g->h_supbig.push_back(y);
        g->h_supbig.push_back(r0);
      }
  if (!g->h_supbig.empty()) {
    PAC_TRY(g->supbig.reserve(sizeof(uint32_t) * g->h_supbig.size()));
    PAC_CHECK_HIP(hipMemcpy(g->supbig.p, g->h_supbig.data(),
                            sizeof(uint32_t) * g->h_supbig.size(), hipMemcpyHostToDevice));
  }
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  g->hitw_built = true;
  return PAC_OK;
}

}  // namespace pac
