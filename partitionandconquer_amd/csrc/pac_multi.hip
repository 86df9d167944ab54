// pac_multi.hip -- multi-way intersections over the symmetric CSR: k-way common
// neighbours, friend triangles and square counting.
//
// Reference counterparts (Byy-0024/PartitionAndConquer):
//   Graph::cnt_common_neis_{galloping,merge}(const vector<VID>&)   Graph.cpp:200-225, 255-270
//   Graph::batch_cnt_common_neis_{galloping,merge}(vector<vector<VID>>)  Graph.cpp:423-429, 441-447
//   Graph::solve_friend_triangle_{galloping,merge}                Graph.cpp:309-338
//   Partitioned_Graph::solve_friend_triangle_*                    Graph.cpp:1222-1264
//   Partitioned_Graph::solve_square_counting                      Graph.cpp:1291-1324
//   Graph::get_common_neis / Partitioned_Graph::get_common_neis, get_neis
//                                                   Graph.cpp:468-493, 1475-1498
//   per-partition k-way counts (Partitioned_Graph)                 Graph.cpp:1087-1130
//
// All three count exactly; their results do not depend on the partitioning (the
// partitioned reference versions sum per-partition partials of the same quantity).
#include <algorithm>

#include "pac_internal.hpp"

namespace pac {
namespace {

// Branch-free lower_bound over a sorted list in global memory: first i in [0,n) with
// a[i] >= x (n if none).  The trip count depends only on n.
__device__ __forceinline__ uint32_t lbound(const uint32_t *a, uint32_t n, uint32_t x) {
  if (n == 0) return 0;
  uint32_t base = 0;
  while (n > 1) {
    const uint32_t half = n >> 1;
    base = (a[base + half] < x) ? base + half : base;
    n -= half;
  }
  return base + (a[base] < x ? 1u : 0u);
}

__device__ __forceinline__ unsigned long long wsum64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// L2-coherent loads and stores of the square-counting scratch (bypass the per-CU L1,
// which the returning atomics of other waves do not update).
__device__ __forceinline__ uint32_t ld_l2(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_l2(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- k-way common neighbours: one wave per query ---------------------------------
// The minimum-degree list of the query drives (as in the galloping reference); each of
// its elements is binary-searched in every other list, and the wave stops searching
// once no lane's element is still present in all lists seen so far.
// With part not null only common neighbours w with part[w] == pid count (the per-partition
// form, Graph.cpp:1087-1130).
__global__ void k_kway(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
                       const uint64_t *__restrict__ qoff, const int32_t *__restrict__ qv,
                       uint64_t nq, unsigned long long *__restrict__ res,
                       const uint32_t *__restrict__ part, uint32_t pid) {
  const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6;
  for (uint64_t q = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); q < nq;
       q += (uint64_t)gridDim.x * wpb) {
    const uint64_t k0 = qoff[q], k1 = qoff[q + 1];
    uint32_t bd = 0xFFFFFFFFu, bv = 0;
    for (uint64_t i = k0 + lane; i < k1; i += 64) {
      const uint32_t v = (uint32_t)qv[i], d = off[v + 1] - off[v];
      if (d < bd || (d == bd && v < bv)) {
        bd = d;
        bv = v;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint32_t od = __shfl_xor(bd, o, 64), ov = __shfl_xor(bv, o, 64);
      if (od < bd || (od == bd && ov < bv)) {
        bd = od;
        bv = ov;
      }
    }
    const uint32_t u = bv, ub = off[u], du = bd;
    uint32_t cnt = 0;
    for (uint32_t j0 = 0; j0 < du; j0 += 64) {
      const uint32_t j = j0 + lane;
      bool alive = j < du;
      const uint32_t x = alive ? adj[ub + j] : 0u;
      if (part && alive) alive = part[x] == pid;
      for (uint64_t i = k0; i < k1; ++i) {  // wave-uniform
        if (!__any(alive)) break;
        const uint32_t v = (uint32_t)qv[i];
        if (v == u) continue;
        const uint32_t vb = off[v], nv = off[v + 1] - vb;
        const uint32_t p = lbound(adj + vb, nv, x);
        alive = alive && p < nv && adj[vb + p] == x;
      }
      cnt += (uint32_t)__popcll(__ballot(alive));
    }
    if (lane == 0) res[q] = cnt;
  }
}

// ---- common-neighbour lists (one workgroup of 256) ---------------------------------
// The values of the sorted list L that are neighbours of u (each value once; adjacent
// duplicates of L are skipped) and, with part not null, lie in partition pid, written in
// ascending order: a block-wide ballot compaction per 256 elements of L.  Matches the
// merges of Graph.cpp:468-493 and 1475-1493 for a sorted L.
__global__ void __launch_bounds__(256)
k_common_list(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj, uint32_t u,
              const int32_t *__restrict__ L, uint64_t n, const uint32_t *__restrict__ part,
              uint32_t pid, int32_t *__restrict__ out, unsigned long long *__restrict__ count) {
  __shared__ uint32_t wcnt[4];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const uint32_t ub = off[u], du = off[u + 1] - ub;
  uint64_t base = 0;
  for (uint64_t i0 = 0; i0 < n; i0 += 256) {
    const uint64_t i = i0 + tid;
    bool keep = false;
    uint32_t x = 0;
    if (i < n) {
      x = (uint32_t)L[i];
      keep = i == 0 || (uint32_t)L[i - 1] != x;
      if (keep) {
        const uint32_t p = lbound(adj + ub, du, x);
        keep = p < du && adj[ub + p] == x;  // x is then a vertex ID
      }
      if (keep && part) keep = part[x] == pid;
    }
    const uint64_t bal = __ballot(keep);
    if (lane == 0) wcnt[w] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      const uint32_t c = wcnt[k];
      pre += k < w ? c : 0u;
      tot += c;
    }
    if (keep) out[base + pre + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = (int32_t)x;
    base += tot;
    __syncthreads();  // wcnt is rewritten by the next chunk
  }
  if (tid == 0) *count = base;
}

// Contiguous partitions: part[v] = v / width.
__global__ void k_part_contig(uint32_t *__restrict__ part, uint32_t V, uint64_t width) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    part[v] = (uint32_t)(v / width);
}

// ---- friend triangles --------------------------------------------------------------
// sum over index pairs i < j (nodes[i], nodes[j] != root) of |N(root) & N(ni) & N(nj)|
//   = sum over x in N(root) of C(c(x), 2), c(x) = #{i : nodes[i] != root, x in N(ni)}.
// One wave per node: the shorter of N(root) and N(node) is streamed and looked up in the
// other; a hit adds 1 to c at x's position in N(root).
__global__ void k_friend_hits(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
                              uint32_t root, const int32_t *__restrict__ nodes, uint64_t n,
                              uint32_t *__restrict__ c) {
  const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6;
  const uint32_t rb = off[root], rn = off[root + 1] - rb;
  for (uint64_t i = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); i < n;
       i += (uint64_t)gridDim.x * wpb) {
    const uint32_t v = (uint32_t)nodes[i];
    if (v == root) continue;
    const uint32_t vb = off[v], vn = off[v + 1] - vb;
    if (vn <= rn) {
      for (uint32_t j = lane; j < vn; j += 64) {
        const uint32_t x = adj[vb + j];
        const uint32_t p = lbound(adj + rb, rn, x);
        if (p < rn && adj[rb + p] == x) atomicAdd(&c[p], 1u);
      }
    } else {
      for (uint32_t p = lane; p < rn; p += 64) {
        const uint32_t x = adj[rb + p];
        const uint32_t j = lbound(adj + vb, vn, x);
        if (j < vn && adj[vb + j] == x) atomicAdd(&c[p], 1u);
      }
    }
  }
}

__global__ void k_friend_sum(const uint32_t *__restrict__ c, uint32_t rn,
                             unsigned long long *__restrict__ res) {
  unsigned long long s = 0;
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < rn;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long x = c[p];
    s += x * (x - (x ? 1ull : 0ull)) / 2ull;
  }
  s = wsum64(s);
  if ((threadIdx.x & 63u) == 0 && s) atomicAdd(res, s);
}

// ---- square counting ---------------------------------------------------------------
// The reference sums, over every u and neighbours w < v < u of u, the common neighbours
// x < w of v and w: each 4-cycle u-v-x-w whose largest and smallest IDs sit at opposite
// corners, once.  Regrouped by v:
//   sum over v, w < v of a(v,w) * b(v,w),
//   a(v,w) = #{u > v : u ~ v, u ~ w},  b(v,w) = #{x < w : x ~ v, x ~ w}.
// A workgroup takes one v at a time from a queue.  Phase A counts a(v,w) for every w
// reached by a path v - u - w (u > v > w) in a per-workgroup array cnt[V] (touched
// entries listed for clearing); phase B adds cnt[w] for every path v - x - w with
// x < w < v, which sums a*b; phase C clears the touched entries.
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_squares(const uint32_t *__restrict__ off,
                                                        const uint32_t *__restrict__ adj,
                                                        uint32_t V, uint32_t *queue,
                                                        uint32_t *scratch,
                                                        unsigned long long *total) {
  constexpr uint32_t BS = 64u * WAVES;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  uint32_t *cnt = scratch + (uint64_t)blockIdx.x * 2u * V;
  uint32_t *touched = cnt + V;
  __shared__ uint32_t sv, ntouch;
  unsigned long long acc = 0;
  while (true) {
    __syncthreads();
    if (tid == 0) {
      sv = atomicAdd(queue, 1u);
      ntouch = 0;
    }
    __syncthreads();
    const uint32_t v = sv;
    if (v >= V) break;
    const uint32_t vb = off[v], dv = off[v + 1] - vb;
    const uint32_t split = lbound(adj + vb, dv, v);  // N(v)[split..] > v
    // phase A: u in N(v), u > v; w in N(u), w < v
    for (uint32_t iu = split + wave; iu < dv; iu += WAVES) {
      const uint32_t u = adj[vb + iu], ub = off[u], ue = off[u + 1];
      for (uint32_t j0 = ub; j0 < ue; j0 += 64) {
        const uint32_t j = j0 + lane;
        const uint32_t w = j < ue ? adj[j] : 0xFFFFFFFFu;
        const bool ok = j < ue && w < v;
        if (ok && atomicAdd(&cnt[w], 1u) == 0u) {
          const uint32_t k = atomicAdd(&ntouch, 1u);
          st_l2(&touched[k], w);
        }
        if (!__any(ok)) break;  // rows are sorted: nothing below v follows
      }
    }
    __syncthreads();
    // phase B: x in N(v), x < v; w in N(x), x < w < v
    for (uint32_t ix = wave; ix < split; ix += WAVES) {
      const uint32_t x = adj[vb + ix], xb = off[x], xe = off[x + 1];
      for (uint32_t j0 = xb; j0 < xe; j0 += 64) {
        const uint32_t j = j0 + lane;
        const uint32_t w = j < xe ? adj[j] : 0xFFFFFFFFu;
        if (w > x && w < v) acc += ld_l2(&cnt[w]);
        if (!__any(w < v)) break;
      }
    }
    __syncthreads();
    // phase C: clear what phase A touched
    const uint32_t nt = ntouch;
    for (uint32_t i = tid; i < nt; i += BS) st_l2(&cnt[ld_l2(&touched[i])], 0u);
  }
  acc = wsum64(acc);
  if (lane == 0 && acc) atomicAdd(total, acc);
}

uint32_t blocks_for(uint64_t n, uint32_t per_block) {
  uint64_t b = (n + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > (1u << 20)) b = 1u << 20;
  return (uint32_t)b;
}

__global__ void k_zero_u32(uint32_t *p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

}  // namespace

int kway_counts(Ctx *ctx, Graph *g, const uint64_t *d_qoff, const int32_t *d_qv, uint64_t nq,
                uint64_t *d_res, const uint32_t *d_part, uint32_t pid) {
  if (nq == 0) return PAC_OK;
  hipLaunchKernelGGL(k_kway, dim3(blocks_for(nq, 4)), dim3(256), 0, ctx->stream,
                     g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(), d_qoff, d_qv, nq,
                     reinterpret_cast<unsigned long long *>(d_res), d_part, pid);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

int common_list(Ctx *ctx, Graph *g, uint32_t u, const int32_t *d_list, uint64_t n,
                const uint32_t *d_part, uint32_t pid, int32_t *d_out,
                unsigned long long *d_count) {
  hipLaunchKernelGGL(k_common_list, dim3(1), dim3(256), 0, ctx->stream,
                     g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(), u, d_list, n,
                     d_part, pid, d_out, d_count);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

int partition_map(Graph *g, const uint32_t **d_part) {
  if (!g->part_ready) {
    PAC_TRY(g->part.reserve(4ull * (g->V ? g->V : 1)));
    if (!g->h_v2p.empty()) {
      PAC_CHECK_HIP(hipMemcpyAsync(g->part.p, g->h_v2p.data(), 4ull * g->V,
                                   hipMemcpyHostToDevice, g->ctx->stream));
      PAC_CHECK_HIP(hipStreamSynchronize(g->ctx->stream));
    } else if (g->V) {
      hipLaunchKernelGGL(k_part_contig, dim3(blocks_for(g->V, 256 * 8)), dim3(256), 0,
                         g->ctx->stream, g->part.as<uint32_t>(), g->V, g->part_width);
      PAC_CHECK_HIP(hipGetLastError());
    }
    g->part_ready = true;
  }
  *d_part = g->part.as<const uint32_t>();
  return PAC_OK;
}

int friend_triangles(Ctx *ctx, Graph *g, uint32_t root, const int32_t *d_nodes, uint64_t n,
                     uint32_t root_deg, uint32_t *d_c, uint64_t *d_res) {
  if (root_deg) {
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(root_deg, 256)), dim3(256), 0, ctx->stream,
                       d_c, (uint64_t)root_deg);
    PAC_CHECK_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream,
                     reinterpret_cast<uint32_t *>(d_res), (uint64_t)2);
  PAC_CHECK_HIP(hipGetLastError());
  if (n == 0 || root_deg == 0) return PAC_OK;
  hipLaunchKernelGGL(k_friend_hits, dim3(blocks_for(n, 4)), dim3(256), 0, ctx->stream,
                     g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(), root, d_nodes, n,
                     d_c);
  PAC_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_friend_sum, dim3(blocks_for(root_deg, 256 * 8)), dim3(256), 0,
                     ctx->stream, d_c, root_deg,
                     reinterpret_cast<unsigned long long *>(d_res));
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

int square_count(Ctx *ctx, Graph *g, uint64_t *h_res) {
  *h_res = 0;
  const uint32_t V = g->V;
  if (V == 0 || g->nnz == 0) return PAC_OK;
  // one cnt[V] + touched[V] scratch pair per workgroup, within a 1 GiB budget
  const uint64_t per_wg = 8ull * V;
  uint64_t nwg = (uint64_t)std::max(1, ctx->num_cus) * 2;
  nwg = std::max<uint64_t>(1, std::min<uint64_t>(nwg, (1ull << 30) / per_wg));
  DevBuf scratch, q;
  PAC_TRY(scratch.reserve(per_wg * nwg));
  PAC_TRY(q.reserve(64));
  hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(2ull * V * nwg, 256 * 8)), dim3(256), 0,
                     ctx->stream, scratch.as<uint32_t>(), 2ull * V * nwg);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, q.as<uint32_t>(),
                     (uint64_t)4);
  PAC_CHECK_HIP(hipGetLastError());
  unsigned long long *tot = reinterpret_cast<unsigned long long *>(q.as<uint32_t>() + 2);
  hipLaunchKernelGGL(k_squares<4>, dim3((uint32_t)nwg), dim3(256), 0, ctx->stream,
                     g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(), V,
                     q.as<uint32_t>(), scratch.as<uint32_t>(), tot);
  PAC_CHECK_HIP(hipGetLastError());
  PAC_CHECK_HIP(hipMemcpyAsync(h_res, tot, 8, hipMemcpyDeviceToHost, ctx->stream));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // scratch and q are freed on return
  return PAC_OK;
}

}  // namespace pac
