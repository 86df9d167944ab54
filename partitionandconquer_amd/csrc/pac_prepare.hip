// pac_prepare.hip -- symmetric CSR upload/validation, the degree-ordered orientation and
// the intersection pairs.
//
// Reference counterparts (Byy-0024/PartitionAndConquer):
//   Graph::Graph(csrdir, bool) / Graph(adj_list)  Graph.cpp:92-149 (CSR members)
//   orientation: the reference orients implicitly by vertex ID (the `u <= v` skip of
//   Graph.cpp:361 and the threshold merge of Graph.cpp:272-285).  Here the orientation
//   is by (degree, id), which bounds every out-list by sqrt(2m); triangle counts do
//   not depend on the orientation.
//
// Orientation = flags / scan / keys / sort into a uint32 CSR, edge-parallel throughout:
//   1. one keep bit per symmetric entry (u, v): (d(u),u) < (d(v),v), written as aligned
//      64-bit words (word k holds entries 64k .. 64k+63); each lane finds its row by a
//      short search between the rows of its chunk's first entry and the next chunk's;
//   2. exclusive scan of the words' popcounts: the kept entry i goes to oadj position
//      wpre[i/64] + (kept entries below it in its word), and ooff[u] is that same
//      ordinal of off[u];
//   3. each kept entry (u, v) writes the key (u, rank(v)) at its position, rank(v) being
//      v's place among all vertices sorted by (degree, id), and one radix sort of the keys
//      puts every row's out-neighbours in rank order -- the order of the orientation
//      itself, so the neighbours of N+(t) that rank above one of them, h, are a suffix of
//      the list: the only part of N+(t) that can meet N+(h).
// The intersection pairs are built the same way from one ownership bit per entry.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

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

// Entries are handled in aligned words of 64 and chunks of kChunkWords words; crow[c]
// is the row holding entry c * kChunk (the first entry of chunk c).
constexpr uint32_t kChunkWords = 4, kChunk = 64 * kChunkWords;

__host__ __device__ __forceinline__ uint64_t num_chunks(uint64_t nnz) { return (nnz + kChunk - 1) / kChunk; }

// crow: each non-empty row writes the chunk starts that fall inside it (thread per row).
__global__ void k_chunk_rows(const uint32_t *__restrict__ off, uint32_t V,
                             uint32_t *__restrict__ crow) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < V;
       u += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t a = off[u], b = off[u + 1];
    for (uint64_t c = (a + kChunk - 1) / kChunk; c * kChunk < b; ++c) crow[c] = (uint32_t)u;
  }
}

// The row holding entry i: the largest t in [lo, hi] with off[t] <= i, given off[lo] <= i
// and that i lies before the first entry of row hi + 1.  (Empty rows t have off[t] =
// off[t + 1], so the largest such t is the non-empty row holding i.)  The trip count
// depends only on hi - lo, which is uniform across a chunk's lanes.
__device__ __forceinline__ uint32_t row_of(const uint32_t *__restrict__ off, uint32_t lo,
                                           uint32_t hi, uint32_t i) {
  uint32_t n = hi - lo + 1;
  while (n > 1) {
    const uint32_t half = n >> 1;
    if (off[lo + half] <= i) {
      lo += half;
      n -= half;
    } else {
      n = half;
    }
  }
  return lo;
}

// The rows bounding chunk c: lo holds its first entry, hi its last (or a later row).
__device__ __forceinline__ void chunk_rows(const uint32_t *__restrict__ crow, uint64_t c,
                                           uint64_t nchunks, uint32_t V, uint32_t &lo,
                                           uint32_t &hi) {
  lo = crow[c];
  hi = c + 1 < nchunks ? crow[c + 1] : V - 1;
}

__device__ __forceinline__ bool bit_of(unsigned long long w, uint32_t b) {
  return (w >> b) & 1ull;
}

// Position of entry i among the set bits of flags: wpre[i/64] + the set bits below it
// in its word (wpre = exclusive scan of the words' popcounts).
__device__ __forceinline__ uint32_t ordinal(const unsigned long long *__restrict__ flags,
                                            const uint32_t *__restrict__ wpre, uint32_t i) {
  const uint32_t k = i >> 6, r = i & 63u;
  return wpre[k] + (uint32_t)__popcll(flags[k] & ((1ull << r) - 1ull));
}

// 1. keep bits: entry (t, v) of row t is kept when (d(t),t) < (d(v),v).  One wave per
// chunk; word k is written once, by the wave of its chunk.
__global__ void __launch_bounds__(256)
k_orient_flags(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
               const uint32_t *__restrict__ deg, const uint32_t *__restrict__ crow, uint32_t V,
               uint32_t nnz, unsigned long long *__restrict__ flags) {
  const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6;
  const uint64_t nchunks = num_chunks(nnz);
  for (uint64_t c = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); c < nchunks;
       c += (uint64_t)gridDim.x * wpb) {
    uint32_t lo, hi;
    chunk_rows(crow, c, nchunks, V, lo, hi);
#pragma unroll
    for (uint32_t j = 0; j < kChunkWords; ++j) {
      const uint64_t i = c * kChunk + 64u * j + lane;
      bool keep = false;
      if (i < nnz) {
        const uint32_t t = row_of(off, lo, hi, (uint32_t)i);
        const uint32_t v = adj[i];
        keep = before(deg[t], t, deg[v], v);
      }
      const unsigned long long mask = __ballot(keep);
      const uint64_t k = c * kChunkWords + j;
      if (lane == 0 && 64u * k < nnz) flags[k] = mask;
    }
  }
}

// cnt[k] = set bits of word k (n words).
__global__ void k_popc_words(const unsigned long long *__restrict__ flags, uint64_t n,
                             uint32_t *__restrict__ cnt) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n;
       k += (uint64_t)gridDim.x * blockDim.x)
    cnt[k] = (uint32_t)__popcll(flags[k]);
}

// Row offsets over the set bits: out_off[u] = ordinal(off[u]) for u <= V, out_len[u] =
// the row's set bits (u < V; may be null).
__global__ void k_ordinal_offsets(const uint32_t *__restrict__ off,
                                  const unsigned long long *__restrict__ flags,
                                  const uint32_t *__restrict__ wpre, uint32_t V,
                                  uint32_t *__restrict__ out_off, uint32_t *__restrict__ out_len) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u <= V;
       u += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t a = ordinal(flags, wpre, off[u]);
    out_off[u] = a;
    if (out_len && u < V) out_len[u] = ordinal(flags, wpre, off[u + 1]) - a;
  }
}

// (degree, id) keys of the vertices; sorted, their low words list the vertices in rank
// order.
__global__ void k_vertex_keys(const uint32_t *__restrict__ deg, uint32_t V,
                              unsigned long long *__restrict__ keys) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < V;
       u += (uint64_t)gridDim.x * blockDim.x)
    keys[u] = ((unsigned long long)deg[u] << 32) | u;
}

// rank[v] = the position of v in (degree, id) order, from the sorted keys.
__global__ void k_rank_scatter(const unsigned long long *__restrict__ keys, uint32_t V,
                               uint32_t *__restrict__ rank) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < V;
       i += (uint64_t)gridDim.x * blockDim.x)
    rank[(uint32_t)keys[i]] = (uint32_t)i;
}

// 3. the key of each kept entry (t, v) at its oadj position: t << L | rank(v).  One wave
// per chunk, as k_orient_flags (the row is needed for the key).
__global__ void __launch_bounds__(256)
k_orient_keys(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
              const uint32_t *__restrict__ rank, const uint32_t *__restrict__ crow,
              const unsigned long long *__restrict__ flags, const uint32_t *__restrict__ wpre,
              uint32_t V, uint32_t nnz, uint32_t L, unsigned long long *__restrict__ keys) {
  const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6;
  const uint64_t nchunks = num_chunks(nnz);
  for (uint64_t c = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); c < nchunks;
       c += (uint64_t)gridDim.x * wpb) {
    uint32_t lo, hi;
    chunk_rows(crow, c, nchunks, V, lo, hi);
#pragma unroll
    for (uint32_t j = 0; j < kChunkWords; ++j) {
      const uint64_t k = c * kChunkWords + j;
      const uint64_t i = 64u * k + lane;
      if (64u * k >= nnz) break;  // uniform
      const unsigned long long bits = flags[k];
      if (i < nnz && bit_of(bits, lane)) {
        const uint32_t t = row_of(off, lo, hi, (uint32_t)i);
        keys[wpre[k] + lanes_below(bits)] = ((unsigned long long)t << L) | rank[adj[i]];
      }
    }
  }
}

// After the sort: okey = the key's rank field, oadj = the vertex of that rank.
__global__ void k_orient_unpack(const unsigned long long *__restrict__ keys, uint64_t m,
                                uint32_t L, const unsigned long long *__restrict__ vkeys,
                                uint32_t *__restrict__ okey, uint32_t *__restrict__ oadj) {
  const unsigned long long mask = (1ull << L) - 1ull;
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < m;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = (uint32_t)(keys[j] & mask);
    okey[j] = r;
    oadj[j] = (uint32_t)vkeys[r];
  }
}

// ---- intersection pairs --------------------------------------------------------------
// The oriented edge t->h stands for the intersection N+(t) & N+(h).  Every element of
// N+(h) ranks above h, and N+(t) is in rank order with h in it, so only the suffix of
// N+(t) after h can meet N+(h): the head h owns the pair, stages N+(h) and streams that
// suffix.  Row x therefore owns the entries (x, u) with u->x (entries the orientation did
// not keep), provided d+(x) > 0 and x is not the last element of N+(u) -- otherwise the
// pair holds no triangle and is dropped.

// out[e] = e: the values of the transposing sort.
__global__ void k_iota(uint32_t *__restrict__ out, uint64_t n) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < n;
       e += (uint64_t)gridDim.x * blockDim.x)
    out[e] = (uint32_t)e;
}

// Ownership bits of the entries (x, u) of row x, the position of x in N+(u) of each owned
// entry (epos[i]), and each row's streamed length rowwork[x] = the sum of its pairs'
// suffix lengths (rowwork zeroed before): the lanes of one row are contiguous, so a
// segmented scan across the wave leaves each row's part of the word in its last lane,
// which adds it.  The position comes from tin, the oriented edges' positions sorted
// (stably) by head: an entry (x, u) whose edge is u -> x is the j-th such entry of the
// symmetric lists, j = i less the kept entries before it, and the j-th position of tin
// is that of its edge -- both orders run by head, then by tail (the rows of the symmetric
// lists and of oadj ascend).
__global__ void __launch_bounds__(256)
k_pair_flags(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
             const uint32_t *__restrict__ ooff, const uint32_t *__restrict__ tin,
             const unsigned long long *__restrict__ oflags, const uint32_t *__restrict__ owpre,
             const uint32_t *__restrict__ crow, uint32_t V, uint32_t nnz,
             unsigned long long *__restrict__ pflags, uint32_t *__restrict__ epos,
             uint32_t *__restrict__ rowwork) {
  const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6;
  const uint64_t nchunks = num_chunks(nnz);
  for (uint64_t c = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); c < nchunks;
       c += (uint64_t)gridDim.x * wpb) {
    uint32_t lo, hi;
    chunk_rows(crow, c, nchunks, V, lo, hi);
#pragma unroll
    for (uint32_t j = 0; j < kChunkWords; ++j) {
      const uint64_t k = c * kChunkWords + j;
      const uint64_t i = 64u * k + lane;
      bool own = false;
      uint32_t x = 0xFFFFFFFFu, w = 0;  // lanes past the end: no row
      if (i < nnz) {
        x = row_of(off, lo, hi, (uint32_t)i);
        const unsigned long long kept = oflags[k];
        if (!bit_of(kept, lane) && ooff[x + 1] > ooff[x]) {  // u -> x, N+(x) not empty
          const uint32_t u = adj[i];
          const uint32_t a = ooff[u], n = ooff[u + 1] - a;
          const uint32_t pos = tin[(uint32_t)i - owpre[k] - lanes_below(kept)] - a;
          if (pos + 1u < n) {
            own = true;
            w = n - pos - 1u;
            epos[i] = pos;
          }
        }
      }
      const unsigned long long mask = __ballot(own);
      if (lane == 0 && 64u * k < nnz) pflags[k] = mask;
#pragma unroll
      for (uint32_t sft = 1; sft < 64; sft <<= 1) {
        const uint32_t wu = (uint32_t)__shfl_up((int)w, sft, 64);
        const uint32_t xu = (uint32_t)__shfl_up((int)x, sft, 64);
        if (lane >= sft && xu == x) w += wu;
      }
      const uint32_t xn = (uint32_t)__shfl_down((int)x, 1, 64);
      if ((lane == 63u || xn != x) && w != 0u) atomicAdd(&rowwork[x], w);
    }
  }
}

// Pair lists: thread per entry (each wave on one aligned word); an owned entry (x, u)
// becomes pair ordinal(i) of the ownership bits, in input order, so each row's pairs
// come out sorted by partner ID.  pedge = the position of u->x in oadj, pinfo = (the
// position after it, the suffix length).
__global__ void __launch_bounds__(256)
k_pair_fill(const uint32_t *__restrict__ adj, const unsigned long long *__restrict__ pflags,
            const uint32_t *__restrict__ ppre, const uint32_t *__restrict__ ooff,
            const uint32_t *__restrict__ epos, uint32_t nnz, uint32_t *__restrict__ ppar,
            uint32_t *__restrict__ pedge, uint2 *__restrict__ pinfo) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nnz;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = i >> 6;
    const uint32_t b = (uint32_t)(i & 63u);
    const unsigned long long pb = pflags[k];
    if (bit_of(pb, b)) {
      const uint32_t p = ppre[k] + lanes_below(pb);
      const uint32_t u = adj[i];
      const uint32_t a = ooff[u], n = ooff[u + 1] - a;  // mostly one line per pair
      const uint32_t e = a + epos[i];
      ppar[p] = u;
      pedge[p] = e;
      pinfo[p] = make_uint2(e + 1u, a + n - e - 1u);
    }
  }
}

// The 3-byte lists: the list of vertex y starts at 4-byte word pack_word(ooff[y], y).
__device__ __forceinline__ uint64_t pack_word(uint32_t o, uint32_t y) {
  return (3 * ((uint64_t)o + y) + 3) / 4;
}

// LPR lanes per row y: lane k of the group writes words k, k + LPR, ... of y's 3-byte
// list.  Word k holds bytes 4k .. 4k+3 of the list: element e = floor(4k/3) from its byte
// r = 4k - 3e on, then the low bytes of element e + 1 (zero past the list's end).
template <uint32_t LPR>
__global__ void __launch_bounds__(256)
k_pack24(const uint32_t *__restrict__ ooff, const uint32_t *__restrict__ oadj, uint32_t V,
         uint32_t *__restrict__ opk) {
  const uint32_t sub = threadIdx.x & (LPR - 1u);
  for (uint64_t y = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / LPR; y < V;
       y += ((uint64_t)gridDim.x * blockDim.x) / LPR) {
    const uint32_t o = ooff[y], n = ooff[y + 1] - o;
    uint32_t *const out = opk + pack_word(o, (uint32_t)y);
    for (uint32_t k = sub; k < (3 * n + 3) / 4; k += LPR) {
      const uint32_t e = 4 * k / 3, r = 4 * k - 3 * e;
      const uint32_t lo = (oadj[o + e] & 0xFFFFFFu) >> (8 * r);
      const uint32_t hi = e + 1 < n ? oadj[o + e + 1] << (24 - 8 * r) : 0u;
      out[k] = lo | hi;
    }
  }
}

// Per pair: where its stream starts in the 3-byte lists.  Element 4j of a list starts
// word 3j of it, so the suffix start s (within the partner's list) is rounded down to a
// multiple of 4 and the length grows by the s mod 4 elements before it, all ranked at most
// the owner and so never found in its list.
__global__ void __launch_bounds__(256)
k_pack_info(const uint32_t *__restrict__ ppar, const uint2 *__restrict__ pinfo,
            const uint32_t *__restrict__ ooff, uint32_t np, uint2 *__restrict__ pkinfo) {
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < np;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint2 pi = pinfo[p];
    const uint32_t y = ppar[p], a = ooff[y], s = pi.x - a;
    pkinfo[p] = make_uint2((uint32_t)(pack_word(a, y) + 3u * (s >> 2)), pi.y + (s & 3u));
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

namespace {

uint32_t grid_for_n(uint64_t n, uint32_t per_block) {
  uint64_t b = (n + per_block - 1) / per_block;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 1u << 20));
}

// wpre = exclusive scan of the popcounts of flags[0 .. n) (n = words + 1; the last word
// is zero, so wpre[n - 1] is the total).  cnt: n uint32 of temporary storage.
int scan_words(Ctx *ctx, const unsigned long long *flags, uint64_t n, uint32_t *cnt,
               uint32_t *wpre) {
  hipLaunchKernelGGL(k_popc_words, dim3(grid_for_n(n, 256)), dim3(256), 0, ctx->stream, flags, n,
                     cnt);
  PAC_CHECK_HIP(hipGetLastError());
  size_t tmp = 0;
  PAC_CHECK_HIP(rocprim::exclusive_scan(nullptr, tmp, cnt, wpre, 0u, (size_t)n,
                                        rocprim::plus<uint32_t>(), ctx->stream));
  PAC_TRY(ctx->scratch.reserve(tmp));
  PAC_CHECK_HIP(rocprim::exclusive_scan(ctx->scratch.p, tmp, cnt, wpre, 0u, (size_t)n,
                                        rocprim::plus<uint32_t>(), ctx->stream));
  return PAC_OK;
}

}  // namespace

int orient(Graph *g) {
  Ctx *ctx = g->ctx;
  const uint32_t V = g->V;
  const uint64_t nnz = g->nnz, nwords = (nnz + 63) / 64, nchunks = num_chunks(nnz);
  PAC_TRY(g->deg.reserve(sizeof(uint32_t) * (V ? V : 1)));
  PAC_TRY(g->odeg.reserve(sizeof(uint32_t) * ((uint64_t)V + 1)));
  PAC_TRY(g->ooff.reserve(sizeof(uint32_t) * ((uint64_t)V + 1)));
  PAC_TRY(g->crow.reserve(sizeof(uint32_t) * (nchunks ? nchunks : 1)));
  PAC_TRY(g->oflags.reserve(sizeof(unsigned long long) * (nwords + 1)));
  PAC_TRY(g->owpre.reserve(sizeof(uint32_t) * (nwords + 1)));
  PAC_TRY(ctx->scratch2.reserve(64));
  uint32_t *maxdeg = ctx->scratch2.as<uint32_t>();
  PAC_TRY(fill_u32(ctx->stream, maxdeg, 1, 0u));
  PAC_TRY(fill_u32(ctx->stream, g->odeg.as<uint32_t>(), (uint64_t)V + 1, 0u));
  // the word past the last entry stays zero: ordinal(nnz) = the kept total
  PAC_TRY(fill_u32(ctx->stream, g->oflags.as<uint32_t>() + 2 * nwords, 2, 0u));
  if (V) {
    hipLaunchKernelGGL(k_deg, dim3(std::max<uint32_t>(1u, std::min<uint32_t>(4096u, (V + 255) / 256))),
                       dim3(256), 0, ctx->stream,
                       g->off.as<const uint32_t>(), V, g->deg.as<uint32_t>(), maxdeg);
    PAC_CHECK_HIP(hipGetLastError());
  }
  if (nnz) {
    hipLaunchKernelGGL(k_chunk_rows, dim3(grid_for_n(V, 256)), dim3(256), 0, ctx->stream,
                       g->off.as<const uint32_t>(), V, g->crow.as<uint32_t>());
    hipLaunchKernelGGL(k_orient_flags, dim3(grid_for_n(nchunks, 4)), dim3(256), 0, ctx->stream,
                       g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(),
                       g->deg.as<const uint32_t>(), g->crow.as<const uint32_t>(), V,
                       (uint32_t)nnz, g->oflags.as<unsigned long long>());
    PAC_CHECK_HIP(hipGetLastError());
  }
  {
    DevBuf cnt;
    PAC_TRY(cnt.reserve(sizeof(uint32_t) * (nwords + 1)));
    PAC_TRY(scan_words(ctx, g->oflags.as<const unsigned long long>(), nwords + 1,
                       cnt.as<uint32_t>(), g->owpre.as<uint32_t>()));
    hipLaunchKernelGGL(k_ordinal_offsets, dim3(grid_for_n((uint64_t)V + 1, 256)), dim3(256), 0,
                       ctx->stream, g->off.as<const uint32_t>(),
                       g->oflags.as<const unsigned long long>(), g->owpre.as<const uint32_t>(), V,
                       g->ooff.as<uint32_t>(), g->odeg.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
    uint32_t hm[2] = {0, 0};
    PAC_CHECK_HIP(hipMemcpyAsync(&hm[0], g->ooff.as<uint32_t>() + V, 4, hipMemcpyDeviceToHost,
                                 ctx->stream));
    PAC_CHECK_HIP(hipMemcpyAsync(&hm[1], maxdeg, 4, hipMemcpyDeviceToHost, ctx->stream));
    PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // cnt is freed on return
    g->m = hm[0];
    g->max_deg = hm[1];
  }
  if (2 * g->m != g->nnz) {
    set_error("orientation kept %llu of %llu entries: the CSR is not symmetric",
              (unsigned long long)g->m, (unsigned long long)g->nnz);
    return PAC_ERR_FORMAT;
  }
  // m entries plus sentinels (kListEnd) that list streams past their end read: one, and
  // three more for the 16-byte loads of the lane path
  PAC_TRY(g->oadj.reserve(sizeof(uint32_t) * (g->m + 4)));
  PAC_TRY(g->okey.reserve(sizeof(uint32_t) * (g->m + 1)));
  PAC_TRY(g->rank.reserve(sizeof(uint32_t) * (V ? V : 1)));
  if (V) {
    // vertex ranks: the (degree, id) keys sorted; L bits hold a vertex or a rank
    uint32_t L = 1;
    while (L < 32 && (1ull << L) < V) ++L;
    uint32_t dbits = 0;
    while (dbits < 32 && (1ull << dbits) <= g->max_deg) ++dbits;
    DevBuf vk, vks, tmp;
    PAC_TRY(vk.reserve(sizeof(unsigned long long) * V));
    PAC_TRY(vks.reserve(sizeof(unsigned long long) * V));
    hipLaunchKernelGGL(k_vertex_keys, dim3(grid_for_n(V, 256 * 4)), dim3(256), 0, ctx->stream,
                       g->deg.as<const uint32_t>(), V, vk.as<unsigned long long>());
    PAC_CHECK_HIP(hipGetLastError());
    size_t tb = 0;
    PAC_CHECK_HIP(rocprim::radix_sort_keys(nullptr, tb, vk.as<uint64_t>(), vks.as<uint64_t>(),
                                           (size_t)V, 0u, 32u + dbits, ctx->stream));
    PAC_TRY(tmp.reserve(tb));
    PAC_CHECK_HIP(rocprim::radix_sort_keys(tmp.p, tb, vk.as<uint64_t>(), vks.as<uint64_t>(),
                                           (size_t)V, 0u, 32u + dbits, ctx->stream));
    hipLaunchKernelGGL(k_rank_scatter, dim3(grid_for_n(V, 256 * 4)), dim3(256), 0, ctx->stream,
                       vks.as<const unsigned long long>(), V, g->rank.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
    if (g->m) {
      // every kept entry's key (row, rank of the neighbour) at its position, one sort
      DevBuf ek, eks;
      PAC_TRY(ek.reserve(sizeof(unsigned long long) * g->m));
      PAC_TRY(eks.reserve(sizeof(unsigned long long) * g->m));
      hipLaunchKernelGGL(k_orient_keys, dim3(grid_for_n(nchunks, 4)), dim3(256), 0, ctx->stream,
                         g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(),
                         g->rank.as<const uint32_t>(), g->crow.as<const uint32_t>(),
                         g->oflags.as<const unsigned long long>(), g->owpre.as<const uint32_t>(),
                         V, (uint32_t)nnz, L, ek.as<unsigned long long>());
      PAC_CHECK_HIP(hipGetLastError());
      tb = 0;
      PAC_CHECK_HIP(rocprim::radix_sort_keys(nullptr, tb, ek.as<uint64_t>(), eks.as<uint64_t>(),
                                             (size_t)g->m, 0u, 2u * L, ctx->stream));
      if (tb > tmp.bytes) {
        PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // the first sort may still run
        tmp.release();
        PAC_TRY(tmp.reserve(tb));
      }
      PAC_CHECK_HIP(rocprim::radix_sort_keys(tmp.p, tb, ek.as<uint64_t>(), eks.as<uint64_t>(),
                                             (size_t)g->m, 0u, 2u * L, ctx->stream));
      hipLaunchKernelGGL(k_orient_unpack, dim3(grid_for_n(g->m, 256 * 4)), dim3(256), 0,
                         ctx->stream, eks.as<const unsigned long long>(), g->m, L,
                         vks.as<const unsigned long long>(), g->okey.as<uint32_t>(),
                         g->oadj.as<uint32_t>());
      PAC_CHECK_HIP(hipGetLastError());
    }
    PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // the temporaries are freed on return
  }
  PAC_TRY(fill_u32(ctx->stream, g->oadj.as<uint32_t>() + g->m, 4, kListEnd));
  PAC_TRY(fill_u32(ctx->stream, g->okey.as<uint32_t>() + g->m, 1, kListEnd));
  g->pairs_built = false;
  g->hitw_built = false;
  return PAC_OK;
}

// Pair lists and row work = ownership bits / scan / fill, like the orientation.
int build_pairs(Graph *g) {
  if (g->pairs_built) return PAC_OK;
  Ctx *ctx = g->ctx;
  const uint32_t V = g->V;
  const uint64_t nnz = g->nnz, nwords = (nnz + 63) / 64, nchunks = num_chunks(nnz);
  PAC_TRY(g->poff.reserve(sizeof(uint32_t) * ((uint64_t)V + 1)));
  PAC_TRY(g->pflags.reserve(sizeof(unsigned long long) * (nwords + 1)));
  PAC_TRY(g->ppre.reserve(sizeof(uint32_t) * (nwords + 1)));
  PAC_TRY(g->rowwork.reserve(sizeof(uint32_t) * (V ? V : 1)));
  PAC_TRY(fill_u32(ctx->stream, g->pflags.as<uint32_t>() + 2 * nwords, 2, 0u));
  PAC_TRY(fill_u32(ctx->stream, g->rowwork.as<uint32_t>(), V, 0u));
  DevBuf epos;  // per symmetric entry owned as a pair: the owner's position in the partner's list
  PAC_TRY(epos.reserve(sizeof(uint32_t) * (nnz ? nnz : 1)));
  if (nnz) {  // then m = nnz / 2 > 0
    // tin: the positions in oadj sorted by head (a stable sort of e by oadj[e])
    DevBuf tin, keys, vals, tmp;
    PAC_TRY(tin.reserve(sizeof(uint32_t) * g->m));
    PAC_TRY(keys.reserve(sizeof(uint32_t) * g->m));
    PAC_TRY(vals.reserve(sizeof(uint32_t) * g->m));
    hipLaunchKernelGGL(k_iota, dim3(grid_for_n(g->m, 256 * 4)), dim3(256), 0, ctx->stream,
                       vals.as<uint32_t>(), (uint64_t)g->m);
    PAC_CHECK_HIP(hipGetLastError());
    uint32_t L = 1;
    while (L < 32 && (1ull << L) < V) ++L;
    size_t tb = 0;
    PAC_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, tb, g->oadj.as<const uint32_t>(),
                                            keys.as<uint32_t>(), vals.as<const uint32_t>(),
                                            tin.as<uint32_t>(), (size_t)g->m, 0u, L,
                                            ctx->stream));
    PAC_TRY(tmp.reserve(tb));
    PAC_CHECK_HIP(rocprim::radix_sort_pairs(tmp.p, tb, g->oadj.as<const uint32_t>(),
                                            keys.as<uint32_t>(), vals.as<const uint32_t>(),
                                            tin.as<uint32_t>(), (size_t)g->m, 0u, L,
                                            ctx->stream));
    hipLaunchKernelGGL(k_pair_flags, dim3(grid_for_n(nchunks, 4)), dim3(256), 0, ctx->stream,
                       g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(),
                       g->ooff.as<const uint32_t>(), tin.as<const uint32_t>(),
                       g->oflags.as<const unsigned long long>(), g->owpre.as<const uint32_t>(),
                       g->crow.as<const uint32_t>(), V, (uint32_t)nnz,
                       g->pflags.as<unsigned long long>(), epos.as<uint32_t>(),
                       g->rowwork.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
    PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // the sort's buffers are freed here
  }
  uint32_t np = 0;
  {
    DevBuf cnt;
    PAC_TRY(cnt.reserve(sizeof(uint32_t) * (nwords + 1)));
    PAC_TRY(scan_words(ctx, g->pflags.as<const unsigned long long>(), nwords + 1,
                       cnt.as<uint32_t>(), g->ppre.as<uint32_t>()));
    hipLaunchKernelGGL(k_ordinal_offsets, dim3(grid_for_n((uint64_t)V + 1, 256)), dim3(256), 0,
                       ctx->stream, g->off.as<const uint32_t>(),
                       g->pflags.as<const unsigned long long>(), g->ppre.as<const uint32_t>(), V,
                       g->poff.as<uint32_t>(), nullptr);
    PAC_CHECK_HIP(hipGetLastError());
    PAC_CHECK_HIP(hipMemcpyAsync(&np, g->poff.as<uint32_t>() + V, 4, hipMemcpyDeviceToHost,
                                 ctx->stream));
    PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // cnt is freed on return
  }
  if (np > g->m) {
    set_error("pair construction produced %u pairs for %llu oriented edges", np,
              (unsigned long long)g->m);
    return PAC_ERR_HIP;
  }
  PAC_TRY(g->ppar.reserve(sizeof(uint32_t) * (np ? np : 1)));
  PAC_TRY(g->pedge.reserve(sizeof(uint32_t) * (np ? np : 1)));
  PAC_TRY(g->pinfo.reserve(sizeof(uint2) * (np ? np : 1)));
  if (nnz) {
    hipLaunchKernelGGL(k_pair_fill, dim3(grid_for_n(nnz, 256 * 4)), dim3(256), 0, ctx->stream,
                       g->adj.as<const uint32_t>(), g->pflags.as<const unsigned long long>(),
                       g->ppre.as<const uint32_t>(), g->ooff.as<const uint32_t>(),
                       epos.as<const uint32_t>(), (uint32_t)nnz, g->ppar.as<uint32_t>(),
                       g->pedge.as<uint32_t>(), g->pinfo.as<uint2>());
    PAC_CHECK_HIP(hipGetLastError());
  }
  if (V <= kPackMaxV) {
    // the lists' words, + 4 words past the end (the last list's word rounding)
    const uint64_t nw = (3 * (g->m + V) + 3) / 4 + 4;
    PAC_TRY(g->opk.reserve(4 * nw));
    PAC_TRY(g->pkinfo.reserve(sizeof(uint2) * (np ? np : 1)));
    PAC_TRY(fill_u32(ctx->stream, g->opk.as<uint32_t>(), nw, 0u));
    if (V)
      // 16 lanes per row: 3.4 ms at R-MAT scale 24 (64: 3.6 ms, 8: 4.0 ms)
      hipLaunchKernelGGL(k_pack24<16>, dim3(grid_for_n(V, 16)), dim3(256), 0, ctx->stream,
                         g->ooff.as<const uint32_t>(), g->oadj.as<const uint32_t>(), V,
                         g->opk.as<uint32_t>());
    if (np)
      hipLaunchKernelGGL(k_pack_info, dim3(grid_for_n(np, 256 * 4)), dim3(256), 0,
                         ctx->stream, g->ppar.as<const uint32_t>(), g->pinfo.as<const uint2>(),
                         g->ooff.as<const uint32_t>(), np, g->pkinfo.as<uint2>());
    PAC_CHECK_HIP(hipGetLastError());
  }
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // epos is freed on return
  g->pairs_built = true;
  g->hitw_built = false;
  return PAC_OK;
}

namespace {

// Per vertex: whether it owns pairs (flag), and over all vertices W = sum of d+(u) d(u)
// (= sum over oriented t->h of d+(t) + d+(h)) and the largest d+ (one atomic per block).
__global__ void __launch_bounds__(256)
k_vertex_stats(const uint32_t *__restrict__ off, const uint32_t *__restrict__ odeg,
               const uint32_t *__restrict__ poff, uint32_t V, uint8_t *__restrict__ owns_pairs,
               unsigned long long *W, uint32_t *maxo) {
  unsigned long long w = 0;
  uint32_t mx = 0;
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < V;
       u += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t o = odeg[u];
    w += (unsigned long long)o * (off[u + 1] - off[u]);
    mx = max(mx, o);
    owns_pairs[u] = poff[u + 1] > poff[u] ? 1 : 0;
  }
  for (int sh = 32; sh > 0; sh >>= 1) {
    w += __shfl_xor(w, sh, 64);
    mx = max(mx, (uint32_t)__shfl_xor((int)mx, sh, 64));
  }
  __shared__ unsigned long long ws[4];
  __shared__ uint32_t ms[4];
  if ((threadIdx.x & 63u) == 0) {
    ws[threadIdx.x >> 6] = w;
    ms[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t i = 1; i < (blockDim.x >> 6); ++i) {
      w += ws[i];
      mx = max(mx, ms[i]);
    }
    if (w) atomicAdd(W, w);
    if (mx) atomicMax(maxo, mx);
  }
}

// The summary of each row that owns pairs: d+(r), its streamed length, its pairs, d(r).
__global__ void k_row_gather(const uint32_t *__restrict__ rows, uint32_t n,
                             const uint32_t *__restrict__ off, const uint32_t *__restrict__ odeg,
                             const uint32_t *__restrict__ poff,
                             const uint32_t *__restrict__ rowwork, uint4 *__restrict__ out) {
  for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < n;
       q += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = rows[q];
    out[q] = make_uint4(odeg[r], rowwork[r], poff[r + 1] - poff[r], off[r + 1] - off[r]);
  }
}

// out[k] = ooff[min(k * width, V)], k = 0..P: the oriented-edge bounds of contiguous
// partitions.
__global__ void k_part_bounds(const uint32_t *__restrict__ ooff, uint32_t V, uint64_t width,
                              uint32_t P, uint32_t *__restrict__ out) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k <= P;
       k += (uint64_t)gridDim.x * blockDim.x)
    out[k] = ooff[std::min<uint64_t>(k * width, V)];
}

}  // namespace

int row_summary(Graph *g, uint64_t part_width, uint32_t P, RowSummary &out) {
  Ctx *ctx = g->ctx;
  const uint32_t V = g->V;
  PAC_TRY(build_pairs(g));
  DevBuf flags, acc, nsel;
  PAC_TRY(flags.reserve(V ? V : 1));
  PAC_TRY(acc.reserve(16));
  PAC_TRY(nsel.reserve(8));
  PAC_TRY(fill_u32(ctx->stream, acc.as<uint32_t>(), 4, 0u));
  if (V) {
    hipLaunchKernelGGL(k_vertex_stats, dim3(grid_for_n(V, 256 * 8)), dim3(256), 0, ctx->stream,
                       g->off.as<const uint32_t>(), g->odeg.as<const uint32_t>(),
                       g->poff.as<const uint32_t>(), V, flags.as<uint8_t>(),
                       acc.as<unsigned long long>(), acc.as<uint32_t>() + 2);
    PAC_CHECK_HIP(hipGetLastError());
  }
  // the rows, in ID order
  PAC_TRY(g->rows.reserve(sizeof(uint32_t) * (V ? V : 1)));
  PAC_TRY(fill_u32(ctx->stream, nsel.as<uint32_t>(), 1, 0u));
  if (V) {
    rocprim::counting_iterator<uint32_t> ids(0u);
    size_t tmp = 0;
    PAC_CHECK_HIP(rocprim::select(nullptr, tmp, ids, flags.as<const uint8_t>(),
                                  g->rows.as<uint32_t>(), nsel.as<uint32_t>(), (size_t)V,
                                  ctx->stream));
    PAC_TRY(ctx->scratch.reserve(tmp));
    PAC_CHECK_HIP(rocprim::select(ctx->scratch.p, tmp, ids, flags.as<const uint8_t>(),
                                  g->rows.as<uint32_t>(), nsel.as<uint32_t>(), (size_t)V,
                                  ctx->stream));
  }
  uint32_t h_acc[4] = {0, 0, 0, 0}, n = 0;
  PAC_CHECK_HIP(hipMemcpyAsync(h_acc, acc.p, 16, hipMemcpyDeviceToHost, ctx->stream));
  PAC_CHECK_HIP(hipMemcpyAsync(&n, nsel.p, 4, hipMemcpyDeviceToHost, ctx->stream));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  out.W = (uint64_t)h_acc[0] | ((uint64_t)h_acc[1] << 32);
  out.max_outdeg = h_acc[2];
  out.rows.resize(n);
  out.stats.resize(n);
  if (n) {
    DevBuf st;
    PAC_TRY(st.reserve(sizeof(uint4) * n));
    hipLaunchKernelGGL(k_row_gather, dim3(grid_for_n(n, 256)), dim3(256), 0, ctx->stream,
                       g->rows.as<const uint32_t>(), n, g->off.as<const uint32_t>(),
                       g->odeg.as<const uint32_t>(), g->poff.as<const uint32_t>(),
                       g->rowwork.as<const uint32_t>(), st.as<uint4>());
    PAC_CHECK_HIP(hipGetLastError());
    PAC_CHECK_HIP(hipMemcpyAsync(out.rows.data(), g->rows.p, 4ull * n, hipMemcpyDeviceToHost,
                                 ctx->stream));
    PAC_CHECK_HIP(hipMemcpyAsync(out.stats.data(), st.p, sizeof(uint4) * n,
                                 hipMemcpyDeviceToHost, ctx->stream));
    PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // st is freed on return
  }
  if (part_width) {  // contiguous partitions: their oriented-edge counts from ooff
    DevBuf b;
    PAC_TRY(b.reserve(4ull * (P + 1)));
    hipLaunchKernelGGL(k_part_bounds, dim3(grid_for_n((uint64_t)P + 1, 256)), dim3(256), 0,
                       ctx->stream, g->ooff.as<const uint32_t>(), V, part_width, P,
                       b.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
    std::vector<uint32_t> h(P + 1);
    PAC_CHECK_HIP(hipMemcpyAsync(h.data(), b.p, 4ull * (P + 1), hipMemcpyDeviceToHost, ctx->stream));
    PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    out.part_edges.resize(P);
    for (uint32_t k = 0; k < P; ++k) out.part_edges[k] = h[k + 1] - h[k];
  }
  return PAC_OK;
}

namespace {

// Per oriented edge y->h: 1 when it is a pair's edge (the pair of owner h and partner y:
// d+(h) > 0 and h is not the last element of N+(y)), else 0 -- here the first condition,
// for every edge (a gather of odeg in edge order) ...
__global__ void k_pair_edge_flags(const uint32_t *__restrict__ oadj,
                                  const uint32_t *__restrict__ odeg, uint64_t m,
                                  uint32_t *__restrict__ flag) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x)
    flag[e] = odeg[oadj[e]] > 0u ? 1u : 0u;
}
// ... then the second: the last edge of every non-empty list is cleared.
__global__ void k_clear_last_edges(const uint32_t *__restrict__ ooff, uint32_t V,
                                   uint32_t *__restrict__ flag) {
  for (uint64_t y = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; y < V;
       y += (uint64_t)gridDim.x * blockDim.x)
    if (ooff[y + 1] > ooff[y]) flag[ooff[y + 1] - 1u] = 0u;
}

// Per vertex y, what its pairs' hit words need: (ooff[y], d+(y)) and the word base shifted
// by the rank of its first pair edge, wbase[y] - prank[ooff[y]] * nw (mod 2^64; nw =
// ceil(d+(y) / 64)), so that a pair's row is prank[e] alone.
__global__ void k_vertex_words(const uint32_t *__restrict__ ooff,
                               const uint32_t *__restrict__ prank,
                               const unsigned long long *__restrict__ wbase, uint32_t V,
                               uint4 *__restrict__ vw) {
  for (uint64_t y = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; y < V;
       y += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t a = ooff[y], d = ooff[y + 1] - a;
    const unsigned long long b =
        wbase[y] - (unsigned long long)prank[a] * (unsigned long long)((d + 63u) >> 6);
    vw[y] = make_uint4(a, d, (uint32_t)b, (uint32_t)(b >> 32));
  }
}

// Pairs streaming each vertex's out-list: R[y] = prank[ooff[y+1]] - prank[ooff[y]].
__global__ void k_partner_streams(const uint32_t *__restrict__ ooff,
                                  const uint32_t *__restrict__ prank, uint32_t V,
                                  uint32_t *__restrict__ R) {
  for (uint64_t y = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; y < V;
       y += (uint64_t)gridDim.x * blockDim.x)
    R[y] = prank[ooff[y + 1]] - prank[ooff[y]];
}

// Hit words per vertex: pairs streaming it x words per stream; entry V is 0, so the
// exclusive scan ends with the total.
__global__ void k_word_counts(const uint32_t *__restrict__ R, const uint32_t *__restrict__ odeg,
                              uint32_t V, unsigned long long *__restrict__ w) {
  for (uint64_t y = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; y <= V;
       y += (uint64_t)gridDim.x * blockDim.x)
    w[y] = y < V ? (unsigned long long)R[y] * ((odeg[y] + 63u) >> 6) : 0ull;
}

// SUP-mode stream of each pair and its first hit word.  The stream starts at the
// 64-element boundary of N+(y) at or before the pair's suffix (the elements before the
// suffix rank at most the owner and are never found), so each 64-element block of the
// stream is a word column c of the list.  The pair's words: its partner's word base,
// plus its rank among the pairs streaming that partner times the words per stream, plus
// the column of its first block -- the columns before it are never written and stay zero.
// The rank is the pair's edge's position among its partner's pair edges (prank), so the
// rows of a list's words come in the order of their first columns: the column sums read,
// for column c, only the rows whose stream reaches back to it.
__global__ void k_pair_words(const uint32_t *__restrict__ ppar, const uint32_t *__restrict__ pedge,
                             const uint2 *__restrict__ pinfo, uint64_t np,
                             const uint4 *__restrict__ vw, const uint32_t *__restrict__ prank,
                             unsigned long long *__restrict__ pword, uint2 *__restrict__ pinfo64) {
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint4 w = vw[ppar[p]];  // one 16-byte read per pair for its partner
    const uint32_t a = w.x, d = w.y;
    const uint2 pi = pinfo[p];
    const uint32_t c0 = (pi.x - a) >> 6;  // the column of the stream's first block
    pword[p] = (((unsigned long long)w.w << 32) | w.z) +
               (unsigned long long)prank[pedge[p]] * ((d + 63u) >> 6) + c0;
    pinfo64[p] = make_uint2(a + 64u * c0, d - 64u * c0);
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
  PAC_TRY(g->pinfo64.reserve(sizeof(uint2) * (np ? np : 1)));
  PAC_TRY(g->wbase.reserve(sizeof(unsigned long long) * ((uint64_t)V + 1)));
  PAC_TRY(g->prank.reserve(sizeof(uint32_t) * ((uint64_t)g->m + 1)));
  DevBuf words;
  PAC_TRY(words.reserve(sizeof(unsigned long long) * ((uint64_t)V + 1)));
  PAC_TRY(fill_u32(ctx->stream, g->pstreams.as<uint32_t>(), (uint64_t)V + 1, 0u));
  {
    // prank = the exclusive scan of the pair-edge flags over oadj
    DevBuf flag;
    PAC_TRY(flag.reserve(sizeof(uint32_t) * ((uint64_t)g->m + 1)));
    PAC_TRY(fill_u32(ctx->stream, flag.as<uint32_t>() + g->m, 1, 0u));
    if (g->m)
      hipLaunchKernelGGL(k_pair_edge_flags, dim3(grid_elems(g->m)), dim3(256), 0, ctx->stream,
                         g->oadj.as<const uint32_t>(), g->odeg.as<const uint32_t>(),
                         (uint64_t)g->m, flag.as<uint32_t>());
    if (V)
      hipLaunchKernelGGL(k_clear_last_edges, dim3(grid_elems(V)), dim3(256), 0, ctx->stream,
                         g->ooff.as<const uint32_t>(), V, flag.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
    size_t tb = 0;
    PAC_CHECK_HIP(rocprim::exclusive_scan(nullptr, tb, flag.as<const uint32_t>(),
                                          g->prank.as<uint32_t>(), 0u, (size_t)g->m + 1,
                                          rocprim::plus<uint32_t>(), ctx->stream));
    PAC_TRY(ctx->scratch.reserve(tb));
    PAC_CHECK_HIP(rocprim::exclusive_scan(ctx->scratch.p, tb, flag.as<const uint32_t>(),
                                          g->prank.as<uint32_t>(), 0u, (size_t)g->m + 1,
                                          rocprim::plus<uint32_t>(), ctx->stream));
    if (V)
      hipLaunchKernelGGL(k_partner_streams, dim3(grid_elems(V)), dim3(256), 0, ctx->stream,
                         g->ooff.as<const uint32_t>(), g->prank.as<const uint32_t>(), V,
                         g->pstreams.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
    uint32_t nflag = 0;
    PAC_CHECK_HIP(hipMemcpyAsync(&nflag, g->prank.as<uint32_t>() + g->m, 4,
                                 hipMemcpyDeviceToHost, ctx->stream));
    PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // flag is freed here
    if (nflag != np) {
      set_error("hit-word layout: %u pair edges for %u pairs", nflag, np);
      return PAC_ERR_HIP;
    }
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
    DevBuf vw;
    PAC_TRY(vw.reserve(sizeof(uint4) * V));
    hipLaunchKernelGGL(k_vertex_words, dim3(grid_elems(V)), dim3(256), 0, ctx->stream,
                       g->ooff.as<const uint32_t>(), g->prank.as<const uint32_t>(),
                       g->wbase.as<const unsigned long long>(), V, vw.as<uint4>());
    hipLaunchKernelGGL(k_pair_words, dim3(grid_elems(np)), dim3(256), 0, ctx->stream,
                       g->ppar.as<const uint32_t>(), g->pedge.as<const uint32_t>(),
                       g->pinfo.as<const uint2>(), (uint64_t)np, vw.as<const uint4>(),
                       g->prank.as<const uint32_t>(), g->pword.as<unsigned long long>(),
                       g->pinfo64.as<uint2>());
    PAC_CHECK_HIP(hipGetLastError());
    PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // vw is freed here
  }
  unsigned long long nwords = 0;
  PAC_CHECK_HIP(hipMemcpyAsync(&nwords, g->wbase.as<unsigned long long>() + V, 8,
                               hipMemcpyDeviceToHost, ctx->stream));
  std::vector<uint32_t> R(V);
  if (V)
    PAC_CHECK_HIP(hipMemcpyAsync(R.data(), g->pstreams.p, 4ull * V, hipMemcpyDeviceToHost,
                                 ctx->stream));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // words is freed on return
  // Every pair writes the words of its stream's columns in each SUP launch; the zero fill
  // gives the columns before its stream their value (no triangle) for good.
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
