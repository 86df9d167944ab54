// pac_gen.hip -- device construction of the symmetric CSR: from a seeded R-MAT
// generator, or from an uploaded undirected edge list.
//
// Edge lists: both directions of every edge become 64-bit keys (u << shift | v,
// self-loops -> sentinel), a device radix sort + unique + compaction yields the sorted,
// duplicate-free rows -- the row building of Graph::Graph(graphfile), Graph.cpp:70-85,
// done on the GPU.
//
// R-MAT has no reference counterpart: the reference only reads graphs from files
// (Graph.cpp:27-137).  The generator is the synthetic-input side of the benchmarks:
// ef*2^scale samples, each drawing `scale` quadrants with probabilities (a,b,c,d)
// from a SplitMix64 stream (Steele, Lea, Flood 2014) seeded by a counter-based hash
// of (seed, sample).  Only integer arithmetic is used, so the host restatement in
// the oracle produces the identical edge list.  Self-loops are dropped, both
// directions are emitted, and a device radix sort + unique + compaction gives the
// symmetric CSR with sorted, duplicate-free rows.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "pac_internal.hpp"

namespace pac {
namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_rmat(uint64_t ns, uint32_t scale, uint64_t t1, uint64_t t2, uint64_t t3,
                       uint64_t key, uint64_t *__restrict__ keys) {
  const uint64_t sentinel = 1ull << (2 * scale);
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < ns;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t st = mix64(key + i * kGamma);
    uint32_t u = 0, v = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      st += kGamma;
      const uint64_t r = mix64(st) >> 11;
      const uint32_t q = r < t1 ? 0u : (r < t2 ? 1u : (r < t3 ? 2u : 3u));
      u = (u << 1) | (q >> 1);
      v = (v << 1) | (q & 1u);
    }
    if (u == v) {
      keys[2 * i] = sentinel;
      keys[2 * i + 1] = sentinel;
    } else {
      keys[2 * i] = ((uint64_t)u << scale) | v;
      keys[2 * i + 1] = ((uint64_t)v << scale) | u;
    }
  }
}

// Directed keys (u << shift | v) of both directions of each input edge; self-loops get
// the sentinel, IDs >= V set the error flag (and the sentinel).
__global__ void k_edge_keys(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst,
                            uint64_t ne, uint32_t shift, uint32_t V,
                            uint64_t *__restrict__ keys, uint32_t *err) {
  const uint64_t sentinel = 1ull << (2 * shift);
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < ne;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = src[i], v = dst[i];
    if (u >= V || v >= V) {
      atomicOr(err, 1u);
      keys[2 * i] = sentinel;
      keys[2 * i + 1] = sentinel;
    } else if (u == v) {
      keys[2 * i] = sentinel;
      keys[2 * i + 1] = sentinel;
    } else {
      keys[2 * i] = ((uint64_t)u << shift) | v;
      keys[2 * i + 1] = ((uint64_t)v << shift) | u;
    }
  }
}

__global__ void k_unique_flags(const uint64_t *__restrict__ keys, uint64_t n, uint64_t sentinel,
                               uint32_t *__restrict__ flags) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    flags[i] = (k != sentinel && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
  }
}

__global__ void k_compact(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ flags,
                          const uint32_t *__restrict__ pos, uint64_t n, uint32_t mask,
                          uint32_t *__restrict__ adj, uint64_t *__restrict__ ckeys) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (flags[i]) {
      const uint64_t k = keys[i];
      adj[pos[i]] = (uint32_t)(k & mask);
      ckeys[pos[i]] = k;
    }
  }
}

// off[u] = first position whose key has source >= u (binary search over sorted keys)
__global__ void k_offsets(const uint64_t *__restrict__ ckeys, uint64_t nnz, uint32_t scale,
                          uint32_t V, uint32_t *__restrict__ off) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u <= V;
       u += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t target = u << scale;
    uint64_t lo = 0, n = nnz;
    while (n > 0) {
      const uint64_t half = n >> 1;
      if (ckeys[lo + half] < target) {
        lo += half + 1;
        n -= half + 1;
      } else {
        n = half;
      }
    }
    off[u] = (uint32_t)lo;
  }
}

uint32_t grid1d(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > (1u << 20)) b = 1u << 20;
  return (uint32_t)b;
}

}  // namespace

// Sort the n2 directed keys in keys_a (key = u << shift | v, sentinel 1 << 2*shift),
// drop duplicates and sentinels, and build the symmetric CSR of V vertices in g.
int csr_from_keys_device(Graph *g, DevBuf &keys_a, uint64_t n2, uint32_t shift, uint32_t V) {
  Ctx *ctx = g->ctx;
  const uint64_t sentinel = 1ull << (2 * shift);
  DevBuf keys_b, flags, pos, tmp;
  PAC_TRY(keys_b.reserve(sizeof(uint64_t) * (n2 ? n2 : 1)));
  uint64_t nnz = 0;
  if (n2) {
    size_t tb = 0;
    PAC_CHECK_HIP(rocprim::radix_sort_keys(nullptr, tb, keys_a.as<uint64_t>(),
                                           keys_b.as<uint64_t>(), n2, 0u, 2 * shift + 1,
                                           ctx->stream));
    PAC_TRY(tmp.reserve(tb));
    PAC_CHECK_HIP(rocprim::radix_sort_keys(tmp.p, tb, keys_a.as<uint64_t>(),
                                           keys_b.as<uint64_t>(), n2, 0u, 2 * shift + 1,
                                           ctx->stream));
    PAC_TRY(flags.reserve(sizeof(uint32_t) * n2));
    PAC_TRY(pos.reserve(sizeof(uint32_t) * n2));
    hipLaunchKernelGGL(k_unique_flags, dim3(grid1d(n2)), dim3(256), 0, ctx->stream,
                       keys_b.as<const uint64_t>(), n2, sentinel, flags.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
    size_t sb = 0;
    PAC_CHECK_HIP(rocprim::exclusive_scan(nullptr, sb, flags.as<uint32_t>(), pos.as<uint32_t>(),
                                          0u, (size_t)n2, rocprim::plus<uint32_t>(), ctx->stream));
    if (sb > tmp.bytes) {
      tmp.release();
      PAC_TRY(tmp.reserve(sb));
    }
    PAC_CHECK_HIP(rocprim::exclusive_scan(tmp.p, sb, flags.as<uint32_t>(), pos.as<uint32_t>(),
                                          0u, (size_t)n2, rocprim::plus<uint32_t>(), ctx->stream));
    uint32_t last[2] = {0, 0};
    PAC_CHECK_HIP(hipMemcpyAsync(&last[0], pos.as<uint32_t>() + (n2 - 1), 4,
                                 hipMemcpyDeviceToHost, ctx->stream));
    PAC_CHECK_HIP(hipMemcpyAsync(&last[1], flags.as<uint32_t>() + (n2 - 1), 4,
                                 hipMemcpyDeviceToHost, ctx->stream));
    PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    nnz = (uint64_t)last[0] + last[1];
  }
  g->V = V;
  g->nnz = nnz;
  PAC_TRY(g->adj.reserve(sizeof(uint32_t) * (nnz ? nnz : 1)));
  PAC_TRY(g->off.reserve(sizeof(uint32_t) * ((uint64_t)V + 1)));
  if (n2) {
    // compacted keys go to keys_a (free after the sort)
    hipLaunchKernelGGL(k_compact, dim3(grid1d(n2)), dim3(256), 0, ctx->stream,
                       keys_b.as<const uint64_t>(), flags.as<const uint32_t>(),
                       pos.as<const uint32_t>(), n2, (uint32_t)((1ull << shift) - 1),
                       g->adj.as<uint32_t>(), keys_a.as<uint64_t>());
  }
  hipLaunchKernelGGL(k_offsets, dim3(grid1d((uint64_t)V + 1)), dim3(256), 0, ctx->stream,
                     keys_a.as<const uint64_t>(), nnz, shift, V, g->off.as<uint32_t>());
  PAC_CHECK_HIP(hipGetLastError());
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return PAC_OK;
}

int generate_rmat(Graph *g, uint32_t scale, uint32_t ef, double a, double b, double c,
                  uint64_t seed) {
  Ctx *ctx = g->ctx;
  if (scale < 1 || scale > 26 || ef < 1 || ef > 1024) {
    set_error("rmat: scale must be in [1,26] and edge factor in [1,1024]");
    return PAC_ERR_ARG;
  }
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) {
    set_error("rmat: probabilities must be >= 0 and a+b+c <= 1");
    return PAC_ERR_ARG;
  }
  const uint64_t ns = (uint64_t)ef << scale;
  const uint64_t n2 = 2 * ns;
  if (n2 >= (1ull << 32)) {
    set_error("rmat: %llu samples exceed the uint32 CSR limit", (unsigned long long)ns);
    return PAC_ERR_UNSUPPORTED;
  }
  const double two53 = 9007199254740992.0;
  const uint64_t t1 = (uint64_t)(a * two53), t2 = (uint64_t)((a + b) * two53),
                 t3 = (uint64_t)((a + b + c) * two53);
  uint64_t z = seed + kGamma;  // host copy of mix64(seed + gamma)
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  const uint64_t key = z ^ (z >> 31);
  DevBuf keys_a;
  PAC_TRY(keys_a.reserve(sizeof(uint64_t) * n2));
  hipLaunchKernelGGL(k_rmat, dim3(grid1d(ns)), dim3(256), 0, ctx->stream, ns, scale, t1, t2, t3,
                     key, keys_a.as<uint64_t>());
  PAC_CHECK_HIP(hipGetLastError());
  return csr_from_keys_device(g, keys_a, n2, scale, 1u << scale);
}

namespace {

// ---- text edge lists (Graph::Graph(graphfile), Graph.cpp:27-90) on the device --------
// The reference reads the file with fgets into a 512-byte buffer: a line of up to 511
// characters is one read.  Its rules per read: '#' or '%' first -> skipped; a number
// (digits from the start), exactly one skipped character, a second number; a 0 in either
// marks the file as 0-based; self-loops are dropped after that check; the largest ID of
// the kept edges sets V.  A read ends at its terminating NUL, so the parse below stops at
// the end of the line -- except when the first number runs into that end (no separator):
// the skip then steps over the NUL into bytes left by earlier reads, and a line longer
// than 511 characters is split into several reads.  Such lines are flagged and the caller
// parses the file on the host instead, with the reference's own buffer discipline.

// Position i starts a line when i == 0 or byte i-1 is a newline.
struct LineStart {
  const char *t;
  __device__ bool operator()(uint64_t i) const { return i == 0 || t[i - 1] == '\n'; }
};

__global__ void k_count_lines(const char *__restrict__ t, uint64_t n,
                              unsigned long long *__restrict__ count) {
  const LineStart is_start{t};
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x; i0 < n;
       i0 += (uint64_t)gridDim.x * blockDim.x) {  // block-uniform trip count
    const uint64_t i = i0 + threadIdx.x;
    const int c = __syncthreads_count(i < n && is_start(i));
    if (threadIdx.x == 0 && c) atomicAdd(count, (unsigned long long)c);
  }
}

// flags[0]: a 0 was read; [1]: a line needs the host parser; [2]: an ID >= 2^31.
// Edge k (line k) is kept as u << 32 | v with keep[k] = 1.
__global__ void k_parse_lines(const char *__restrict__ t, uint64_t n,
                              const uint64_t *__restrict__ ls, uint64_t nl,
                              unsigned long long *__restrict__ edge, uint8_t *__restrict__ keep,
                              uint32_t *__restrict__ flags, uint32_t *__restrict__ maxid) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < nl;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = ls[k], e = k + 1 < nl ? ls[k + 1] : n;  // [s, e), e - s >= 1
    keep[k] = 0;
    if (e - s > 511u) {
      atomicOr(&flags[1], 1u);
      continue;
    }
    if (t[s] == '#' || t[s] == '%') continue;
    // (the accumulators saturate above 2^31, which is rejected below, so that no digit
    // string wraps around)
    uint64_t p = s, u = 0, v = 0;
    while (p < e && t[p] >= '0' && t[p] <= '9') {
      u = u * 10 + (uint64_t)(t[p++] - '0');
      u = u > 0xFFFFFFFFull ? 0xFFFFFFFFull : u;
    }
    if (p == e) {  // no separator before the end of the read
      atomicOr(&flags[1], 1u);
      continue;
    }
    ++p;  // exactly one separator character
    while (p < e && t[p] >= '0' && t[p] <= '9') {
      v = v * 10 + (uint64_t)(t[p++] - '0');
      v = v > 0xFFFFFFFFull ? 0xFFFFFFFFull : v;
    }
    if (u > 0x7FFFFFFFull || v > 0x7FFFFFFFull) {
      atomicOr(&flags[2], 1u);
      continue;
    }
    if (u == 0 || v == 0) atomicOr(&flags[0], 1u);
    if (u == v) continue;
    atomicMax(maxid, (uint32_t)(u > v ? u : v));
    edge[k] = (u << 32) | v;
    keep[k] = 1;
  }
}

__global__ void k_split_edges(const unsigned long long *__restrict__ edge, uint64_t ne,
                              uint32_t shift, uint32_t *__restrict__ src,
                              uint32_t *__restrict__ dst) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < ne;
       i += (uint64_t)gridDim.x * blockDim.x) {
    src[i] = (uint32_t)(edge[i] >> 32) - shift;
    dst[i] = (uint32_t)edge[i] - shift;
  }
}

}  // namespace

int parse_text_device(Ctx *ctx, const char *text, uint64_t n, DevBuf &d_src, DevBuf &d_dst,
                      uint64_t *ne, uint64_t *V, bool *host_needed) {
  *ne = 0;
  *V = 0;
  *host_needed = false;
  if (n == 0) return PAC_OK;
  DevBuf dt, ls, cnt, edge, keep, edges, flags;
  PAC_TRY(dt.reserve(n));
  PAC_TRY(cnt.reserve(16));
  PAC_CHECK_HIP(hipMemcpy(dt.p, text, n, hipMemcpyHostToDevice));
  // line starts: count them, then select the positions that start a line
  uint64_t nl = 0;
  PAC_TRY(fill_u32(ctx->stream, cnt.as<uint32_t>(), 4, 0u));
  hipLaunchKernelGGL(k_count_lines, dim3(grid1d(n)), dim3(256), 0, ctx->stream,
                     dt.as<const char>(), n, reinterpret_cast<unsigned long long *>(cnt.p));
  PAC_CHECK_HIP(hipGetLastError());
  PAC_CHECK_HIP(hipMemcpyAsync(&nl, cnt.p, 8, hipMemcpyDeviceToHost, ctx->stream));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  PAC_TRY(ls.reserve(sizeof(uint64_t) * (nl ? nl : 1)));
  {
    size_t tmp = 0;
    uint64_t got = 0;
    rocprim::counting_iterator<uint64_t> pos(0);
    const LineStart pred{dt.as<const char>()};
    PAC_CHECK_HIP(rocprim::select(nullptr, tmp, pos, ls.as<uint64_t>(), cnt.as<uint64_t>(),
                                  (size_t)n, pred, ctx->stream));
    PAC_TRY(ctx->scratch.reserve(tmp));
    PAC_CHECK_HIP(rocprim::select(ctx->scratch.p, tmp, pos, ls.as<uint64_t>(), cnt.as<uint64_t>(),
                                  (size_t)n, pred, ctx->stream));
    PAC_CHECK_HIP(hipMemcpyAsync(&got, cnt.p, 8, hipMemcpyDeviceToHost, ctx->stream));
    PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    if (got != nl) {
      set_error("line count mismatch (%llu vs %llu)", (unsigned long long)got,
                (unsigned long long)nl);
      return PAC_ERR_HIP;
    }
  }
  PAC_TRY(edge.reserve(sizeof(unsigned long long) * nl));
  PAC_TRY(keep.reserve(nl));
  PAC_TRY(flags.reserve(16));
  PAC_TRY(fill_u32(ctx->stream, flags.as<uint32_t>(), 4, 0u));
  hipLaunchKernelGGL(k_parse_lines, dim3(grid1d(nl)), dim3(256), 0, ctx->stream,
                     dt.as<const char>(), n, ls.as<const uint64_t>(), nl,
                     edge.as<unsigned long long>(), keep.as<uint8_t>(), flags.as<uint32_t>(),
                     flags.as<uint32_t>() + 3);
  PAC_CHECK_HIP(hipGetLastError());
  uint32_t hf[4] = {0, 0, 0, 0};
  PAC_CHECK_HIP(hipMemcpyAsync(hf, flags.p, 16, hipMemcpyDeviceToHost, ctx->stream));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (hf[1]) {
    *host_needed = true;
    return PAC_OK;
  }
  if (hf[2]) {
    set_error("vertex ID exceeds int32");
    return PAC_ERR_UNSUPPORTED;
  }
  dt.release();
  ls.release();
  PAC_TRY(edges.reserve(sizeof(unsigned long long) * (nl ? nl : 1)));
  {
    size_t tmp = 0;
    PAC_CHECK_HIP(rocprim::select(nullptr, tmp, edge.as<const unsigned long long>(),
                                  keep.as<const uint8_t>(), edges.as<unsigned long long>(),
                                  cnt.as<uint64_t>(), (size_t)nl, ctx->stream));
    PAC_TRY(ctx->scratch.reserve(tmp));
    PAC_CHECK_HIP(rocprim::select(ctx->scratch.p, tmp, edge.as<const unsigned long long>(),
                                  keep.as<const uint8_t>(), edges.as<unsigned long long>(),
                                  cnt.as<uint64_t>(), (size_t)nl, ctx->stream));
  }
  uint64_t m = 0;
  PAC_CHECK_HIP(hipMemcpyAsync(&m, cnt.p, 8, hipMemcpyDeviceToHost, ctx->stream));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  const bool zero_based = hf[0] != 0;
  *V = zero_based ? (uint64_t)hf[3] + 1 : (uint64_t)hf[3];
  *ne = m;
  PAC_TRY(d_src.reserve(4 * (m ? m : 1)));
  PAC_TRY(d_dst.reserve(4 * (m ? m : 1)));
  if (m) {
    hipLaunchKernelGGL(k_split_edges, dim3(grid1d(m)), dim3(256), 0, ctx->stream,
                       edges.as<const unsigned long long>(), m, zero_based ? 0u : 1u,
                       d_src.as<uint32_t>(), d_dst.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
  }
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // the temporaries are freed on return
  return PAC_OK;
}

int graph_from_device_edges(Graph *g, const uint32_t *d_src, const uint32_t *d_dst, uint64_t ne,
                            uint32_t V) {
  Ctx *ctx = g->ctx;
  uint32_t shift = 1;
  while ((1ull << shift) < (uint64_t)V) ++shift;
  const uint64_t n2 = 2 * ne;
  DevBuf keys_a;
  PAC_TRY(keys_a.reserve(sizeof(uint64_t) * (n2 ? n2 : 1)));
  PAC_TRY(ctx->scratch2.reserve(64));
  uint32_t *err = ctx->scratch2.as<uint32_t>();
  PAC_TRY(fill_u32(ctx->stream, err, 1, 0u));
  if (ne) {
    hipLaunchKernelGGL(k_edge_keys, dim3(grid1d(ne)), dim3(256), 0, ctx->stream, d_src, d_dst, ne,
                       shift, V, keys_a.as<uint64_t>(), err);
    PAC_CHECK_HIP(hipGetLastError());
  }
  uint32_t h = 0;
  PAC_CHECK_HIP(hipMemcpyAsync(&h, err, 4, hipMemcpyDeviceToHost, ctx->stream));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (h) {
    set_error("edge endpoint >= number of vertices (%u)", V);
    return PAC_ERR_FORMAT;
  }
  return csr_from_keys_device(g, keys_a, n2, shift, V);
}

}  // namespace pac
