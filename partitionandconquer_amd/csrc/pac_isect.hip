// pac_isect.hip -- gfx950 intersection kernels for triangle counting, LCC and
// per-edge support.
//
// Counterparts of the reference's hot loops (Byy-0024/PartitionAndConquer):
//   Graph::cnt_common_neis_merge / _less_than_merge / _galloping (Graph.cpp:186-285)
//   Graph::solve_triangle_counting_merge (Graph.cpp:355-368)
//   Partitioned_Graph::solve_triangle_counting_merge (Graph.cpp:1346-1362)
//   Graph::solve_lcc_merge (Graph.cpp:828-840)
//
// Work decomposition.  The graph is oriented by degree: u->v iff (d(u),u) < (d(v),v),
// so every triangle {t,h,w} is found exactly once, at the oriented edge t->h whose
// tail t has the smallest (degree, id), as w in N+(t) & N+(h).  Every out-list is kept in
// that (degree, id) order (oadj holds the IDs, okey their ranks), so the elements of N+(t)
// that can meet N+(h) -- those ranked above h -- are the suffix of N+(t) after h.  Each
// oriented edge t->h is an intersection pair owned by its head h (pairs whose suffix or
// N+(h) is empty are dropped): a workgroup owns a row x, stages N+(x) (the list A) once in
// LDS, and streams for each pair (y, x) the suffix of N+(y) above x (the list B) from HBM,
// looking each element up in A.  The suffixes total 0.6 of what streaming the shorter whole
// list of every pair would.  The three strategies differ only in the lookup:
//   kMerge  : wave64 merge-path merge of A (LDS) and B (HBM) by rank: each lane walks an
//             equal share of the merged sequence after a diagonal binary search.
//   kGallop : the shorter list's ranks are binary-searched in the longer one
//             (B elements in A held in LDS, or A elements in B in HBM).
//   kBitmap : A is held in a two-choice cuckoo hash table in LDS (two buckets of two
//             slots per key, load at most 1/2); each B element is looked up with two
//             8-byte LDS reads and four compares, no loop (exact); a table that
//             leaves a key homeless is built again with another hash multiplier.  Pairs of
//             at most kChunkMax elements are looked up in 4-element chunks spread over the
//             lanes.
// Counting is exact by construction in all three: a hit is only recorded for an
// element equal to a list entry.
//
// Per-vertex and per-edge counts.  A hit w of pair (y, x) is the triangle {y, x, w}:
// t(x) += 1 (a register per row); t(w) goes to the LDS counter of w's place in A;
// t(y) goes to one global atomic per pair (the pair's edge y->x enters the owner).  After
// the row each nonzero LDS counter is added to the global count of its vertex -- about
// one global atomic per pair instead of one per triangle.  SUP mode (per-edge support,
// for Jaccard) keeps the same counters per edge: x->w at its place in A, y->x per pair;
// the third edge y->w lies in the streamed list, and the hash path records it as one hit
// bit per streamed element (the pair's hit words), summed per edge after the launch
// (k_sup_small, k_sup_big), where the sorted and HBM paths add one atomic per triangle
// instead.
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "pac_internal.hpp"

// The temporaries of pac_tc_noalloc: one copy per GPU, held until the process ends.  Host
// code never uses this name as an address (on the host it is a stand-in symbol): it takes
// the device address from hipGetSymbolAddress, once per device, in noalloc_scratch_init.
// Layout: [0, 256) work-queue heads, [256, 512) the graph descriptor, then the
// deferred-row queue.
static __device__ __align__(256) unsigned char scratch[PAC_NOALLOC_SCRATCH_BYTES];

namespace pac {

// waves per row, LDS list capacity (uint32 entries)
const TileCfg kTiles[PAC_NUM_TILES] = {{2, 1024}, {4, 2048}, {4, 4096}, {8, 4096}};

namespace {

// Device view of a GraphDev with its pointers in the global address space.  The
// descriptors are read from memory, so without this the compiler cannot prove where
// the pointers point and emits flat loads, which also count against the LDS wait
// counter: every wait for an LDS result would then wait for the outstanding HBM loads.
#define PAC_GLOBAL __attribute__((address_space(1)))
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
// four IDs at any 4-byte boundary (a typedef may lower the alignment)
typedef unsigned int u32x4a __attribute__((ext_vector_type(4), aligned(4)));
struct GDev {
  const PAC_GLOBAL uint32_t *ooff, *oadj, *okey, *poff, *ppar, *pedge, *rows;
  const PAC_GLOBAL u32x2 *pinfo;
  const PAC_GLOBAL uint8_t *sadj;  // the streamed lists, selem bytes per element
  const PAC_GLOBAL u32x2 *sinfo;   // per pair: 4-byte word of sadj where its list starts
  PAC_GLOBAL unsigned long long *total, *tvert;
  PAC_GLOBAL uint32_t *tlow;
  PAC_GLOBAL uint32_t *sup;
  const PAC_GLOBAL unsigned long long *pword;
  PAC_GLOBAL unsigned long long *hitw;
  uint32_t V, m, part_base, selem;
};

__device__ __forceinline__ GDev to_dev(const GraphDev &d) {
  GDev g;
  g.ooff = (const PAC_GLOBAL uint32_t *)d.ooff;
  g.oadj = (const PAC_GLOBAL uint32_t *)d.oadj;
  g.okey = (const PAC_GLOBAL uint32_t *)d.okey;
  g.poff = (const PAC_GLOBAL uint32_t *)d.poff;
  g.ppar = (const PAC_GLOBAL uint32_t *)d.ppar;
  g.pedge = (const PAC_GLOBAL uint32_t *)d.pedge;
  g.pinfo = (const PAC_GLOBAL u32x2 *)d.pinfo;
  g.sadj = (const PAC_GLOBAL uint8_t *)d.sadj;
  g.sinfo = (const PAC_GLOBAL u32x2 *)d.sinfo;
  g.selem = d.selem;
  g.rows = (const PAC_GLOBAL uint32_t *)d.rows;
  g.total = (PAC_GLOBAL unsigned long long *)d.total;
  g.tvert = (PAC_GLOBAL unsigned long long *)d.tvert;
  g.tlow = (PAC_GLOBAL uint32_t *)d.tlow;
  g.sup = (PAC_GLOBAL uint32_t *)d.sup;
  g.pword = (const PAC_GLOBAL unsigned long long *)d.pword;
  g.hitw = (PAC_GLOBAL unsigned long long *)d.hitw;
  g.V = d.V;
  g.m = d.m;
  g.part_base = d.part_base;
  return g;
}

// Relaxed device-scope atomic adds on global memory.
__device__ __forceinline__ void gadd(PAC_GLOBAL uint32_t *p, uint32_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gadd(PAC_GLOBAL unsigned long long *p, unsigned long long v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LCC: a pair's credit to its partner y (the triangles in which y ranks lowest): 32-bit
// atomics into tlow when the launch has it, else into the 64-bit counts.
__device__ __forceinline__ void credit_low(const GDev &G, uint32_t y, uint32_t v) {
  if (G.tlow)
    gadd(&G.tlow[y], v);
  else
    gadd(&G.tvert[y], (unsigned long long)v);
}

constexpr uint32_t kAllC = 0xFFFFFFFFu;
constexpr uint32_t kPassPairs = 57344;  // pairs per hash pass (bounds the 16-bit counters)
// TC/LCC hash path: pairs streaming at most this many elements go through the chunk path.
constexpr uint32_t kChunkMax = 64;

constexpr uint32_t pow2_at_least(uint32_t x) { return x <= 1 ? 1u : 2u * pow2_at_least((x + 1) / 2); }

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Branch-free lower_bound: first index i in [0,n) with a[i] >= x (n if none).  The
// trip count depends only on n, so lanes searching the same list stay converged.
template <class P>
__device__ __forceinline__ uint32_t lower_bound_u32(P a, uint32_t n, uint32_t x) {
  if (n == 0) return 0;
  uint32_t base = 0;
  while (n > 1) {
    const uint32_t half = n >> 1;
    base = (a[base + half] < x) ? base + half : base;
    n -= half;
  }
  return base + (a[base] < x ? 1u : 0u);
}

constexpr uint32_t kEmpty = 0xFFFFFFFFu;  // empty hash-table slot (IDs are < 2^31)
// Word 3 of a raw (stride 0) buffer resource on this target: 32-bit data format, which
// untyped buffer loads require; out-of-range lanes read 0.
constexpr int kRawBufferFlags = 0x00020000;

// LDS of one workgroup.  The sorted-list strategies (merge, gallop) keep A and one
// 32-bit counter per position; the hash strategy keeps the cuckoo table of A with packed
// 16-bit counters per slot (or, for rows one wave does alone, one table per wave in a
// 1/WAVES share of both arrays).  Both live in one union: a workgroup runs one strategy
// at a time.  (8-byte aligned: a bucket's two slots are read with one 64-bit LDS load.)
// `homeless` counts the keys a workgroup's hash-table build could not place.
template <int CAP, int MODE, int WAVES>
struct alignas(8) Smem {
  static constexpr uint32_t kSlots = 2u * pow2_at_least(CAP);
  union {
    struct {
      uint32_t A[CAP];
      uint32_t cnt[MODE == kModeTC ? 1 : CAP];
    } s;  // merge / gallop
    struct {
      uint32_t tab[kSlots];
      uint32_t cnt16[MODE == kModeTC ? 1 : kSlots / 2];   // 2 x u16 per word
    } h;  // hash
  } u;
  uint32_t item;
  uint32_t defer_slot;
  uint32_t homeless;
  uint32_t rowtot;  // the row's triangle count (LCC mode)
  // LCC mode: per wave, the hit counts of its batch's small pairs (chunk path of row_hash)
  uint32_t pcnt[MODE == kModeLCC ? WAVES : 1][MODE == kModeLCC ? 64 : 1];
  // SUP mode: per wave, the hit words of its batch's small pairs (chunk path of row_hash)
  unsigned long long phw[MODE == kModeSUP ? WAVES : 1][MODE == kModeSUP ? 64 : 1];
};

__device__ __forceinline__ uint32_t readlane(uint32_t v, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

__device__ __forceinline__ unsigned long long readlane64(unsigned long long v, uint32_t lane) {
  return (unsigned long long)readlane((uint32_t)v, lane) |
         ((unsigned long long)readlane((uint32_t)(v >> 32), lane) << 32);
}

// 16-bit counters packed two per word.  A hash pass covers at most kPassPairs pairs,
// so a counter stays below 2^16 (at most one hit per pair for an element of A plus at
// most nA <= CAP <= 4096 hits of the pair x->y at y's place) and halves never carry.
__device__ __forceinline__ uint32_t get16(const uint32_t *w, uint32_t slot) {
  return (w[slot >> 1] >> (16u * (slot & 1u))) & 0xFFFFu;
}

// Record a confirmed triangle (r, c, w) in the sorted-list strategies: w = A[i] = B[j].
template <int MODE>
__device__ __forceinline__ void on_hit(uint32_t *cnt, uint32_t i, PAC_GLOBAL uint32_t *sup,
                                       uint32_t bj) {
  if (MODE != kModeTC) atomicAdd(&cnt[i], 1u);
  if (MODE == kModeSUP) gadd(&sup[bj], 1u);
}

// ---- the three lookup engines: one wave, one pair (x, y) ---------------------------
// The sorted strategies compare ranks: A = okey of N+(x) (staged, ascending) and
// B = okey[b0 .. b0+nb), the ranks of the pair's stream (ascending; rank order is each
// list's order).  Each returns this lane's number of confirmed common neighbours.

template <int MODE>
__device__ __forceinline__ uint32_t probe_gallop(const PAC_GLOBAL uint32_t *__restrict__ okey,
                                                 const uint32_t *A, uint32_t nA, uint32_t *cnt,
                                                 uint32_t b0, uint32_t nb,
                                                 PAC_GLOBAL uint32_t *sup, uint32_t lane) {
  uint32_t hits = 0;
  if (nb <= nA) {
    // B is the shorter list: each B element is binary-searched in A (LDS).
    for (uint32_t j0 = lane; j0 < nb; j0 += 128) {
      uint32_t w[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const uint32_t j = j0 + 64u * u;
        w[u] = j < nb ? okey[b0 + j] : 0u;
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const uint32_t j = j0 + 64u * u;
        if (j < nb) {
          const uint32_t i = lower_bound_u32(A, nA, w[u]);
          if (i < nA && A[i] == w[u]) {
            ++hits;
            on_hit<MODE>(cnt, i, sup, b0 + j);
          }
        }
      }
    }
  } else {
    // A is the shorter list: each A element is binary-searched in B (HBM), skipping
    // A elements outside B's value range.
    const PAC_GLOBAL uint32_t *B = okey + b0;
    const uint32_t lo = B[0], hi = B[nb - 1];
    for (uint32_t i = lane; i < nA; i += 64) {
      const uint32_t x = A[i];
      if (x < lo || x > hi) continue;
      const uint32_t j = lower_bound_u32(B, nb, x);
      if (j < nb && B[j] == x) {
        ++hits;
        on_hit<MODE>(cnt, i, sup, b0 + j);
      }
    }
  }
  return hits;
}

// Merge path: the merged order takes A[i] before B[j] when A[i] <= B[j].  Lane l
// owns merged positions [l*L, (l+1)*L).  Its start (i,j) comes from a binary search
// on its diagonal; then it walks.  When B[j] is taken, every A element <= B[j] has
// already been taken, so B[j] is a common element iff A[i-1] == B[j].
template <int MODE>
__device__ __forceinline__ uint32_t probe_merge(const PAC_GLOBAL uint32_t *__restrict__ okey,
                                                const uint32_t *A, uint32_t nA, uint32_t *cnt,
                                                uint32_t b0, uint32_t nb,
                                                PAC_GLOBAL uint32_t *sup, uint32_t lane) {
  const PAC_GLOBAL uint32_t *B = okey + b0;
  const uint32_t total = nA + nb;
  const uint32_t L = (total + 63u) >> 6;
  uint32_t d0 = lane * L;
  if (d0 >= total) return 0;
  const uint32_t d1 = min(d0 + L, total);
  uint32_t lo = d0 > nb ? d0 - nb : 0u, hi = min(d0, nA);
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (A[mid] <= B[d0 - mid - 1]) lo = mid + 1;
    else hi = mid;
  }
  uint32_t i = lo, j = d0 - lo;
  uint32_t ai = i < nA ? A[i] : 0xFFFFFFFFu;
  uint32_t bj = j < nb ? B[j] : 0xFFFFFFFFu;
  uint32_t hits = 0;
  for (uint32_t d = d0; d < d1; ++d) {
    if (i < nA && (j >= nb || ai <= bj)) {
      ++i;
      ai = i < nA ? A[i] : 0xFFFFFFFFu;
    } else {
      if (i > 0 && A[i - 1] == bj) {
        ++hits;
        on_hit<MODE>(cnt, i - 1, sup, b0 + j);
      }
      ++j;
      bj = j < nb ? B[j] : 0xFFFFFFFFu;
    }
  }
  return hits;
}

// The pairs of a row are visited in batches of 64 per wave: lane l loads the bounds of
// N+(y) stored with the wave's l-th pair (and, when counts are kept, its partner y and edge
// index e), so the offsets of 64 lists arrive with one round trip; the wave then walks the
// batch, broadcasting each pair with readlane.  Wave w takes pair positions k_lo + w, k_lo + w + WAVES, ...
// body(y, e, b0, nb, p) runs once per pair, p being its index in the pair lists (e is only
// loaded when counts are kept).
template <int WAVES, int MODE, class Body>
__device__ __forceinline__ void for_each_pair(const GDev &G, uint32_t p0, uint32_t k_lo,
                                              uint32_t k_hi, uint32_t wave, uint32_t lane,
                                              Body body) {
  for (uint32_t t0 = 0;; t0 += 64) {
    const uint32_t kl = k_lo + wave + (uint32_t)WAVES * (t0 + lane);
    uint32_t yl = 0, el = 0, bl = 0, nl = 0;
    if (kl < k_hi) {
      const u32x2 pi = G.pinfo[p0 + kl];
      bl = pi.x;
      nl = pi.y;
      if (MODE != kModeTC) {
        yl = G.ppar[p0 + kl];
        el = G.pedge[p0 + kl];
      }
    }
    const uint32_t nvalid = (uint32_t)__popcll(__ballot(kl < k_hi));
    for (uint32_t jj = 0; jj < nvalid; ++jj)
      body(readlane(yl, jj), readlane(el, jj), readlane(bl, jj), readlane(nl, jj),
           p0 + k_lo + wave + (uint32_t)WAVES * (t0 + jj));
    if (nvalid < 64) break;
  }
}

// SUP mode, paths that credit the streamed list's edges directly: the pair's hit words
// are written as zero, so the column sums after the launch add nothing for it.
__device__ __forceinline__ void clear_hit_words(const GDev &G, uint32_t nb, uint32_t p,
                                                uint32_t lane) {
  const uint32_t nw = (nb + 63u) >> 6;
  const unsigned long long base = G.pword[p];
  for (uint32_t c = lane; c < nw; c += 64) G.hitw[base + c] = 0ull;
}

// ---- one row, sorted-list strategies (merge, gallop): A staged in LDS ---------------
// (Inlined, like row_hbm: as calls, which kept their code out of the persistent kernel's
// body, the LDS arrays reached them as generic pointers, and the launches whose dispatch
// picks these strategies slowed down -- 256 mixed graphs 32.3 -> 36.5 ms -- while R-MAT
// scale 24 LCC moved within its run-to-run spread.)
template <int STRAT, int MODE, int WAVES, int CAP>
__device__ __forceinline__ void row_sorted(const GDev &G, uint32_t r, uint32_t k_lo,
                                           uint32_t k_hi, Smem<CAP, MODE, WAVES> &sm,
                                           unsigned long long &acc) {
  constexpr uint32_t BS = 64u * WAVES;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t a0 = G.ooff[r], nA = G.ooff[r + 1] - a0;
  const uint32_t p0 = G.poff[r], np = G.poff[r + 1] - p0;
  uint32_t *A = sm.u.s.A;
  uint32_t *cnt = sm.u.s.cnt;
  for (uint32_t i = tid; i < nA; i += BS) {
    A[i] = G.okey[a0 + i];
    if (MODE != kModeTC) cnt[i] = 0;
  }
  if (tid == 0) sm.rowtot = 0;
  __syncthreads();
  if (k_hi > np) k_hi = np;
  uint32_t row_hits = 0;
  for_each_pair<WAVES, MODE>(G, p0, k_lo, k_hi, wave, lane,
                             [&](uint32_t y, uint32_t e, uint32_t b0, uint32_t nb, uint32_t p) {
    uint32_t h = 0;
    if (MODE == kModeSUP) clear_hit_words(G, nb, p, lane);
    if (nb != 0) {
      if (STRAT == kGallop)
        h = probe_gallop<MODE>(G.okey, A, nA, cnt, b0, nb, G.sup, lane);
      else
        h = probe_merge<MODE>(G.okey, A, nA, cnt, b0, nb, G.sup, lane);
    }
    if (MODE != kModeTC) {
      const uint32_t ch = wave_sum(h);
      if (lane == 0 && ch) {
        if (MODE == kModeLCC)
          credit_low(G, y, ch);  // the pair's edge y->r: t(y)
        else
          gadd(&G.sup[e], ch);  // the pair's edge y->r
      }
    }
    row_hits += h;
  });
  acc += row_hits;
  if (MODE == kModeLCC) {
    const uint32_t wh = wave_sum(row_hits);
    if (lane == 0 && wh) atomicAdd(&sm.rowtot, wh);
  }
  __syncthreads();
  if (MODE != kModeTC) {
    for (uint32_t i = tid; i < nA; i += BS) {
      const uint32_t v = cnt[i];
      if (v && MODE == kModeLCC) gadd(&G.tvert[G.oadj[a0 + i]], (unsigned long long)v);  // t(w)
      else if (v) gadd(&G.sup[a0 + i], v);  // the edge r->w
    }
    // an atomic even for a whole row: other rows' tail credits to r arrive meanwhile
    if (MODE == kModeLCC && tid == 0 && sm.rowtot) gadd(&G.tvert[r], (unsigned long long)sm.rowtot);
  }
  __syncthreads();
}

// ---- one row, hash strategy: two-choice cuckoo table of A in LDS ---------------------
// A is held in a table of 2^lg buckets of two slots (2^lg >= nA, so the load factor is
// at most 1/2, far below the ~0.89 limit of this layout).  Key x may sit in bucket
// b1(x) or b2(x); a lookup reads both buckets (two 8-byte LDS reads) and compares four
// slots -- a fixed cost with no loop and no divergence.  Insertion is a bounded walk: a
// key takes a free slot of its two buckets, else evicts a resident of the bucket it was
// not itself evicted from and carries the evicted key on (always evicting from the
// other bucket is what keeps two keys from trading one slot back and forth: on R-MAT
// out-lists it leaves about 2 rows in 100 000 with a key still homeless after
// kMaxKicks evictions, where alternating between the two buckets by kick number left
// 2 in 1 000, most of them rows of 129-512 keys).  A row with a homeless key is built
// again with the next of kHashTries hash multipliers (the row's lookups then use that
// one), and only a row homeless under all of them is done by another path (exact either
// way: the sorted lists for a workgroup's row, HBM lookups for a one-wave row), so lookups
// never scan an overflow area.
// Each pair's list N+(y) is streamed by one wave, 4 x 64 elements per step with
// coalesced loads, software-pipelined one step ahead across the wave's pairs; the
// pair's hit count is a ballot popcount held in scalar registers.
constexpr uint32_t kMaxKicks = 48;
// The two buckets of x among 2^lg (sh = 32 - lg), from one full-rate 24-bit multiply-add
// and no shift: h = (x mod 2^24) * M + x, which for IDs below 2^24 is x times the odd
// constant M + 1 (the + x also carries ID bits above 23 into the top of h); b1 is the top
// lg bits of h and b2 = b1 ^ d with d = bits [3, 3 + lg) of h made odd (lg <= 12, so the
// two fields are disjoint) -- so the two choices always differ (one bucket of each parity,
// the two tables of classic cuckoo hashing).  The streamed lookups need two vector
// instructions fewer per element than with d taken elsewhere: its bucket field sits where a
// bucket's byte address has it (see probe in row_hash).  M is the row's multiplier, a
// wave-uniform value, first 0x9E3778: on the out-lists of R-MAT scales 22, 24 and 25 a
// sequential replay of the table build with it leaves 111, 297 and 843 rows with a homeless
// key (0.003-0.005 %); at scale 25 those include hub rows owning up to 93 245 pairs, whose
// fallback cost the per-edge Jaccard launch about a third of its time.  With the next
// multipliers 7 of those rows still fail the second and none the third.  (__umul24 is
// declared returning int: its result is taken as unsigned before any shift, so the bucket
// indices stay below 2^lg.)
constexpr uint32_t kHashTries = 4;
__device__ __forceinline__ uint32_t hash_mul(uint32_t t) {
  return t == 0 ? 0x9E3778u : t == 1 ? 0xC2B2AEu : t == 2 ? 0x85EBCAu : 0x27D4EAu;
}
__device__ __forceinline__ uint32_t hash_key(uint32_t x, uint32_t mul) {
  return (uint32_t)__umul24(x, mul) + x;
}
__device__ __forceinline__ void buckets(uint32_t x, uint32_t sh, uint32_t lg, uint32_t mul,
                                        uint32_t &b1, uint32_t &b2) {
  const uint32_t h = hash_key(x, mul);
  b1 = h >> sh;
  b2 = b1 ^ (__builtin_amdgcn_ubfe(h, 3u, lg) | 1u);
}

// LDS byte addresses: a shared-memory pointer's offset in the workgroup's LDS, and 8-byte
// loads / 32-bit atomic adds at such an offset.
typedef __attribute__((address_space(3))) uint32_t LdsU32;
typedef __attribute__((address_space(3))) const u32x2 LdsU2;
__device__ __forceinline__ uint32_t lds_addr(const void *p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}
__device__ __forceinline__ uint2 lds_load_u2(uint32_t a) {
  const u32x2 v = *(LdsU2 *)(uintptr_t)a;
  return make_uint2(v.x, v.y);
}
// base + off, base a constant where possible: the compiler then folds it into the
// instruction's offset field.
__device__ __forceinline__ void lds_add_u32(uint32_t base, uint32_t off, uint32_t v) {
  __attribute__((address_space(3))) char *p =
      (__attribute__((address_space(3))) char *)(uintptr_t)base;
  __hip_atomic_fetch_add((LdsU32 *)(p + off), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// SOLO: one wave does the whole row on its own share (1/WAVES) of the table space, with
// wave-level ordering instead of workgroup barriers, so the workgroup's waves work on
// different small rows at once (the caller keeps the row within that share and to one
// batch of pairs).  Otherwise the workgroup's waves share the row, taking its pairs in turn.
// ESF: the streamed lists' element width, fixed at compile time (3 or 4) or, with 0, read
// from the descriptor (then both stream loops are compiled in).
template <int MODE, int WAVES, int CAP, bool SOLO = false, int ESF = 0>
__device__ __forceinline__ bool row_hash(const GDev &G, uint32_t r, uint32_t c_lo,
                                         uint32_t c_hi, Smem<CAP, MODE, WAVES> &sm,
                                         unsigned long long &acc) {
  using S = Smem<CAP, MODE, WAVES>;
  constexpr uint32_t BS = SOLO ? 64u : 64u * WAVES;  // threads on the row
  constexpr uint32_t PS = SOLO ? 1u : (uint32_t)WAVES;  // pair stride between its waves
  const uint32_t tid = SOLO ? (threadIdx.x & 63u) : threadIdx.x;  // rank on the row
  const uint32_t lane = threadIdx.x & 63u, wave = SOLO ? 0u : threadIdx.x >> 6;
  auto wave_sync = []() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  auto sync = [&]() {
    if (SOLO)
      wave_sync();
    else
      __syncthreads();
  };
  const uint32_t a0 = G.ooff[r], nA = G.ooff[r + 1] - a0;
  const uint32_t p0 = G.poff[r], np = G.poff[r + 1] - p0;
  auto &H = sm.u.h;
  const uint32_t share = SOLO ? threadIdx.x >> 6 : 0u;  // the wave's share of the table
  uint32_t *const tab = H.tab + share * (S::kSlots / WAVES);
  uint32_t *const cnt16 = MODE == kModeTC ? H.cnt16 : H.cnt16 + share * (S::kSlots / WAVES / 2);
  if (c_hi > np) c_hi = np;
  // The wave's first batch of pairs is loaded before the table is built, so the latency
  // overlaps it: per pair the bounds of its stream (stored with the pair: no gather
  // through ooff), and when counts are kept the partner; SUP mode also needs the edge
  // index and the first hit word.
  uint32_t kl = c_lo + wave + PS * lane, yl = 0, el = 0, bl = 0, nl = 0;
  uint32_t b4 = 0, n4 = 0;    // the pair's exact stream (its suffix) in oadj
  unsigned long long wl = 0;  // SUP mode: the pair's first hit word
  auto load_pair = [&]() {
    if (kl < c_hi) {
      const u32x2 pi = G.sinfo[p0 + kl];
      bl = pi.x;
      nl = pi.y;
      b4 = bl;
      n4 = nl;
      // 3-byte streams start on a 4-element boundary, SUP-mode streams on a 64-element one
      if (MODE == kModeSUP || ESF == 3 || (ESF == 0 && G.selem == 3u)) {
        const u32x2 q = G.pinfo[p0 + kl];
        b4 = q.x;
        n4 = q.y;
      }
      if (MODE != kModeTC) yl = G.ppar[p0 + kl];
      if (MODE == kModeSUP) {
        el = G.pedge[p0 + kl];
        wl = G.pword[p0 + kl];
      }
    }
  };
  load_pair();
  uint32_t lg = 4;
  while ((1u << lg) < nA) ++lg;  // 2^lg >= nA: load factor <= 1/2
  const uint32_t nslot = 2u << lg, sh = 32u - lg;
  if (MODE != kModeTC)
    for (uint32_t i = tid; i < nslot / 2; i += BS) cnt16[i] = 0;
  if (!SOLO && tid == 0) sm.rowtot = 0;
  // The thread's first element of A, loaded here and kept for the builds and, in SUP mode,
  // the row's edge credits at the end (most rows have at most one element per thread).
  uint32_t key0 = 0;
  if (tid < nA) key0 = G.oadj[a0 + tid];
  uint32_t mul = 0;  // the row's hash multiplier (uniform)
  for (uint32_t attempt = 0;; ++attempt) {
    mul = hash_mul(attempt);
    for (uint32_t i = tid; i < nslot; i += BS) tab[i] = kEmpty;
    if (!SOLO && tid == 0) sm.homeless = 0;
    bool lost = false;  // SOLO: this lane left a key homeless
    sync();
    for (uint32_t i = tid; i < nA; i += BS) {
      uint32_t x = i == tid ? key0 : G.oadj[a0 + i];
      uint32_t from = kEmpty;  // the bucket x was just evicted from (none yet)
      for (uint32_t kick = 0;; ++kick) {
        uint32_t b1, b2;
        buckets(x, sh, lg, mul, b1, b2);
        if (atomicCAS(&tab[2u * b1], kEmpty, x) == kEmpty) break;
        if (atomicCAS(&tab[2u * b1 + 1u], kEmpty, x) == kEmpty) break;
        if (atomicCAS(&tab[2u * b2], kEmpty, x) == kEmpty) break;
        if (atomicCAS(&tab[2u * b2 + 1u], kEmpty, x) == kEmpty) break;
        if (kick == kMaxKicks) {
          if (SOLO)
            lost = true;
          else
            atomicAdd(&sm.homeless, 1u);
          break;
        }
        // evict from the bucket x did not just come from, alternating slots every two kicks
        const uint32_t vb = from == b1 ? b2 : b1;
        x = atomicExch(&tab[2u * vb + ((kick >> 1) & 1u)], x);
        from = vb;
        if (x == kEmpty) break;
      }
    }
    sync();
    if (SOLO ? __ballot(lost) == 0ull : sm.homeless == 0) break;  // uniform
    if (attempt + 1 == kHashTries) return false;
    sync();  // every thread has read homeless before the next build resets it
  }
  // The streamed lookups address the table by byte: bucket b is at tb + 8b, and the table
  // starts at a multiple of its size (the workgroup's LDS begins with it, and a wave's share
  // of it is a multiple of the largest table one wave builds alone), so bucket b1 ^ f lies at
  // (tb + 8 b1) ^ 8f -- checked here, uniformly, and never false in practice: a table placed
  // otherwise is done by the exact fallback paths.  Counter word b is at
  // cb + (bucket address >> 1), cb = (counters' address) - tb / 2: a wave's share of the
  // counters starts half as far into them as its share of the table does into the table,
  // so cb is the same for every share, taken from the workgroup's arrays (a constant).
  const uint32_t tb = lds_addr(tab), m3 = ((1u << lg) - 1u) << 3;
  if ((tb & (m3 | 7u)) != 0) return false;  // uniform
  const uint32_t cb = MODE == kModeTC ? 0u : lds_addr(H.cnt16) - (lds_addr(H.tab) >> 1);
  // The counters' base as an instruction offset: a constant when the workgroup's LDS (one
  // Smem) starts at address 0, checked here (uniform; never false in practice).
  constexpr uint32_t kCntBase = 4u * S::kSlots;
  if (MODE != kModeTC && cb != kCntBase) return false;
  // Lookup of w.  A bucket's two slots share one word of the packed 16-bit counters, so
  // a hit is reported as that word (the bucket) and the increment for w's half of it
  // (1 or 1 << 16); word and increment are only meaningful on a hit.  Both buckets are
  // always read (two 8-byte LDS loads, no branch): with 64-bit bucket loads LDS has
  // headroom, and a lookup that skips the second bucket when the first has a free slot --
  // exact, since a key only goes to its second bucket when the first is full -- costs
  // more in divergent scalar work than it saves.
  auto find = [&](uint32_t w, uint32_t &word, uint32_t &inc) -> bool {
    uint32_t b1, b2;
    buckets(w, sh, lg, mul, b1, b2);
    const uint2 q1 = reinterpret_cast<const uint2 *>(tab)[b1];
    const uint32_t k10 = q1.x, k11 = q1.y;
    const uint2 q2 = reinterpret_cast<const uint2 *>(tab)[b2];
    const uint32_t k20 = q2.x, k21 = q2.y;
    // bitwise | (no short-circuit): all four slots are read up front, no branches
    const bool in1 = (k10 == w) | (k11 == w), hi = (k11 == w) | (k21 == w);
    inc = hi ? 0x10000u : 1u;
    word = in1 ? b1 : b2;
    return in1 | (k20 == w) | hi;
  };
  // The streamed lookups: the four slot compares of w are balloted one by one (each
  // ballot is the compare's own lane mask), so the wave's hit mask never passes through
  // a vector register; the hit's counter word and increment are only formed under it.
  // The buckets of buckets() by byte address: a1 = tb + 8 b1, a2 = a1 ^ 8 (bits [3, 3 + lg)
  // of h, made odd) = a1 ^ ((h | 8) & m3): a multiply-add, a shift, a shift-add, a bitop3 and
  // an xor.
  struct Probe {
    bool e10, e11, e20, e21;
    uint32_t a1, a2;
  };
  auto probe = [&](uint32_t w) -> Probe {
    Probe q;
    const uint32_t h = hash_key(w, mul);
    q.a1 = tb + ((h >> sh) << 3);
    q.a2 = q.a1 ^ ((h | 8u) & m3);
    const uint2 q1 = lds_load_u2(q.a1);
    const uint2 q2 = lds_load_u2(q.a2);
    q.e10 = q1.x == w;
    q.e11 = q1.y == w;
    q.e20 = q2.x == w;
    q.e21 = q2.y == w;
    return q;
  };
  uint32_t row_hits = 0;  // wave-uniform
  // Blocks of up to 4 x 64 elements of one pair's list N+(y), in pair order; the loads
  // of the next block are issued before the current one is probed.  The loads go through
  // a buffer resource that starts at the block and ends with N+(y): the per-lane offsets
  // are fixed for the pair (no vector register is rewritten between the pipelined loads,
  // no per-element address arithmetic; the four sub-blocks are immediate offsets), and
  // lanes past the list's end read 0 from the hardware's range check or, with 3-byte
  // elements, bytes that follow the list -- either way their lookups are masked.  The
  // two register sets swap roles every step, so the waits cover only the block being
  // probed.  (Unconditional loads: the wait for a block then counts exactly the loads
  // issued after it, which keeps the next block's loads in flight.)  The stream position
  // is a scalar cursor: the block's byte address and the bytes left in the pair's list
  // from there.
  //   ES = 4: lane l reads element 64u + l of sub-block u with one 4-byte load.
  //   ES = 3: each list starts on a 4-byte boundary; element 64u + l lies in bytes
  // 192u + 3l .. +2, within the aligned word 192u + 4 floor(3l / 4) and, for l = 1, 2
  // (mod 4), the next word, which is the word lane l + 1 reads: lane l loads its word and
  // takes lane l + 1's with a lane shift inside the 16-lane row (lanes 15 (mod 16) never
  // need it), then picks the element's three bytes with one byte permute.  Aligned loads
  // throughout: unaligned 4-byte loads of the elements themselves measured far slower.
  // The range covers the list rounded up to whole words.
  auto stream_batch = [&](auto es_c, uint32_t nvalid, uint32_t &phl) {
    constexpr uint32_t ES = decltype(es_c)::value;
    constexpr uint32_t SSTEP = 64u * ES, BLK = 4u * SSTEP;  // sub-block, block (bytes)
    // ES = 3: the lane's word in a sub-block and the permute that picks its element
    // (bytes s, s+1, s+2 of the word pair, s = 3l mod 4, and a zero top byte)
    const uint32_t lw = ES == 3u ? 4u * (3u * lane / 4u) : 4u * lane;
    const uint32_t sel = 0x0C020100u + ((3u * lane) & 3u) * 0x00010101u;
    auto load_blk = [&](unsigned long long p, uint32_t nbytes, uint32_t *w) {
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          (void *)(uintptr_t)p, (short)0, (int)(ES == 3u ? (nbytes + 3u) & ~3u : nbytes),
          kRawBufferFlags);
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u)
        w[u] = __builtin_amdgcn_raw_buffer_load_b32(rs, lw + SSTEP * u, 0, 0);
    };
    auto elem = [&](const uint32_t *w, uint32_t u) -> uint32_t {
      if constexpr (ES == 4u) {
        return w[u];
      } else {
        // row_shl:1 -- lane l reads lane l + 1 of its row (lane 15 of a row reads 0)
        const uint32_t nx = (uint32_t)__builtin_amdgcn_mov_dpp((int)w[u], 0x101, 0xF, 0xF, true);
        return __builtin_amdgcn_perm(nx, w[u], sel);
      }
    };
    // per pair (lane): the address of N+(y) and its length in bytes
    const unsigned long long pa = (unsigned long long)(uintptr_t)G.sadj + 4ull * bl;
    const uint32_t pn = nl * ES;
    uint32_t jj = 0, nrem = readlane(pn, 0), ph = 0;
    unsigned long long ptr = readlane64(pa, 0);
    unsigned long long hb = MODE == kModeSUP ? readlane64(wl, 0) : 0ull;
    uint32_t sa[4], sb[4];
    load_blk(ptr, nrem, sa);
    auto step = [&](const uint32_t *cur, uint32_t *nxt) {
      const uint32_t rem = nrem;  // bytes left in the pair's list (0 if it is empty)
      // The next block: the pair's next 256 elements, or the first block of the next pair
      // (past the batch's last pair, that pair's first block again: in bounds, unused).
      uint32_t jn = jj, nremn = rem - BLK;
      unsigned long long ptrn = ptr + BLK, hbn = hb + 4u;
      if (rem <= BLK) {  // wave-uniform
        jn = jj + 1u;
        const uint32_t jl = min(jn, nvalid - 1u);
        ptrn = readlane64(pa, jl);
        nremn = readlane(pn, jl);
        if (MODE == kModeSUP) hbn = readlane64(wl, jl);
      }
      load_blk(ptrn, nremn, nxt);
      unsigned long long hw = 0;  // SUP mode: this lane's hit word (lanes 0..3)
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        if (SSTEP * u >= rem) break;  // wave-uniform
        const bool live = lane * ES + SSTEP * u < rem;  // lanes past the end are masked
        const Probe q = probe(elem(cur, u));
        const unsigned long long bal = (__ballot(q.e10) | __ballot(q.e11) | __ballot(q.e20) |
                                        __ballot(q.e21)) & __ballot(live);
        if (MODE == kModeTC) {
          row_hits += (uint32_t)__popcll(bal);  // only the row's total is kept
          continue;
        }
        ph += (uint32_t)__popcll(bal);
        if (MODE == kModeSUP && lane == u) hw = bal;
        if ((q.e10 | q.e11 | q.e20 | q.e21) & live) {
          // t(w) / edge r->w
          // (written out: the compiler adds the base with a vector instruction instead of
          // folding it into the atomic's offset; the atomic returns nothing, and LDS
          // operations complete in order, so later waits on other LDS accesses cover it)
          const uint32_t ca = ((q.e10 | q.e11) ? q.a1 : q.a2) >> 1;
          const uint32_t inc = (q.e11 | q.e21) ? 0x10000u : 1u;
          asm volatile("ds_add_u32 %0, %1 offset:%2" : : "v"(ca), "v"(inc), "i"(kCntBase)
                       : "memory");
        }
      }
      if (MODE == kModeSUP) {
        // edges y->w: one hit word per 64 streamed elements, summed after the launch.  A
        // zero word is not stored: the words start zeroed when the layout is built, and a
        // pair's words are the same in every launch (or zeroed by the paths that credit the
        // supports directly), so a word with no hit now never held one.
        const uint32_t nsub = min(4u, (rem + SSTEP - 1u) / SSTEP);
        if (lane < nsub && hw != 0ull) G.hitw[hb + lane] = hw;
      }
      if (MODE != kModeTC && jn != jj) {
        // pair jj is done: its count goes to its lane, credited with the batch's others
        row_hits += ph;
        phl = lane == jj ? ph : phl;
        ph = 0;
      }
      jj = jn;
      ptr = ptrn;
      nrem = nremn;
      hb = hbn;
    };
    // Two steps per trip (the register sets swap roles) and a single exit, at the bottom:
    // with an exit after each step the compiler materialises undefined values for the
    // loop's scalar state on every trip.  (The exit test reads jj through readfirstlane so
    // that it is not folded into the guard of the second step.)
    do {  // wave-uniform
      step(sa, sb);
      if (jj < nvalid) step(sb, sa);
    } while ((uint32_t)__builtin_amdgcn_readfirstlane((int)jj) < nvalid);
  };
  // Small pairs (at most kChunkMax elements) in chunks of four elements spread over the
  // lanes: the batch's small pairs are cut into 4-element chunks of their 4-byte lists,
  // numbered pair after pair (an exclusive scan of the chunk counts over the pair lanes),
  // and lane l takes chunks l, l + 64, ...: its pair is the last pair lane whose first chunk
  // is at most its chunk (a 6-step binary search across the lanes), its elements one
  // 16-byte load, each looked up in the table as in the wave path.  Such pairs are half of
  // the pairs but 7 % of the streamed elements at R-MAT scale 24 (the hubs' in-edges,
  // whose suffixes above the hub are short); in the wave path each would take a whole step
  // of four 64-element sub-blocks for a few lookups.  Returns the pair hit count of the
  // lane's own pair (LCC mode; summed per pair in the wave's pcnt row in LDS).
  auto chunk_path = [&](bool small) -> uint32_t {
    const uint32_t c = small ? (n4 + 3u) >> 2 : 0u;
    uint32_t incl = c;
#pragma unroll
    for (uint32_t sft = 1; sft < 64; sft <<= 1) {
      const uint32_t o = (uint32_t)__shfl_up((int)incl, sft, 64);
      if (lane >= sft) incl += o;
    }
    const uint32_t P = incl - c;  // the pair's first chunk
    const uint32_t C = readlane(incl, 63);
    uint32_t *pc = nullptr;
    if constexpr (MODE == kModeLCC) {
      pc = sm.pcnt[threadIdx.x >> 6];
      pc[lane] = 0;
      wave_sync();
    }
    auto bperm = [](uint32_t from, uint32_t v) -> uint32_t {
      return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(4u * from), (int)v);
    };
    // SUP mode: a small pair's suffix lies in one 64-element block of its partner's list
    // (its SUP-mode stream, from the block's start, has at most 64 elements), so its hits are
    // one hit word, at the suffix's offset in the block (b4 - bl); they are gathered in the
    // wave's phw row by LDS atomic ORs and stored, with the pair-edge credit = its popcount,
    // by the pair's own lane after the last trip.  Only the suffix is looked up: at R-MAT
    // scale 20 the blocks hold twice its elements.
    unsigned long long *pw = nullptr;
    if constexpr (MODE == kModeSUP) {
      pw = sm.phw[threadIdx.x >> 6];
      pw[lane] = 0ull;
      wave_sync();
    }
    for (uint32_t q0 = 0; q0 < C; q0 += 64) {  // wave-uniform
      const uint32_t q = q0 + lane;
      uint32_t i = 0;
#pragma unroll
      for (uint32_t st = 32; st; st >>= 1)
        if (bperm(i + st, P) <= q) i += st;
      const uint32_t Pi = bperm(i, P);
      const uint32_t e0 = 4u * (q - Pi), bi = bperm(i, b4), ni = bperm(i, n4);
      // SUP mode: the chunk's first bit in its pair's hit word (read in every lane: a lane
      // permute inside a branch would read lanes the branch left out)
      const uint32_t bit0 = MODE == kModeSUP ? e0 + bperm(i, b4 - bl) : 0u;
      const bool act = q < C;
      u32x4a v = {kListEnd, kListEnd, kListEnd, kListEnd};
      if (act) v = *(const PAC_GLOBAL u32x4a *)(G.oadj + bi + e0);  // oadj is padded
      uint32_t h = 0, hm = 0;
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        const Probe q4 = probe(act && e0 + u < ni ? v[u] : kListEnd);  // kListEnd: no key
        const bool hit = q4.e10 | q4.e11 | q4.e20 | q4.e21;
        row_hits += (uint32_t)__popcll(__ballot(hit));
        if (MODE != kModeTC) {
          h += hit ? 1u : 0u;
          hm |= hit ? 1u << u : 0u;
          if (hit) {  // t(w) / edge r->w
            const uint32_t ca = ((q4.e10 | q4.e11) ? q4.a1 : q4.a2) >> 1;
            const uint32_t inc = (q4.e11 | q4.e21) ? 0x10000u : 1u;
            asm volatile("ds_add_u32 %0, %1 offset:%2" : : "v"(ca), "v"(inc), "i"(kCntBase)
                         : "memory");
          }
        }
      }
      if (MODE == kModeLCC && h != 0u) atomicAdd(&pc[i], h);
      if (MODE == kModeSUP && hm != 0u)  // its bits land below 64: the suffix ends in the block
        atomicOr(&pw[i], (unsigned long long)hm << bit0);
    }
    if constexpr (MODE == kModeSUP) {
      wave_sync();
      if (small) {
        const unsigned long long w = pw[lane];
        if (w != 0ull) G.hitw[wl] = w;  // edges y->w, summed after the launch (zero: see stream_batch)
        const uint32_t cnt = (uint32_t)__popcll(w);
        if (cnt) gadd(&G.sup[el], cnt);  // the pair's edge y->r
      }
    }
    if constexpr (MODE == kModeLCC) {
      wave_sync();
      return small ? pc[lane] : 0u;
    }
    return 0u;
  };
  for (uint32_t t0 = 0;; t0 += 64) {
    if (t0 != 0) {
      kl = c_lo + wave + PS * (t0 + lane);
      yl = el = bl = nl = b4 = n4 = 0;
      wl = 0;
      load_pair();
    }
    const bool valid = kl < c_hi;
    const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid));
    if (nvalid == 0) break;
    uint32_t phl = 0;  // lane j: the hit count of the batch's pair j, once it is done
    uint32_t nwave = nvalid;  // pairs left to the wave path (the batch's first lanes)
    {
      const bool small = valid && (MODE == kModeSUP ? nl : n4) <= kChunkMax;
      if (__ballot(small) != 0ull) {  // uniform
        const uint32_t h = chunk_path(small);
        if (MODE == kModeLCC && h != 0u) credit_low(G, yl, h);  // t(y)
        // the other pairs move to the first lanes, in order (a permutation of the lanes)
        const bool longp = valid && !small;
        const unsigned long long lm = __ballot(longp);
        nwave = (uint32_t)__popcll(lm);
        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(lm >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)lm, 0u));
        const int to = (int)(4u * (longp ? below : nwave + lane - below));
        bl = (uint32_t)__builtin_amdgcn_ds_permute(to, (int)bl);
        nl = (uint32_t)__builtin_amdgcn_ds_permute(to, (int)nl);
        if (MODE != kModeTC) yl = (uint32_t)__builtin_amdgcn_ds_permute(to, (int)yl);
        if (MODE == kModeSUP) {
          el = (uint32_t)__builtin_amdgcn_ds_permute(to, (int)el);
          const uint32_t wlo = (uint32_t)__builtin_amdgcn_ds_permute(to, (int)(uint32_t)wl);
          const uint32_t whi = (uint32_t)__builtin_amdgcn_ds_permute(to, (int)(uint32_t)(wl >> 32));
          wl = ((unsigned long long)whi << 32) | wlo;
        }
      }
    }
    if (nwave != 0u) {  // uniform
      static_assert(!(MODE == kModeSUP && ESF == 3), "per-edge supports stream 4-byte lists");
      if constexpr (ESF == 3) {
        stream_batch(std::integral_constant<uint32_t, 3u>(), nwave, phl);
      } else if constexpr (ESF == 4 || MODE == kModeSUP) {
        stream_batch(std::integral_constant<uint32_t, 4u>(), nwave, phl);
      } else {
        if (G.selem == 3u)  // uniform
          stream_batch(std::integral_constant<uint32_t, 3u>(), nwave, phl);
        else
          stream_batch(std::integral_constant<uint32_t, 4u>(), nwave, phl);
      }
    }
    if (MODE != kModeTC && phl != 0) {
      // The batch's pair credits, one lane per pair, all at once: the pair's edge y->r
      // goes to t(y) in LCC mode and to the edge's support in SUP mode, atomically (other
      // rows credit the same vertex meanwhile).
      if (MODE == kModeLCC)
        credit_low(G, yl, phl);
      else
        gadd(&G.sup[el], phl);
    }
    if (nvalid < 64) break;
  }
  // row_hits is the wave's total (uniform): counted once into acc, which is summed over
  // the wave's lanes at the end, so lane 0 carries it
  if (lane == 0) acc += row_hits;
  if (MODE == kModeLCC && !SOLO && lane == 0 && row_hits) atomicAdd(&sm.rowtot, row_hits);
  sync();
  if (MODE == kModeLCC) {
    // t(w) for each w of A: its slot's count, read in slot order (a slot with a nonzero
    // count holds the key it was counted for), so no key is hashed and found again
    for (uint32_t s = tid; s < nslot; s += BS) {
      const uint32_t v = (cnt16[s >> 1] >> (16u * (s & 1u))) & 0xFFFFu;
      if (v) gadd(&G.tvert[tab[s]], (unsigned long long)v);
    }
  } else if (MODE == kModeSUP) {
    // per edge r->A[i], in A order: the credits of A[i]'s slot (coalesced)
    for (uint32_t i = tid; i < nA; i += BS) {
      const uint32_t x = i == tid ? key0 : G.oadj[a0 + i];
      uint32_t word, inc;
      find(x, word, inc);
      const uint32_t v = (cnt16[word] >> (inc == 1u ? 0u : 16u)) & 0xFFFFu;
      if (v) gadd(&G.sup[a0 + i], v);  // the edge r->A[i]
    }
  }
  if (MODE == kModeLCC) {
    const uint32_t rt = SOLO ? row_hits : sm.rowtot;  // the row's triangles (uniform)
    if (tid == 0 && rt) gadd(&G.tvert[r], (unsigned long long)rt);  // atomic: see row_sorted
  }
  sync();
  return true;
}

// ---- one row whose N+(r) exceeds the LDS capacity: lookups in HBM ------------------
// (SOLO: one wave does the whole row, as a fallback of the one-wave hash path.)
template <int MODE, int WAVES, bool SOLO = false>
__device__ __forceinline__ void row_hbm(const GDev &G, uint32_t r, uint32_t c_lo,
                                        uint32_t c_hi, unsigned long long &acc) {
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = SOLO ? 0u : tid >> 6;
  const uint32_t a0 = G.ooff[r], nA = G.ooff[r + 1] - a0;
  const uint32_t p0 = G.poff[r], np = G.poff[r + 1] - p0;
  const PAC_GLOBAL uint32_t *A = G.okey + a0;  // ranks: binary-searched
  if (c_hi > np) c_hi = np;
  uint32_t row_hits = 0;
  for_each_pair<SOLO ? 1 : WAVES, MODE>(G, p0, c_lo, c_hi, wave, lane,
                             [&](uint32_t y, uint32_t e, uint32_t b0, uint32_t nb, uint32_t p) {
    uint32_t h = 0;
    if (MODE == kModeSUP) clear_hit_words(G, nb, p, lane);
    for (uint32_t j = lane; j < nb; j += 64) {
      const uint32_t w = G.okey[b0 + j];
      const uint32_t i = lower_bound_u32(A, nA, w);
      if (i < nA && A[i] == w) {
        ++h;
        if (MODE == kModeLCC) gadd(&G.tvert[G.oadj[b0 + j]], 1ull);  // t(w)
        if (MODE == kModeSUP) gadd(&G.sup[a0 + i], 1u);  // edge r->w
        if (MODE == kModeSUP) gadd(&G.sup[b0 + j], 1u);  // edge y->w
      }
    }
    if (MODE != kModeTC) {
      const uint32_t ch = wave_sum(h);
      if (lane == 0 && ch) {
        if (MODE == kModeLCC)
          credit_low(G, y, ch);  // t(y)
        else
          gadd(&G.sup[e], ch);  // the pair's edge y->r
      }
    }
    row_hits += h;
  });
  acc += row_hits;
  if (MODE == kModeLCC) {
    const uint32_t wh = wave_sum(row_hits);
    if (lane == 0 && wh) gadd(&G.tvert[r], (unsigned long long)wh);
  }
}

// A row one wave can do alone: its pairs fit one batch and its hash table one wave's
// share (1/WAVES) of the table space.
template <int CAP, int MODE, int WAVES>
__device__ __forceinline__ bool solo_row(const GDev &G, uint32_t r) {
  const uint32_t nA = G.ooff[r + 1] - G.ooff[r], np = G.poff[r + 1] - G.poff[r];
  uint32_t lg = 4;
  while ((1u << lg) < nA) ++lg;
  return np <= 64u && (2u << lg) <= Smem<CAP, MODE, WAVES>::kSlots / WAVES;
}

// One row (or a pair range of one) with the given strategy.
template <int MODE, int WAVES, int CAP, int ESF = 0>
__device__ __forceinline__ void do_row(const GDev &G, uint32_t r, uint32_t c_lo,
                                       uint32_t c_hi, int st, Smem<CAP, MODE, WAVES> &sm,
                                       unsigned long long &acc) {
  const uint32_t nA = G.ooff[r + 1] - G.ooff[r];
  if (nA > (uint32_t)CAP) {
    row_hbm<MODE, WAVES>(G, r, c_lo, c_hi, acc);
  } else if (st == kBitmap) {
    // passes of at most kPassPairs pairs keep the 16-bit slot counters from carrying
    const uint32_t np = G.poff[r + 1] - G.poff[r];
    const uint32_t hi = c_hi < np ? c_hi : np;
    for (uint32_t lo = c_lo; lo < hi;) {
      const uint32_t e = (MODE == kModeTC || hi - lo <= kPassPairs) ? hi : lo + kPassPairs;
      if (!row_hash<MODE, WAVES, CAP, false, ESF>(G, r, lo, e, sm, acc))
        row_sorted<kGallop, MODE, WAVES, CAP>(G, r, lo, e, sm, acc);
      lo = e;
    }
  } else if (st == kGallop) {
    row_sorted<kGallop, MODE, WAVES, CAP>(G, r, c_lo, c_hi, sm, acc);
  } else {
    row_sorted<kMerge, MODE, WAVES, CAP>(G, r, c_lo, c_hi, sm, acc);
  }
}

template <int WAVES>
__device__ __forceinline__ void flush_total(PAC_GLOBAL unsigned long long *total,
                                            unsigned long long acc) {
  const unsigned long long s = wave_sum64(acc);
  if ((threadIdx.x & 63u) == 0 && s) gadd(total, s);
}

// ---- plain kernel: one workgroup per row, one launch per partition group -----------
template <int STRAT, int MODE, int WAVES, int CAP>
__global__ __launch_bounds__(64 * WAVES) void k_plain(GraphDev Gd, uint32_t row_begin) {
  __shared__ Smem<CAP, MODE, WAVES> sm;
  const GDev G = to_dev(Gd);
  const uint32_t r = G.rows[row_begin + blockIdx.x];
  unsigned long long acc = 0;
  do_row<MODE, WAVES, CAP>(G, r, 0, kAllC, STRAT, sm, acc);
  flush_total<WAVES>(G.total, acc);
}

// ---- persistent kernel: every workgroup pulls work items from one atomic queue --------
// With DEFER, a row whose out-list exceeds this tile's capacity is not looked up in HBM
// but queued (queue[1] counts them, at most defer_cap) for k_deferred, which runs with the
// widest tile; should the queue be full the row is done here after all.  Launches whose
// rows all fit use the variant without it.
// Waves per SIMD that the workgroup's LDS allows, at most 8 in TC mode and 6 otherwise:
// the launch bound asks the compiler to keep the registers within it too.  TC fits 8 waves
// in 64 VGPRs without spilling (R-MAT scale 24: 30.3 ms against 33.9 ms at the 7 waves of
// its natural 66 VGPRs; 9 waves are not reached); LCC at 5 waves (no spills) took 46.0 ms
// against 42.2 ms at 6, and the SUP-mode variants would otherwise take 85 VGPRs.
template <int CAP, int MODE, int WAVES>
constexpr int lds_waves() {
  return std::max(1, std::min(MODE == kModeTC ? 8 : 6,
                              (int)(163840 / sizeof(Smem<CAP, MODE, WAVES>)) * WAVES / 4));
}

// With stamps (profiling only; null otherwise) the workgroup records, per item it, the
// real-time clock when it took the item and when it asked for the next one:
// stamps[2 it], stamps[2 it + 1] (rows it deferred are not included).
template <int MODE, int WAVES, int CAP, bool DEFER, int ESF>
__global__ __launch_bounds__(64 * WAVES, (lds_waves<CAP, MODE, WAVES>()))
void k_persistent(const GraphDev *__restrict__ Gs, const WorkItem *__restrict__ items,
                  uint32_t nitems, const uint8_t *__restrict__ strat, uint32_t *queue,
                  DeferItem *__restrict__ defer, uint32_t defer_cap,
                  unsigned long long *__restrict__ stamps) {
  __shared__ Smem<CAP, MODE, WAVES> s;
  unsigned long long acc = 0;
  PAC_GLOBAL unsigned long long *cur_total = nullptr;
  uint32_t prev = 0xFFFFFFFFu;       // thread 0, with stamps: the item being timed
  unsigned long long t_start = 0;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) {
      if (stamps && prev != 0xFFFFFFFFu) {
        stamps[2ull * prev] = t_start;
        stamps[2ull * prev + 1] = wall_clock64();
      }
      s.item = atomicAdd(queue, 1u);
    }
    __syncthreads();
    const uint32_t it = s.item;
    if (it >= nitems) break;
    if (stamps && threadIdx.x == 0) {
      prev = it;
      t_start = wall_clock64();
    }
    const WorkItem w = items[it];
    const GDev G = to_dev(Gs[w.g]);
    if (G.total != cur_total) {
      if (cur_total) flush_total<WAVES>(cur_total, acc);
      acc = 0;
      cur_total = G.total;
    }
    const uint8_t st = strat[G.part_base + w.part];
    const bool single = (w.hi == w.lo + 1);
    const uint32_t c_lo = single ? w.c_lo : 0u, c_hi = single ? w.c_hi : kAllC;
    // Hash strategy: the item's rows that one wave can do alone go first, one per wave,
    // with no workgroup barrier between them (a small row is a chain of dependent loads
    // that a whole workgroup would wait out together); a row whose table cannot be built
    // falls back to the HBM lookups.  Then the workgroup does the others one by one.
    const bool solo = !single && st == kBitmap;
    if (solo) {
      for (uint32_t q = w.lo + (threadIdx.x >> 6); q < w.hi; q += WAVES) {
        const uint32_t r = G.rows[q];
        if (solo_row<CAP, MODE, WAVES>(G, r) &&
            !row_hash<MODE, WAVES, CAP, true, ESF>(G, r, 0u, kAllC, s, acc))
          row_hbm<MODE, WAVES, true>(G, r, 0u, kAllC, acc);
      }
      __syncthreads();
    }
    for (uint32_t q = w.lo; q < w.hi; ++q) {
      const uint32_t r = G.rows[q];
      if (solo && solo_row<CAP, MODE, WAVES>(G, r)) continue;  // uniform: done above
      if (DEFER && G.ooff[r + 1] - G.ooff[r] > (uint32_t)CAP) {  // uniform
        __syncthreads();  // every thread has read the previous defer_slot
        if (threadIdx.x == 0) {
          const uint32_t k = atomicAdd(&queue[1], 1u);
          if (k < defer_cap) defer[k] = DeferItem{w.g, r, c_lo, c_hi, (uint32_t)st, 0u};
          s.defer_slot = k;
        }
        __syncthreads();
        if (s.defer_slot < defer_cap) continue;
      }
      do_row<MODE, WAVES, CAP, ESF>(G, r, c_lo, c_hi, st, s, acc);
    }
  }
  if (cur_total) flush_total<WAVES>(cur_total, acc);
}

// The rows queued by k_persistent (queue[1] of them, at most defer_cap), pulled through
// queue[2] by workgroups of the widest tile.
template <int MODE, int WAVES, int CAP, int ESF>
__global__ __launch_bounds__(64 * WAVES) void k_deferred(const GraphDev *__restrict__ Gs,
                                                         const DeferItem *__restrict__ defer,
                                                         uint32_t defer_cap, uint32_t *queue) {
  __shared__ Smem<CAP, MODE, WAVES> s;
  const uint32_t n = min(queue[1], defer_cap);  // written by the previous launch
  unsigned long long acc = 0;
  PAC_GLOBAL unsigned long long *cur_total = nullptr;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) s.item = atomicAdd(&queue[2], 1u);
    __syncthreads();
    const uint32_t it = s.item;
    if (it >= n) break;
    const DeferItem d = defer[it];
    const GDev G = to_dev(Gs[d.g]);
    if (G.total != cur_total) {
      if (cur_total) flush_total<WAVES>(cur_total, acc);
      acc = 0;
      cur_total = G.total;
    }
    do_row<MODE, WAVES, CAP, ESF>(G, d.r, d.c_lo, d.c_hi, (int)d.st, s, acc);
  }
  if (cur_total) flush_total<WAVES>(cur_total, acc);
}

__global__ void k_reset_queue(uint32_t *q) {
  if (threadIdx.x < 4) q[threadIdx.x] = 0;
}

// Writes a graph descriptor to device memory (a kernel rather than a host copy, so that
// a no-allocation solve makes no host-to-device transfer).
__global__ void k_put_desc(GraphDev d, GraphDev *out) {
  if (threadIdx.x == 0) *out = d;
}

// Streaming copy, the bandwidth reference: each workgroup copies one contiguous chunk of
// kCopyChunk 16-byte vectors, four vectors per thread in flight, with non-temporal loads
// and stores (the data is touched once).  Of the variants measured on MI355X over 1 GiB
// (grid-stride with 1/4/8 vectors per thread, chunked with 4096/16384/65536 vectors per
// workgroup, with and without the non-temporal hint, and hipMemcpy device to device:
// 4.2-5.7 TB/s), this one was the fastest.
typedef unsigned int v4u32 __attribute__((ext_vector_type(4)));
constexpr uint64_t kCopyChunk = 4096;
__global__ void __launch_bounds__(256)
k_stream_copy(const v4u32 *__restrict__ src, v4u32 *__restrict__ dst, uint64_t n) {
  const uint64_t b0 = blockIdx.x * kCopyChunk;
  const uint64_t b1 = b0 + kCopyChunk < n ? b0 + kCopyChunk : n;
  for (uint64_t i = b0 + threadIdx.x; i < b1; i += 4u * blockDim.x) {
    v4u32 v[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint64_t j = i + u * blockDim.x;
      if (j < b1) v[u] = __builtin_nontemporal_load(&src[j]);
    }
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint64_t j = i + u * blockDim.x;
      if (j < b1) __builtin_nontemporal_store(v[u], &dst[j]);
    }
  }
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// ---- epilogues -----------------------------------------------------------------------
// LCC: the lowest-vertex credits gathered in 32 bits during the launch, into the counts.
__global__ void k_fold_low(const uint32_t *__restrict__ tlow, unsigned long long *tvert,
                           uint32_t V) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < V;
       u += (uint64_t)gridDim.x * blockDim.x)
    tvert[u] += tlow[u];
}

__global__ void k_lcc(const unsigned long long *__restrict__ tvert,
                      const uint32_t *__restrict__ off, uint32_t V, float *lcc32,
                      double *lcc64) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < V;
       u += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t d = off[u + 1] - off[u];
    double r = 0.0;
    if (d >= 1) r = (double)(2ull * tvert[u]) / (double)(d * (d - 1));
    if (lcc32) lcc32[u] = (float)r;
    if (lcc64) lcc64[u] = r;
  }
}

// Per-edge Jaccard (pair formula of Graph.cpp:808-809 on every oriented edge).
// The support of edge e (in sup): its credits as an edge of the staged list and as a
// pair's edge, added during the launch, and as an edge of the streamed lists, added from
// the hit words after it.  The union is dsum[e] - sup[e], dsum[e] = d(t) + d(h) of the
// edge t->h (built once per graph, k_edge_head_deg + k_edge_add_tail_deg).
// k_jaccard_flat: thread per edge, when the edges' tails are not asked for.
__global__ void k_jaccard_flat(const uint32_t *__restrict__ oadj, const uint32_t *__restrict__ sup,
                               const uint32_t *__restrict__ dsum, uint64_t m, int32_t *dst,
                               uint32_t *inter, uint32_t *uni, float *ratio) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t c = sup[e], un = dsum[e] - c;
    if (dst) dst[e] = (int32_t)oadj[e];
    if (inter) inter[e] = c;
    if (uni) uni[e] = un;
    if (ratio) ratio[e] = (float)((double)c / (double)un);
  }
}

// dsum[e] = d(y) + d(oadj[e]) for the out-edges e of every vertex y: the heads' degrees
// gathered in edge order, then each tail's degree added along its out-edges -- each wave
// takes 64 consecutive vertices (lane i: vertex v0 + i, its first out-edge and degree) and
// walks their out-edges 64 at a time, an edge's tail found by a 6-step search across the
// lanes' first edges.
__global__ void k_edge_head_deg(const uint32_t *__restrict__ oadj, const uint32_t *__restrict__ deg,
                                uint64_t m, uint32_t *__restrict__ dsum) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x)
    dsum[e] = deg[oadj[e]];
}
__global__ void __launch_bounds__(256)
k_edge_add_tail_deg(const uint32_t *__restrict__ ooff, const uint32_t *__restrict__ deg,
                    uint32_t V, uint32_t *__restrict__ dsum) {
  const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6;
  for (uint64_t v0 = ((uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6)) * 64u; v0 < V;
       v0 += (uint64_t)gridDim.x * wpb * 64u) {
    const uint64_t vi = v0 + lane;
    const uint32_t a = ooff[vi < V ? vi : V];  // first out-edge (the end past the last vertex)
    const uint32_t du = vi < V ? deg[vi] : 0u;
    const uint32_t E0 = readlane(a, 0);
    const uint32_t E1 = ooff[v0 + 64u < V ? v0 + 64u : V];
    for (uint32_t e = E0 + lane; e - lane < E1; e += 64) {  // wave-uniform trip count
      uint32_t t = 0;  // the last lane whose vertex starts at or before e
#pragma unroll
      for (uint32_t step = 32; step; step >>= 1) {
        const uint32_t at = (uint32_t)__shfl((int)a, (int)(t + step), 64);
        t = at <= e ? t + step : t;
      }
      const uint32_t dt = (uint32_t)__shfl((int)du, (int)t, 64);
      if (e < E1) dsum[e] += dt;
    }
  }
}

// k_jaccard: with the tails (src) as well.  Each wave takes 64 consecutive vertices and
// walks their out-edges (contiguous in the oriented CSR) 64 at a time; an edge's tail is
// found by a binary search over the vertices' first edges held one per lane.
__global__ void __launch_bounds__(256)
k_jaccard(const uint32_t *__restrict__ ooff, const uint32_t *__restrict__ oadj,
          const uint32_t *__restrict__ sup, const uint32_t *__restrict__ dsum, uint32_t V,
          int32_t *src, int32_t *dst, uint32_t *inter, uint32_t *uni, float *ratio) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wpb = blockDim.x >> 6;
  for (uint64_t v0 = ((uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6)) * 64u; v0 < V;
       v0 += (uint64_t)gridDim.x * wpb * 64u) {
    const uint64_t vi = v0 + lane;
    const uint32_t a = ooff[vi < V ? vi : V];  // first out-edge (the end past the last vertex)
    const uint32_t E0 = readlane(a, 0);
    const uint32_t E1 = ooff[v0 + 64u < V ? v0 + 64u : V];
    // Four chunks of 64 edges per trip, each step over all four before the next, so that
    // a wave has four loads of each kind in flight instead of one.
    for (uint32_t eb = E0; eb < E1; eb += 256) {  // wave-uniform
      uint32_t e[4], t[4], v[4], c[4], ds[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        e[u] = eb + 64u * u + lane;
        t[u] = 0;  // last lane whose vertex starts at or before e
      }
#pragma unroll
      for (uint32_t step = 32; step; step >>= 1) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t at = (uint32_t)__shfl((int)a, (int)(t[u] + step), 64);
          t[u] = at <= e[u] ? t[u] + step : t[u];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        v[u] = c[u] = ds[u] = 0;
        if (e[u] < E1) {
          if (dst) v[u] = oadj[e[u]];
          c[u] = sup[e[u]];
          ds[u] = dsum[e[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (e[u] < E1) {
          const uint32_t un = ds[u] - c[u];
          if (src) src[e[u]] = (int32_t)(v0 + t[u]);
          if (dst) dst[e[u]] = (int32_t)v[u];
          if (inter) inter[e[u]] = c[u];
          if (uni) uni[e[u]] = un;
          if (ratio) ratio[e[u]] = (float)((double)c[u] / (double)un);
        }
      }
    }
  }
}

// Chunk merge of two sorted lists by one wave (Graph::cnt_common_neis_merge,
// Graph.cpp:227-239, 64 elements a side at a time): lane k holds A[ia + k] and B[ib + k]
// (0xFFFFFFFF past a list's end, above every ID); each lane looks its B element up in the
// A chunk by a 6-step search across the lanes, and the chunk whose last element is smaller
// advances (both when they are equal), so each common element meets its match exactly
// once.  Returns the lane's share of the hits that pass the filters of k_pairs.
__device__ __forceinline__ uint32_t chunk_merge_hits(const uint32_t *A, uint32_t na,
                                                     const uint32_t *B, uint32_t nb,
                                                     uint32_t lane, int32_t lim,
                                                     const uint32_t *part, uint32_t pid) {
  uint32_t ia = 0, ib = 0, h = 0;
  uint32_t av = lane < na ? A[lane] : 0xFFFFFFFFu;
  uint32_t bv = lane < nb ? B[lane] : 0xFFFFFFFFu;
  while (ia < na && ib < nb) {  // wave-uniform
    const uint32_t alast = min(64u, na - ia) - 1u, blast = min(64u, nb - ib) - 1u;
    const uint32_t amax = __shfl(av, (int)alast), bmax = __shfl(bv, (int)blast);
    uint32_t pos = 0;  // lower bound of bv among the A chunk's lanes (at most 63)
#pragma unroll
    for (uint32_t st = 32; st; st >>= 1)
      if (__shfl(av, (int)(pos + st - 1u)) < bv) pos += st;
    // every lane takes part in the cross-lane read (under a short-circuit it would run
    // with the lanes past blast switched off, and a read from those returns no data)
    const uint32_t at = __shfl(av, (int)pos);
    bool hit = lane <= blast && at == bv && (int32_t)bv < lim;
    if (part && hit) hit = part[bv] == pid;
    h += hit ? 1u : 0u;
    if (amax <= bmax) {
      ia += 64u;
      av = ia + lane < na ? A[ia + lane] : 0xFFFFFFFFu;
    }
    if (bmax <= amax) {
      ib += 64u;
      bv = ib + lane < nb ? B[ib + lane] : 0xFFFFFFFFu;
    }
  }
  return h;
}

// |N(u) & N(v)| over the full symmetric lists, one wave per pair, by whichever of two
// strategies takes fewer dependent load steps: the chunk merge above, about (na + nb) / 64
// steps, or each of the shorter list's elements binary-searched in the longer one
// (Graph::cnt_common_neis_galloping, Graph.cpp:186-198), ceil(na / 64) x log2(nb) steps;
// the merge is taken when 1.5x its steps is below the search's.  Optional Jaccard ratio
// c / ((d_u + d_v) - c) in double (Graph.cpp:808-809).
// A common neighbour w counts when w < thr[p] (thr null: no bound) and, when part is
// not null, part[w] == pid.
__global__ void k_pairs(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
                        const int32_t *__restrict__ pu, const int32_t *__restrict__ pv,
                        uint64_t n, unsigned long long *res, double *jac, const int32_t *q,
                        uint32_t nq, const int32_t *__restrict__ thr,
                        const uint32_t *__restrict__ part, uint32_t pid) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wpb = blockDim.x >> 6;
  // Each wave takes 64 consecutive pairs, one per lane.  A lane whose two lists both have
  // at most kLanePair entries merges them itself (most pairs of random vertices); the
  // others are then done one at a time by the whole wave.
  constexpr uint32_t kLanePair = 16;
  for (uint64_t p0 = ((uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6)) * 64u; p0 < n;
       p0 += (uint64_t)gridDim.x * wpb * 64u) {
    const uint64_t p = p0 + lane;
    const bool valid = p < n;
    uint32_t a0 = 0, na = 0, b0 = 0, nb = 0;
    int32_t lim = 0x7FFFFFFF;
    if (valid) {
      uint32_t u, v;
      if (q) {
        u = (uint32_t)q[p / nq];
        v = (uint32_t)q[p % nq];
      } else {
        u = (uint32_t)pu[p];
        v = (uint32_t)pv[p];
      }
      a0 = off[u];
      na = off[u + 1] - a0;
      b0 = off[v];
      nb = off[v + 1] - b0;
      if (thr) lim = thr[p];
    }
    const uint32_t du = na, dv = nb;
    if (na > nb) {
      uint32_t t = a0; a0 = b0; b0 = t;
      t = na; na = nb; nb = t;
    }
    const bool alone = valid && nb <= kLanePair;
    uint32_t h = 0;
    if (alone) {  // two-pointer merge of at most 2 x kLanePair entries
      uint32_t i = 0, j = 0;
      while (i < na && j < nb) {
        const uint32_t x = adj[a0 + i], y = adj[b0 + j];
        if (x == y) {
          bool hit = (int32_t)x < lim;
          if (part && hit) hit = part[x] == pid;
          h += hit ? 1u : 0u;
        }
        i += x <= y ? 1u : 0u;
        j += y <= x ? 1u : 0u;
      }
    }
    unsigned long long rest = __ballot(valid && !alone);
    while (rest) {  // wave-uniform
      const uint32_t k = (uint32_t)__builtin_ctzll(rest);
      rest &= rest - 1ull;
      const uint32_t A0 = readlane(a0, k), NA = readlane(na, k);
      const uint32_t B0 = readlane(b0, k), NB = readlane(nb, k);
      const int32_t L = (int32_t)readlane((uint32_t)lim, k);
      const uint32_t *B = adj + B0;
      uint32_t hk = 0;
      const uint64_t merge_steps = ((uint64_t)NA + NB + 63u) >> 6;
      const uint64_t search_steps = (((uint64_t)NA + 63u) >> 6) * (32u - __clz(NB));
      if (3u * merge_steps < 2u * search_steps) {  // wave-uniform
        hk = chunk_merge_hits(adj + A0, NA, B, NB, lane, L, part, pid);
      } else {
        for (uint32_t i = lane; i < NA; i += 64) {
          const uint32_t x = adj[A0 + i];
          const uint32_t j = lower_bound_u32(B, NB, x);
          bool hit = j < NB && B[j] == x && (int32_t)x < L;
          if (part && hit) hit = part[x] == pid;
          hk += hit ? 1u : 0u;
        }
      }
      hk = wave_sum(hk);
      h = lane == k ? hk : h;
    }
    if (valid) {
      if (res) res[p] = h;
      if (jac) {
        const double c = (double)h;
        jac[p] = c / ((double)((uint64_t)du + dv) - c);
      }
    }
  }
}

template <class K>
uint32_t persistent_grid(Ctx *ctx, K kernel, int threads, uint32_t cap) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(kernel),
                                                   threads, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  uint32_t grid = (uint32_t)(per_cu * ctx->num_cus);
  if (grid > cap) grid = cap ? cap : 1;
  return grid;
}

// The widest tile: the second pass over rows that exceed the launch tile's capacity.
constexpr int kWideWaves = 8, kWideCap = 4096;

// Device buffers of one persistent launch: graph descriptors, per-partition strategy
// table, work-queue heads ([0] items, [1] deferred rows, [2] their queue) and the
// deferred-row queue (defer_cap entries).
struct PersistBufs {
  const GraphDev *gdesc;
  const uint8_t *strat;
  uint32_t *queue;
  DeferItem *defer;
  uint32_t defer_cap;
  unsigned long long *stamps = nullptr;  // per-item clock stamps (profiling)
};

// The persistent launch (and the wide second pass when rows were deferred).
template <int MODE, int WAVES, int CAP, int ES>
int run_persistent_cfg(Ctx *ctx, const WorkItem *items, uint32_t nitems, const PersistBufs &b,
                       uint64_t *launches) {
  constexpr bool kCanDefer = CAP < kWideCap;
  const uint32_t defer_cap = kCanDefer ? b.defer_cap : 0u;
  hipLaunchKernelGGL((k_reset_queue), dim3(1), dim3(64), 0, ctx->stream, b.queue);
  if (ctx->time_kernel) PAC_CHECK_HIP(hipEventRecord(ctx->kev[0], ctx->stream));
  if (defer_cap) {
    const uint32_t grid =
        persistent_grid(ctx, k_persistent<MODE, WAVES, CAP, kCanDefer, ES>, 64 * WAVES, nitems);
    hipLaunchKernelGGL((k_persistent<MODE, WAVES, CAP, kCanDefer, ES>), dim3(grid), dim3(64 * WAVES),
                       0, ctx->stream, b.gdesc, items, nitems, b.strat, b.queue, b.defer,
                       defer_cap, b.stamps);
  } else {
    const uint32_t grid =
        persistent_grid(ctx, k_persistent<MODE, WAVES, CAP, false, ES>, 64 * WAVES, nitems);
    hipLaunchKernelGGL((k_persistent<MODE, WAVES, CAP, false, ES>), dim3(grid), dim3(64 * WAVES), 0,
                       ctx->stream, b.gdesc, items, nitems, b.strat, b.queue, nullptr, 0u,
                       b.stamps);
  }
  PAC_CHECK_HIP(hipGetLastError());
  if (defer_cap) {
    const uint32_t g2 = persistent_grid(ctx, k_deferred<MODE, kWideWaves, kWideCap, ES>,
                                        64 * kWideWaves, defer_cap);
    hipLaunchKernelGGL((k_deferred<MODE, kWideWaves, kWideCap, ES>), dim3(g2), dim3(64 * kWideWaves),
                       0, ctx->stream, b.gdesc, b.defer, defer_cap, b.queue);
    PAC_CHECK_HIP(hipGetLastError());
    if (launches) *launches += 1;
  }
  if (ctx->time_kernel) PAC_CHECK_HIP(hipEventRecord(ctx->kev[1], ctx->stream));
  if (launches) *launches += 2;
  return PAC_OK;
}

template <int MODE, int ES>
int run_persistent_mode(Ctx *ctx, const WorkItem *items, uint32_t nitems, const PersistBufs &b,
                        uint64_t *launches) {
  switch (ctx->tile) {
    case 0: return run_persistent_cfg<MODE, 2, 1024, ES>(ctx, items, nitems, b, launches);
    case 1: return run_persistent_cfg<MODE, 4, 2048, ES>(ctx, items, nitems, b, launches);
    case 2: return run_persistent_cfg<MODE, 4, 4096, ES>(ctx, items, nitems, b, launches);
    default: return run_persistent_cfg<MODE, 8, 4096, ES>(ctx, items, nitems, b, launches);
  }
}

// es: the element width of the lists the launch's descriptors stream (launch_elem_bytes).
int run_persistent_any(Ctx *ctx, int mode, uint32_t es, const WorkItem *items, uint32_t nitems,
                       const PersistBufs &b, uint64_t *launches) {
  if (mode == kModeSUP) return run_persistent_mode<kModeSUP, 4>(ctx, items, nitems, b, launches);
  if (mode == kModeLCC)
    return es == 3u ? run_persistent_mode<kModeLCC, 3>(ctx, items, nitems, b, launches)
                    : run_persistent_mode<kModeLCC, 4>(ctx, items, nitems, b, launches);
  return es == 3u ? run_persistent_mode<kModeTC, 3>(ctx, items, nitems, b, launches)
                  : run_persistent_mode<kModeTC, 4>(ctx, items, nitems, b, launches);
}

// The context's own buffers for a persistent launch (the deferred-row queue is reserved
// here for long_rows entries when the tile can defer).
int ctx_bufs(Ctx *ctx, uint32_t long_rows, PersistBufs &b) {
  const bool can_defer = kTiles[ctx->tile].cap < kWideCap;
  const uint32_t defer_cap = can_defer ? long_rows : 0u;
  if (defer_cap) PAC_TRY(ctx->defer.reserve(sizeof(DeferItem) * defer_cap));
  b.gdesc = ctx->gdesc.as<const GraphDev>();
  b.strat = ctx->dispatch.as<const uint8_t>();
  b.queue = ctx->queue.as<uint32_t>();
  b.defer = ctx->defer.as<DeferItem>();
  b.defer_cap = defer_cap;
  b.stamps = ctx->item_stamps;
  return PAC_OK;
}

template <int STRAT, int MODE, int WAVES, int CAP>
int run_plain_cfg(Ctx *ctx, const GraphDev &G, uint32_t row_begin, uint32_t nrows) {
  hipLaunchKernelGGL((k_plain<STRAT, MODE, WAVES, CAP>), dim3(nrows), dim3(64 * WAVES), 0,
                     ctx->stream, G, row_begin);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

template <int STRAT, int MODE>
int run_plain_tile(Ctx *ctx, const GraphDev &G, uint32_t row_begin, uint32_t nrows) {
  switch (ctx->tile) {
    case 0: return run_plain_cfg<STRAT, MODE, 2, 1024>(ctx, G, row_begin, nrows);
    case 1: return run_plain_cfg<STRAT, MODE, 4, 2048>(ctx, G, row_begin, nrows);
    case 2: return run_plain_cfg<STRAT, MODE, 4, 4096>(ctx, G, row_begin, nrows);
    default: return run_plain_cfg<STRAT, MODE, 8, 4096>(ctx, G, row_begin, nrows);
  }
}

template <int MODE>
int run_plain(Ctx *ctx, const GraphDev &G, int strat, uint32_t row_begin, uint32_t nrows) {
  if (strat == kBitmap) return run_plain_tile<kBitmap, MODE>(ctx, G, row_begin, nrows);
  if (strat == kGallop) return run_plain_tile<kGallop, MODE>(ctx, G, row_begin, nrows);
  return run_plain_tile<kMerge, MODE>(ctx, G, row_begin, nrows);
}

int run_plain_any(Ctx *ctx, const GraphDev &G, int mode, int strat, uint32_t row_begin,
                  uint32_t nrows) {
  if (nrows == 0) return PAC_OK;
  if (mode == kModeLCC) return run_plain<kModeLCC>(ctx, G, strat, row_begin, nrows);
  if (mode == kModeSUP) return run_plain<kModeSUP>(ctx, G, strat, row_begin, nrows);
  return run_plain<kModeTC>(ctx, G, strat, row_begin, nrows);
}

// ---- SUP mode: per-edge credits of the streamed lists from the hit words ----------
// 64 x 64 bit transpose across the wave: lane L holds row L as (lo, hi) = bits 0..31,
// 32..63; afterwards lane j holds column j (its bit L = bit j of row L).  At each scale s
// (32, 16, ..., 1) the two off-diagonal s x s blocks of every 2s x 2s block are swapped:
// a lane in the upper half of its 2s-group (lane & s == 0) replaces its right bits
// (j & s != 0) with its partner's left bits moved up by s, and the partner the reverse --
// each lane sends the bits it replaces (x & mask), rotates what it receives by s into
// place, and merges it under the same mask.
__device__ __forceinline__ void transpose64(uint32_t &lo, uint32_t &hi, uint32_t lane) {
  {
    const bool top = (lane & 32u) == 0;
    const uint32_t recv = (uint32_t)__shfl_xor((int)(top ? hi : lo), 32, 64);
    if (top)
      hi = recv;
    else
      lo = recv;
  }
  constexpr uint32_t kLeft[5] = {0x0000FFFFu, 0x00FF00FFu, 0x0F0F0F0Fu, 0x33333333u,
                                 0x55555555u};
#pragma unroll
  for (uint32_t k = 0; k < 5; ++k) {
    const uint32_t s = 16u >> k;
    const bool top = (lane & s) == 0;
    const uint32_t mask = top ? ~kLeft[k] : kLeft[k];  // the bits this lane replaces
    const uint32_t rot = top ? 32u - s : s;            // right-rotate amount into place
    const uint32_t rl = (uint32_t)__shfl_xor((int)(lo & mask), (int)s, 64);
    const uint32_t rh = (uint32_t)__shfl_xor((int)(hi & mask), (int)s, 64);
    lo = (lo & ~mask) | (__builtin_amdgcn_alignbit(rl, rl, rot) & mask);
    hi = (hi & ~mask) | (__builtin_amdgcn_alignbit(rh, rh, rot) & mask);
  }
}

// Vertices streamed by at most 64 pairs (one load per word column): each wave loads the
// metadata of 64 consecutive vertices (lane i: vertex v0 + i) and walks their (vertex,
// column) units in turn, the next unit's loads issued before the current transpose.  It
// is the only writer of those vertices' out-edge supports here, so plain read-modify-writes.
__global__ void __launch_bounds__(256)
k_sup_small(const uint32_t *__restrict__ ooff, const uint32_t *__restrict__ pstreams,
            const unsigned long long *__restrict__ wbase,
            const unsigned long long *__restrict__ hitw, uint32_t V, uint32_t *__restrict__ sup) {
  const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6;
  const PAC_GLOBAL unsigned long long *W = (const PAC_GLOBAL unsigned long long *)hitw;
  for (uint64_t v0 = ((uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6)) * 64u; v0 < V;
       v0 += (uint64_t)gridDim.x * wpb * 64u) {
    const uint64_t vi = v0 + lane;
    uint32_t Rl = 0, a0l = 0, ndl = 0;
    unsigned long long wbl = 0;
    if (vi < V) {
      Rl = pstreams[vi];
      if (Rl > 64u) Rl = 0;  // summed by k_sup_big
      if (Rl) {
        a0l = ooff[vi];
        ndl = ooff[vi + 1] - a0l;
        wbl = wbase[vi];
      }
    }
    unsigned long long todo = __ballot(Rl != 0);
    if (todo) {
      uint32_t i = (uint32_t)__builtin_ctzll(todo), c = 0;
      todo &= todo - 1ull;
      uint32_t R = readlane(Rl, i), a0 = readlane(a0l, i), nd = readlane(ndl, i);
      uint32_t nw = (nd + 63u) >> 6;
      unsigned long long base = readlane64(wbl, i);
      unsigned long long w = lane < R ? W[base + (unsigned long long)lane * nw] : 0ull;
      uint32_t old = lane < nd ? sup[a0 + lane] : 0u;  // the support being added to
      while (true) {  // wave-uniform
        uint32_t R2 = R, a02 = a0, nd2 = nd, nw2 = nw, c2 = c + 1;
        unsigned long long base2 = base;
        bool more = true;
        if (c2 >= nw) {
          more = todo != 0ull;
          if (more) {
            const uint32_t i2 = (uint32_t)__builtin_ctzll(todo);
            todo &= todo - 1ull;
            R2 = readlane(Rl, i2);
            a02 = readlane(a0l, i2);
            nd2 = readlane(ndl, i2);
            nw2 = (nd2 + 63u) >> 6;
            base2 = readlane64(wbl, i2);
            c2 = 0;
          }
        }
        const unsigned long long w2 =
            more && lane < R2 ? W[base2 + (unsigned long long)lane * nw2 + c2] : 0ull;
        const uint32_t e2 = 64u * c2 + lane;
        const uint32_t old2 = more && e2 < nd2 ? sup[a02 + e2] : 0u;
        uint32_t lo = (uint32_t)w, hi = (uint32_t)(w >> 32);
        transpose64(lo, hi, lane);
        const uint32_t acc = (uint32_t)__popc(lo) + (uint32_t)__popc(hi);
        const uint32_t e = 64u * c + lane;
        if (e < nd && acc) sup[a0 + e] = old + acc;
        if (!more) break;
        old = old2;
        R = R2;
        a0 = a02;
        nd = nd2;
        nw = nw2;
        c = c2;
        base = base2;
        w = w2;
      }
    }
  }
}

// Vertices streamed by more than 64 pairs, one workgroup per unit (y, first stream r0) of
// kSupRun streams; each wave takes bands of kBand word columns, its lanes reading kBand
// consecutive words of one stream each (a row's words share cache lines, so a band costs
// one line request per row instead of kBand); units add atomically.  A band reads only
// the rows whose streams reach it: the rows come in the order of the pairs' positions in
// N+(y), and those with a position below 64 (c + 1) - 1 start at or before column c, so
// they are the first prank[a0 + 64 (c + 1) - 1] - prank[a0] rows (R-MAT scale 25: about
// half of the words).
__global__ void __launch_bounds__(256)
k_sup_big(const uint32_t *__restrict__ units, const uint32_t *__restrict__ ooff,
          const uint32_t *__restrict__ pstreams, const unsigned long long *__restrict__ wbase,
          const uint32_t *__restrict__ prank, const unsigned long long *__restrict__ hitw,
          uint32_t *__restrict__ sup) {
  constexpr uint32_t kBand = 8;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, wpb = blockDim.x >> 6;
  const PAC_GLOBAL unsigned long long *W = (const PAC_GLOBAL unsigned long long *)hitw;
  const uint32_t y = units[2u * blockIdx.x], r0 = units[2u * blockIdx.x + 1u];
  const uint32_t R = pstreams[y];
  const uint32_t a0 = ooff[y], nd = ooff[y + 1] - a0, nw = (nd + 63u) >> 6;
  const unsigned long long base = wbase[y];
  const uint32_t r1 = min(R, r0 + kSupRun), pr0 = prank[a0];
  for (uint32_t c0 = wave * kBand; c0 < nw; c0 += wpb * kBand) {  // wave-uniform
    const uint32_t nc = min(kBand, nw - c0);
    // the rows reaching the band's last column
    const uint32_t rb = min(r1, prank[a0 + min(64u * (c0 + nc) - 1u, nd)] - pr0);
    if (rb <= r0) continue;  // wave-uniform: no row reaches the band
    uint32_t acc[kBand];
#pragma unroll
    for (uint32_t k = 0; k < kBand; ++k) acc[k] = 0;
    for (uint32_t r = r0; r < rb; r += 64) {  // wave-uniform
      const uint32_t row = r + lane;
      const PAC_GLOBAL unsigned long long *src = W + base + (unsigned long long)row * nw + c0;
      unsigned long long w[kBand];
#pragma unroll
      for (uint32_t k = 0; k < kBand; ++k) w[k] = 0ull;
      if (nc == kBand) {  // wave-uniform: whole band, read as 16-byte loads
        if (row < rb) {
#pragma unroll
          for (uint32_t k = 0; k < kBand; k += 2) {
            HIP_vector_type<unsigned long long, 2> t;
            __builtin_memcpy(&t, (const unsigned long long *)(src + k), sizeof(t));
            w[k] = t.x;
            w[k + 1] = t.y;
          }
        }
      } else {
#pragma unroll
        for (uint32_t k = 0; k < kBand; ++k) w[k] = row < rb && k < nc ? src[k] : 0ull;
      }
#pragma unroll
      for (uint32_t k = 0; k < kBand; ++k) {
        if (k < nc) {  // wave-uniform
          uint32_t lo = (uint32_t)w[k], hi = (uint32_t)(w[k] >> 32);
          transpose64(lo, hi, lane);
          acc[k] += (uint32_t)__popc(lo) + (uint32_t)__popc(hi);
        }
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < kBand; ++k) {
      const uint32_t e = 64u * (c0 + k) + lane;
      if (k < nc && e < nd && acc[k]) gadd((PAC_GLOBAL uint32_t *)(sup + a0 + e), acc[k]);
    }
  }
}


// Device descriptor of a graph; its output pointers must already be allocated.
GraphDev make_desc(Graph *g, uint32_t part_base, uint32_t es, int mode) {
  GraphDev d;
  d.ooff = g->ooff.as<const uint32_t>();
  d.oadj = g->oadj.as<const uint32_t>();
  d.okey = g->okey.as<const uint32_t>();
  d.poff = g->poff.as<const uint32_t>();
  d.ppar = g->ppar.as<const uint32_t>();
  d.pedge = g->pedge.as<const uint32_t>();
  // SUP mode streams from 64-element boundaries (its hit words are list columns)
  d.pinfo = mode == kModeSUP ? g->pinfo64.as<const uint2>() : g->pinfo.as<const uint2>();
  // the 3-byte lists where the graph has them and the context streams them
  const bool packed = es == 3u;  // launch_elem_bytes checked that the lists exist
  d.sadj = packed ? g->opk.as<const uint8_t>() : g->oadj.as<const uint8_t>();
  d.sinfo = packed ? g->pkinfo.as<const uint2>() : d.pinfo;
  d.selem = packed ? 3u : 4u;
  d.rows = g->rows.as<const uint32_t>();
  d.total = g->total.as<unsigned long long>();
  d.tvert = g->tvert.as<unsigned long long>();
  d.tlow = mode == kModeLCC ? g->tlow.as<uint32_t>() : nullptr;
  d.sup = g->sup.as<uint32_t>();
  d.pword = g->pword.as<const unsigned long long>();
  d.hitw = g->hitw.as<unsigned long long>();
  d.V = g->V;
  d.m = (uint32_t)g->m;
  d.part_base = part_base;
  return d;
}

uint32_t grid_for(uint64_t n, uint32_t per_block) {
  uint64_t b = (n + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > (1u << 20)) b = 1u << 20;
  return (uint32_t)b;
}

}  // namespace

int stream_copy(Ctx *ctx, uint64_t bytes, int reps, double *ms) {
  if (bytes < 16 || bytes % 16 || reps < 1) {
    set_error("stream copy needs a multiple of 16 bytes and reps >= 1");
    return PAC_ERR_ARG;
  }
  DevBuf a, b;
  PAC_TRY(a.reserve(bytes));
  PAC_TRY(b.reserve(bytes));
  PAC_TRY(fill_u32(ctx->stream, a.as<uint32_t>(), bytes / 4, 1u));
  const uint64_t n = bytes / 16;
  const uint64_t nblk = (n + kCopyChunk - 1) / kCopyChunk;
  if (nblk > 0x7FFFFFFFull) {
    set_error("stream copy of %llu bytes exceeds one launch", (unsigned long long)bytes);
    return PAC_ERR_ARG;
  }
  const uint32_t grid = (uint32_t)nblk;
  std::vector<double> t;
  for (int r = 0; r <= reps; ++r) {
    PAC_CHECK_HIP(hipEventRecord(ctx->kev[0], ctx->stream));
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, ctx->stream,
                       a.as<const v4u32>(), b.as<v4u32>(), n);
    PAC_CHECK_HIP(hipGetLastError());
    PAC_CHECK_HIP(hipEventRecord(ctx->kev[1], ctx->stream));
    PAC_CHECK_HIP(hipEventSynchronize(ctx->kev[1]));
    float x = 0.f;
    PAC_CHECK_HIP(hipEventElapsedTime(&x, ctx->kev[0], ctx->kev[1]));
    if (r) t.push_back(x);  // first launch is a warm-up
  }
  std::sort(t.begin(), t.end());
  *ms = t[t.size() / 2];
  return PAC_OK;
}

int fill_u32(hipStream_t s, uint32_t *p, uint64_t n, uint32_t val) {
  if (n == 0) return PAC_OK;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256 * 8)), dim3(256), 0, s, p, n, val);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

int alloc_outputs(Graph *const *gs, uint32_t G, int mode) {
  for (uint32_t i = 0; i < G; ++i) {
    Graph *g = gs[i];
    PAC_TRY(g->total.reserve(sizeof(unsigned long long)));
    if (mode == kModeLCC) {
      PAC_TRY(g->tvert.reserve(sizeof(unsigned long long) * (size_t)(g->V ? g->V : 1)));
      PAC_TRY(g->tlow.reserve(sizeof(uint32_t) * (size_t)(g->V ? g->V : 1)));
    }
    if (mode == kModeSUP) PAC_TRY(g->sup.reserve(sizeof(uint32_t) * (size_t)(g->m ? g->m : 1)));
  }
  return PAC_OK;
}

// Zero-fill kernels (not memset nodes, so the same sequence can be graph-captured).
int zero_outputs(Ctx *ctx, Graph *const *gs, uint32_t G, int mode) {
  PAC_TRY(alloc_outputs(gs, G, mode));
  for (uint32_t i = 0; i < G; ++i) {
    Graph *g = gs[i];
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream,
                       g->total.as<unsigned long long>(), (uint64_t)1);
    if (mode == kModeLCC && g->V) {
      hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(g->V, 256 * 8)), dim3(256), 0, ctx->stream,
                         g->tvert.as<unsigned long long>(), (uint64_t)g->V);
      PAC_TRY(fill_u32(ctx->stream, g->tlow.as<uint32_t>(), g->V, 0u));
    }
    if (mode == kModeSUP) PAC_TRY(fill_u32(ctx->stream, g->sup.as<uint32_t>(), g->m, 0u));
    PAC_CHECK_HIP(hipGetLastError());
  }
  return PAC_OK;
}

// One launch per partition group: a maximal run of consecutive partitions with the
// same strategy (their rows are contiguous in the row list), one workgroup per row.
int launch_plain(Ctx *ctx, Graph *g, int mode, const uint8_t *strategy_per_part,
                 uint64_t *launches) {
  Graph *gs[1] = {g};
  PAC_TRY(alloc_outputs(gs, 1, mode));
  const GraphDev d = make_desc(g, 0, launch_elem_bytes(ctx, gs, 1, mode), mode);
  uint32_t k = 0;
  while (k < g->P) {
    uint32_t e = k + 1;
    while (e < g->P && strategy_per_part[e] == strategy_per_part[k]) ++e;
    const uint32_t rb = g->h_pstart[k], re = g->h_pstart[e];
    if (re > rb) {
      PAC_TRY(run_plain_any(ctx, d, mode, strategy_per_part[k], rb, re - rb));
      if (launches) ++*launches;
    }
    k = e;
  }
  return PAC_OK;
}

int launch_persistent(Ctx *ctx, Graph *const *gs, uint32_t G, int mode,
                      const uint8_t *strategy_per_part_concat, bool use_graph,
                      uint64_t *launches) {
  // Output buffers must exist before the descriptors that point at them are built
  // (on the hipGraph path nothing has allocated them yet).
  PAC_TRY(alloc_outputs(gs, G, mode));
  std::vector<GraphDev> descs(G);
  const uint32_t es = launch_elem_bytes(ctx, gs, G, mode);
  uint64_t nparts = 0, nitems = 0, long_rows = 0;
  bool any_long = false;
  const uint32_t cap = (uint32_t)kTiles[ctx->tile].cap;
  for (uint32_t i = 0; i < G; ++i) {
    descs[i] = make_desc(gs[i], (uint32_t)nparts, es, mode);
    nparts += gs[i]->P;
    nitems += gs[i]->h_items.size();
    // rows over the tile's capacity, each queued at most once per item that holds it
    long_rows += gs[i]->h_items.size() + gs[i]->n_long_rows;
    any_long = any_long || gs[i]->max_outdeg > cap;
  }
  if (!any_long) long_rows = 0;
  if (long_rows > 0xFFFFFFF0ull) long_rows = 0xFFFFFFF0ull;
  if (nitems > 0xFFFFFFF0ull) {
    set_error("too many work items (%llu)", (unsigned long long)nitems);
    return PAC_ERR_UNSUPPORTED;
  }
  PAC_TRY(ctx->gdesc.reserve(sizeof(GraphDev) * (G ? G : 1)));
  PAC_TRY(ctx->dispatch.reserve(nparts ? nparts : 1));
  PAC_TRY(ctx->queue.reserve(64));
  // Descriptors and the strategy table are plain uploads before the launch (and before
  // any graph replay), so a captured graph always reads the current table.
  PAC_CHECK_HIP(hipMemcpyAsync(ctx->gdesc.p, descs.data(), sizeof(GraphDev) * G,
                               hipMemcpyHostToDevice, ctx->stream));
  PAC_CHECK_HIP(hipMemcpyAsync(ctx->dispatch.p, strategy_per_part_concat, nparts,
                               hipMemcpyHostToDevice, ctx->stream));
  const WorkItem *items = nullptr;
  std::vector<const void *> key;
  for (uint32_t i = 0; i < G; ++i) key.push_back(gs[i]);
  if (G == 1) {
    items = gs[0]->items.as<const WorkItem>();
  } else {
    if (ctx->batch_key != key) {
      PAC_TRY(ctx->batch_items.reserve(sizeof(WorkItem) * (nitems ? nitems : 1)));
      std::vector<WorkItem> all;
      all.reserve(nitems);
      for (uint32_t i = 0; i < G; ++i)
        for (WorkItem w : gs[i]->h_items) {
          w.g = i;
          all.push_back(w);
        }
      PAC_CHECK_HIP(hipMemcpyAsync(ctx->batch_items.p, all.data(), sizeof(WorkItem) * nitems,
                                   hipMemcpyHostToDevice, ctx->stream));
      PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
      ctx->batch_key = key;
    }
    items = ctx->batch_items.as<const WorkItem>();
  }
  if (!use_graph) {
    // the caller has zeroed the outputs
    if (nitems == 0) return PAC_OK;
    PersistBufs b;
    PAC_TRY(ctx_bufs(ctx, (uint32_t)long_rows, b));
    return run_persistent_any(ctx, mode, es, items, (uint32_t)nitems, b, launches);
  }
  // hipGraph: zero-fill kernels + queue reset + persistent kernel, captured once per
  // (graph set, mode, tile, buffer addresses) and replayed.
  key.push_back(items);
  key.push_back(ctx->gdesc.p);
  key.push_back(ctx->dispatch.p);
  key.push_back(ctx->queue.p);
  if (long_rows) PAC_TRY(ctx->defer.reserve(sizeof(DeferItem) * long_rows));
  key.push_back(ctx->defer.p);
  key.push_back(reinterpret_cast<const void *>((uintptr_t)es));  // the kernels' list width
  for (uint32_t i = 0; i < G; ++i) {
    key.push_back(gs[i]->tvert.p);
    key.push_back(gs[i]->tlow.p);
    key.push_back(gs[i]->sup.p);
    key.push_back(gs[i]->total.p);
  }
  if (!(ctx->hexec && ctx->hgraph_key == key && ctx->hgraph_mode == mode &&
        ctx->hgraph_tile == ctx->tile)) {
    if (ctx->hexec) (void)hipGraphExecDestroy(ctx->hexec);
    if (ctx->hgraph) (void)hipGraphDestroy(ctx->hgraph);
    ctx->hexec = nullptr;
    ctx->hgraph = nullptr;
    PAC_CHECK_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const bool timing = ctx->time_kernel;
    ctx->time_kernel = false;  // no event records inside the captured graph
    int rc = zero_outputs(ctx, gs, G, mode);
    PersistBufs b;
    if (rc == PAC_OK) rc = ctx_bufs(ctx, (uint32_t)long_rows, b);  // reserved above: no allocation
    if (rc == PAC_OK && nitems)
      rc = run_persistent_any(ctx, mode, es, items, (uint32_t)nitems, b, nullptr);
    ctx->time_kernel = timing;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != PAC_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    PAC_CHECK_HIP(e);
    ctx->hgraph = graph;
    PAC_CHECK_HIP(hipGraphInstantiate(&ctx->hexec, graph, nullptr, nullptr, 0));
    ctx->hgraph_key = key;
    ctx->hgraph_mode = mode;
    ctx->hgraph_tile = ctx->tile;
  }
  PAC_CHECK_HIP(hipGraphLaunch(ctx->hexec, ctx->stream));
  if (launches) *launches += 1;
  return PAC_OK;
}

namespace {
constexpr size_t kScratchQueue = 0, kScratchDesc = 256, kScratchDefer = 512;
static_assert(4 * sizeof(uint32_t) <= kScratchDesc - kScratchQueue &&
                  sizeof(GraphDev) <= kScratchDefer - kScratchDesc &&
                  kScratchDefer % alignof(DeferItem) == 0,
              "the queue heads and the descriptor fit their slots of the no-allocation scratch");
constexpr uint64_t kNoAllocDeferCap =
    (PAC_NOALLOC_SCRATCH_BYTES - kScratchDefer) / sizeof(DeferItem);
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_scratch_init(Ctx *ctx) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (ctx->device < 0 || ctx->device >= kMaxDevices) {
    set_error("device %d: no-allocation scratch supports devices 0..%d", ctx->device,
              kMaxDevices - 1);
    return PAC_ERR_UNSUPPORTED;
  }
  if (!g_scratch_ptr[ctx->device]) {
    // the caller has made ctx->device current: each GPU holds its own copy of the array
    void *p = nullptr;
    PAC_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("hipGetSymbolAddress returned no address for the no-allocation scratch");
      return PAC_ERR_HIP;
    }
    g_scratch_ptr[ctx->device] = p;
  }
  ctx->noalloc_scratch = g_scratch_ptr[ctx->device];
  return PAC_OK;
}

int launch_persistent_noalloc(Ctx *ctx, Graph *g, int mode, const uint8_t *d_strat,
                              unsigned long long *d_total, unsigned long long *d_tvert) {
  unsigned char *base = static_cast<unsigned char *>(ctx->noalloc_scratch);
  if (!base) {
    set_error("no-allocation scratch is not set up for device %d", ctx->device);
    return PAC_ERR_STATE;
  }
  const uint64_t nitems = g->h_items.size();
  const uint32_t cap = (uint32_t)kTiles[ctx->tile].cap;
  const bool defers = cap < (uint32_t)kWideCap && g->max_outdeg > cap;
  const uint64_t long_rows = defers ? nitems + g->n_long_rows : 0;
  if (long_rows > kNoAllocDeferCap) {
    set_error("the deferred-row queue needs %llu entries; the no-allocation scratch holds %llu "
              "(%u bytes): use pac_tc, or the widest tile",
              (unsigned long long)long_rows, (unsigned long long)kNoAllocDeferCap,
              (unsigned)PAC_NOALLOC_SCRATCH_BYTES);
    return PAC_ERR_UNSUPPORTED;
  }
  // outputs zeroed by kernels (no memset nodes should a caller capture the sequence)
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, d_total, (uint64_t)1);
  if (mode == kModeLCC && g->V)
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(g->V, 256 * 8)), dim3(256), 0, ctx->stream,
                       d_tvert, (uint64_t)g->V);
  PAC_CHECK_HIP(hipGetLastError());
  Graph *gs1[1] = {g};
  const uint32_t es = launch_elem_bytes(ctx, gs1, 1, mode);
  GraphDev d = make_desc(g, 0, es, mode);
  d.total = d_total;
  d.tvert = mode == kModeLCC ? d_tvert : nullptr;
  d.tlow = nullptr;  // every credit straight into the caller's 64-bit counts
  d.sup = nullptr;
  d.pword = nullptr;
  d.hitw = nullptr;
  GraphDev *desc = reinterpret_cast<GraphDev *>(base + kScratchDesc);
  hipLaunchKernelGGL(k_put_desc, dim3(1), dim3(64), 0, ctx->stream, d, desc);
  PAC_CHECK_HIP(hipGetLastError());
  if (nitems == 0) return PAC_OK;
  PersistBufs b;
  b.gdesc = desc;
  b.strat = d_strat;
  b.queue = reinterpret_cast<uint32_t *>(base + kScratchQueue);
  b.defer = reinterpret_cast<DeferItem *>(base + kScratchDefer);
  b.defer_cap = (uint32_t)long_rows;
  return run_persistent_any(ctx, mode, es, g->items.as<const WorkItem>(), (uint32_t)nitems, b,
                            nullptr);
}

int fold_low_credits(Ctx *ctx, Graph *g) {
  if (g->V == 0 || !g->tlow.p) return PAC_OK;
  hipLaunchKernelGGL(k_fold_low, dim3(grid_for(g->V, 256 * 8)), dim3(256), 0, ctx->stream,
                     g->tlow.as<const uint32_t>(), g->tvert.as<unsigned long long>(), g->V);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

int sup_from_hit_words(Ctx *ctx, Graph *g) {
  if (g->V == 0 || !g->hitw_built) return PAC_OK;
  hipLaunchKernelGGL(k_sup_small, dim3(grid_for(g->V, 4 * 64)), dim3(256), 0, ctx->stream,
                     g->ooff.as<const uint32_t>(), g->pstreams.as<const uint32_t>(),
                     g->wbase.as<const unsigned long long>(),
                     g->hitw.as<const unsigned long long>(), g->V, g->sup.as<uint32_t>());
  PAC_CHECK_HIP(hipGetLastError());
  if (!g->h_supbig.empty()) {
    hipLaunchKernelGGL(k_sup_big, dim3((uint32_t)(g->h_supbig.size() / 2)), dim3(256), 0,
                       ctx->stream, g->supbig.as<const uint32_t>(),
                       g->ooff.as<const uint32_t>(), g->pstreams.as<const uint32_t>(),
                       g->wbase.as<const unsigned long long>(), g->prank.as<const uint32_t>(),
                       g->hitw.as<const unsigned long long>(), g->sup.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
  }
  return PAC_OK;
}

int epilogue_lcc(Ctx *ctx, Graph *g, bool want32, bool want64) {
  if (want32) PAC_TRY(g->lcc32.reserve(sizeof(float) * (size_t)(g->V ? g->V : 1)));
  if (want64) PAC_TRY(g->lcc64.reserve(sizeof(double) * (size_t)(g->V ? g->V : 1)));
  if (g->V == 0) return PAC_OK;
  hipLaunchKernelGGL(k_lcc, dim3(grid_for(g->V, 256 * 4)), dim3(256), 0, ctx->stream,
                     g->tvert.as<const unsigned long long>(), g->off.as<const uint32_t>(), g->V,
                     want32 ? g->lcc32.as<float>() : nullptr,
                     want64 ? g->lcc64.as<double>() : nullptr);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

int epilogue_jaccard(Ctx *ctx, Graph *g, int32_t *src, int32_t *dst, uint32_t *inter,
                     uint32_t *uni, float *ratio) {
  if (g->V == 0) return PAC_OK;
  if (!g->dsum.p && g->m) {
    // d(t) + d(h) per oriented edge, once per graph: a gather of deg over oadj that the
    // per-solve epilogue would otherwise repeat (a cache line per 4-byte read)
    PAC_TRY(g->dsum.reserve(sizeof(uint32_t) * g->m));
    hipLaunchKernelGGL(k_edge_head_deg, dim3(grid_for(g->m, 256)), dim3(256), 0, ctx->stream,
                       g->oadj.as<const uint32_t>(), g->deg.as<const uint32_t>(), (uint64_t)g->m,
                       g->dsum.as<uint32_t>());
    hipLaunchKernelGGL(k_edge_add_tail_deg, dim3(grid_for(g->V, 4 * 64)), dim3(256), 0,
                       ctx->stream, g->ooff.as<const uint32_t>(), g->deg.as<const uint32_t>(),
                       g->V, g->dsum.as<uint32_t>());
    PAC_CHECK_HIP(hipGetLastError());
  }
  if (src)  // the tails need the per-vertex walk
    hipLaunchKernelGGL(k_jaccard, dim3(grid_for(g->V, 4 * 64)), dim3(256), 0, ctx->stream,
                       g->ooff.as<const uint32_t>(), g->oadj.as<const uint32_t>(),
                       g->sup.as<const uint32_t>(), g->dsum.as<const uint32_t>(), g->V, src,
                       dst, inter, uni, ratio);
  else if (g->m)
    hipLaunchKernelGGL(k_jaccard_flat, dim3(grid_for(g->m, 256 * 4)), dim3(256), 0, ctx->stream,
                       g->oadj.as<const uint32_t>(), g->sup.as<const uint32_t>(),
                       g->dsum.as<const uint32_t>(), (uint64_t)g->m, dst, inter, uni, ratio);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

int pair_counts(Ctx *ctx, Graph *g, const int32_t *d_u, const int32_t *d_v, uint64_t n,
                uint64_t *d_res, const int32_t *d_thr, const uint32_t *d_part, uint32_t pid) {
  if (n == 0) return PAC_OK;
  hipLaunchKernelGGL(k_pairs, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream,  // 64 pairs a wave
                     g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(), d_u, d_v, n,
                     reinterpret_cast<unsigned long long *>(d_res), nullptr, nullptr, 1u, d_thr,
                     d_part, pid);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

int query_matrix(Ctx *ctx, Graph *g, const int32_t *d_q, uint32_t nq, double *d_res) {
  const uint64_t n = (uint64_t)nq * nq;
  if (n == 0) return PAC_OK;
  hipLaunchKernelGGL(k_pairs, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream,  // 64 pairs a wave
                     g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(), nullptr, nullptr,
                     n, nullptr, d_res, d_q, nq, nullptr, nullptr, 0u);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

}  // namespace pac
