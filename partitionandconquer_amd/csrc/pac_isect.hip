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
// tail t has the smallest (degree, id), as w in N+(t) & N+(h).  Each oriented edge is
// an intersection pair owned by the endpoint whose out-list is longer (ties: the
// tail); pairs whose other list is empty are dropped.  A workgroup owns a row x:
// N+(x) (the list A) is staged once in LDS and each of x's pairs is given to one wave,
// which streams the partner y's out-list N+(y) (the list B, never longer than A) from
// HBM with coalesced loads and looks each element up in A.  Streaming the shorter list
// of every pair moves sum(min(d+(t), d+(h))) elements from HBM rather than
// sum(d+(h)).  The three strategies differ only in the lookup:
//   kMerge  : wave64 merge-path merge of A (LDS) and B (HBM): each lane walks an
//             equal share of the merged sequence after a diagonal binary search.
//   kGallop : the shorter list's elements are binary-searched in the longer one
//             (B elements in A held in LDS, or A elements in B in HBM).
//   kBitmap : a hashed bitmap of A in LDS filters B elements with one LDS read; the
//             survivors are confirmed in an LDS hash table of A (exact).
// Counting is exact by construction in all three: a hit is only recorded for an
// element equal to a list entry.
//
// Per-vertex and per-edge counts.  A hit w of pair (x, y) is the triangle {x, y, w}:
// t(x) += 1 (a register per row); t(w) goes to the LDS counter of w's place in A;
// t(y) goes to the LDS counter of y's place in A when x->y, otherwise (y->x) to one
// global atomic per pair.  After the row each nonzero LDS counter is added to the
// global count of its vertex -- about one global atomic per pair instead of one per
// triangle.  SUP mode (per-edge support, for Jaccard) keeps the same counters per
// edge: x->w and x->y at their places in A, y->x per pair, and y->w with one atomic
// per triangle at w's position in B.
#include <algorithm>

#include "pac_internal.hpp"

namespace pac {

// waves per row, LDS list capacity (uint32 entries), LDS bitmap bits
const TileCfg kTiles[PAC_NUM_TILES] = {
    {2, 1024, 16384}, {4, 2048, 32768}, {4, 4096, 65536}, {8, 4096, 65536}};

namespace {

constexpr uint32_t kHashMul = 0x9E3779B1u;  // Fibonacci hashing multiplier
constexpr uint32_t kAllC = 0xFFFFFFFFu;
constexpr uint32_t kPassPairs = 16384;  // pairs per bitmap pass (bounds the 16-bit counters)

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x >> 1); }

template <int NBITS>
__device__ __forceinline__ uint32_t bit_of(uint32_t w) {
  return (w * kHashMul) >> (32 - ilog2(NBITS));
}

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
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *a, uint32_t n, uint32_t x) {
  if (n == 0) return 0;
  uint32_t base = 0;
  while (n > 1) {
    const uint32_t half = n >> 1;
    base = (a[base + half] < x) ? base + half : base;
    n -= half;
  }
  return base + (a[base] < x ? 1u : 0u);
}

constexpr uint32_t kTabMul = 0x85EBCA6Bu;   // hash-table multiplier (independent of kHashMul)
constexpr uint32_t kEmpty = 0xFFFFFFFFu;     // empty hash-table slot (IDs are < 2^31)

// LDS of one workgroup.  The sorted-list strategies (merge, gallop) keep A and one
// 32-bit counter per position; the bitmap strategy keeps a linear-probing hash table
// of A (1.5*CAP slots) with packed 16-bit counters and, in SUP mode, 16-bit positions,
// plus one counter per pair of each wave's current batch.  Both live in one union: a
// workgroup runs one strategy per row.  The filter bitmap sits outside the union and
// is all-zero between bitmap rows.
template <int CAP, int NBITS, int MODE, int WAVES>
struct Smem {
  static constexpr uint32_t kTabCap = CAP + CAP / 2;  // hash slots: load factor <= 2/3
  union {
    struct {
      uint32_t A[CAP];
      uint32_t cnt[MODE == kModeTC ? 1 : CAP];
    } s;  // merge / gallop
    struct {
      uint32_t tab[kTabCap];
      uint32_t cnt16[MODE == kModeTC ? 1 : kTabCap / 2];   // 2 x u16 per word
      uint32_t pos16[MODE == kModeSUP ? kTabCap / 2 : 1];  // 2 x u16 per word
      uint32_t pcnt[MODE == kModeTC ? 1 : 64 * WAVES];     // hits per pair of a batch
    } h;  // bitmap
  } u;
  uint32_t bits[NBITS / 32];
  uint32_t item;
};

__device__ __forceinline__ uint32_t readlane(uint32_t v, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

// 16-bit counters packed two per word.  A bitmap pass covers at most kPassPairs pairs,
// so a counter stays below 2^16 (at most one hit per pair for an element of A plus at
// most nA <= CAP hits of the pair x->y at y's place) and halves never carry.
__device__ __forceinline__ void add16(uint32_t *w, uint32_t slot, uint32_t v) {
  atomicAdd(&w[slot >> 1], v << (16u * (slot & 1u)));
}
__device__ __forceinline__ uint32_t get16(const uint32_t *w, uint32_t slot) {
  return (w[slot >> 1] >> (16u * (slot & 1u))) & 0xFFFFu;
}

// Home slot of x in a table of tsize slots: high half of (hash * tsize).
__device__ __forceinline__ uint32_t tab_home(uint32_t x, uint32_t tsize) {
  return (uint32_t)(((uint64_t)(x * kTabMul) * tsize) >> 32);
}

// Slot of x in the table (kEmpty if absent); linear probing with wrap-around.  The
// table always has empty slots (load factor <= 2/3), so the probe ends.
__device__ __forceinline__ uint32_t tab_find(const uint32_t *tab, uint32_t tsize, uint32_t x) {
  uint32_t h = tab_home(x, tsize);
  while (true) {
    const uint32_t k = tab[h];
    if (k == x) return h;
    if (k == kEmpty) return kEmpty;
    h = (h + 1u == tsize) ? 0u : h + 1u;
  }
}

// Record a confirmed triangle (r, c, w) in the sorted-list strategies: w = A[i] = B[j].
template <int MODE>
__device__ __forceinline__ void on_hit(uint32_t *cnt, uint32_t i, uint32_t *sup, uint32_t bj) {
  if (MODE != kModeTC) atomicAdd(&cnt[i], 1u);
  if (MODE == kModeSUP) atomicAdd(&sup[bj], 1u);
}

// ---- the three lookup engines: one wave, one pair (x, y) ---------------------------
// B = oadj[b0 .. b0+nb): sorted N+(y) in HBM.  Each returns this lane's number of
// confirmed common neighbours.

template <int MODE>
__device__ __forceinline__ uint32_t probe_gallop(const uint32_t *__restrict__ oadj,
                                                 const uint32_t *A, uint32_t nA, uint32_t *cnt,
                                                 uint32_t b0, uint32_t nb, uint32_t *sup,
                                                 uint32_t lane) {
  uint32_t hits = 0;
  if (nb <= nA) {
    // B is the shorter list: each B element is binary-searched in A (LDS).
    for (uint32_t j0 = lane; j0 < nb; j0 += 128) {
      uint32_t w[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const uint32_t j = j0 + 64u * u;
        w[u] = j < nb ? oadj[b0 + j] : 0u;
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
    const uint32_t *B = oadj + b0;
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
__device__ __forceinline__ uint32_t probe_merge(const uint32_t *__restrict__ oadj,
                                                const uint32_t *A, uint32_t nA, uint32_t *cnt,
                                                uint32_t b0, uint32_t nb, uint32_t *sup,
                                                uint32_t lane) {
  const uint32_t *B = oadj + b0;
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

// The pairs of a row are visited in batches of 64 per wave: lane l loads the partner y
// and edge index e of the wave's l-th pair and the bounds of N+(y), so the offsets of
// 64 lists arrive with one round trip; the wave then walks the batch, broadcasting
// each pair with readlane.  Wave w takes pair positions k_lo + w, k_lo + w + WAVES, ...
// body(y, e, b0, nb) runs once per pair (e is only loaded when counts are kept).
template <int WAVES, int MODE, class Body>
__device__ __forceinline__ void for_each_pair(const GraphDev &G, uint32_t p0, uint32_t k_lo,
                                              uint32_t k_hi, uint32_t wave, uint32_t lane,
                                              Body body) {
  for (uint32_t t0 = 0;; t0 += 64) {
    const uint32_t kl = k_lo + wave + (uint32_t)WAVES * (t0 + lane);
    uint32_t yl = 0, el = 0, bl = 0, nl = 0;
    if (kl < k_hi) {
      yl = G.ppar[p0 + kl];
      if (MODE != kModeTC) el = G.pedge[p0 + kl];
      bl = G.ooff[yl];
      nl = G.ooff[yl + 1] - bl;
    }
    const uint32_t nvalid = (uint32_t)__popcll(__ballot(kl < k_hi));
    for (uint32_t jj = 0; jj < nvalid; ++jj)
      body(readlane(yl, jj), readlane(el, jj), readlane(bl, jj), readlane(nl, jj));
    if (nvalid < 64) break;
  }
}

// Completes this wave's LDS atomics before its lanes read what other lanes added.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// ---- one row, sorted-list strategies (merge, gallop): A staged in LDS ---------------
template <int STRAT, int MODE, int WAVES, int CAP, int NBITS>
__device__ __forceinline__ void row_sorted(const GraphDev &G, uint32_t r, uint32_t k_lo,
                                           uint32_t k_hi, Smem<CAP, NBITS, MODE, WAVES> &sm,
                                           unsigned long long &acc) {
  constexpr uint32_t BS = 64u * WAVES;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t a0 = G.ooff[r], nA = G.ooff[r + 1] - a0;
  const uint32_t p0 = G.poff[r], np = G.poff[r + 1] - p0;
  uint32_t *A = sm.u.s.A;
  uint32_t *cnt = sm.u.s.cnt;
  for (uint32_t i = tid; i < nA; i += BS) {
    A[i] = G.oadj[a0 + i];
    if (MODE != kModeTC) cnt[i] = 0;
  }
  __syncthreads();
  if (k_hi > np) k_hi = np;
  uint32_t row_hits = 0;
  for_each_pair<WAVES, MODE>(G, p0, k_lo, k_hi, wave, lane,
                             [&](uint32_t y, uint32_t e, uint32_t b0, uint32_t nb) {
    uint32_t h = 0;
    if (nb != 0) {
      if (STRAT == kGallop)
        h = probe_gallop<MODE>(G.oadj, A, nA, cnt, b0, nb, G.sup, lane);
      else
        h = probe_merge<MODE>(G.oadj, A, nA, cnt, b0, nb, G.sup, lane);
    }
    if (MODE != kModeTC) {
      const uint32_t ch = wave_sum(h);
      if (lane == 0 && ch) {
        if (e - a0 < nA)
          atomicAdd(&cnt[e - a0], ch);  // r->y: y = A[e - a0]
        else if (MODE == kModeLCC)
          atomicAdd(&G.tvert[y], (unsigned long long)ch);
        else
          atomicAdd(&G.sup[e], ch);
      }
    }
    row_hits += h;
  });
  acc += row_hits;
  if (MODE == kModeLCC) {
    const uint32_t wh = wave_sum(row_hits);
    if (lane == 0 && wh) atomicAdd(&G.tvert[r], (unsigned long long)wh);
  }
  __syncthreads();
  if (MODE != kModeTC) {
    for (uint32_t i = tid; i < nA; i += BS) {
      const uint32_t v = cnt[i];
      if (v) {
        if (MODE == kModeLCC) atomicAdd(&G.tvert[A[i]], (unsigned long long)v);
        if (MODE == kModeSUP) atomicAdd(&G.sup[a0 + i], v);
      }
    }
  }
  __syncthreads();
}

// ---- one row, bitmap strategy: hashed filter bitmap + LDS hash table of A -------------
// The table has tsize = max(64, min(2*nA, TABCAP)) slots (load factor <= 2/3), indexed
// by the high half of (hash * tsize).  A wave takes its pairs in batches of 64 (lane l
// loads partner y_l, the bounds of N+(y_l) and, when y_l is in A, its table slot), then
// streams the concatenation of the batch's 64 lists 256 elements at a time, dense
// across lanes and coalesced; each lane finds the list its element belongs to by
// walking the batch's inclusive prefix sums (held one per lane, read with ds_bpermute),
// which only moves forward.
template <int MODE, int WAVES, int CAP, int NBITS>
__device__ __forceinline__ void row_bitmap(const GraphDev &G, uint32_t r, uint32_t c_lo,
                                           uint32_t c_hi, Smem<CAP, NBITS, MODE, WAVES> &sm,
                                           unsigned long long &acc) {
  constexpr uint32_t BS = 64u * WAVES;
  constexpr uint32_t TABCAP = Smem<CAP, NBITS, MODE, WAVES>::kTabCap;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t a0 = G.ooff[r], nA = G.ooff[r + 1] - a0;
  const uint32_t p0 = G.poff[r], np = G.poff[r + 1] - p0;
  uint32_t *tab = sm.u.h.tab;
  uint32_t *cnt16 = sm.u.h.cnt16;
  uint32_t *pos16 = sm.u.h.pos16;
  uint32_t *pc = sm.u.h.pcnt + 64u * wave;
  uint32_t tsize = 2u * nA;
  if (tsize > TABCAP) tsize = TABCAP;
  if (tsize < 64u) tsize = 64u;
  for (uint32_t i = tid; i < tsize; i += BS) tab[i] = kEmpty;
  if (MODE != kModeTC) {
    for (uint32_t i = tid; i < (tsize + 1) / 2; i += BS) cnt16[i] = 0;
    for (uint32_t i = tid; i < BS; i += BS) sm.u.h.pcnt[i] = 0;
  }
  if (MODE == kModeSUP)
    for (uint32_t i = tid; i < (tsize + 1) / 2; i += BS) pos16[i] = 0;
  __syncthreads();
  for (uint32_t i = tid; i < nA; i += BS) {
    const uint32_t w = G.oadj[a0 + i];
    uint32_t h = tab_home(w, tsize);
    while (atomicCAS(&tab[h], kEmpty, w) != kEmpty) h = (h + 1u == tsize) ? 0u : h + 1u;
    if (MODE == kModeSUP) add16(pos16, h, i);
    const uint32_t hb = bit_of<NBITS>(w);
    atomicOr(&sm.bits[hb >> 5], 1u << (hb & 31u));
  }
  __syncthreads();
  if (c_hi > np) c_hi = np;
  uint32_t row_hits = 0;
  for (uint32_t t0 = 0;; t0 += 64) {
    const uint32_t kl = c_lo + wave + (uint32_t)WAVES * (t0 + lane);
    uint32_t yl = 0, el = 0, bl = 0, nl = 0, sc = kEmpty;
    if (kl < c_hi) {
      yl = G.ppar[p0 + kl];
      bl = G.ooff[yl];
      nl = G.ooff[yl + 1] - bl;
      if (MODE != kModeTC) {
        el = G.pedge[p0 + kl];
        if (el - a0 < nA) sc = tab_find(tab, tsize, yl);  // r->y: y is in A
      }
    }
    const uint32_t nvalid = (uint32_t)__popcll(__ballot(kl < c_hi));
    // inclusive prefix of the list lengths over the batch's lanes
    uint32_t incl = nl;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = __shfl_up(incl, o, 64);
      if (lane >= (uint32_t)o) incl += v;
    }
    const uint32_t total = readlane(incl, 63);
    const uint32_t base = bl - (incl - nl);  // flat index q of list l sits at oadj[base_l + q]
    uint32_t cur = 0;  // owner walk, monotone in q; always < 64 (see below)
    uint32_t hits = 0;
    for (uint32_t q0 = 0; q0 < total; q0 += 256u) {  // wave-uniform loop
      uint32_t w[4], addr[4], so[4], own[4];
      bool ok[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t q = q0 + 64u * u + lane;
        ok[u] = q < total;
        // The owner is the first lane whose inclusive prefix exceeds q.  Lanes past the
        // end use q = total-1, so every walk stops at a lane <= 63 (incl[63] = total).
        // Every shuffle runs with all 64 lanes active (bpermute reads of inactive lanes
        // are not defined), and the walk continues while any lane still has to advance.
        const uint32_t qc = ok[u] ? q : total - 1u;
        while (true) {
          const bool adv = (uint32_t)__shfl(incl, (int)cur, 64) <= qc;
          if (!__any(adv)) break;
          cur += adv ? 1u : 0u;
        }
        own[u] = cur;
        addr[u] = (uint32_t)__shfl(base, (int)cur, 64) + qc;
        so[u] = (MODE != kModeTC) ? (uint32_t)__shfl(sc, (int)cur, 64) : 0u;
        w[u] = ok[u] ? G.oadj[addr[u]] : 0u;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (!ok[u]) continue;
        const uint32_t hb = bit_of<NBITS>(w[u]);
        if ((sm.bits[hb >> 5] >> (hb & 31u)) & 1u) {
          const uint32_t slot = tab_find(tab, tsize, w[u]);
          if (slot != kEmpty) {
            ++hits;
            if (MODE != kModeTC) {
              add16(cnt16, slot, 1u);  // t(w) / edge r->w
              if (so[u] != kEmpty)
                add16(cnt16, so[u], 1u);  // t(y) / edge r->y, y in A
              else
                atomicAdd(&pc[own[u]], 1u);  // t(y) / edge y->r
            }
            if (MODE == kModeSUP) atomicAdd(&G.sup[addr[u]], 1u);  // edge y->w
          }
        }
      }
    }
    if (MODE != kModeTC) {
      // pairs y->r: one global atomic per pair with hits, then the counter is reset
      wave_lds_sync();
      if (kl < c_hi && sc == kEmpty) {
        const uint32_t v = pc[lane];
        if (v) {
          pc[lane] = 0;
          if (MODE == kModeLCC) atomicAdd(&G.tvert[yl], (unsigned long long)v);
          if (MODE == kModeSUP) atomicAdd(&G.sup[el], v);
        }
      }
      wave_lds_sync();
    }
    row_hits += hits;
    if (nvalid < 64) break;
  }
  acc += row_hits;
  if (MODE == kModeLCC) {
    const uint32_t wh = wave_sum(row_hits);
    if (lane == 0 && wh) atomicAdd(&G.tvert[r], (unsigned long long)wh);
  }
  __syncthreads();
  for (uint32_t i = tid; i < tsize; i += BS) {
    const uint32_t key = tab[i];
    if (key == kEmpty) continue;
    if (MODE != kModeTC) {
      const uint32_t v = get16(cnt16, i);
      if (v) {
        if (MODE == kModeLCC) atomicAdd(&G.tvert[key], (unsigned long long)v);
        if (MODE == kModeSUP) atomicAdd(&G.sup[a0 + get16(pos16, i)], v);
      }
    }
    sm.bits[bit_of<NBITS>(key) >> 5] = 0u;
  }
  __syncthreads();
}

// ---- one row whose N+(r) exceeds the LDS capacity: lookups in HBM ------------------
template <int MODE, int WAVES>
__device__ __forceinline__ void row_hbm(const GraphDev &G, uint32_t r, uint32_t c_lo,
                                        uint32_t c_hi, unsigned long long &acc) {
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t a0 = G.ooff[r], nA = G.ooff[r + 1] - a0;
  const uint32_t p0 = G.poff[r], np = G.poff[r + 1] - p0;
  const uint32_t *A = G.oadj + a0;
  if (c_hi > np) c_hi = np;
  uint32_t row_hits = 0;
  for_each_pair<WAVES, MODE>(G, p0, c_lo, c_hi, wave, lane,
                             [&](uint32_t y, uint32_t e, uint32_t b0, uint32_t nb) {
    uint32_t h = 0;
    for (uint32_t j = lane; j < nb; j += 64) {
      const uint32_t w = G.oadj[b0 + j];
      const uint32_t i = lower_bound_u32(A, nA, w);
      if (i < nA && A[i] == w) {
        ++h;
        if (MODE == kModeLCC) atomicAdd(&G.tvert[w], 1ull);
        if (MODE == kModeSUP) {
          atomicAdd(&G.sup[a0 + i], 1u);
          atomicAdd(&G.sup[b0 + j], 1u);
        }
      }
    }
    if (MODE != kModeTC) {
      const uint32_t ch = wave_sum(h);
      if (lane == 0 && ch) {
        if (MODE == kModeLCC) atomicAdd(&G.tvert[y], (unsigned long long)ch);
        if (MODE == kModeSUP) atomicAdd(&G.sup[e], ch);
      }
    }
    row_hits += h;
  });
  acc += row_hits;
  if (MODE == kModeLCC) {
    const uint32_t wh = wave_sum(row_hits);
    if (lane == 0 && wh) atomicAdd(&G.tvert[r], (unsigned long long)wh);
  }
}

// One row (or a pair range of one) with the given strategy.
template <int MODE, int WAVES, int CAP, int NBITS>
__device__ __forceinline__ void do_row(const GraphDev &G, uint32_t r, uint32_t c_lo,
                                       uint32_t c_hi, int st, Smem<CAP, NBITS, MODE, WAVES> &sm,
                                       unsigned long long &acc) {
  const uint32_t nA = G.ooff[r + 1] - G.ooff[r];
  if (nA > (uint32_t)CAP) {
    row_hbm<MODE, WAVES>(G, r, c_lo, c_hi, acc);
  } else if (st == kBitmap) {
    if (MODE == kModeTC) {
      row_bitmap<MODE, WAVES, CAP, NBITS>(G, r, c_lo, c_hi, sm, acc);
    } else {
      const uint32_t np = G.poff[r + 1] - G.poff[r];
      const uint32_t hi = c_hi < np ? c_hi : np;
      for (uint32_t lo = c_lo; lo < hi; lo = (hi - lo > kPassPairs) ? lo + kPassPairs : hi)
        row_bitmap<MODE, WAVES, CAP, NBITS>(
            G, r, lo, (hi - lo > kPassPairs) ? lo + kPassPairs : hi, sm, acc);
    }
  } else if (st == kGallop) {
    row_sorted<kGallop, MODE, WAVES, CAP, NBITS>(G, r, c_lo, c_hi, sm, acc);
  } else {
    row_sorted<kMerge, MODE, WAVES, CAP, NBITS>(G, r, c_lo, c_hi, sm, acc);
  }
}

template <int WAVES>
__device__ __forceinline__ void flush_total(unsigned long long *total, unsigned long long acc) {
  const unsigned long long s = wave_sum64(acc);
  if ((threadIdx.x & 63u) == 0 && s) atomicAdd(total, s);
}

// ---- plain kernel: one workgroup per row, one launch per partition group -----------
template <int STRAT, int MODE, int WAVES, int CAP, int NBITS>
__global__ __launch_bounds__(64 * WAVES) void k_plain(GraphDev G, uint32_t row_begin) {
  __shared__ Smem<CAP, NBITS, MODE, WAVES> sm;
  if (STRAT == kBitmap)
    for (uint32_t i = threadIdx.x; i < NBITS / 32; i += 64 * WAVES) sm.bits[i] = 0;
  __syncthreads();
  const uint32_t r = G.rows[row_begin + blockIdx.x];
  unsigned long long acc = 0;
  do_row<MODE, WAVES, CAP, NBITS>(G, r, 0, kAllC, STRAT, sm, acc);
  flush_total<WAVES>(G.total, acc);
}

// ---- persistent kernel: every workgroup pulls work items from one atomic queue --------
template <int MODE, int WAVES, int CAP, int NBITS>
__global__ __launch_bounds__(64 * WAVES) void k_persistent(const GraphDev *__restrict__ Gs,
                                                           const WorkItem *__restrict__ items,
                                                           uint32_t nitems,
                                                           const uint8_t *__restrict__ strat,
                                                           uint32_t *queue) {
  __shared__ Smem<CAP, NBITS, MODE, WAVES> s;
  for (uint32_t i = threadIdx.x; i < NBITS / 32; i += 64 * WAVES) s.bits[i] = 0;
  unsigned long long acc = 0;
  unsigned long long *cur_total = nullptr;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) s.item = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t it = s.item;
    if (it >= nitems) break;
    const WorkItem w = items[it];
    const GraphDev G = Gs[w.g];
    if (G.total != cur_total) {
      if (cur_total) flush_total<WAVES>(cur_total, acc);
      acc = 0;
      cur_total = G.total;
    }
    const uint8_t st = strat[G.part_base + w.part];
    const bool single = (w.hi == w.lo + 1);
    const uint32_t c_lo = single ? w.c_lo : 0u, c_hi = single ? w.c_hi : kAllC;
    for (uint32_t q = w.lo; q < w.hi; ++q)
      do_row<MODE, WAVES, CAP, NBITS>(G, G.rows[q], c_lo, c_hi, st, s, acc);
  }
  if (cur_total) flush_total<WAVES>(cur_total, acc);
}

__global__ void k_reset_queue(uint32_t *q) { *q = 0; }

// Plain streaming copy (16-byte vectors, grid-stride): the bandwidth reference.
__global__ void k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
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

// work[u] = sum of d+(partner) over u's pairs: the streamed length of row u (wave per row).
__global__ void k_rowwork(const uint32_t *__restrict__ poff, const uint32_t *__restrict__ ppar,
                          const uint32_t *__restrict__ odeg, uint32_t V,
                          uint32_t *__restrict__ work) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wpb = blockDim.x >> 6;
  for (uint64_t u = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); u < V;
       u += (uint64_t)gridDim.x * wpb) {
    const uint32_t p0 = poff[u], p1 = poff[u + 1];
    uint64_t s = 0;
    for (uint32_t i = p0 + lane; i < p1; i += 64) s += odeg[ppar[i]];
    s = wave_sum64(s);
    if (lane == 0) work[u] = s > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)s;
  }
}

// ---- epilogues -----------------------------------------------------------------------
// LCC of Graph::solve_lcc_merge (Graph.cpp:828-840): numerator 2t(u) and denominator
// d*(d-1) are exact integers converted to double, one IEEE division, then one
// rounding to float.  d = 0 keeps 0.0, d = 1 gives 0/0 = NaN, as the reference.
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
__global__ void k_jaccard(const uint32_t *__restrict__ ooff, const uint32_t *__restrict__ oadj,
                          const uint32_t *__restrict__ sup, const uint32_t *__restrict__ off,
                          uint32_t V, int32_t *src, int32_t *dst, uint32_t *inter,
                          uint32_t *uni, float *ratio) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wpb = blockDim.x >> 6;
  for (uint64_t u = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); u < V;
       u += (uint64_t)gridDim.x * wpb) {
    const uint32_t a0 = ooff[u], a1 = ooff[u + 1];
    const uint32_t du = off[u + 1] - off[u];
    for (uint32_t e = a0 + lane; e < a1; e += 64) {
      const uint32_t v = oadj[e];
      const uint32_t c = sup[e];
      const uint32_t un = du + (off[v + 1] - off[v]) - c;
      if (src) src[e] = (int32_t)u;
      if (dst) dst[e] = (int32_t)v;
      if (inter) inter[e] = c;
      if (uni) uni[e] = un;
      if (ratio) ratio[e] = (float)((double)c / (double)un);
    }
  }
}

// |N(u) & N(v)| over the full symmetric lists, one wave per pair: the shorter list's
// elements are binary-searched in the longer one (Graph::cnt_common_neis_galloping,
// Graph.cpp:186-198).  Optional Jaccard ratio c / ((d_u + d_v) - c) in double
// (Graph.cpp:808-809).
__global__ void k_pairs(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj,
                        const int32_t *__restrict__ pu, const int32_t *__restrict__ pv,
                        uint64_t n, unsigned long long *res, double *jac, const int32_t *q,
                        uint32_t nq) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wpb = blockDim.x >> 6;
  for (uint64_t p = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); p < n;
       p += (uint64_t)gridDim.x * wpb) {
    uint32_t u, v;
    if (q) {
      u = (uint32_t)q[p / nq];
      v = (uint32_t)q[p % nq];
    } else {
      u = (uint32_t)pu[p];
      v = (uint32_t)pv[p];
    }
    uint32_t a0 = off[u], na = off[u + 1] - a0, b0 = off[v], nb = off[v + 1] - b0;
    const uint32_t du = na, dv = nb;
    if (na > nb) {
      uint32_t t = a0; a0 = b0; b0 = t;
      t = na; na = nb; nb = t;
    }
    uint32_t h = 0;
    if (nb) {
      const uint32_t *B = adj + b0;
      for (uint32_t i = lane; i < na; i += 64) {
        const uint32_t x = adj[a0 + i];
        const uint32_t j = lower_bound_u32(B, nb, x);
        h += (j < nb && B[j] == x) ? 1u : 0u;
      }
    }
    h = wave_sum(h);
    if (lane == 0) {
      if (res) res[p] = h;
      if (jac) {
        const double c = (double)h;
        jac[p] = c / ((double)((uint64_t)du + dv) - c);
      }
    }
  }
}

template <int MODE, int WAVES, int CAP, int NBITS>
int run_persistent_cfg(Ctx *ctx, const WorkItem *items, uint32_t nitems, uint64_t *launches) {
  int per_cu = 0;
  PAC_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void *>(k_persistent<MODE, WAVES, CAP, NBITS>),
      64 * WAVES, 0));
  if (per_cu < 1) per_cu = 1;
  uint32_t grid = (uint32_t)(per_cu * ctx->num_cus);
  if (grid > nitems) grid = nitems ? nitems : 1;
  hipLaunchKernelGGL((k_reset_queue), dim3(1), dim3(1), 0, ctx->stream, ctx->queue.as<uint32_t>());
  if (ctx->time_kernel) PAC_CHECK_HIP(hipEventRecord(ctx->kev[0], ctx->stream));
  hipLaunchKernelGGL((k_persistent<MODE, WAVES, CAP, NBITS>), dim3(grid), dim3(64 * WAVES), 0,
                     ctx->stream, ctx->gdesc.as<const GraphDev>(), items, nitems,
                     ctx->dispatch.as<const uint8_t>(), ctx->queue.as<uint32_t>());
  PAC_CHECK_HIP(hipGetLastError());
  if (ctx->time_kernel) PAC_CHECK_HIP(hipEventRecord(ctx->kev[1], ctx->stream));
  if (launches) *launches += 2;
  return PAC_OK;
}

template <int MODE>
int run_persistent_mode(Ctx *ctx, const WorkItem *items, uint32_t nitems, uint64_t *launches) {
  switch (ctx->tile) {
    case 0: return run_persistent_cfg<MODE, 2, 1024, 16384>(ctx, items, nitems, launches);
    case 1: return run_persistent_cfg<MODE, 4, 2048, 32768>(ctx, items, nitems, launches);
    case 2: return run_persistent_cfg<MODE, 4, 4096, 65536>(ctx, items, nitems, launches);
    default: return run_persistent_cfg<MODE, 8, 4096, 65536>(ctx, items, nitems, launches);
  }
}

int run_persistent_any(Ctx *ctx, int mode, const WorkItem *items, uint32_t nitems,
                       uint64_t *launches) {
  if (mode == kModeLCC) return run_persistent_mode<kModeLCC>(ctx, items, nitems, launches);
  if (mode == kModeSUP) return run_persistent_mode<kModeSUP>(ctx, items, nitems, launches);
  return run_persistent_mode<kModeTC>(ctx, items, nitems, launches);
}

template <int STRAT, int MODE, int WAVES, int CAP, int NBITS>
int run_plain_cfg(Ctx *ctx, const GraphDev &G, uint32_t row_begin, uint32_t nrows) {
  hipLaunchKernelGGL((k_plain<STRAT, MODE, WAVES, CAP, NBITS>), dim3(nrows), dim3(64 * WAVES), 0,
                     ctx->stream, G, row_begin);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

template <int STRAT, int MODE>
int run_plain_tile(Ctx *ctx, const GraphDev &G, uint32_t row_begin, uint32_t nrows) {
  switch (ctx->tile) {
    case 0: return run_plain_cfg<STRAT, MODE, 2, 1024, 16384>(ctx, G, row_begin, nrows);
    case 1: return run_plain_cfg<STRAT, MODE, 4, 2048, 32768>(ctx, G, row_begin, nrows);
    case 2: return run_plain_cfg<STRAT, MODE, 4, 4096, 65536>(ctx, G, row_begin, nrows);
    default: return run_plain_cfg<STRAT, MODE, 8, 4096, 65536>(ctx, G, row_begin, nrows);
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

// Device descriptor of a graph; its output pointers must already be allocated.
GraphDev make_desc(Graph *g, uint32_t part_base) {
  GraphDev d;
  d.ooff = g->ooff.as<const uint32_t>();
  d.oadj = g->oadj.as<const uint32_t>();
  d.poff = g->poff.as<const uint32_t>();
  d.ppar = g->ppar.as<const uint32_t>();
  d.pedge = g->pedge.as<const uint32_t>();
  d.rows = g->rows.as<const uint32_t>();
  d.total = g->total.as<unsigned long long>();
  d.tvert = g->tvert.as<unsigned long long>();
  d.sup = g->sup.as<uint32_t>();
  d.V = g->V;
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
  const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)ctx->num_cus * 64);
  std::vector<double> t;
  for (int r = 0; r <= reps; ++r) {
    PAC_CHECK_HIP(hipEventRecord(ctx->kev[0], ctx->stream));
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, ctx->stream, a.as<const uint4>(),
                       b.as<uint4>(), n);
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
    if (mode == kModeLCC) PAC_TRY(g->tvert.reserve(sizeof(unsigned long long) * (size_t)(g->V ? g->V : 1)));
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
    if (mode == kModeLCC && g->V)
      hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(g->V, 256 * 8)), dim3(256), 0, ctx->stream,
                         g->tvert.as<unsigned long long>(), (uint64_t)g->V);
    if (mode == kModeSUP) PAC_TRY(fill_u32(ctx->stream, g->sup.as<uint32_t>(), g->m, 0u));
    PAC_CHECK_HIP(hipGetLastError());
  }
  return PAC_OK;
}

int compute_rowwork(Graph *g) {
  PAC_TRY(build_pairs(g));
  PAC_TRY(g->rowwork.reserve(sizeof(uint32_t) * (size_t)(g->V ? g->V : 1)));
  if (g->V == 0) return PAC_OK;
  hipLaunchKernelGGL(k_rowwork, dim3(grid_for(g->V, 4)), dim3(256), 0, g->ctx->stream,
                     g->poff.as<const uint32_t>(), g->ppar.as<const uint32_t>(),
                     g->odeg.as<const uint32_t>(), g->V, g->rowwork.as<uint32_t>());
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

// One launch per partition group: a maximal run of consecutive partitions with the
// same strategy (their rows are contiguous in the row list), one workgroup per row.
int launch_plain(Ctx *ctx, Graph *g, int mode, const uint8_t *strategy_per_part,
                 uint64_t *launches) {
  Graph *gs[1] = {g};
  PAC_TRY(alloc_outputs(gs, 1, mode));
  const GraphDev d = make_desc(g, 0);
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

int launch_partition(Ctx *ctx, Graph *g, int mode, uint32_t part, int strat) {
  Graph *gs[1] = {g};
  PAC_TRY(alloc_outputs(gs, 1, mode));
  const GraphDev d = make_desc(g, 0);
  const uint32_t b = g->h_pstart[part], e = g->h_pstart[part + 1];
  return run_plain_any(ctx, d, mode, strat, b, e - b);
}

int launch_persistent(Ctx *ctx, Graph *const *gs, uint32_t G, int mode,
                      const uint8_t *strategy_per_part_concat, bool use_graph,
                      uint64_t *launches) {
  // Output buffers must exist before the descriptors that point at them are built
  // (on the hipGraph path nothing has allocated them yet).
  PAC_TRY(alloc_outputs(gs, G, mode));
  std::vector<GraphDev> descs(G);
  uint64_t nparts = 0, nitems = 0;
  for (uint32_t i = 0; i < G; ++i) {
    descs[i] = make_desc(gs[i], (uint32_t)nparts);
    nparts += gs[i]->P;
    nitems += gs[i]->h_items.size();
  }
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
    return run_persistent_any(ctx, mode, items, (uint32_t)nitems, launches);
  }
  // hipGraph: zero-fill kernels + queue reset + persistent kernel, captured once per
  // (graph set, mode, tile, buffer addresses) and replayed.
  key.push_back(items);
  key.push_back(ctx->gdesc.p);
  key.push_back(ctx->dispatch.p);
  key.push_back(ctx->queue.p);
  for (uint32_t i = 0; i < G; ++i) {
    key.push_back(gs[i]->tvert.p);
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
    if (rc == PAC_OK && nitems) rc = run_persistent_any(ctx, mode, items, (uint32_t)nitems, nullptr);
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
  hipLaunchKernelGGL(k_jaccard, dim3(grid_for(g->V, 4)), dim3(256), 0, ctx->stream,
                     g->ooff.as<const uint32_t>(), g->oadj.as<const uint32_t>(),
                     g->sup.as<const uint32_t>(), g->off.as<const uint32_t>(), g->V, src, dst,
                     inter, uni, ratio);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

int pair_counts(Ctx *ctx, Graph *g, const int32_t *d_u, const int32_t *d_v, uint64_t n,
                uint64_t *d_res) {
  if (n == 0) return PAC_OK;
  hipLaunchKernelGGL(k_pairs, dim3(grid_for(n, 4)), dim3(256), 0, ctx->stream,
                     g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(), d_u, d_v, n,
                     reinterpret_cast<unsigned long long *>(d_res), nullptr, nullptr, 1u);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

int query_matrix(Ctx *ctx, Graph *g, const int32_t *d_q, uint32_t nq, double *d_res) {
  const uint64_t n = (uint64_t)nq * nq;
  if (n == 0) return PAC_OK;
  hipLaunchKernelGGL(k_pairs, dim3(grid_for(n, 4)), dim3(256), 0, ctx->stream,
                     g->off.as<const uint32_t>(), g->adj.as<const uint32_t>(), nullptr, nullptr,
                     n, nullptr, d_res, d_q, nq);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

}  // namespace pac
