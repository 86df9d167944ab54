// asan_check.cpp -- runs the oracle's functions on small graphs under AddressSanitizer
// and UndefinedBehaviorSanitizer (host code only; `make -C oracle asan`).  TEST
// INFRASTRUCTURE ONLY.  The reference's partitioned threshold merge reads one element past
// a sub-list (Graph.cpp:1136-1149); the restatement checks bounds first, and this run is
// what shows it.  Exits non-zero when two formulations of a count disagree.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pac_oracle.h"

namespace {
int failures = 0;
void expect(bool ok, const char *what) {
  if (!ok) {
    std::printf("FAIL %s\n", what);
    ++failures;
  }
}
}  // namespace

int main() {
  for (uint32_t scale : {6u, 8u, 10u}) {
    uint64_t V = 0, nnz = 0;
    uint64_t *vl = nullptr;
    int32_t *el = nullptr;
    if (orc_gen_rmat(scale, 8, 0.57, 0.19, 0.19, scale, &V, &vl, &el, &nnz) != 0) return 2;
    const uint64_t T = orc_tc_merge(vl, el, V, 2);
    expect(orc_tc_galloping(vl, el, V, 2) == 6 * T, "galloping = 6T");
    expect(orc_tc_oriented(vl, el, V, 2) == T, "oriented = merge");
    std::vector<double> lcc(V), lccg(V);
    orc_lcc_merge(vl, el, V, lcc.data(), 2);
    orc_lcc_galloping(vl, el, V, lccg.data(), 2);
    std::vector<uint64_t> numer(V);
    orc_lcc_numer(vl, el, V, numer.data(), 2);
    uint64_t sum = 0;
    for (uint64_t x : numer) sum += x;
    expect(sum == 6 * T, "sum of LCC numerators = 6T");
    std::vector<int32_t> q;
    for (uint64_t u = 0; u < V && q.size() < 40; u += 3) q.push_back((int32_t)u);
    std::vector<double> js(q.size() * q.size());
    orc_jsim_merge(vl, el, q.data(), q.size(), js.data(), 2);
    orc_jsim_galloping(vl, el, q.data(), q.size(), js.data(), 2);
    std::vector<int32_t> src, dst;
    for (uint64_t u = 0; u < V; ++u)
      for (uint64_t i = vl[u]; i < vl[u + 1]; ++i)
        if ((int64_t)u < el[i]) {
          src.push_back((int32_t)u);
          dst.push_back(el[i]);
        }
    std::vector<uint32_t> inter(src.size()), uni(src.size());
    std::vector<float> ratio(src.size());
    orc_jaccard_edges(vl, el, src.data(), dst.data(), src.size(), inter.data(), uni.data(),
                      ratio.data(), 2);
    uint64_t si = 0;
    for (uint32_t x : inter) si += x;
    expect(si == 3 * T, "sum of edge intersections = 3T");
    for (uint64_t P : {1ull, 5ull}) {
      std::vector<int32_t> v2p(V);
      orc_random_partition(V, P, 7, v2p.data());
      orc_pgraph *pg = orc_pgraph_build(vl, el, V, v2p.data(), 2);
      expect(orc_pgraph_tc_merge(pg) == T, "partitioned merge = T");
      expect(orc_pgraph_tc_galloping(pg) == 6 * T, "partitioned galloping = 6T");
      std::vector<double> plcc(V);
      orc_pgraph_lcc_merge(pg, plcc.data());
      for (uint64_t u = 0; u + 1 < V && u < 200; ++u)
        for (uint64_t k = 0; k < P; ++k) {
          const int32_t v = (int32_t)((u * 7 + 3) % V);
          (void)orc_pgraph_cnt(pg, (int32_t)u, v, k, 0);
          (void)orc_pgraph_cnt(pg, (int32_t)u, v, k, 1);
          (void)orc_pgraph_cnt_less_than(pg, (int32_t)u, v, v, k);  // the reference's OOB case
        }
      uint64_t mp = 0, md = 0;
      (void)orc_pgraph_evaluate_partition(pg, &mp, &md);
      (void)orc_pgraph_square_counting(pg);
      orc_pgraph_free(pg);
      orc_greedy_partition_2(vl, el, V, P, 3, v2p.data());
    }
    orc_pairs ps;
    orc_pair_stats(vl, el, V, &ps);
    orc_stats st;
    orc_oriented_stats(vl, el, V, &st);
    expect(ps.pairs <= st.m, "pairs <= oriented edges");
    orc_free(vl);
    orc_free(el);
  }
  std::printf("%s\n", failures ? "asan check FAILED" : "asan check ok");
  return failures ? 1 : 0;
}
