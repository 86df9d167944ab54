// pac_main.cpp -- command-line driver with the reference's argv protocol for this path
// (src/test.cpp:301-389 of Byy-0024/PartitionAndConquer) plus new verbs.
//
//   pac_main tc <graphdir> <merge|galloping> <is_partition> <method> <P> <threads>
//       test.cpp:235-265, 371-379: loads <graphdir>/origin, optionally
//       <graphdir>/partition/<method>/P<P>.bin, times the solve call only and prints
//       "Time used for ..." and "#Triangles: N." (6T for galloping, as the reference).
//   pac_main partition <graphdir> <P> <outdir>             test.cpp:303-320
//   pac_main lcc <graphdir> <merge|galloping> <is_partition> <method> <P> <threads> [out.bin]
//   pac_main jaccard <graphdir> <is_partition> <method> <P> <threads> [outdir]
//   pac_main preprocess <graphdir> <graphfile>             preprocess.cpp:39-45 (no weights)
//   pac_main gen-rmat <scale> <ef> <a> <b> <c> <seed> <graphdir>
//   pac_main generate_queries <query_num> <outputfile> [num_vertices]
//       test.cpp:321-330: 2 * query_num values of the C library's rand() (unseeded) as
//       int32, written raw; with num_vertices each value is taken modulo it, so that the
//       pairs are vertex IDs of a graph of that size.
//   pac_main common <graphdir> <queryfile> <merge|galloping> <is_partition> <method> <P> [out]
//       batch_cnt_common_neis_* over the (u, v) pairs of a query file (Graph.cpp:413-447,
//       1382-1458); prints the sum of the counts and writes them as uint64 to [out].
// With PAC_JSON=1 every solve also prints one JSON line with the device timing.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <numeric>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "pac_graph.hpp"

using namespace pac_ref;

namespace {

// time_counter::print, Graph.cpp:16-21
void print_time(const std::string &what, unsigned long long ns) {
  if (ns < 1000)
    printf("Time used for %s: %llu ns.\n", what.c_str(), ns);
  else if (ns < 1000000)
    printf("Time used for %s: %.3f us.\n", what.c_str(), (float)ns / 1000);
  else if (ns < 1000000000ull)
    printf("Time used for %s: %.3f ms.\n", what.c_str(), (float)ns / 1000000);
  else
    printf("Time used for %s: %.3f s.\n", what.c_str(), (float)ns / (float)1000000000);
}

template <class F>
unsigned long long timed(F f) {
  auto a = std::chrono::high_resolution_clock::now();
  f();
  auto b = std::chrono::high_resolution_clock::now();
  return (unsigned long long)std::chrono::duration_cast<std::chrono::nanoseconds>(b - a).count();
}

std::vector<int> read_partition(const std::string &file) {
  std::ifstream in(file, std::ios::binary);
  if (!in) {
    fprintf(stderr, "Error: cannot open binfile %s\n", file.c_str());
    exit(1);
  }
  in.seekg(0, std::ios::end);
  const std::streamoff n = in.tellg();
  in.seekg(0, std::ios::beg);
  std::vector<int> v((size_t)n / sizeof(int));
  in.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * sizeof(int)));
  return v;
}

void json_line(const char *verb, const Graph &g) {
  if (!getenv("PAC_JSON")) return;
  pac_timing t;
  if (pac_ctx_last_timing(g.context(), &t) != PAC_OK) return;
  printf("{\"verb\": \"%s\", \"device_ms\": %.6f, \"isect_ms\": %.6f, \"epilogue_ms\": %.6f, "
         "\"launches\": %llu}\n",
         verb, t.total_ms, t.isect_ms, t.epilogue_ms, (unsigned long long)t.launches);
}

int usage() {
  fprintf(stderr,
          "usage: pac_main tc <graphdir> <merge|galloping> <is_partition> <method> <P> <threads>\n"
          "       pac_main partition <graphdir> <P> <outdir>\n"
          "       pac_main lcc <graphdir> <merge|galloping> <is_partition> <method> <P> <threads> [out]\n"
          "       pac_main jaccard <graphdir> <is_partition> <method> <P> <threads> [outdir]\n"
          "       pac_main preprocess <graphdir> <graphfile>\n"
          "       pac_main gen-rmat <scale> <ef> <a> <b> <c> <seed> <graphdir>\n"
          "       pac_main generate_queries <query_num> <outputfile> [num_vertices]\n"
          "       pac_main common <graphdir> <queryfile> <merge|galloping> <is_partition> "
          "<method> <P> [out]\n");
  return 2;
}

void mkdirs(const std::string &d) {
  std::string cur;
  for (size_t i = 0; i <= d.size(); ++i) {
    if (i == d.size() || d[i] == '/') {
      if (!cur.empty()) mkdir(cur.c_str(), 0755);
    }
    if (i < d.size()) cur += d[i];
  }
}

int run(int argc, char **argv) {
  if (argc < 2) return usage();
  const std::string task = argv[1];
  if (task == "tc" || task == "lcc") {
    if (argc < 8) return usage();
    const std::string dir = argv[2], method = argv[3], pmethod = argv[5];
    const bool is_partition = std::stod(argv[4]) != 0;
    const int P = (int)std::stod(argv[6]);
    Graph g(dir + "/origin", false);
    const bool gallop = method == "galloping";
    if (!gallop && method != "merge") {
      fprintf(stderr, "unknown intersection method %s\n", method.c_str());
      return 2;
    }
    size_t res = 0;
    std::vector<double> lcc;
    unsigned long long ns = 0;
    if (is_partition) {
      const std::vector<int> v2p =
          read_partition(dir + "/partition/" + pmethod + "/P" + std::to_string(P) + ".bin");
      Partitioned_Graph G(g, v2p, P);
      if (task == "tc") {
        ns = timed([&] {
          res = gallop ? G.solve_triangle_counting_galloping() : G.solve_triangle_counting_merge();
        });
      } else {
        ns = timed([&] {
          if (gallop)
            G.solve_lcc_galloping(lcc);
          else
            G.solve_lcc_merge(lcc);
        });
      }
      print_time("partitioned " + method, ns);
    } else {
      if (task == "tc") {
        ns = timed([&] {
          res = gallop ? g.solve_triangle_counting_galloping(0) : g.solve_triangle_counting_merge(0);
        });
      } else {
        ns = timed([&] {
          if (gallop)
            g.solve_lcc_galloping(lcc, 0);
          else
            g.solve_lcc_merge(lcc, 0);
        });
      }
      print_time("paralle " + method, ns);
    }
    if (task == "tc") {
      printf("#Triangles: %lu.\n", res);
    } else {
      double s = 0;
      size_t nn = 0;
      for (double x : lcc)
        if (x == x) {
          s += x;
          ++nn;
        }
      printf("Average LCC: %.9f over %zu vertices.\n", nn ? s / (double)nn : 0.0, nn);
      if (argc > 8) {
        std::ofstream out(argv[8], std::ios::binary);
        out.write(reinterpret_cast<const char *>(lcc.data()),
                  (std::streamsize)(lcc.size() * sizeof(double)));
      }
    }
    json_line(task.c_str(), g);
    return 0;
  }
  if (task == "partition") {
    if (argc < 5) return usage();
    const std::string dir = argv[2], outdir = argv[4];
    const size_t P = std::stoul(argv[3]);
    Graph g(dir + "/origin", false);
    std::vector<int> v2p;
    const unsigned long long ns = timed([&] { g.greedy_partition(P, v2p); });
    print_time("greedy partition", ns);
    Partitioned_Graph G(g, v2p, (int)P);
    G.evaluate_partition();
    mkdirs(outdir);
    std::ofstream out(outdir + "/P" + std::to_string(P) + ".bin", std::ios::binary);
    out.write(reinterpret_cast<const char *>(v2p.data()), (std::streamsize)(v2p.size() * sizeof(int)));
    return 0;
  }
  if (task == "jaccard") {
    if (argc < 7) return usage();
    const std::string dir = argv[2], pmethod = argv[4];
    const bool is_partition = std::stod(argv[3]) != 0;
    const int P = (int)std::stod(argv[5]);
    Graph g(dir + "/origin", false);
    std::vector<int32_t> src, dst;
    std::vector<uint32_t> inter, uni;
    std::vector<float> ratio;
    unsigned long long ns = 0;
    if (is_partition) {
      const std::vector<int> v2p =
          read_partition(dir + "/partition/" + pmethod + "/P" + std::to_string(P) + ".bin");
      Partitioned_Graph G(g, v2p, P);
      ns = timed([&] {
        G.per_edge_jaccard(src, dst, inter, uni, ratio, PAC_DISPATCH_FIXED, PAC_LAUNCH_PERSISTENT);
      });
    } else {
      ns = timed([&] {
        g.per_edge_jaccard(src, dst, inter, uni, ratio, PAC_DISPATCH_FIXED, PAC_LAUNCH_PERSISTENT);
      });
    }
    print_time("per-edge jaccard", ns);
    const double mean = ratio.empty() ? 0.0
                                      : std::accumulate(ratio.begin(), ratio.end(), 0.0) /
                                            (double)ratio.size();
    printf("#Edges: %zu, mean Jaccard: %.9f.\n", ratio.size(), mean);
    if (argc > 7) {
      const std::string od = argv[7];
      mkdirs(od);
      auto dump = [&](const char *name, const void *p, size_t bytes) {
        std::ofstream o(od + "/" + name, std::ios::binary);
        o.write(reinterpret_cast<const char *>(p), (std::streamsize)bytes);
      };
      dump("src.bin", src.data(), src.size() * 4);
      dump("dst.bin", dst.data(), dst.size() * 4);
      dump("inter.bin", inter.data(), inter.size() * 4);
      dump("union.bin", uni.data(), uni.size() * 4);
      dump("ratio.bin", ratio.data(), ratio.size() * 4);
    }
    json_line("jaccard", g);
    return 0;
  }
  if (task == "preprocess") {
    if (argc < 4) return usage();
    const std::string dir = argv[2], file = argv[3];
    Graph g(dir + file);  // graphdir + graphfile, no separator (preprocess.cpp:43)
    mkdirs(dir + "/origin");
    g.save_as_csr_binfile(dir + "/origin/csr_vlist.bin", dir + "/origin/csr_elist.bin");
    return 0;
  }
  if (task == "generate_queries") {
    if (argc < 4) return usage();
    const size_t n = std::stoul(argv[2]);
    const long long nv = argc > 4 ? std::stoll(argv[4]) : 0;
    std::vector<int> q;
    q.reserve(2 * n);
    for (size_t i = 0; i < 2 * n; ++i) q.push_back(nv > 0 ? (int)(rand() % nv) : rand());
    std::ofstream out(argv[3], std::ios::binary);
    out.write(reinterpret_cast<const char *>(q.data()), (std::streamsize)(q.size() * sizeof(int)));
    return 0;
  }
  if (task == "common") {
    if (argc < 8) return usage();
    const std::string dir = argv[2], method = argv[4], pmethod = argv[6];
    const bool is_partition = std::stod(argv[5]) != 0;
    const int P = (int)std::stod(argv[7]);
    const bool gallop = method == "galloping";
    if (!gallop && method != "merge") {
      fprintf(stderr, "unknown intersection method %s\n", method.c_str());
      return 2;
    }
    const std::vector<int> raw = read_partition(argv[3]);  // raw int32 file
    std::vector<std::pair<VID, VID>> queries(raw.size() / 2);
    for (size_t i = 0; i < queries.size(); ++i) queries[i] = {raw[2 * i], raw[2 * i + 1]};
    Graph g(dir + "/origin", false);
    std::vector<size_t> cnt;
    unsigned long long ns = 0;
    if (is_partition) {
      const std::vector<int> v2p =
          read_partition(dir + "/partition/" + pmethod + "/P" + std::to_string(P) + ".bin");
      Partitioned_Graph G(g, v2p, P);
      ns = timed([&] {
        if (gallop)
          G.batch_cnt_common_neis_galloping(queries, cnt);
        else
          G.batch_cnt_common_neis_merge(queries, cnt);
      });
      print_time("partitioned " + method + " queries", ns);
    } else {
      cnt.assign(queries.size(), 0);
      ns = timed([&] {
        if (gallop)
          g.batch_cnt_common_neis_galloping(queries, cnt, 0);
        else
          g.batch_cnt_common_neis_merge(queries, cnt, 0);
      });
      print_time("paralle " + method + " queries", ns);
    }
    printf("#Queries: %zu, #Common neighbours: %llu.\n", cnt.size(),
           (unsigned long long)std::accumulate(cnt.begin(), cnt.end(), 0ull));
    if (argc > 8) {
      std::vector<uint64_t> c64(cnt.begin(), cnt.end());
      std::ofstream out(argv[8], std::ios::binary);
      out.write(reinterpret_cast<const char *>(c64.data()),
                (std::streamsize)(c64.size() * sizeof(uint64_t)));
    }
    return 0;
  }
  if (task == "gen-rmat") {
    if (argc < 9) return usage();
    pac_graph *h = nullptr;
    pac_ctx *ctx = default_context();
    const int rc = pac_graph_generate_rmat(ctx, (uint32_t)std::stoul(argv[2]),
                                           (uint32_t)std::stoul(argv[3]), std::stod(argv[4]),
                                           std::stod(argv[5]), std::stod(argv[6]),
                                           std::stoull(argv[7]), &h);
    if (rc != PAC_OK) throw Error(rc, pac_last_error());
    Graph g(ctx, h);
    const std::string dir = std::string(argv[8]) + "/origin";
    mkdirs(dir);
    g.save_as_csr_binfile(dir + "/csr_vlist.bin", dir + "/csr_elist.bin");
    return 0;
  }
  return usage();
}

}  // namespace

int main(int argc, char **argv) {
  try {
    return run(argc, argv);
  } catch (const Error &e) {
    fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 1;
  } catch (const std::exception &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
