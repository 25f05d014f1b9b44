// Driver for the reference's own host path (graph.hpp + cliques.hpp + bucket_ex.hpp),
// compiled from the headers where they lie under the reference tree by oracle/Makefile
// (target `ref`). The reference ships these templates but no main that composes
// them; this file composes them in the order the reference's CUDA mains run the same steps:
//   graph_t::from_edges/deserialize -> make_oriented -> make_split_parts
//   -> cliques_tri_par -> graph_t<..,2>(tris) -> cliques_qud_par
//   -> graph_t<..,3>(quds, T, undirected) -> bucket_ex::peeling_par -> support()
// Output: a JSON line on stdout (counts, K_max, phase seconds) and, if asked, the
// k file: one "(a,b,c)\tk" row per triangle in (a,b) order.
// Test infrastructure only: used to cross-check the oracle restatement.
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>
#include <omp.h>

#include "graph/graph.hpp"
#include "cliques.hpp"
#include "bucket_ex.hpp"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: ref34 <dataset> [kfile] [parts]\n";
    return 2;
  }
  std::string file = argv[1];
  std::string kfile = argc > 2 ? argv[2] : "";
  size_t parts = argc > 3 ? std::stoul(argv[3]) : static_cast<size_t>(omp_get_num_procs());
  using clk = std::chrono::steady_clock;
  auto secs = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double>(b - a).count();
  };
  try {
    auto t0 = clk::now();
    graph_t<EDGE_T, VERTEX_T, 1> g;
    std::string ext = file.substr(file.find_last_of('.'));
    if (ext == ".mtx") { g.from_edges(file, true); g.make_oriented(); }
    else if (ext == ".edges" || ext == ".txt") { g.from_edges(file, false); g.make_oriented(); }
    else if (ext == ".grh") { g.deserialize(file); g.build_offset(); }
    else { std::cerr << "Unknown dataset format: " << ext << "\n"; return 3; }
    g.make_split_parts(parts);
    auto t1 = clk::now();
    auto tris = cliques_tri_par(g);
    auto t2 = clk::now();
    graph_t<EDGE_T, VERTEX_T, 2> tg(tris, g.size_edges(), true);
    tg.make_split_parts(parts);
    auto t3 = clk::now();
    auto quds = cliques_qud_par(tg);
    auto t4 = clk::now();
    size_t T = tris.size(), Q = quds.size();
    long kmax = 0;
    std::vector<MTYPE> k;
    double t_qg = 0, t_peel = 0;
    if (T) {
      auto t5 = clk::now();
      graph_t<EDGE_T, VERTEX_T, 3> qg(quds, T, false);
      auto t6 = clk::now();
      bucket_ex<VERTEX_T, MTYPE> b(qg.D);
      kmax = static_cast<long>(b.peeling_par(qg));
      k = b.support();
      auto t7 = clk::now();
      t_qg = secs(t5, t6);
      t_peel = secs(t6, t7);
      if (Q == 0) kmax = 0;
    }
    if (!kfile.empty()) {
      std::ofstream out(kfile);
      for (size_t i = 0; i < T; ++i)
        out << '(' << tris[i][0] << ',' << tris[i][1] << ',' << tris[i][2] << ")\t" << k[i] << '\n';
    }
    printf("{\"n\": %zu, \"m\": %zu, \"T\": %zu, \"Q\": %zu, \"K_max\": %ld, "
           "\"load_s\": %.6f, \"tri_s\": %.6f, \"tri_graph_s\": %.6f, \"qud_s\": %.6f, "
           "\"qud_graph_s\": %.6f, \"peel_s\": %.6f, \"threads\": %zu}\n",
           g.size_vertices(), g.size_edges(), T, Q, kmax, secs(t0, t1), secs(t1, t2),
           secs(t2, t3), secs(t3, t4), t_qg, t_peel, parts);
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
