// ORACLE — test infrastructure only. Command-line front end of the restatement:
//   oracle34 <dataset> [--mode host|ondemand] [--parts N] [--kfile PATH] [--bin PREFIX]
// Prints one JSON line (counts, K_max, phase seconds, threads) and optionally
// writes the k file: "(a,b,c)\tk" per triangle in (a,b) order. --bin writes
// PREFIX.k (int32 k) and PREFIX.tri (uint32 a,b,c) per triangle, raw
// little-endian in (a,b) order, for hashing results too large for text.
#include <omp.h>

#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

#include "oracle34.hpp"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: oracle34 <dataset> [--mode host|ondemand] [--parts N] [--kfile PATH]\n";
    return 2;
  }
  std::string file = argv[1], mode = "host", kfile, binprefix;
  size_t parts = static_cast<size_t>(omp_get_num_procs());
  for (int i = 2; i + 1 < argc; i += 2) {
    std::string a = argv[i];
    if (a == "--mode") mode = argv[i + 1];
    else if (a == "--parts") parts = std::stoul(argv[i + 1]);
    else if (a == "--kfile") kfile = argv[i + 1];
    else if (a == "--bin") binprefix = argv[i + 1];
  }
  try {
    auto t0 = std::chrono::steady_clock::now();
    bool pre = false;
    o34::Graph g = o34::load_dataset(file, &pre);
    if (!pre) g = o34::make_oriented(g);
    auto t1 = std::chrono::steady_clock::now();
    o34::Result r = mode == "ondemand" ? o34::ondemand(g, parts) : o34::host_path(g, parts);
    const double load = std::chrono::duration<double>(t1 - t0).count();
    const double total = r.t_tri + r.t_trigraph + r.t_qud + r.t_qudgraph + r.t_peel + r.t_incidence + r.t_support;
    if (!kfile.empty()) {
      std::ofstream out(kfile);
      for (size_t i = 0; i < r.T; ++i)
        out << '(' << r.tris[i][0] << ',' << r.tris[i][1] << ',' << r.tris[i][2] << ")\t" << r.k[i] << '\n';
    }
    if (!binprefix.empty()) {
      std::ofstream fk(binprefix + ".k", std::ios::binary), ft(binprefix + ".tri", std::ios::binary);
      fk.write(reinterpret_cast<const char*>(r.k.data()), static_cast<std::streamsize>(r.k.size() * 4));
      ft.write(reinterpret_cast<const char*>(r.tris.data()), static_cast<std::streamsize>(r.tris.size() * 12));
    }
    std::printf(
        "{\"mode\": \"%s\", \"n\": %u, \"m\": %zu, \"T\": %llu, \"Q\": %llu, \"K_max\": %d, "
        "\"rounds\": %llu, \"load_s\": %.6f, \"tri_s\": %.6f, \"tri_graph_s\": %.6f, \"qud_s\": %.6f, "
        "\"qud_graph_s\": %.6f, \"incidence_s\": %.6f, \"support_s\": %.6f, \"peel_s\": %.6f, "
        "\"decomp_s\": %.6f, \"threads\": %zu}\n",
        mode.c_str(), g.n, g.E.size(), static_cast<unsigned long long>(r.T),
        static_cast<unsigned long long>(r.Q), r.kmax, static_cast<unsigned long long>(r.rounds), load, r.t_tri,
        r.t_trigraph, r.t_qud, r.t_qudgraph, r.t_incidence, r.t_support, r.t_peel, total, parts);
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
