// nuc34_gen: writes seeded synthetic graphs as bare .edges files.
//   nuc34_gen er-p   <n> <p> <seed> <out>
//   nuc34_gen er-deg <n> <avg_deg> <seed> <out>
//   nuc34_gen rmat   <scale> <edge_factor> <a> <b> <c> <seed> <out> [<cliques> <size>]
//   nuc34_gen ba     <n> <m> <seed> <out>
//   nuc34_gen batch  <seed> <index> <out>        one graph of the mixed batch
//   nuc34_gen preset <name> <out>                er1000 | rmat16 | rmat20k30 | rmat18
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "edge_writer.hpp"
#include "gen.hpp"
#include "presets.hpp"

using namespace nuc34;

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cerr << "usage: nuc34_gen er-p|er-deg|rmat|ba|batch|preset ...\n";
    return 2;
  }
  const std::string kind = argv[1];
  try {
    gen::Edges e;
    std::string out;
    if (kind == "er-p" && argc == 6) {
      e = gen::erdos_renyi_p(std::stoul(argv[2]), std::stod(argv[3]), std::stoull(argv[4]));
      out = argv[5];
    } else if (kind == "er-deg" && argc == 6) {
      gen::Rng r(std::stoull(argv[4]));
      e = gen::erdos_renyi_deg(std::stoul(argv[2]), std::stod(argv[3]), r);
      out = argv[5];
    } else if (kind == "rmat" && (argc == 9 || argc == 11)) {
      const uint64_t seed = std::stoull(argv[7]);
      e = gen::rmat(std::stoul(argv[2]), std::stoul(argv[3]), std::stod(argv[4]), std::stod(argv[5]),
                    std::stod(argv[6]), seed);
      out = argv[8];
      if (argc == 11) add_planted(e, std::stoul(argv[2]), seed, std::stoul(argv[9]), std::stoul(argv[10]));
    } else if (kind == "ba" && argc == 6) {
      gen::Rng r(std::stoull(argv[4]));
      e = gen::barabasi_albert(std::stoul(argv[2]), std::stoul(argv[3]), r);
      out = argv[5];
    } else if (kind == "batch" && argc == 5) {
      e = gen::batch_graph(std::stoull(argv[2]), std::stoull(argv[3]), kBatchNLo, kBatchNHi);
      out = argv[4];
    } else if (kind == "preset" && argc == 4) {
      e = preset_edges(argv[2]);
      out = argv[3];
    } else {
      std::cerr << "bad arguments for " << kind << "\n";
      return 2;
    }
    write_edges_file(out, e);
    std::printf("{\"file\": \"%s\", \"pairs\": %zu}\n", out.c_str(), e.size());
  } catch (const std::exception& ex) {
    std::cerr << "error: " << ex.what() << "\n";
    return 1;
  }
  return 0;
}
