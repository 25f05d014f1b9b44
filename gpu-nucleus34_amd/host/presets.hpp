// Named synthetic workloads:
//   er1000     G(n=1000, p=0.1), seed 0
//   rmat16     R-MAT scale 16, edge factor 16, (a,b,c,d) = (.57,.19,.19,.05), seed 1
//   rmat20k30  R-MAT scale 20, edge factor 16, same (a,b,c,d), seed 2, plus 64
//              planted cliques of 30 vertices
//   rmat18     R-MAT scale 18, edge factor 16, same (a,b,c,d), seed 1
// and the mixed batch: n log-uniform in [200, 50000].
#pragma once
#include <stdexcept>
#include <string>

#include "gen.hpp"

namespace nuc34 {

constexpr uint32_t kBatchNLo = 200;
constexpr uint32_t kBatchNHi = 50000;

// Planted cliques draw from their own stream so the R-MAT samples are the same
// with and without them.
inline void add_planted(gen::Edges& e, unsigned scale, uint64_t seed, uint32_t count, uint32_t size) {
  gen::Rng r(seed * 0x2545F4914F6CDD1Dull + 0x632BE59BD9B4E019ull);
  gen::plant_cliques(e, 1u << scale, count, size, r);
}

inline gen::Edges preset_edges(const std::string& name) {
  if (name == "er1000") return gen::erdos_renyi_p(1000, 0.1, 0);
  if (name == "rmat16") return gen::rmat(16, 16, 0.57, 0.19, 0.19, 1);
  if (name == "rmat18") return gen::rmat(18, 16, 0.57, 0.19, 0.19, 1);
  if (name == "rmat20k30") {
    auto e = gen::rmat(20, 16, 0.57, 0.19, 0.19, 2);
    add_planted(e, 20, 2, 64, 30);
    return e;
  }
  throw std::invalid_argument("unknown preset " + name);
}

}  // namespace nuc34
