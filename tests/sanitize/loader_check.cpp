// Test harness for the sanitizer run of the dataset loaders (built by
// tests/test_sanitize_cpu.py with -fsanitize=address,undefined, next to
// gpu-nucleus34_amd/host/graph_io.cpp): for every path argument, loads it as
// the library does and prints one line,
//   OK <n> <m> <oriented> <fnv64 of O> <fnv64 of E>   or   ERR <message>
// then, for undirected results, takes them through the host orientation, the
// oriented-CSR check, the parts split and a .grh write / read round trip, and
// rebuilds the CSR from its own pairs; any disagreement prints BAD and exits 1.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../gpu-nucleus34_amd/host/graph_io.hpp"

namespace {

uint64_t fnv(const std::vector<uint32_t>& v) {
  uint64_t h = 1469598103934665603ull;
  for (uint32_t x : v)
    for (int b = 0; b < 4; ++b) {
      h ^= (x >> (8 * b)) & 0xFFu;
      h *= 1099511628211ull;
    }
  return h;
}

bool same(const nuc34::Csr& a, const nuc34::Csr& b) { return a.n == b.n && a.O == b.O && a.E == b.E; }

}  // namespace

int main(int argc, char** argv) {
  const std::string tmp = argc > 1 ? argv[1] : "/tmp";
  int bad = 0;
  for (int i = 2; i < argc; ++i) {
    try {
      const nuc34::LoadedGraph lg = nuc34::load_dataset(argv[i]);
      const nuc34::Csr& g = lg.g;
      std::printf("OK %u %llu %d %llu %llu\n", g.n, static_cast<unsigned long long>(g.m()), lg.oriented ? 1 : 0,
                  static_cast<unsigned long long>(fnv(g.O)), static_cast<unsigned long long>(fnv(g.E)));
      if (lg.oriented) continue;
      const nuc34::Csr o = nuc34::orient_host(g);
      const std::string why = nuc34::check_oriented(o);
      if (!why.empty() || 2 * o.m() != g.m()) {
        std::printf("BAD orientation of %s: %s\n", argv[i], why.c_str());
        ++bad;
      }
      const auto parts = nuc34::split_parts(o, 3);
      if (parts.size() != 3 || parts[0] != 0) {
        std::printf("BAD parts of %s\n", argv[i]);
        ++bad;
      }
      const std::string grh = tmp + "/roundtrip.grh";
      nuc34::write_grh(grh, o);
      const nuc34::LoadedGraph back = nuc34::load_dataset(grh);
      if (!back.oriented || !same(back.g, o)) {
        std::printf("BAD .grh round trip of %s\n", argv[i]);
        ++bad;
      }
      std::vector<uint32_t> s, d;
      for (uint32_t u = 0; u < g.n; ++u)
        for (uint32_t k = g.O[u]; k < g.O[u + 1]; ++k) {
          s.push_back(u);
          d.push_back(g.E[k]);
        }
      const nuc34::Csr p = nuc34::csr_from_pairs(s.data(), d.data(), s.size());
      // isolated vertices above the largest id are not in a pairs rebuild
      if (p.E != g.E || p.n > g.n) {
        std::printf("BAD pairs rebuild of %s\n", argv[i]);
        ++bad;
      }
    } catch (const std::exception& e) {
      std::printf("ERR %s\n", e.what());
    }
  }
  return bad ? 1 : 0;
}
