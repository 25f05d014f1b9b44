// Seeded synthetic graph generators: Erdos-Renyi (per-pair Bernoulli and G(n,M)),
// R-MAT, Barabasi-Albert and planted cliques, plus the mixed batch used for
// many-graph runs. Integer-only sampling on a splitmix64-seeded xoshiro256**,
// so every platform draws the same edges. Edge lists are raw (u, v) pairs:
// duplicates, reversed duplicates and self-loops are left for the loader, which
// removes them exactly as the reference's graph_t::from_edges does
// (nucleus/graph/graph.hpp:356-389).
#pragma once
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>

namespace nuc34 {
namespace gen {

inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& w : s) w = splitmix64(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // Unbiased integer in [0, n) (Lemire's multiply-and-reject).
  uint64_t below(uint64_t n) {
    if (n <= 1) return 0;
    for (;;) {
      unsigned __int128 m = static_cast<unsigned __int128>(next()) * n;
      uint64_t lo = static_cast<uint64_t>(m);
      if (lo >= n || lo >= (0 - n) % n) return static_cast<uint64_t>(m >> 64);
    }
  }
  // Uniform integer in [lo, hi].
  uint64_t range(uint64_t lo, uint64_t hi) { return lo + below(hi - lo + 1); }
};

using Edges = std::vector<std::pair<uint32_t, uint32_t>>;

// Probability p in [0,1] as a 64-bit acceptance threshold for next().
inline uint64_t prob_threshold(double p) {
  if (p <= 0) return 0;
  if (p >= 1) return ~0ull;
  return static_cast<uint64_t>(std::ldexp(static_cast<long double>(p), 64));
}

// G(n, p): one Bernoulli draw per unordered pair u < v, in (u, v) order.
inline Edges erdos_renyi_p(uint32_t n, double p, uint64_t seed) {
  Rng r(seed);
  const uint64_t th = prob_threshold(p);
  Edges e;
  for (uint32_t u = 0; u < n; ++u)
    for (uint32_t v = u + 1; v < n; ++v)
      if (r.next() < th) e.emplace_back(u, v);
  return e;
}

// G(n, M) with M = round(n * avg_deg / 2) uniformly drawn pairs u != v
// (duplicates are removed by the loader, so the mean degree is marginally lower).
inline Edges erdos_renyi_deg(uint32_t n, double avg_deg, Rng& r) {
  Edges e;
  if (n < 2) return e;
  const uint64_t M = static_cast<uint64_t>(std::llround(n * avg_deg / 2.0));
  e.reserve(M);
  for (uint64_t i = 0; i < M; ++i) {
    uint32_t u = static_cast<uint32_t>(r.below(n));
    uint32_t v = static_cast<uint32_t>(r.below(n - 1));
    if (v >= u) ++v;
    e.emplace_back(u, v);
  }
  return e;
}

// R-MAT: edge_factor * 2^scale samples, each descending `scale` quadrant levels
// with probabilities (a, b, c, d = 1 - a - b - c); no noise, no relabelling.
inline Edges rmat(unsigned scale, unsigned edge_factor, double a, double b, double c,
                  uint64_t seed) {
  if (scale > 31) throw std::invalid_argument("rmat scale must be <= 31");
  Rng r(seed);
  const uint64_t ta = static_cast<uint64_t>(std::ldexp(a, 32));
  const uint64_t tb = static_cast<uint64_t>(std::ldexp(a + b, 32));
  const uint64_t tc = static_cast<uint64_t>(std::ldexp(a + b + c, 32));
  const uint64_t count = static_cast<uint64_t>(edge_factor) << scale;
  Edges e;
  e.reserve(count);
  for (uint64_t i = 0; i < count; ++i) {
    uint32_t u = 0, v = 0;
    for (unsigned l = 0; l < scale; ++l) {
      const uint64_t x = r.next() >> 32;
      u <<= 1;
      v <<= 1;
      if (x < ta) {
      } else if (x < tb) {
        v |= 1;
      } else if (x < tc) {
        u |= 1;
      } else {
        u |= 1;
        v |= 1;
      }
    }
    e.emplace_back(u, v);
  }
  return e;
}

// Barabasi-Albert: a clique on m+1 seed vertices, then each new vertex attaches
// to m distinct targets drawn proportionally to degree (endpoint-list sampling).
inline Edges barabasi_albert(uint32_t n, uint32_t m, Rng& r) {
  Edges e;
  if (m == 0 || n <= m + 1) {
    for (uint32_t u = 0; u < n; ++u)
      for (uint32_t v = u + 1; v < n; ++v) e.emplace_back(u, v);
    return e;
  }
  std::vector<uint32_t> ends;
  ends.reserve(2ull * n * m);
  for (uint32_t u = 0; u <= m; ++u)
    for (uint32_t v = u + 1; v <= m; ++v) {
      e.emplace_back(u, v);
      ends.push_back(u);
      ends.push_back(v);
    }
  std::vector<uint32_t> pick;
  for (uint32_t v = m + 1; v < n; ++v) {
    pick.clear();
    while (pick.size() < m) {
      uint32_t t = ends[r.below(ends.size())];
      bool dup = false;
      for (uint32_t q : pick) dup |= (q == t);
      if (!dup) pick.push_back(t);
    }
    for (uint32_t t : pick) {
      e.emplace_back(v, t);
      ends.push_back(v);
      ends.push_back(t);
    }
  }
  return e;
}

// Adds `count` cliques of `size` vertices, each on distinct vertices drawn
// uniformly from [0, n).
inline void plant_cliques(Edges& e, uint32_t n, uint32_t count, uint32_t size, Rng& r) {
  if (size > n) size = n;
  std::vector<uint32_t> vs;
  for (uint32_t c = 0; c < count; ++c) {
    vs.clear();
    while (vs.size() < size) {
      uint32_t x = static_cast<uint32_t>(r.below(n));
      bool dup = false;
      for (uint32_t q : vs) dup |= (q == x);
      if (!dup) vs.push_back(x);
    }
    for (uint32_t i = 0; i < size; ++i)
      for (uint32_t j = i + 1; j < size; ++j) e.emplace_back(vs[i], vs[j]);
  }
}

// One graph of the mixed batch: graph `index` of a batch drawn from `seed`.
// n is log-uniform in [n_lo, n_hi]; the family is index mod 3:
//   0: G(n, M) with mean degree uniform in [8, 64]
//   1: Barabasi-Albert with m uniform in [4, 16]
//   2: G(n, M) with mean degree 8 plus 1 + n/5000 planted cliques of size
//      uniform in [10, 40]
struct BatchSpec {
  uint32_t n;
  int family;
  double param;
};

inline uint64_t batch_graph_seed(uint64_t seed, uint64_t index) {
  uint64_t s = seed * 0x100000001B3ull + index * 0x9E3779B97F4A7C15ull + 1;
  return splitmix64(s);
}

inline Edges batch_graph(uint64_t seed, uint64_t index, uint32_t n_lo, uint32_t n_hi,
                         BatchSpec* spec_out = nullptr) {
  Rng r(batch_graph_seed(seed, index));
  // log-uniform n from a 53-bit uniform fraction
  const double u = static_cast<double>(r.next() >> 11) * (1.0 / 9007199254740992.0);
  uint32_t n = static_cast<uint32_t>(std::floor(n_lo * std::exp(u * std::log(double(n_hi) / n_lo))));
  if (n < n_lo) n = n_lo;
  if (n > n_hi) n = n_hi;
  BatchSpec spec{n, static_cast<int>(index % 3), 0};
  Edges e;
  if (spec.family == 0) {
    spec.param = static_cast<double>(r.range(8, 64));
    e = erdos_renyi_deg(n, spec.param, r);
  } else if (spec.family == 1) {
    spec.param = static_cast<double>(r.range(4, 16));
    e = barabasi_albert(n, static_cast<uint32_t>(spec.param), r);
  } else {
    spec.param = 8;
    e = erdos_renyi_deg(n, 8.0, r);
    const uint32_t cliques = 1 + n / 5000;
    for (uint32_t c = 0; c < cliques; ++c) {
      uint32_t size = static_cast<uint32_t>(r.range(10, 40));
      plant_cliques(e, n, 1, size, r);
    }
  }
  if (spec_out) *spec_out = spec;
  return e;
}

}  // namespace gen
}  // namespace nuc34
