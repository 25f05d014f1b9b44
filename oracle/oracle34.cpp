// ORACLE — test infrastructure only (see oracle34.hpp for the rule and the
// per-function citations of the reference code restated here).
#include "oracle34.hpp"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <limits>
#include <numeric>
#include <parallel/algorithm>
#include <sstream>
#include <stdexcept>

namespace o34 {

namespace {
using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

Graph from_sorted_pairs(std::vector<std::array<uint32_t, 2>>& edges, uint32_t max_node) {
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
  Graph g;
  g.n = max_node + 1;
  g.O.assign(static_cast<size_t>(g.n) + 1, 0);
  g.E.resize(edges.size());
  for (size_t i = 0; i < edges.size(); ++i) {
    g.O[edges[i][0] + 1]++;
    g.E[i] = edges[i][1];
  }
  for (uint32_t v = 0; v < g.n; ++v) g.O[v + 1] += g.O[v];
  return g;
}

// graph.hpp:286-411, with the stream semantics spelled out: a line whose first
// extraction fails, or that has no second token, is skipped (the reference
// reads an uninitialised destination there); a second token that is not a
// number reads as 0, as a failed num_get conversion stores 0.
Graph from_edges(const std::string& file, bool mtx) {
  std::ifstream fs(file);
  if (!fs.is_open()) throw std::runtime_error("Dataset " + file + " is not accessible!");
  std::string line;
  while (std::getline(fs, line)) {
    if (std::string("#%").find(line[0]) == std::string::npos) break;
  }
  if (mtx) {
    std::istringstream iss(line);
    std::vector<std::string> tok{std::istream_iterator<std::string>{iss},
                                 std::istream_iterator<std::string>{}};
    size_t nV = 0, nE = 0;
    if (tok.size() == 2) {
      nV = std::stoul(tok[0], nullptr, 0);
      nE = std::stoul(tok[1], nullptr, 0);
    } else if (tok.size() == 3) {
      nV = std::max(std::stoul(tok[0], nullptr, 0), std::stoul(tok[1], nullptr, 0));
      nE = std::stoul(tok[2], nullptr, 0);
    } else {
      throw std::runtime_error("Wrong header for mtx format!");
    }
    if (nE > std::numeric_limits<uint32_t>::max())
      throw std::runtime_error("Edges size are larger than defined type.");
    if (nV > std::numeric_limits<uint32_t>::max())
      throw std::runtime_error("Vertices id are larger than defined type.");
  }
  std::vector<std::array<uint32_t, 2>> edges;
  uint32_t max_node = 0;
  while (std::getline(fs, line)) {
    std::stringstream ss(line);
    uint32_t s = 0, d = 0;
    if (!(ss >> s)) continue;
    ss >> std::ws;
    if (ss.eof()) continue;
    ss >> d;  // on a failed conversion d == 0, on overflow d == UINT32_MAX
    if (s != d) {
      max_node = std::max(max_node, std::max(s, d));
      edges.push_back({s, d});
      edges.push_back({d, s});
    }
  }
  return from_sorted_pairs(edges, max_node);
}

Graph read_grh(const std::string& file) {
  std::ifstream f(file, std::ios::binary);
  if (!f.is_open()) throw std::runtime_error("Can't open file for reading.");
  size_t sv = 0, nV = 0, nE = 0;
  f.read(reinterpret_cast<char*>(&sv), sizeof(size_t));
  if (sv != sizeof(uint32_t)) throw std::runtime_error("Incompatible types in dataset.");
  f.read(reinterpret_cast<char*>(&nV), sizeof(size_t));
  f.read(reinterpret_cast<char*>(&nE), sizeof(size_t));
  Graph g;
  g.n = static_cast<uint32_t>(nV);
  std::vector<uint32_t> D(nV);
  g.E.resize(nE);
  f.read(reinterpret_cast<char*>(D.data()), static_cast<std::streamsize>(nV * 4));
  f.read(reinterpret_cast<char*>(g.E.data()), static_cast<std::streamsize>(nE * 4));
  g.O.assign(nV + 1, 0);
  for (size_t v = 0; v < nV; ++v) g.O[v + 1] = g.O[v] + D[v];
  return g;
}
}  // namespace

Graph load_dataset(const std::string& path, bool* oriented) {
  const auto p = path.find_last_of('.');
  if (p == std::string::npos) throw std::runtime_error("Unknown dataset format: (no extension)");
  const std::string ext = path.substr(p);
  if (oriented) *oriented = false;
  if (ext == ".mtx") return from_edges(path, true);
  if (ext == ".edges" || ext == ".txt") return from_edges(path, false);
  if (ext == ".grh") {
    if (oriented) *oriented = true;
    return read_grh(path);
  }
  throw std::runtime_error("Unknown dataset format: " + ext);
}

Graph make_oriented(const Graph& und) {
  Graph o;
  o.n = und.n;
  o.O.assign(static_cast<size_t>(und.n) + 1, 0);
  auto D = [&](uint32_t v) { return und.O[v + 1] - und.O[v]; };
  auto vertex_less = [&](uint32_t s, uint32_t d) { return D(s) < D(d) || (D(s) == D(d) && s < d); };
  for (uint32_t u = 0; u < und.n; ++u) {
    for (uint32_t i = und.O[u]; i < und.O[u + 1]; ++i)
      if (vertex_less(u, und.E[i])) o.E.push_back(und.E[i]);
    o.O[u + 1] = static_cast<uint32_t>(o.E.size());
  }
  return o;
}

// graph.hpp:581-605 over an offset vector of the reference's length n (no
// trailing m): boundaries default to n-1, the total is O[n-1].
static std::vector<size_t> split_offsets(const std::vector<uint32_t>& O_np1, size_t nverts, size_t parts) {
  if (parts == 0) parts = static_cast<size_t>(omp_get_num_procs());
  std::vector<size_t> sp(parts, nverts ? nverts - 1 : 0);
  sp[0] = 0;
  const double scale = 1.0 / static_cast<double>(parts);
  const double deg_sum = nverts ? static_cast<double>(O_np1[nverts - 1]) : 0.0;
  for (size_t i = 0, j = 1; i < nverts && j < parts; ++i)
    if (O_np1[i] / deg_sum > static_cast<double>(j) * scale) sp[j++] = i;
  return sp;
}

std::vector<size_t> split_parts(const Graph& g, size_t parts) {
  return split_offsets(g.O, g.n, parts);
}

std::vector<Tri> triangles_par(const Graph& g, const std::vector<size_t>& sp) {
  std::vector<Tri> all;
  const size_t P = sp.size();
  omp_set_dynamic(0);
  omp_set_num_threads(static_cast<int>(P));
#pragma omp parallel
  {
    const size_t th = static_cast<size_t>(omp_get_thread_num());
    const size_t bg = sp[th];
    const size_t ed = (th + 1 < P) ? sp[th + 1] : g.n;
    std::vector<Tri> local;
    for (size_t u = bg; u < ed; ++u) {
      const uint32_t d1 = g.O[u], d2 = g.O[u + 1];
      for (uint32_t j = d1; j < d2; ++j) {
        const uint32_t v = g.E[j];
        uint32_t i1 = d1, j1 = g.O[v];
        const uint32_t j2 = g.O[v + 1];
        while (i1 < d2 && j1 < j2) {
          const uint32_t p = g.E[i1], q = g.E[j1];
          if (p < q) {
            ++i1;
          } else {
            if (!(q < p)) {
              local.push_back({i1, j, j1});
              ++i1;
            }
            ++j1;
          }
        }
      }
    }
#pragma omp critical
    all.insert(all.end(), local.begin(), local.end());
  }
  return all;
}

TriGraph::TriGraph(std::vector<Tri>& tris, size_t m) {
  // (a,b) keys are unique, so any correct sort gives the same order; the
  // parallel one stands for the reference's std::execution::par sort
  __gnu_parallel::sort(tris.begin(), tris.end(), [](const Tri& x, const Tri& y) {
    return x[0] != y[0] ? x[0] < y[0] : x[1] < y[1];
  });
  D.assign(m, 0);
  E.resize(tris.size());
  for (size_t i = 0; i < tris.size(); ++i) {
    D[tris[i][0]]++;
    E[i] = {tris[i][1], tris[i][2]};
  }
  O.assign(m + 1, 0);
  for (size_t i = 0; i < m; ++i) O[i + 1] = O[i] + D[i];
}

std::vector<Qud> fourcliques_par(const TriGraph& tg, const std::vector<size_t>& sp) {
  std::vector<Qud> all;
  const size_t P = sp.size();
  const size_t nv = tg.D.size();
  omp_set_dynamic(0);
  omp_set_num_threads(static_cast<int>(P));
#pragma omp parallel
  {
    const size_t th = static_cast<size_t>(omp_get_thread_num());
    const size_t bg = sp[th];
    const size_t ed = (th + 1 < P) ? sp[th + 1] : nv;
    std::vector<Qud> local;
    for (size_t i = bg; i < ed; ++i) {
      const uint32_t d1 = tg.O[i], d2 = tg.O[i + 1];
      for (uint32_t j = d1; j < d2; ++j) {
        uint32_t i1 = d1;
        const uint32_t bj = tg.E[j][0];
        uint32_t j1 = tg.O[bj];
        const uint32_t j2 = tg.O[bj + 1];
        while (i1 < d2 && j1 < j2) {
          const uint32_t p = tg.E[i1][0];
          const uint32_t o1 = j1;
          const uint32_t q = tg.E[o1][0];
          if (p < q) {
            ++i1;
          } else {
            if (!(q < p)) {
              const uint32_t ci = tg.E[i1][1];
              uint32_t k = tg.O[ci];
              const uint32_t k2 = tg.O[ci + 1];
              for (; k < k2; ++k) {
                if (tg.E[j][1] == tg.E[k][1]) {
                  local.push_back({i1, o1, j, k});
                  ++i1;
                  break;
                }
              }
              if (k >= k2) throw std::logic_error("fourth triangle of a four-clique not found");
            }
            ++j1;
          }
        }
      }
    }
#pragma omp critical
    all.insert(all.end(), local.begin(), local.end());
  }
  return all;
}

namespace {

// graph_t<U,V,3> undirected ctor: each K4 lists, under each member, the other
// three members in tuple order.
struct QudGraph {
  std::vector<uint32_t> O, D;
  std::vector<std::array<uint32_t, 3>> E;
  QudGraph(const std::vector<Qud>& q, size_t T) {
    D.assign(T, 0);
#pragma omp parallel for
    for (size_t i = 0; i < q.size(); ++i)
      for (int s = 0; s < 4; ++s) __sync_fetch_and_add(&D[q[i][s]], 1u);
    O.assign(T + 1, 0);
    for (size_t i = 0; i < T; ++i) O[i + 1] = O[i] + D[i];
    std::vector<uint32_t> TO(O.begin(), O.end() - 1);
    E.resize(4 * q.size());
#pragma omp parallel for
    for (size_t i = 0; i < q.size(); ++i) {
      for (int nn = 0; nn < 4; ++nn) {
        const uint32_t id = q[i][nn];
        const uint32_t slot = __sync_fetch_and_add(&TO[id], 1u);
        int dim = 0;
        for (int s = 0; s < 4; ++s)
          if (q[i][s] != id) E[slot][dim++] = q[i][s];
      }
    }
  }
};

// bucket_ex<T=uint32_t, L=int32_t> (bucket_ex.hpp:7-360).
class Bucket {
 public:
  explicit Bucket(const std::vector<uint32_t>& vals) {
    if (vals.empty()) return;
    const uint32_t max_value = *std::max_element(vals.begin(), vals.end());
    std::vector<uint32_t> cnt(static_cast<size_t>(max_value) + 1, 0);
    for (uint32_t v : vals) cnt[v]++;
    offset_.resize(cnt.size());
    std::inclusive_scan(cnt.begin(), cnt.end(), offset_.begin());
    perm_.resize(vals.size());
    std::iota(perm_.begin(), perm_.end(), 0u);
    std::sort(perm_.begin(), perm_.end(), [&](uint32_t a, uint32_t b) { return vals[a] < vals[b]; });
    perm_ex_.resize(vals.size());
    for (size_t i = 0; i < perm_.size(); ++i) perm_ex_[perm_[i]] = static_cast<uint32_t>(i);
    values_.resize(vals.size());
    for (size_t i = 0; i < vals.size(); ++i) values_[i] = vals[i] == 0 ? -1 : static_cast<int32_t>(vals[i]);
    for (; min_id_ < perm_.size(); ++min_id_)
      if (vals[perm_[min_id_]]) break;
  }

  std::vector<int32_t> support() const {
    std::vector<int32_t> s(values_.size());
    for (size_t i = 0; i < s.size(); ++i) s[i] = values_[i] + 1;
    return s;
  }

  int32_t peeling_par(const QudGraph& qg, uint64_t* batches) {
    in_batch_.assign(values_.size(), 0);
    std::vector<int> dec_req(values_.size(), 0);
    std::vector<uint32_t> ids;
    uint64_t nb = 0;
    while (pop_min(ids)) {
      ++nb;
      if (ids.size() < 256) {
        for (uint32_t id : ids) {
          for (uint32_t d = qg.O[id]; d < qg.O[id + 1]; ++d) peel(qg.E[d]);
          values_[id]--;
        }
        continue;
      }
      for (uint32_t id : ids) in_batch_[id] = 1;
      std::vector<uint32_t> touched;
#pragma omp parallel
      {
        std::vector<uint32_t> lt;
#pragma omp for schedule(static)
        for (size_t i = 0; i < ids.size(); ++i) {
          const uint32_t id = ids[i];
          for (uint32_t d = qg.O[id]; d < qg.O[id + 1]; ++d) peel_count(id, qg.E[d], dec_req, lt);
        }
#pragma omp critical
        touched.insert(touched.end(), lt.begin(), lt.end());
      }
      for (uint32_t key : touched) {
        int& req = dec_req[key];
        if (!req) continue;
        const int32_t allowed = values_[key] - last_;
        if (allowed > 0) {
          const int32_t times = std::min<int32_t>(req, allowed);
          for (int32_t t = 0; t < times; ++t) decrease(key);
        }
        req = 0;
      }
      for (uint32_t id : ids) {
        values_[id]--;
        in_batch_[id] = 0;
      }
    }
    if (batches) *batches = nb;
    return last_;
  }

 private:
  void decrease(uint32_t key) {
    int32_t& val = values_[key];
    const uint32_t pos = perm_ex_[key];
    uint32_t& off = offset_[static_cast<size_t>(val - 1)];
    const uint32_t other = perm_[off];
    std::swap(perm_[pos], perm_[off]);
    std::swap(perm_ex_[key], perm_ex_[other]);
    ++off;
    --val;
  }
  void peel(const std::array<uint32_t, 3>& k) {
    for (uint32_t x : k)
      if (values_[x] < last_) return;
    for (uint32_t x : k)
      if (values_[x] != last_) decrease(x);
  }
  void peel_count(uint32_t id, const std::array<uint32_t, 3>& k, std::vector<int>& dec_req,
                  std::vector<uint32_t>& lt) {
    for (uint32_t x : k)
      if (values_[x] < last_) return;
    for (uint32_t x : k)
      if (in_batch_[x] && x < id) return;
    for (uint32_t x : k)
      if (!in_batch_[x])
        if (__sync_fetch_and_add(&dec_req[x], 1) == 0) lt.push_back(x);
  }
  bool pop_min(std::vector<uint32_t>& keys) {
    if (min_id_ >= perm_.size()) return false;
    last_ = values_[perm_[min_id_]];
    keys.resize(offset_[static_cast<size_t>(last_)] - min_id_);
    for (size_t i = 0; i < keys.size(); ++i, ++min_id_) keys[i] = perm_[min_id_];
    return true;
  }

  size_t min_id_ = 0;
  int32_t last_ = 0;
  std::vector<int32_t> values_;
  std::vector<uint32_t> perm_, perm_ex_, offset_;
  std::vector<char> in_batch_;
};

}  // namespace

Result host_path(const Graph& g, size_t parts) {
  Result r;
  const size_t P = parts ? parts : static_cast<size_t>(omp_get_num_procs());
  auto t0 = clk::now();
  r.tris = triangles_par(g, split_parts(g, P));
  auto t1 = clk::now();
  TriGraph tg(r.tris, g.E.size());
  auto sp2 = split_offsets(tg.O, tg.D.size(), P);
  auto t2 = clk::now();
  std::vector<Qud> quds = fourcliques_par(tg, sp2);
  auto t3 = clk::now();
  r.T = r.tris.size();
  r.Q = quds.size();
  r.t_tri = secs(t0, t1);
  r.t_trigraph = secs(t1, t2);
  r.t_qud = secs(t2, t3);
  if (!r.T) return r;
  QudGraph qg(quds, r.T);
  std::vector<Qud>().swap(quds);
  auto t4 = clk::now();
  Bucket b(qg.D);
  omp_set_num_threads(static_cast<int>(P));
  const int32_t last = b.peeling_par(qg, &r.rounds);
  auto t5 = clk::now();
  r.k = b.support();
  r.support.assign(qg.D.begin(), qg.D.end());
  r.kmax = r.Q ? last : 0;
  r.t_qudgraph = secs(t3, t4);
  r.t_peel = secs(t4, t5);
  return r;
}

Result ondemand(const Graph& g, size_t parts) {
  // ORACLE34_PROGRESS set: one stderr line per phase and every 256 peel rounds
  // (long runs at full size stay visibly alive)
  const bool progress = std::getenv("ORACLE34_PROGRESS") != nullptr;
  const auto t_start = clk::now();
  auto note = [&](const char* what, uint64_t a, uint64_t b) {
    if (progress)
      std::fprintf(stderr, "[oracle34 ondemand] %.1f s: %s %llu %llu\n", secs(t_start, clk::now()), what,
                   static_cast<unsigned long long>(a), static_cast<unsigned long long>(b));
  };
  Result r;
  const size_t P = parts ? parts : static_cast<size_t>(omp_get_num_procs());
  auto t0 = clk::now();
  r.tris = triangles_par(g, split_parts(g, P));
  TriGraph tg(r.tris, g.E.size());  // sorts r.tris into (a,b) order
  auto t1 = clk::now();
  r.t_tri = secs(t0, t1);
  const size_t T = r.tris.size();
  r.T = T;
  note("triangles", T, 0);
  if (!T) return r;
  const size_t m = g.E.size();
  omp_set_num_threads(static_cast<int>(P));
  // src[e]: owner vertex of oriented edge e (cuNucleus34_ondemand.cu:514-518)
  std::vector<uint32_t> src(m);
  for (uint32_t u = 0; u < g.n; ++u)
    for (uint32_t i = g.O[u]; i < g.O[u + 1]; ++i) src[i] = u;
  // Incidence: under each edge, (third vertex, triangle) sorted by third vertex
  // (build_incidence_entries + sort_by_key, cuNucleus34_ondemand.cu:92-116, 524-549).
  std::vector<uint64_t> incO(m + 1, 0);
  for (const Tri& t : r.tris) {
    incO[t[0] + 1]++;
    incO[t[1] + 1]++;
    incO[t[2] + 1]++;
  }
  for (size_t e = 0; e < m; ++e) incO[e + 1] += incO[e];
  std::vector<uint64_t> inc(3 * T);
  {
    std::vector<uint64_t> cur(incO.begin(), incO.end() - 1);
    for (size_t i = 0; i < T; ++i) {
      const Tri& t = r.tris[i];
      const uint32_t u = src[t[1]], v = g.E[t[1]], w = g.E[t[2]];
      inc[cur[t[1]]++] = (static_cast<uint64_t>(w) << 32) | i;  // (u,v): third w
      inc[cur[t[0]]++] = (static_cast<uint64_t>(v) << 32) | i;  // (u,w): third v
      inc[cur[t[2]]++] = (static_cast<uint64_t>(u) << 32) | i;  // (v,w): third u
    }
  }
#pragma omp parallel for schedule(dynamic, 1024)
  for (size_t e = 0; e < m; ++e) std::sort(inc.begin() + incO[e], inc.begin() + incO[e + 1]);
  auto t2 = clk::now();
  r.t_incidence = secs(t1, t2);
  note("incidence", 3 * T, 0);
  // 3-way zipper over the lists of edges b, a, c (for_each_clique_of, 136-170).
  auto for_each_clique = [&](size_t ti, auto&& fn) {
    const Tri& t = r.tris[ti];
    uint64_t i = incO[t[1]], iE = incO[t[1] + 1];
    uint64_t j = incO[t[0]], jE = incO[t[0] + 1];
    uint64_t k = incO[t[2]], kE = incO[t[2] + 1];
    while (i < iE && j < jE && k < kE) {
      const uint32_t x1 = static_cast<uint32_t>(inc[i] >> 32);
      const uint32_t x2 = static_cast<uint32_t>(inc[j] >> 32);
      const uint32_t x3 = static_cast<uint32_t>(inc[k] >> 32);
      if (x1 == x2 && x2 == x3) {
        fn(static_cast<uint32_t>(inc[i]), static_cast<uint32_t>(inc[j]), static_cast<uint32_t>(inc[k]));
        ++i;
        ++j;
        ++k;
      } else {
        const uint32_t mx = std::max(x1, std::max(x2, x3));
        if (x1 < mx) ++i;
        if (x2 < mx) ++j;
        if (x3 < mx) ++k;
      }
    }
  };
  std::vector<int32_t> DD(T);
  uint64_t total = 0;
  std::atomic<uint64_t> counted{0};
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total)
  for (size_t t = 0; t < T; ++t) {
    int32_t c = 0;
    for_each_clique(t, [&](uint32_t, uint32_t, uint32_t) { ++c; });
    DD[t] = c;
    total += static_cast<uint64_t>(c);
    if (progress && ((counted.fetch_add(1, std::memory_order_relaxed) + 1) & ((1u << 24) - 1)) == 0)
      note("support triangles done", counted.load(), T);
  }
  auto t3 = clk::now();
  r.t_support = secs(t2, t3);
  r.support = DD;
  r.Q = total / 4;
  note("support", r.Q, 0);
  if (!r.Q) {
    r.k.assign(T, 0);
    return r;
  }
  // Level-synchronous peel (cuNucleus34_ondemand.cu:599-690): level = max(level,
  // min alive support); dying = alive with support <= level; a K4 belongs to its
  // first member to die, same-round ties to the lowest index; owners debit the
  // surviving members. Rounds after the first of a level only need the
  // triangles whose support fell from level+1 to level in the previous round:
  // every other alive triangle still has support > level, so this is the same
  // dying set the reference's full rescan finds, round for round.
  std::vector<int32_t> kv(T, -1);
  std::vector<uint32_t> round_of(T, 0xFFFFFFFFu);
  std::vector<uint32_t> alive(T), dying, crossed;
  std::iota(alive.begin(), alive.end(), 0u);
  int32_t level = 0;
  size_t peeled = 0;
  uint32_t round = 0;
  const int nth = omp_get_max_threads();
  std::vector<std::vector<uint32_t>> local(static_cast<size_t>(nth));
  auto last_note = clk::now();
  while (peeled < T) {
    if (crossed.empty()) {
      int32_t mn = std::numeric_limits<int32_t>::max();
#pragma omp parallel for reduction(min : mn)
      for (size_t i = 0; i < alive.size(); ++i) {
        const uint32_t t = alive[i];
        if (kv[t] == -1) mn = std::min(mn, std::max(DD[t], 0));
      }
      level = std::max(level, mn);
      dying.clear();
      std::vector<uint32_t> keep;
      for (uint32_t t : alive) {
        if (kv[t] != -1) continue;
        if (DD[t] <= level) dying.push_back(t);
        else keep.push_back(t);
      }
      alive.swap(keep);
    } else {
      dying.swap(crossed);
      crossed.clear();
    }
    for (uint32_t t : dying) {
      kv[t] = level;
      round_of[t] = round;
    }
    for (auto& l : local) l.clear();
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t di = 0; di < dying.size(); ++di) {
      const uint32_t t = dying[di];
      auto& out = local[static_cast<size_t>(omp_get_thread_num())];
      for_each_clique(t, [&](uint32_t p1, uint32_t p2, uint32_t p3) {
        const uint32_t p[3] = {p1, p2, p3};
        for (uint32_t x : p) {
          if (kv[x] == -1) continue;
          if (round_of[x] < round) return;
          if (x < t) return;
        }
        for (uint32_t x : p)
          if (kv[x] == -1 && __atomic_fetch_sub(&DD[x], 1, __ATOMIC_RELAXED) == level + 1) out.push_back(x);
      });
    }
    for (auto& l : local) crossed.insert(crossed.end(), l.begin(), l.end());
    peeled += dying.size();
    ++round;
    const auto now = clk::now();
    if (round % 256 == 0 || secs(last_note, now) > 20) {
      note("peel round / peeled", round, peeled);
      last_note = now;
    }
  }
  auto t4 = clk::now();
  r.t_peel = secs(t3, t4);
  r.rounds = round;
  r.k = kv;
  r.kmax = level;
  return r;
}

}  // namespace o34
