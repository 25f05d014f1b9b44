// ORACLE — test infrastructure only. A CPU restatement of the reference's host
// path for the (3,4)-nucleus decomposition, used as the checker for the HIP
// path. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load or run anything under oracle/; the product never links it.
//
// Functions and the reference code each one restates:
//   o34::load_dataset      graph_t::from_edges / deserialize / build_offset
//                          (nucleus/graph/graph.hpp:286-411, 491-536)
//   o34::make_oriented     graph_t::make_oriented + vertex_less
//                          (graph.hpp:176-218, graph_api.hpp:124-127)
//   o34::split_parts       graph_t::make_split_parts (graph.hpp:581-605)
//   o34::triangles_par     cliques_tri_par (nucleus/cliques.hpp:98-226): per part,
//                          zipper of E[O[u]..] against v's list, emits {i1, d, o}
//   o34::TriGraph          graph_t<U,V,2> tuple ctor, oriented (graph.hpp:100-131):
//                          sort by (a,b), D[a]++, E[t] = {b, c}, build_offset
//   o34::fourcliques_par   cliques_qud_par (cliques.hpp:570-711): emits
//                          {i1, o1, j, o2} = {T(pqs), T(pqr), T(prs), T(qrs)}
//   o34::QudGraph          graph_t<U,V,3> tuple ctor, undirected (graph.hpp:132-173)
//   o34::Bucket            bucket_ex ctor/decrease/peel/peel_count/pop_min/
//                          peeling_par/support (nucleus/bucket_ex.hpp:31-360)
//   o34::ondemand          cuNucleus34_ondemand.cu's algorithm on the CPU:
//                          edge->triangle incidence keyed by third vertex (92-128,
//                          502-555), 3-way zipper support (136-194), level-synchronous
//                          peel with first-dying/lowest-index K4 ownership (203-305,
//                          599-690). Same k semantics; O(T) memory.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace o34 {

struct Graph {
  uint32_t n = 0;
  std::vector<uint32_t> O;  // n + 1 (O[n] == m)
  std::vector<uint32_t> E;  // m
};

using Tri = std::array<uint32_t, 3>;
using Qud = std::array<uint32_t, 4>;

// Undirected graph as from_edges builds it; `oriented` is true for .grh input.
Graph load_dataset(const std::string& path, bool* oriented);
Graph make_oriented(const Graph& und);
std::vector<size_t> split_parts(const Graph& g, size_t parts);

// Triangles in (a,b) order (the TriGraph ctor sorts them).
std::vector<Tri> triangles_par(const Graph& g, const std::vector<size_t>& parts);

struct TriGraph {
  std::vector<uint32_t> O, D;           // per oriented edge (vertex of this graph)
  std::vector<std::array<uint32_t, 2>> E;  // {b, c} per triangle, triangle id = index
  TriGraph(std::vector<Tri>& tris, size_t m);
};

std::vector<Qud> fourcliques_par(const TriGraph& tg, const std::vector<size_t>& parts);

struct Result {
  uint64_t T = 0, Q = 0;
  int32_t kmax = 0;
  std::vector<Tri> tris;         // (a,b) order
  std::vector<int32_t> support;  // K4s per triangle
  std::vector<int32_t> k;        // per triangle, (a,b) order
  double t_tri = 0, t_trigraph = 0, t_qud = 0, t_qudgraph = 0, t_peel = 0, t_support = 0, t_incidence = 0;
  uint64_t rounds = 0;
};

// Full host path: triangles -> triangle graph -> K4 list -> K4 graph -> bucket peel.
Result host_path(const Graph& oriented, size_t parts);
// On-demand algorithm on the CPU (no K4 list).
Result ondemand(const Graph& oriented, size_t parts);

}  // namespace o34
