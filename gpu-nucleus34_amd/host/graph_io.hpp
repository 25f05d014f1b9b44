// Graph input for the (3,4)-nucleus path: the reference's dataset loaders and
// CSR construction, with the same observable results.
//   .mtx          -> from_edges(file, mtx header)       (nucleus/graph/graph.hpp:286-411)
//   .edges / .txt -> from_edges(file, no header: the first non-comment line is dropped)
//   .grh          -> deserialize + build_offset, already oriented (graph.hpp:491-536)
// and the degree orientation make_oriented (graph.hpp:176-218, vertex_less at
// graph_api.hpp:124-127). Edge index = position in E; lists sorted by vertex id.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace nuc34 {

// Compressed sparse rows: O has n+1 entries (O[n] == E.size()).
struct Csr {
  uint32_t n = 0;
  std::vector<uint32_t> O;  // n + 1
  std::vector<uint32_t> E;  // m
  uint64_t m() const { return E.size(); }
};

// What a dataset file holds after loading: the undirected CSR from from_edges
// (oriented == false), or the pre-oriented CSR of a .grh file (oriented == true).
struct LoadedGraph {
  Csr g;
  bool oriented = false;
  uint64_t lines = 0;  // edge lines read after the header / dropped first line
};

// Extension as the reference computes it: substring from the last '.'
// (nucleus/utility/def_system.hpp:56-58). Throws if the name has no '.'.
std::string file_ext(const std::string& path);
std::string strip_path(const std::string& path);

// from_edges semantics on an in-memory text buffer.
LoadedGraph parse_edge_text(const char* data, size_t len, bool mtx_header);
// Loads by extension (.mtx, .edges, .txt, .grh); throws std::runtime_error with
// the reference's messages ("Unknown dataset format: <ext>", ...).
LoadedGraph load_dataset(const std::string& path);

// Undirected CSR from raw (s, d) pairs: self-loops dropped, both directions,
// sorted, deduplicated, n = max id + 1 (graph.hpp:356-411).
Csr csr_from_pairs(const uint32_t* s, const uint32_t* d, size_t count);

// Degree orientation on the host: keep u->v iff (deg u, u) < (deg v, v) with the
// undirected degree; list order preserved.
Csr orient_host(const Csr& und);

// .grh binary (graph.hpp:465-520): size_t sizeof(V)=4 | size_t nV | size_t nE | V D[nV] | V E[nE].
void write_grh(const std::string& path, const Csr& g);

// Vertex-range split balanced by edge prefix, as make_split_parts
// (graph.hpp:581-605): `parts` entries, first is 0.
std::vector<uint64_t> split_parts(const Csr& g, size_t parts);

// Checks a CSR that claims to be oriented: offsets monotone, ids < n, every
// list strictly increasing, no self-loops, no reciprocal pair. Returns an
// empty string when valid, else the first violation.
std::string check_oriented(const Csr& g);

}  // namespace nuc34
