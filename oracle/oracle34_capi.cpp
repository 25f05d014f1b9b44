// ORACLE — test infrastructure only. C entry points over the restatement in
// oracle34.cpp, for the pytest harness (ctypes) and bench.py's cpu_baseline leg.
#include <cstring>
#include <string>

#include "oracle34.hpp"

namespace {
struct Handle {
  o34::Graph g;
  o34::Result r;
};
void set_err(char* err, int errlen, const char* msg) {
  if (err && errlen > 0) {
    std::strncpy(err, msg, static_cast<size_t>(errlen) - 1);
    err[errlen - 1] = 0;
  }
}
}  // namespace

extern "C" {

// Loads a dataset by extension. orient != 0 applies make_oriented to the
// undirected result of .mtx/.edges/.txt (a .grh is returned as stored).
void* o34_load(const char* path, int orient, char* err, int errlen) {
  try {
    bool pre = false;
    auto* h = new Handle;
    h->g = o34::load_dataset(path, &pre);
    if (orient && !pre) h->g = o34::make_oriented(h->g);
    return h;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return nullptr;
  }
}

void* o34_from_csr(const uint32_t* O, const uint32_t* E, uint32_t n, uint64_t m, int orient) {
  auto* h = new Handle;
  h->g.n = n;
  h->g.O.assign(O, O + n + 1);
  h->g.E.assign(E, E + m);
  if (orient) h->g = o34::make_oriented(h->g);
  return h;
}

uint32_t o34_n(void* p) { return static_cast<Handle*>(p)->g.n; }
uint64_t o34_m(void* p) { return static_cast<Handle*>(p)->g.E.size(); }
void o34_copy_csr(void* p, uint32_t* O, uint32_t* E) {
  auto* h = static_cast<Handle*>(p);
  std::memcpy(O, h->g.O.data(), h->g.O.size() * 4);
  if (!h->g.E.empty()) std::memcpy(E, h->g.E.data(), h->g.E.size() * 4);
}

// mode 0: full host path (triangles, K4 list, bucket peel); mode 1: on-demand.
int o34_run(void* p, int mode, int parts, char* err, int errlen) {
  auto* h = static_cast<Handle*>(p);
  try {
    h->r = mode == 0 ? o34::host_path(h->g, static_cast<size_t>(parts))
                     : o34::ondemand(h->g, static_cast<size_t>(parts));
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

uint64_t o34_T(void* p) { return static_cast<Handle*>(p)->r.T; }
uint64_t o34_Q(void* p) { return static_cast<Handle*>(p)->r.Q; }
int32_t o34_kmax(void* p) { return static_cast<Handle*>(p)->r.kmax; }
uint64_t o34_rounds(void* p) { return static_cast<Handle*>(p)->r.rounds; }
void o34_copy_tris(void* p, uint32_t* abc) {
  auto* h = static_cast<Handle*>(p);
  if (!h->r.tris.empty()) std::memcpy(abc, h->r.tris.data(), h->r.tris.size() * 12);
}
void o34_copy_k(void* p, int32_t* k) {
  auto* h = static_cast<Handle*>(p);
  if (!h->r.k.empty()) std::memcpy(k, h->r.k.data(), h->r.k.size() * 4);
}
void o34_copy_support(void* p, int32_t* s) {
  auto* h = static_cast<Handle*>(p);
  if (!h->r.support.empty()) std::memcpy(s, h->r.support.data(), h->r.support.size() * 4);
}
// 0 triangles, 1 triangle graph, 2 K4 list, 3 K4 graph, 4 peel, 5 incidence, 6 support
double o34_time(void* p, int which) {
  const auto& r = static_cast<Handle*>(p)->r;
  const double v[7] = {r.t_tri, r.t_trigraph, r.t_qud, r.t_qudgraph, r.t_peel, r.t_incidence, r.t_support};
  return (which >= 0 && which < 7) ? v[which] : 0.0;
}
void o34_free(void* p) { delete static_cast<Handle*>(p); }
}
