// (c), (d), (e): four-clique support, four-clique list and triangle->K4
// incidence (direct), and the level-synchronous peel with frontiers.
#pragma once
#include "common.hpp"

namespace nuc34 {

// Read-only view of the triangle list and its incidence.
struct TriView {
  const uint32_t* triA;   // a-edge per triangle
  const uint64_t* triBC;  // (b << 32) | c per triangle
  uint64_t T;
  const uint64_t* incO;   // m + 1
  const uint32_t* third;  // 3T
  const uint32_t* tri;    // 3T
};

// Triangles whose three incidence lists hold more than this many entries in
// total are intersected by a whole wave; shorter ones by one lane.
constexpr uint32_t kSmallL = 128;

// support[t] = K4s on t, owned[t] = K4s whose lowest three vertices are t,
// *total += sum of supports. Triangles above kSmallL are appended to bigq.
void launch_support(const TriView& v, int32_t* support, uint32_t* owned, uint64_t* total, uint32_t* bigq,
                    uint32_t* bigcount, hipStream_t s);
// Writes the K4s owned by each triangle at k4off[t]... as
// {T(pqs), T(pqr), T(prs), T(qrs)}.
void launch_k4_emit(const TriView& v, const uint32_t* owned, const uint64_t* k4off, uint32_t* qud,
                    const uint32_t* bigq, const uint32_t* bigcount, hipStream_t s);
// Triangle -> K4 lists (4 entries per K4) at incK[koff[t]..koff[t+1]).
void launch_k4_incidence(const uint32_t* qud, uint64_t Q, const uint64_t* koff, uint32_t* cur, uint32_t* incK,
                         hipStream_t s);

// Peel state shared by the round kernels (device memory).
struct PeelCounters {
  uint32_t front;      // size of the current frontier
  uint32_t next;       // size of the frontier being built
  uint32_t big;        // frontier triangles queued for the wave path
  uint32_t minval;     // min over alive supports (level selection)
  uint32_t alive;      // size of the compacted alive list being built
  uint32_t pad[3];
};

void launch_alive_min(const uint32_t* alive, uint32_t nalive, const int32_t* kv, const int32_t* dd,
                      PeelCounters* pc, hipStream_t s);
void launch_collect(const uint32_t* alive, uint32_t nalive, int32_t* kv, const int32_t* dd, uint32_t* rnd,
                    int32_t level, uint32_t round, uint32_t* front, PeelCounters* pc, hipStream_t s);
void launch_mark(const uint32_t* front, const PeelCounters* pc, int32_t* kv, uint32_t* rnd, int32_t level,
                 uint32_t round, hipStream_t s);
// On-demand debit: each dying triangle re-enumerates its K4s.
void launch_debit_ondemand(const TriView& v, const uint32_t* front, PeelCounters* pc, const int32_t* kv,
                           const uint32_t* rnd, int32_t* dd, int32_t level, uint32_t round, uint32_t* next,
                           uint32_t* bigq, hipStream_t s);
// Direct debit: each dying triangle walks its stored K4 list.
void launch_debit_direct(const uint64_t* koff, const uint32_t* incK, const uint32_t* qud, const uint32_t* front,
                         PeelCounters* pc, const int32_t* kv, const uint32_t* rnd, int32_t* dd, int32_t level,
                         uint32_t round, uint32_t* next, hipStream_t s);
void launch_compact_alive(const uint32_t* alive, uint32_t nalive, const int32_t* kv, uint32_t* out,
                          PeelCounters* pc, hipStream_t s);

}  // namespace nuc34
