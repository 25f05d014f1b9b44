// (c), (d), (e): four-clique support, four-clique list and triangle->K4
// incidence (direct), and the level-synchronous peel with frontiers.
#pragma once
#include "bitrows.hpp"
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
// Support from the bit rows: sup = K4s on t, sup_up = those whose fourth vertex
// ranks above t's lowest vertex, own = those whose fourth vertex ranks above all
// of t (t holds the K4's lowest three vertices); *total += sum of sup.
void launch_support_bits(const uint32_t* triA, const uint64_t* triBC, uint64_t T, const BitRows& rows, int32_t* sup,
                         uint32_t* sup_up, uint32_t* own, uint64_t* total, hipStream_t s);
// Direct variant: per-triangle K4 lists, three partner triangle ids per entry,
// at lists[3*koff[t] ..]. Each K4 is enumerated once, from the triangle on its
// lowest three vertices (k4owned: (t << 32 | x position) at k4off[t] from a scan
// of `owned`), and written into the lists of its four triangles at positions
// given by prefix counts (up-K4s first, in fourth-vertex order, then down-K4s).
void launch_k4_owned(const uint32_t* triA, const uint64_t* triBC, uint64_t T, const BitRows& rows,
                     const uint64_t* k4off, uint64_t* k4owned, hipStream_t s);
void launch_k4_place(const uint64_t* k4owned, uint64_t Q, const uint32_t* triA, const uint64_t* triBC,
                     const uint32_t* O, const uint32_t* E, const uint32_t* src, const uint32_t* inpos,
                     const BitRows& rows, const uint64_t* offA, const uint64_t* koff, const uint32_t* sup_up,
                     uint32_t* lists, hipStream_t s);
// The K4 list in canonical order (by the triangle on the three lowest vertices,
// then by fourth vertex) as {T(pqs), T(pqr), T(prs), T(qrs)} for p<q<r<s, for
// owned-K4 records [first, first + count).
void launch_k4_quads(const uint64_t* k4owned, uint64_t first, uint64_t count, const uint32_t* triA,
                     const uint64_t* triBC, const uint32_t* O, const uint32_t* E, const uint32_t* src,
                     const BitRows& rows, const uint64_t* offA, uint32_t* quads, hipStream_t s);

// ---- peel ------------------------------------------------------------------
//
// Per-triangle peel state: dd = current support (INT32_MAX once the triangle's
// k is fixed), rnd = round of death (all ones while alive), kv = k (-1 while
// alive; the result).
// Round control lives on the device so rounds run back to back without host
// round trips; a one-workgroup kernel closes each round.
struct PeelCtl {
  uint32_t count[2];    // sizes of the two frontier buffers
  uint32_t parity;      // buffer holding the current frontier
  uint32_t need_level;  // frontier empty: pick the next level
  uint32_t minval;      // min alive support (level selection)
  int32_t level;
  uint32_t round;
  uint32_t done;
  uint64_t peeled;
  uint32_t big;         // frontier triangles queued for the wave path (on-demand)
  uint32_t error;       // nonzero: inconsistent state detected
  uint32_t nalive[2];   // sizes of the two alive-list buffers
  uint32_t apar;        // buffer holding the current alive list
};

struct PeelBufs {
  int32_t* dd;
  uint32_t* rnd;
  int32_t* kv;
  uint32_t* buf0;
  uint32_t* buf1;
  uint32_t* alive0;  // triangles not yet fixed at the last level change
  uint32_t* alive1;
  uint32_t* bigq;
  PeelCtl* ctl;
  uint64_t T;
};

// One peel round: level selection (only when the frontier is empty), then the
// debit of every K4 owned by a dying triangle. Every kernel returns at once
// when ctl->done is set, so extra rounds are no-ops.
void launch_peel_round_direct(const PeelBufs& p, const uint64_t* koff, const uint32_t* lists, hipStream_t s);
void launch_peel_round_ondemand(const PeelBufs& p, const TriView& v, hipStream_t s);
void launch_peel_init(const PeelBufs& p, const int32_t* sup, hipStream_t s);


}  // namespace nuc34
