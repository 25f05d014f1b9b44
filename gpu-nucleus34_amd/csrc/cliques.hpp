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
// Per-triangle peel state: dd = current support, kv = k (-1 until fixed; the
// result) and a 2-bit state, 16 triangles per 32-bit word of st:
//   0 alive: k not fixed yet
//   1 dies in the current round
//   2 k fixed in the current round (pushed); dies in the next one
//   3 died in an earlier round (also the padding fields past T)
// The ownership test of a K4 reads only these states (4 MB per 16 M
// triangles), so its random reads stay in L2.
//
// Round control lives on the device, so rounds run back to back without host
// round trips. Round r reads ctl[r & 1] and its last kernel writes the next
// round's control into ctl[(r + 1) & 1], so no kernel writes a control field
// that another workgroup of the same kernel reads. Rounds are enqueued in
// groups of even size, so a round's parity is its slot parity in the group.
struct PeelCtl {
  uint32_t ncur;        // size of this round's frontier
  uint32_t nnext;       // size of the next round's frontier (the debit appends)
  uint32_t sel;         // buffer holding this round's frontier
  uint32_t need_level;  // frontier empty: this round picks the next level
  uint32_t minval;      // min support over alive triangles (when need_level)
  int32_t level;
  uint32_t round;
  uint32_t done;
  uint32_t big;         // frontier triangles queued for the wave path (on-demand)
  uint32_t error;       // nonzero: inconsistent state detected
  uint64_t peeled;
};

struct PeelBufs {
  int32_t* dd;      // padded to a multiple of 16 entries
  uint32_t* st;     // ceil(T / 16) state words
  int32_t* kv;
  uint32_t* buf0;   // frontier buffers, T entries each
  uint32_t* buf1;
  uint32_t* bigq;   // on-demand: frontier triangles for the wave path
  PeelCtl* ctl;     // two slots
  uint64_t T;
};

// One peel round at position `slot` of its group: level selection when the
// frontier is empty, the debit of every K4 owned by a dying triangle, and the
// state advance. Every kernel returns at once when the round's control says
// done, so extra rounds are no-ops.
void launch_peel_round_direct(const PeelBufs& p, int slot, const uint64_t* koff, const uint32_t* lists,
                              hipStream_t s);
void launch_peel_round_ondemand(const PeelBufs& p, int slot, const TriView& v, hipStream_t s);
void launch_peel_init(const PeelBufs& p, const int32_t* sup, hipStream_t s);


}  // namespace nuc34
