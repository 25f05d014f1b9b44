// (c), (d), (e): four-clique support, the per-triangle K4 lists (direct) and
// the level-synchronous peel, all driven by one K4 enumeration from the ranked
// bit rows (bitrows.hpp).
#pragma once
#include "bitrows.hpp"
#include "common.hpp"

namespace nuc34 {

// Everything a K4 enumeration from the bit rows reads.
struct K4Src {
  const uint32_t* O;      // oriented CSR offsets (n + 1)
  const uint32_t* E;      // edge heads (m)
  const uint32_t* src;    // edge tails (m)
  const uint32_t* inpos;  // position of each edge in its head's in-list
  BitRows rows;
  const uint32_t* triA;   // a-edge per triangle
  const uint64_t* triBC;  // (b << 32) | c per triangle
  // per c-slot: the triangle's id and the in-list position of its a-edge (a
  // down K4's third partner is one rank lookup from the second, which this
  // pair gives with its id in one load)
  const uint2* cSlot;
  uint64_t T;
  // Peel rows for the on-demand peel (null elsewhere), word for word beside
  // (Aout|Ain) and Cm: each PA word repeats the Aout bits and running rank
  // next to two planes, `live` (the triangles not yet retired) and `dying`
  // (the current round's frontier); PC holds the c-role rows' two planes. The on-demand peel numbers triangles in
  // peel order, by b-edge and then a-edge, which is the running rank of the
  // Aout bits: a dying triangle's K4s with their fourth vertex above its own
  // three find all three partners at consecutive ids along the rows of its
  // own edges, so their decrements share cache lines. triA, triBC and cSlot
  // are in that order everywhere. A triangle's three bits of both planes are
  // cleared after the round it dies in, so a dying triangle enumerates only
  // K4s whose two partners on its lowest vertex's rows are still alive or dying
  // in the same round, and it drops at once, before any partner lookup, every
  // K4 one of those two partners owns (dying now with a lower index). The
  // third partner's rank lookup reads the word that holds its planes too, so
  // its state costs no load of its own.
  struct PeelWord* PA;
  struct PlaneWord* PC;  // the c-role rows' planes only: their bits and ranks are Cm's
  // per vertex u: in-neighbours with a lower vertex id than u (a prefix of
  // N-(u), whose triangles precede u's in the peel order); on-demand peel only
  const uint32_t* icut;
};

struct alignas(16) PeelWord {
  uint32_t bits;  // Aout bits
  uint32_t rank;  // their running rank: peel-order ids
  uint32_t live;  // (Aout|Ain) bits of the triangles not yet retired
  uint32_t dying; // bits of the current round's frontier
};

// The planes of a c-role row word (beside Cm's bits and c-slot ranks): most
// down words hold no K4, so the debit reads these 8 bytes first and the ranked
// word only where the planes leave K4 bits.
struct alignas(8) PlaneWord {
  uint32_t live;   // Cm bits of the triangles not yet retired
  uint32_t dying;  // bits of the current round's frontier
};

// sup[t] = K4s on t, own[t] = K4s whose fourth vertex ranks above all of t
// (t holds the K4's lowest three vertices); *total += sum of sup. ab (WA + 1
// entries) and cb (WC + 1) are scratch: bits-only copies of the rows.
void launch_support_bits(const K4Src& s, uint64_t WA, uint64_t WC, uint2* ab, uint32_t* cb, int32_t* sup,
                         uint32_t* own, uint64_t* total, hipStream_t st);
// Direct variant: the K4 list of every triangle, three partner triangle ids per
// entry, at lists[3 * koff[t] ..] (koff = exclusive scan of sup): the K4s whose
// fourth vertex ranks above t's lowest vertex first, in order of its position
// in that vertex's out-list, then the others in order of their fourth vertex's
// position in the lowest vertex's in-list.
// With `live`, s is the peel order's (peel rows with fresh planes, triangles in
// peel order), t and the partners are peel ids and koff scans the supports in
// peel order.
void launch_k4_lists(const K4Src& s, const uint64_t* koff, uint32_t* lists, unsigned grid_cap, bool live,
                     hipStream_t st);
// Owned K4s: (t << 32 | position of the fourth vertex in N+(u)); the records
// [first, first + count) of the list laid out at k4off[t] .. (k4off = exclusive
// scan of own) go to k4owned[0, count), walking triangles [t_begin, t_end).
void launch_k4_owned(const K4Src& s, const uint64_t* k4off, uint64_t t_begin, uint64_t t_end, uint64_t first,
                     uint64_t count, uint64_t* k4owned, hipStream_t st);
// out[0..1] = the triangles holding owned records first and first + count - 1.
void launch_k4_owned_bounds(const uint64_t* k4off, uint64_t T, uint64_t first, uint64_t count, uint64_t* out,
                            hipStream_t st);
// The K4 list in canonical order (by the triangle on the three lowest vertices,
// then by fourth vertex) as {T(pqs), T(pqr), T(prs), T(qrs)} for p<q<r<s, for
// owned-K4 records [first, first + count).
void launch_k4_quads(const K4Src& s, const uint64_t* k4owned, uint64_t first, uint64_t count, uint32_t* quads,
                     hipStream_t st);

// ---- peel ------------------------------------------------------------------
//
// Per-triangle peel state: dd = current support, kv = k (-1 until fixed; the
// result) and a 2-bit state, 16 triangles per 32-bit word of st:
//   0 alive: k not fixed yet
//   1 dies in the current round
//   2 k fixed in the current round (pushed); dies in the next one
//   3 died in an earlier round (also the padding fields past T)
// The on-demand peel keeps the dying and retired triangles in the live and
// dying planes of its rows instead; its state words only tell alive (0) from
// not alive (any other value), which is all its level scans read.
// The ownership test of a K4 reads only these states (4 MB per 16 M
// triangles), so its random reads stay in L2.
//
// Round control lives on the device, so rounds run back to back without host
// round trips. Round r reads ctl[r & 1] and its last kernel writes the next
// round's control into ctl[(r + 1) & 1], so no kernel writes a control field
// that another workgroup of the same kernel reads. Rounds are enqueued in
// groups of even size, so a round's parity is its slot parity in the group.
struct PeelCtl {
  uint32_t ncur;        // this round's frontier: pushed by the previous round ...
  uint32_t nspec;       // ... or collected by the previous round's finish
  uint32_t nnext;       // next round's frontier (the debit appends)
  uint32_t sel;         // buffer holding this round's frontier
  uint32_t need_level;  // the frontier came from a level change: level = max(level, minval)
  uint32_t minval;      // min support over alive triangles at that level change
  int32_t level;
  uint32_t round;       // rounds with a non-empty frontier so far
  uint32_t seq;         // rounds of any kind so far (the later slot has the larger value)
  uint32_t done;
  uint32_t error;       // nonzero: inconsistent state detected
  uint64_t peeled;
};

struct PeelBufs {
  int32_t* dd;      // padded to a multiple of 16 entries
  uint32_t* st;     // ceil(T / 16) state words
  int32_t* kv;
  uint32_t* buf0;   // frontier buffers, T entries each
  uint32_t* buf1;
  PeelCtl* ctl;     // two slots
  uint64_t T;
  // A round's pushes (survivors falling to the level): workgroup w of the
  // debit grid (nwg workgroups) writes its own to wslots[w * wcap, ...) and
  // their count to wcnt[w]; what does not fit goes to the front of the next
  // frontier buffer (count in the control's nnext). The finish gathers both
  // into the next frontier. The per-graph peel does not use them.
  uint32_t* wslots;
  uint32_t* wcnt;
  uint32_t* woff;   // offset of region w among its group's (the group counter's old value)
  uint32_t* gcnt;   // push group counters: 2 round parities x kPushGroups, kGroupStride apart
  uint32_t wcap;
  uint32_t nwg;
  uint32_t gsize;   // regions per group: ceil(nwg / kPushGroups)
};
// The debit workgroups' push regions form kPushGroups groups of consecutive
// workgroups; each workgroup with pushes takes its offset from its group's
// counter (one atomic, on a line of its own), so the finish scans only the
// group totals.
constexpr uint32_t kPushGroups = 256;
constexpr uint32_t kGroupStride = 32;

// One peel round at position `slot` of its group: the debit of every K4 owned
// by a dying triangle, then the state advance and, when the frontier ran
// empty, the next level's collection. Both kernels return at once when the
// round's control says done, so extra rounds are no-ops. The direct round
// reads the K4 lists; the on-demand round enumerates each dying triangle's K4s
// from the bit rows.
// Launch shape of the rounds: workgroups of the debit kernel and the largest
// number of dying triangles one wave takes at a time.
struct PeelShape {
  unsigned debit_grid = 2048;
  unsigned tile_cap = 64;
  unsigned finish_grid = 512;  // any size is correct: the finish loops are grid-strided
};
void launch_peel_round_direct(const PeelBufs& p, int slot, const uint64_t* koff, const uint32_t* lists,
                              const PeelShape& shape, hipStream_t s);
void launch_peel_round_ondemand(const PeelBufs& p, int slot, const K4Src& src, const PeelShape& shape, hipStream_t s);
void launch_peel_init(const PeelBufs& p, const int32_t* sup, hipStream_t s);

// A batch peeled one workgroup per graph (k_peel_graphs): each workgroup takes
// the next graph of `order` and runs all of that graph's rounds, with the
// same rounds, k and ownership as the graph decomposed alone.
struct GraphQueue {
  const uint64_t* tri_first;  // count + 1 triangle offsets of the batch's graphs
  const uint32_t* order;      // the graphs in the order workgroups take them
  uint32_t count;             // graphs in `order`
  uint32_t* next;             // queue position, 0 at launch
  uint32_t* rounds;           // per graph: its rounds (with a non-empty frontier)
  int32_t* level;             // per graph: its last level, its K_max
  uint32_t* error;            // nonzero: inconsistent state (0 at launch)
};
// One launch for the whole peel of the graphs in q (after launch_peel_init,
// and with `live`, launch_live_rows): the direct debit from koff/lists, or
// with `live` the on-demand debit from src's live rows. Up to
// shape.debit_grid workgroups; shape.tile_cap dying triangles per wave.
void launch_peel_graphs(const PeelBufs& p, const K4Src& src, const uint64_t* koff, const uint32_t* lists,
                        const GraphQueue& q, const PeelShape& shape, bool live, hipStream_t s);
// Sum of the residual supports after the peel into *total (zeroed by the
// caller): the peel applied 4Q minus this many support decrements.
void launch_residual_sum(const PeelBufs& p, int64_t* total, hipStream_t s);
// The on-demand peel's rows: PA[i] = {Aout[i], AoutR[i], Aout[i] | Ain[i].bits,
// 0} for i < WA, PC[i] = {Cm[i].bits, 0} for i < WC, and icut for the n
// vertices of the in-CSR (IO, IE).
void launch_live_rows(const K4Src& src, uint64_t WA, uint64_t WC, const uint32_t* IO, const uint32_t* IE, uint32_t n,
                      uint32_t* icut, hipStream_t s);

}  // namespace nuc34
