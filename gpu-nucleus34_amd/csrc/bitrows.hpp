// Bit-packed edge -> triangle incidence.
//
// For an oriented edge e with source u (out-list N+(u), in-list N-(u)):
//   Aout row of e=(u,v): bit j set  <=>  triangle (u, v, N+(u)[j])   (e is its b-edge)
//   Ain  row of e=(u,w): bit j set  <=>  triangle (u, N+(u)[j], w)   (e is its a-edge)
//   Cm   row of e=(v,w): bit i set  <=>  triangle (N-(v)[i], v, w)   (e is its c-edge)
// (vertices named by rank as in the reference's tri_t: a=(u,w), b=(u,v),
// c=(v,w), u<v<w; nucleus/utils.cuh:93-102). Aout/Ain rows of e hold
// ceil(d+(u)/32) words, Cm rows ceil(d-(u)/32) words, at per-edge word offsets.
// (Aout|Ain) of e=(u,y) is y's adjacency inside N+(u).
//
// Ain and Cm words carry a running rank beside their bits: the rank of an Ain
// word of edge e is offA[e] plus the set bits of e's row before that word, so
// the id of the triangle at bit j of the row is rank + popc(bits below j) —
// one 8-byte read; likewise a Cm word's rank is the c-slot offset cOff[e] plus
// the bits before it, and cSlot[c-slot] holds the triangle id. With these:
//   * triangle index of (u, N+(u)[j], w) = row_rank(Ain row of (u,w), j)
//   * K4s of t=(u,v,w) above u: bits of (Aout|Ain)[a] & (Aout|Ain)[b]
//     K4s of t below u:          bits of Cm[a] & Cm[b]
//   * each partner triangle of each K4 is one such rank lookup away.
#pragma once
#include "common.hpp"

namespace nuc34 {

struct RankWord {
  uint32_t bits;
  uint32_t rank;
};

struct BitRows {
  const uint64_t* rowA;  // m + 1 word offsets of the Aout / Ain rows
  const uint64_t* rowC;  // m + 1 word offsets of the Cm rows
  uint32_t* Aout;
  RankWord* Ain;
  RankWord* Cm;
  // running rank of the Aout bits at each Aout word: a triangle's peel id is
  // the rank of its bit in its b-edge's row (set after the bits pass)
  const uint32_t* AoutR = nullptr;
};

// A bare bit word, scanned by its set bits (exclusive_scan over BitWord).
struct BitWord {
  uint32_t bits;
};

// Rank of bit j (j < 32) of a ranked word: its rank plus the set bits below j.
__device__ __forceinline__ uint32_t rank_in_word(RankWord w, uint32_t j) {
  return w.rank + static_cast<uint32_t>(__popc(w.bits & ((1u << j) - 1u)));
}

// Rank of bit position j of a ranked row.
__device__ __forceinline__ uint32_t row_rank(const RankWord* row, uint32_t j) {
  return rank_in_word(row[j >> 5], j & 31);
}

// Set bits of (x & y) strictly below bit position j.
__device__ __forceinline__ uint32_t and_bits_below(const uint32_t* x, const uint32_t* y, uint32_t j) {
  uint32_t n = 0;
  const uint32_t full = j >> 5;
  for (uint32_t i = 0; i < full; ++i) n += static_cast<uint32_t>(__popc(x[i] & y[i]));
  if (j & 31) n += static_cast<uint32_t>(__popc(x[full] & y[full] & ((1u << (j & 31)) - 1u)));
  return n;
}

// In-CSR (transpose of the oriented graph) with lists sorted by source id,
// plus inpos[e] = position of edge e in its target's in-list.
void build_in_csr(const uint32_t* O, const uint32_t* E, const uint32_t* src, uint32_t n, uint64_t m, uint32_t* IO,
                  uint32_t* IE, uint32_t* IEe, uint32_t* inpos, class Allocator& alloc, hipStream_t s);
// Per-edge row lengths in words: wa[e] = ceil(d+(src e)/32), wc[e] = ceil(d-(src e)/32).
void launch_row_words(const uint32_t* O, const uint32_t* IO, const uint32_t* src, uint64_t m, uint32_t* wa,
                      uint32_t* wc, hipStream_t s);
// Triangles on each edge, in any role: cnt[e] = |Aout| + |Ain| + |Cm| of e.
void launch_edge_triangles(const BitRows& r, uint64_t m, uint32_t* cnt, hipStream_t s);
// Running ranks: w[i].rank = P[i], P the exclusive scan of the words' set bits
// (exclusive_scan over RankWord). Rows lie edge after edge, so P at a row's
// first word is the edge's first triangle id (Ain) or first c-slot (Cm).
void launch_set_ranks(RankWord* w, const uint64_t* P, uint64_t n, hipStream_t s);
// r[i] = P[i] for i < n (ranks below 2^32).
void launch_set_ranks32(uint32_t* r, const uint64_t* P, uint64_t n, hipStream_t s);

}  // namespace nuc34
