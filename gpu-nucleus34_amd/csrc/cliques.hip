// (c) four-clique support, (d) the per-triangle K4 lists and (e) the
// level-synchronous peel, over one K4 enumeration from the ranked bit rows.
//
// K4 enumeration. For triangle t = (u, v, w) with a = (u,w), b = (u,v),
// c = (v,w) (ranks u < v < w), the fourth vertices x are
//   up:   bits jx of (Aout|Ain)[a] & (Aout|Ain)[b]  (x = N+(u)[jx], x adjacent
//         to v and w inside N+(u));
//   down: bits i of Cm[a] & Cm[b]                   (x = N-(u)[i]).
// The partner triangles of each K4 are rank lookups in the rows:
//   up, ex = (u,x):
//     T{u,v,x} = rank of jv in Ain[ex] if v < x (b-role bit jx of b), else rank
//                of jx in Ain[b];      T{u,w,x} likewise with w, a;
//     T{v,w,x}: (v,x) or (x,v) is the c-edge of T{u,v,x}; with lowest vertex v
//                it is the rank of w's or x's position in the a-role row of the
//                higher of the two edges from v; with lowest x (x < v) the rank
//                of v's position in N+(x) in the a-role row of (x,w), the
//                c-edge of T{u,x,w};
//   down (x < u): T{x,u,v} and T{x,u,w} are the c-slots of bit i in Cm[b] and
//                Cm[a]; T{x,v,w} is the c-slot of x's in-list position at v
//                (inpos of (x,v), the a-edge of T{x,u,v}) in Cm[c].
// This replaces the reference's K4 listing (cu_cliques_quds, compute_fourcliques,
// nucleus/utils.cuh:550-619, 2173-2277), its triangle->K4 incidence sort
// (nucleus/cuNucleus34.cu:609-673) and the on-demand re-enumeration by 3-way
// list intersection (for_each_clique_of, nucleus/cuNucleus34_ondemand.cu:136-170):
// one wave takes 64 triangles, flattens their row words over its lanes and
// every lane expands the set bits of one word at a time.
//
// Peel (cuNucleus34_ondemand.cu:203-305, 599-690; cuNucleus34.cu:107-141):
// level = max(level, min alive support); the triangles with support <= level
// die this round (kv = level). A K4 is destroyed by its first member to die,
// same-round ties going to the lowest triangle index, and its owner debits the
// surviving members. A survivor whose support falls from level+1 to level is
// pushed onto the next round's frontier, so rounds after the first of a level
// touch only the triangles that crossed.
#include "cliques.hpp"

namespace nuc34 {

namespace {

constexpr int kWPB = kBlock / kWave;

// ---- K4 enumeration ----------------------------------------------------------

// A triangle's rows. Every Aout/Ain row of a vertex's out-edges has the same
// length, so the row of (u, y) is rowA[first out-edge of u] + pos(y) * W.
struct TriRows {
  uint64_t ru0;  // Aout / Ain row of u's first out-edge (rows of a, b, (u,x) follow)
  uint64_t rv0;  // Aout / Ain row of v's first out-edge
  uint64_t ca;   // Cm row of a (over N-(u)); b's is (jv - jw) * WC words further
  uint64_t cc;   // Cm row of c (over N-(v))
  uint32_t c, ou, ov, jv, jw;  // c-edge; O[u], O[v]; positions of v and w in N+(u)
  uint32_t cut;  // on-demand peel: in-neighbours of u with a lower id (icut[u]); 0 elsewhere
  uint32_t W;    // words of the rows of u's out-edges
  uint32_t Wv;   // words of the rows of v's out-edges
  uint32_t WC;   // words of the a / b c-role rows
};

__device__ __forceinline__ TriRows tri_rows(const K4Src& s, uint32_t t) {
  TriRows r;
  const uint32_t a = s.triA[t];
  const uint64_t bc = s.triBC[t];
  const uint32_t b = static_cast<uint32_t>(bc >> 32);
  r.c = static_cast<uint32_t>(bc);
  const uint32_t u = s.src[b];
  r.ou = s.O[u];
  r.cut = s.icut ? s.icut[u] : 0u;
  r.ov = s.O[s.E[b]];
  r.jv = b - r.ou;
  r.jw = a - r.ou;
  const uint64_t ra = s.rows.rowA[a];
  r.W = static_cast<uint32_t>(s.rows.rowA[a + 1] - ra);
  r.ru0 = ra - static_cast<uint64_t>(r.jw) * r.W;
  const uint64_t rc = s.rows.rowA[r.c];
  r.Wv = static_cast<uint32_t>(s.rows.rowA[r.c + 1] - rc);
  r.rv0 = rc - static_cast<uint64_t>(r.c - r.ov) * r.Wv;
  r.ca = s.rows.rowC[a];
  r.WC = static_cast<uint32_t>(s.rows.rowC[a + 1] - r.ca);
  r.cc = s.rows.rowC[r.c];
  return r;
}

__device__ __forceinline__ uint32_t shfl_u32(uint32_t v, uint32_t src) { return __shfl(v, static_cast<int>(src)); }
__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, uint32_t src) {
  return static_cast<uint64_t>(shfl_u32(static_cast<uint32_t>(v), src)) |
         static_cast<uint64_t>(shfl_u32(static_cast<uint32_t>(v >> 32), src)) << 32;
}

// Lane src's rows. Every lane of the wave must be active (a lane reads
// another lane's registers only while that lane is active).
__device__ __forceinline__ TriRows shfl_rows(const TriRows& r, uint32_t src) {
  TriRows o;
  o.ru0 = shfl_u64(r.ru0, src);
  o.rv0 = shfl_u64(r.rv0, src);
  o.ca = shfl_u64(r.ca, src);
  o.cc = shfl_u64(r.cc, src);
  o.c = shfl_u32(r.c, src);
  o.ou = shfl_u32(r.ou, src);
  o.ov = shfl_u32(r.ov, src);
  o.jv = shfl_u32(r.jv, src);
  o.jw = shfl_u32(r.jw, src);
  o.cut = shfl_u32(r.cut, src);
  o.W = shfl_u32(r.W, src);
  o.Wv = shfl_u32(r.Wv, src);
  o.WC = shfl_u32(r.WC, src);
  return o;
}

// Word wi of t's K4 row (W up words, then WC down words) and the row words the
// partner lookups of its bits use.
struct K4Word {
  uint32_t bits;
  uint32_t Aa, Ab;  // up: b-role words of a and b
  RankWord Ia, Ib;  // up: a-role words of a and b; down: c-role words of a and b
  uint32_t Da, Db;  // LIVE: dying-plane words of a and b (0 elsewhere)
};

// Bits j of a row word starting at bit position `base` with base + j < p.
__device__ __forceinline__ uint32_t bits_below(uint32_t p, uint32_t base) {
  if (p <= base) return 0u;
  const uint32_t d = p - base;
  return d >= 32 ? ~0u : (1u << d) - 1u;
}

// LIVE: the K4 bits come from the live rows (the on-demand peel), without the
// K4s that a partner on t's own rows owns: a partner dying in this round with
// a lower triangle index. The on-demand peel numbers triangles in peel order:
// by b-edge, then a-edge, i.e. those sharing their lowest vertex u by the
// positions in N+(u) of their middle, then highest vertex, so for t at
// (jv, jw) and x at bit position jx of an up word:
//   T{u,v,x} (row b) precedes t iff  v < x ? jx < jw : jx < jv;
//   T{u,w,x} (row a) precedes t iff  w < x ? jw < jv : jx < jv;
// and for a down word (x at position i of N-(u), x below u) both row partners
// have lowest vertex x, and precede t iff id x < id u, i.e. i < icut[u].
template <bool LIVE>
__device__ __forceinline__ K4Word k4_word(const K4Src& s, const TriRows& r, uint32_t wi) {
  K4Word k;
  if (wi < r.W) {
    const uint64_t ra = r.ru0 + static_cast<uint64_t>(r.jw) * r.W + wi;
    const uint64_t rb = r.ru0 + static_cast<uint64_t>(r.jv) * r.W + wi;
    if (LIVE) {
      // peel rows: Aout bits with their peel-order rank, so one load per row
      const PeelWord la = s.PA[ra], lb = s.PA[rb];
      k.Aa = la.bits;
      k.Ab = lb.bits;
      k.Ia = RankWord{la.bits, la.rank};
      k.Ib = RankWord{lb.bits, lb.rank};
      const uint32_t bw = bits_below(r.jw, wi * 32), bv = bits_below(r.jv, wi * 32);
      const uint32_t lower_b = (k.Ab & bw) | (~k.Ab & bv);
      const uint32_t lower_a = (k.Aa & (r.jw < r.jv ? ~0u : 0u)) | (~k.Aa & bv);
      k.bits = la.live & lb.live & ~((lb.dying & lower_b) | (la.dying & lower_a));
      k.Da = la.dying;
      k.Db = lb.dying;
    } else {
      k.Aa = s.rows.Aout[ra];
      k.Ab = s.rows.Aout[rb];
      k.Ia = s.rows.Ain[ra];
      k.Ib = s.rows.Ain[rb];
      k.bits = (k.Aa | k.Ia.bits) & (k.Ab | k.Ib.bits);
      k.Da = k.Db = 0;
    }
  } else {
    // b's row: (jv - jw) * WC words from a's (either sign; 64-bit wraparound)
    const uint64_t ca = r.ca + (wi - r.W);
    const uint64_t cb = ca + static_cast<uint64_t>(r.jv) * r.WC - static_cast<uint64_t>(r.jw) * r.WC;
    k.Aa = k.Ab = 0;
    if (LIVE) {
      // the planes first (most down words hold no K4); the ranked words only
      // where the word has K4 bits
      const PlaneWord la = s.PC[ca], lb = s.PC[cb];
      k.bits = la.live & lb.live & ~((la.dying | lb.dying) & bits_below(r.cut, (wi - r.W) * 32));
      k.Da = la.dying;
      k.Db = lb.dying;
      k.Ia = k.Ib = RankWord{0u, 0u};
      if (k.bits) {
        k.Ia = s.rows.Cm[ca];
        k.Ib = s.rows.Cm[cb];
      }
    } else {
      k.Ia = s.rows.Cm[ca];
      k.Ib = s.rows.Cm[cb];
      k.bits = k.Ia.bits & k.Ib.bits;
      k.Da = k.Db = 0;
    }
  }
  return k;
}

// Rank of bit position j of a peel row; *state gets that bit's live (bit 0)
// and dying (bit 1) planes.
__device__ __forceinline__ uint32_t peel_rank(const PeelWord* row, uint32_t j, uint32_t* state) {
  const PeelWord w = row[j >> 5];
  const uint32_t b = j & 31u;
  *state = ((w.live >> b) & 1u) | (((w.dying >> b) & 1u) << 1);
  return w.rank + static_cast<uint32_t>(__popc(w.bits & ((1u << b) - 1u)));
}

// Rank of bit position j of the c-role row at word offset `row` (its c-slot);
// *state gets that bit's live (bit 0) and dying (bit 1) planes.
__device__ __forceinline__ uint32_t c_peel_rank(const K4Src& s, uint64_t row, uint32_t j, uint32_t* state) {
  const PlaneWord w = s.PC[row + (j >> 5)];
  const uint32_t b = j & 31u;
  *state = ((w.live >> b) & 1u) | (((w.dying >> b) & 1u) << 1);
  return row_rank(s.rows.Cm + row, j);
}

// Rank of bit position j of the peel row at word offset `row` (Aout bits, peel
// order): the peel-order id of the triangle whose b-edge owns the row.
__device__ __forceinline__ uint32_t b_rank(const K4Src& s, uint64_t row, uint32_t j) {
  const PeelWord& w = s.PA[row + (j >> 5)];
  return rank_in_word(RankWord{w.bits, w.rank}, j & 31u);
}

// The three partner triangles of the K4 at bit j of word wi. With LIVE (peel
// order ids from the peel rows), *s2 gets q[2]'s live and dying planes
// (peel_rank) from the word its rank comes from.
template <bool LIVE>
__device__ __forceinline__ void k4_partners(const K4Src& s, const TriRows& r, uint32_t wi, const K4Word& k,
                                            uint32_t j, uint32_t q[3], uint32_t* s2) {
  if (wi < r.W) {
    const uint32_t jx = wi * 32 + j;
    const bool vx = (k.Ab >> j) & 1u;  // v ranks below x
    const bool wx = (k.Aa >> j) & 1u;  // w ranks below x
    const uint64_t ix = r.ru0 + static_cast<uint64_t>(jx) * r.W;  // row of (u, x)
    uint64_t row2;
    uint32_t pos2;
    if constexpr (LIVE) {
      // An id is the rank in the row of the triangle's b-edge, at its highest
      // vertex. x above v (w): T{u,v,x} (T{u,w,x}) sits at bit j of this word;
      // neighbouring lanes take consecutive ids. x below: row (u, x).
      q[0] = vx ? rank_in_word(k.Ib, j) : b_rank(s, ix, r.jv);  // T{u,v,x}
      q[1] = wx ? rank_in_word(k.Ia, j) : b_rank(s, ix, r.jw);  // T{u,w,x}
      const uint32_t e0 = static_cast<uint32_t>(s.triBC[q[0]]);  // c-edge of T{u,v,x}: v-x (x-v if x<v)
      if (vx) {
        // v<w<x: b = (v,w) = t's c-edge, x's position in N+(v);
        // v<x<w: b = (v,x), w's position in N+(v)
        row2 = r.rv0 + static_cast<uint64_t>(wx ? r.c - r.ov : e0 - r.ov) * r.Wv;
        pos2 = wx ? e0 - r.ov : r.c - r.ov;
      } else {
        // x<v<w: b = (x,v), w's position in N+(x) from the edge id of (x,w),
        // the c-edge of T{u,x,w}
        const uint32_t exw = static_cast<uint32_t>(s.triBC[q[1]]);
        const uint32_t x = s.E[r.ou + jx];
        row2 = s.rows.rowA[e0];
        pos2 = exw - s.O[x];
      }
      q[2] = peel_rank(s.PA + row2, pos2, s2);
    } else {
      q[0] = vx ? row_rank(s.rows.Ain + ix, r.jv) : rank_in_word(k.Ib, j);  // T{u,v,x}
      q[1] = wx ? row_rank(s.rows.Ain + ix, r.jw) : rank_in_word(k.Ia, j);  // T{u,w,x}
      const uint32_t evx = static_cast<uint32_t>(s.triBC[q[0]]);  // c-edge of T{u,v,x}: v-x
      if (vx) {
        row2 = r.rv0 + static_cast<uint64_t>(wx ? evx - r.ov : r.c - r.ov) * r.Wv;
        pos2 = wx ? r.c - r.ov : evx - r.ov;  // v<w<x: a=(v,x), b=(v,w); v<x<w: a=(v,w), b=(v,x)
      } else {
        // x<v<w: a=(x,w), b=(x,v); x = N+(u)[jx] (consecutive for neighbouring
        // lanes), v's position in N+(x) from the edge id of (x,v)
        const uint32_t exw = static_cast<uint32_t>(s.triBC[q[1]]);
        const uint32_t x = s.E[r.ou + jx];
        row2 = s.rows.rowA[exw];
        pos2 = evx - s.O[x];
      }
      q[2] = row_rank(s.rows.Ain + row2, pos2);
    }
  } else {
    const uint2 xuv = s.cSlot[rank_in_word(k.Ib, j)];  // T{x,u,v}: c-edge b
    q[0] = xuv.x;
    q[1] = s.cSlot[rank_in_word(k.Ia, j)].x;  // T{x,u,w}: c-edge a
    // T{x,v,w}: c-edge c, at x's position in N-(v), the in-list position of
    // (x,v), the a-edge of T{x,u,v}
    q[2] = s.cSlot[LIVE ? c_peel_rank(s, r.cc, xuv.y, s2) : row_rank(s.rows.Cm + r.cc, xuv.y)].x;
  }
}

// Position of the r-th (from 0) set bit of x; x has more than r set bits.
__device__ __forceinline__ uint32_t select_bit(uint32_t x, uint32_t r) {
  uint32_t pos = 0;
#pragma unroll
  for (uint32_t w = 16; w >= 1; w >>= 1) {
    const uint32_t lo = static_cast<uint32_t>(__popc(x & ((1u << w) - 1u)));
    if (r >= lo) {
      r -= lo;
      x >>= w;
      pos += w;
    }
  }
  return pos;
}

// Largest index i < 64 with pre[i] <= key (pre non-decreasing, pre[0] = 0).
__device__ __forceinline__ uint32_t find_slot(const uint32_t* pre, uint32_t key) {
  uint32_t lo = 0, hi = kWave;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pre[mid] <= key) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <bool LIVE>
__device__ __forceinline__ K4Word shfl_word(const K4Word& k, uint32_t src) {
  K4Word o;
  o.bits = shfl_u32(k.bits, src);
  o.Aa = shfl_u32(k.Aa, src);
  o.Ab = shfl_u32(k.Ab, src);
  o.Ia.bits = shfl_u32(k.Ia.bits, src);
  o.Ia.rank = shfl_u32(k.Ia.rank, src);
  o.Ib.bits = shfl_u32(k.Ib.bits, src);
  o.Ib.rank = shfl_u32(k.Ib.rank, src);
  o.Da = LIVE ? shfl_u32(k.Da, src) : 0u;
  o.Db = LIVE ? shfl_u32(k.Db, src) : 0u;
  return o;
}

// Per-wave LDS of the tile walk: the prefix arrays every lane searches. The
// triangles' rows and the chunk's words stay in the registers of the lane
// that loaded them and travel by cross-lane shuffles, which keeps the LDS
// per wave small enough for eight waves per SIMD.
struct TileLds {
  uint32_t pre[kWave + 1];  // first word item of each triangle
  uint32_t run[kWave];      // K4s of each triangle listed so far
  uint32_t epre[kWave];     // first K4 of each word of the current chunk
};

// One tile: lane l brings triangle t (when valid) and a value aux of its own
// that reaches emit with each of t's K4s. LIVE: K4 bits from the live rows
// (the on-demand peel). The tile's row words are
// flattened over the lanes 64 at a time, then the set bits of those 64 words
// (the K4s) are flattened again, one K4 per lane and step, and
// emit(t, aux, pos, q, die) runs once per K4 of each triangle. With POS, pos is the K4's index in the
// triangle's list (up words first, word and bit order); without, pos is 0. With
// LIVE, bit k of die says that q[k] (k = 0, 1: the partners on t's own rows)
// dies in the current round, bit 2 that q[2] is not yet retired and bit 3
// that q[2] dies in the current round; 0 elsewhere.
template <bool POS, bool LIVE, typename Emit>
__device__ __forceinline__ void k4_tile(const K4Src& s, uint32_t t, bool valid, uint64_t aux, TileLds& L,
                                        Emit&& emit) {
  const unsigned lane = lane_id();
  uint32_t words = 0;
  TriRows own{};
  if (valid) {
    own = tri_rows(s, t);
    words = own.W + own.WC;
  }
  L.run[lane] = 0;
  const uint32_t inc = wave_inclusive_sum(words);
  const uint32_t total = __shfl(inc, kWave - 1);
  L.pre[lane] = inc - words;
  if (lane == kWave - 1) L.pre[kWave] = inc;
  wave_sync();
  for (uint32_t base = 0; base < total; base += kWave) {
    // 1. one word per lane
    const uint32_t item = base + lane;
    const bool act = item < total;
    const uint32_t i = find_slot(L.pre, act ? item : total - 1);
    const TriRows ri = shfl_rows(own, i);  // every lane takes part
    uint32_t wi = 0;
    K4Word k{};
    if (act) {
      wi = item - L.pre[i];
      k = k4_word<LIVE>(s, ri, wi);
    }
    const uint32_t pc = static_cast<uint32_t>(__popc(k.bits));
    const uint32_t incl = wave_inclusive_sum(pc);
    const uint32_t nk = __shfl(incl, kWave - 1);
    const uint32_t epre = incl - pc;
    L.epre[lane] = epre;
    uint32_t lpos = 0;
    if constexpr (POS) {
      // list position of the word's first K4: a segmented scan per triangle
      const uint32_t first = act ? (L.pre[i] > base ? L.pre[i] - base : 0u) : lane;
      const uint32_t before = __shfl(incl, static_cast<int>(first > 0 ? first - 1 : 0));
      const uint32_t seg0 = first > 0 ? before : 0u;
      lpos = act ? L.run[i] + epre - seg0 : 0u;
      wave_sync();
      const bool last = act && (lane == kWave - 1 || item + 1 == total || item + 1 == L.pre[i + 1]);
      if (last) L.run[i] += incl - seg0;
    }
    wave_sync();
    // 2. one K4 per lane; the shuffles run on every lane
    for (uint32_t eb = 0; eb < nk; eb += kWave) {
      const uint32_t e = eb + lane;
      const uint32_t sl = find_slot(L.epre, e < nk ? e : nk - 1);
      const K4Word kw = shfl_word<LIVE>(k, sl);
      const uint32_t ti = shfl_u32(i, sl);
      const uint32_t wsl = shfl_u32(wi, sl);
      const uint32_t esl = shfl_u32(epre, sl);
      const uint32_t psl = POS ? shfl_u32(lpos, sl) : 0u;
      const TriRows rt = shfl_rows(own, ti);
      const uint32_t tt = shfl_u32(t, ti);
      const uint64_t ta = shfl_u64(aux, ti);
      if (e < nk) {
        const uint32_t r = e - esl;
        const uint32_t j = select_bit(kw.bits, r);
        uint32_t q[3], s2 = 0;
        k4_partners<LIVE>(s, rt, wsl, kw, j, q, &s2);
        // q[0] sits on b's row, q[1] on a's, at bit j of this word
        const uint32_t die = LIVE ? (((kw.Db >> j) & 1u) | (((kw.Da >> j) & 1u) << 1) | (s2 << 2)) : 0u;
        emit(tt, ta, POS ? psl + r : 0u, q, die);
      }
    }
    wave_sync();
  }
}

// ---- support, K4 lists, canonical K4 list -------------------------------------

// Support and owned counts, a wave per 64 triangles: the triangles' row words
// (W up words of a and b, then WC down words) are flattened over the lanes in
// chunks of kSupChunk consecutive words, 64 chunks at a time, so a triangle
// whose lowest vertex has thousands of in-neighbours (thousands of down words)
// is spread over many lanes instead of one, and each lane's slot search and
// shuffles serve several words; the per-triangle sums gather in LDS. Chunks of
// 2 measured best on R-MAT s20 + 64 K30 (53.6 ms; 1, 4 and 8 words: 82, 71 and
// 71 ms): longer chunks spread a load instruction's lanes over more lines.
constexpr uint32_t kSupChunk = 2;

// Compact rows for the support count: ab[i] = {Aout, Aout | Ain} bits of A-row
// word i, cb[i] = the Cm bits of C-row word i. The count reads only bits, and
// the ranked words would fetch their ranks beside them.
__global__ void k_support_rows(BitRows r, uint64_t WA, uint64_t WC, uint2* __restrict__ ab,
                               uint32_t* __restrict__ cb) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t tid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (uint64_t i = tid; i < WA; i += stride) {
    const uint32_t o = r.Aout[i];
    ab[i] = make_uint2(o, o | r.Ain[i].bits);
  }
  for (uint64_t i = tid; i < WC; i += stride) cb[i] = r.Cm[i].bits;
}

__global__ __launch_bounds__(kBlock) void k_support_bits(K4Src s, const uint2* __restrict__ abw,
                                                         const uint32_t* __restrict__ cbw, int32_t* __restrict__ sup,
                                                         uint32_t* __restrict__ own, uint64_t* __restrict__ total) {
  __shared__ uint32_t s_pre[kWPB][kWave + 1];
  __shared__ uint32_t s_sup[kWPB][kWave];
  __shared__ uint32_t s_own[kWPB][kWave];
  const unsigned w = threadIdx.x / kWave;
  const unsigned lane = lane_id();
  uint64_t acc = 0;
  const uint64_t nw = static_cast<uint64_t>(gridDim.x) * kWPB;
  for (uint64_t tile = static_cast<uint64_t>(blockIdx.x) * kWPB + w; tile * kWave < s.T; tile += nw) {
    const uint64_t t = tile * kWave + lane;
    const bool valid = t < s.T;
    uint64_t ra = 0, rb = 0, ca = 0, cb = 0;
    uint32_t W = 0, WC = 0;
    if (valid) {
      const uint32_t a = s.triA[t];
      const uint32_t b = static_cast<uint32_t>(s.triBC[t] >> 32);
      ra = s.rows.rowA[a];
      rb = s.rows.rowA[b];
      W = static_cast<uint32_t>(s.rows.rowA[a + 1] - ra);
      ca = s.rows.rowC[a];
      cb = s.rows.rowC[b];
      WC = static_cast<uint32_t>(s.rows.rowC[a + 1] - ca);  // a and b are out-edges of one vertex
    }
    // chunks: the up words and the down words are chunked separately
    const uint32_t cu = (W + kSupChunk - 1) / kSupChunk;
    const uint32_t chunks = cu + (WC + kSupChunk - 1) / kSupChunk;
    const uint32_t inc = wave_inclusive_sum(chunks);
    const uint32_t nchunks = __shfl(inc, kWave - 1);
    s_pre[w][lane] = inc - chunks;
    if (lane == kWave - 1) s_pre[w][kWave] = inc;
    s_sup[w][lane] = 0;
    s_own[w][lane] = 0;
    wave_sync();
    for (uint32_t base = 0; base < nchunks; base += kWave) {
      const uint32_t item = base + lane;
      const bool act = item < nchunks;
      const uint32_t i = find_slot(s_pre[w], act ? item : nchunks - 1);
      // triangle i's rows, from lane i (every lane takes part)
      const uint64_t rai = shfl_u64(ra, i), rbi = shfl_u64(rb, i), cai = shfl_u64(ca, i), cbi = shfl_u64(cb, i);
      const uint32_t Wi = shfl_u32(W, i), WCi = shfl_u32(WC, i);
      if (act) {
        const uint32_t ci = item - s_pre[w][i];
        const uint32_t cui = (Wi + kSupChunk - 1) / kSupChunk;
        uint32_t c = 0, o = 0;
        // every load of the chunk is issued before any is used
        if (ci < cui) {
          const uint32_t w0 = ci * kSupChunk;
          uint2 xa[kSupChunk], xb[kSupChunk];
#pragma unroll
          for (uint32_t j = 0; j < kSupChunk; ++j) {
            const bool in = w0 + j < Wi;
            xa[j] = in ? abw[rai + w0 + j] : make_uint2(0u, 0u);
            xb[j] = in ? abw[rbi + w0 + j] : make_uint2(0u, 0u);
          }
#pragma unroll
          for (uint32_t j = 0; j < kSupChunk; ++j) {
            c += static_cast<uint32_t>(__popc(xa[j].y & xb[j].y));
            o += static_cast<uint32_t>(__popc(xa[j].x & xb[j].x));
          }
        } else {
          const uint32_t w0 = (ci - cui) * kSupChunk;
          uint32_t xa[kSupChunk], xb[kSupChunk];
#pragma unroll
          for (uint32_t j = 0; j < kSupChunk; ++j) {
            const bool in = w0 + j < WCi;
            xa[j] = in ? cbw[cai + w0 + j] : 0u;
            xb[j] = in ? cbw[cbi + w0 + j] : 0u;
          }
#pragma unroll
          for (uint32_t j = 0; j < kSupChunk; ++j) c += static_cast<uint32_t>(__popc(xa[j] & xb[j]));
        }
        if (c) atomicAdd(&s_sup[w][i], c);
        if (o) atomicAdd(&s_own[w][i], o);
      }
    }
    wave_sync();
    if (valid) {
      sup[t] = static_cast<int32_t>(s_sup[w][lane]);
      own[t] = s_own[w][lane];
      acc += s_sup[w][lane];
    }
    wave_sync();
  }
  acc = wave_sum_u64(acc);
  if (lane_id() == 0 && acc) atomic_add_u64(total, acc);
}

// LIVE: from the peel rows (fresh planes, so every K4), partners by peel id.
template <bool LIVE>
__global__ __launch_bounds__(kBlock) void k_k4_lists(K4Src s, const uint64_t* __restrict__ koff,
                                                     uint32_t* __restrict__ lists) {
  __shared__ TileLds lds[kWPB];
  const unsigned w = threadIdx.x / kWave;
  const uint64_t nw = static_cast<uint64_t>(gridDim.x) * kWPB;
  // tiles in XCD-contiguous runs: consecutive triangles share their lowest
  // vertex's rows, which then stay in one XCD's L2
  for (uint64_t tile = static_cast<uint64_t>(xcd_block()) * kWPB + w; tile * kWave < s.T; tile += nw) {
    const uint64_t t = tile * kWave + lane_id();
    const bool valid = t < s.T;
    // the lists stream out past the cache (they are read back a round at a
    // time much later); the row words of the lookups keep it
    k4_tile<true, LIVE>(s, static_cast<uint32_t>(valid ? t : 0), valid, valid ? koff[t] : 0, lds[w],
                         [&](uint32_t, uint64_t first, uint32_t pos, const uint32_t q[3], uint32_t) {
                           uint32_t* o = lists + 3 * (first + pos);
                           __builtin_nontemporal_store(q[0], o);
                           __builtin_nontemporal_store(q[1], o + 1);
                           __builtin_nontemporal_store(q[2], o + 2);
                         });
  }
}

// Owned K4s: fourth vertex above all of t, i.e. b-role bits of both a and b.
// Only the records [first, first + count) of the canonical list are written,
// at k4[pos - first], and only triangles [t_begin, t_end) are walked: the
// ones whose owned K4s reach that window.
__global__ __launch_bounds__(kBlock) void k_k4_owned(K4Src s, const uint64_t* __restrict__ k4off, uint64_t t_begin,
                                                     uint64_t t_end, uint64_t first, uint64_t count,
                                                     uint64_t* __restrict__ k4) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t last = first + count;
  for (uint64_t t = t_begin + static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < t_end; t += stride) {
    const uint32_t a = s.triA[t];
    const uint32_t b = static_cast<uint32_t>(s.triBC[t] >> 32);
    const uint32_t* Aa = s.rows.Aout + s.rows.rowA[a];
    const uint32_t* Ab = s.rows.Aout + s.rows.rowA[b];
    const uint32_t W = static_cast<uint32_t>(s.rows.rowA[a + 1] - s.rows.rowA[a]);
    uint64_t pos = k4off[t];
    for (uint32_t wi = 0; wi < W && pos < last; ++wi) {
      uint32_t bits = Aa[wi] & Ab[wi];
      while (bits && pos < last) {
        const uint32_t jx = wi * 32 + static_cast<uint32_t>(__builtin_ctz(bits));
        bits &= bits - 1;
        if (pos >= first) k4[pos - first] = (t << 32) | jx;
        ++pos;
      }
    }
  }
}

// out[0], out[1]: the triangles holding owned records `first` and
// `first + count - 1` (the last t with k4off[t] <= index; count >= 1 and
// first + count <= k4off[T]).
__global__ void k_k4_owned_bounds(const uint64_t* __restrict__ k4off, uint64_t T, uint64_t first, uint64_t count,
                                  uint64_t* __restrict__ out) {
  if (blockIdx.x || threadIdx.x) return;
  const uint64_t key[2] = {first, first + count - 1};
  for (int k = 0; k < 2; ++k) {
    uint64_t lo = 0, hi = T + 1;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (k4off[mid] <= key[k]) lo = mid + 1;
      else hi = mid;
    }
    out[k] = lo - 1;
  }
}

// {T(pqs), T(pqr), T(prs), T(qrs)} of the owned K4 (t = T(u,v,w), x at
// position jx of N+(u), u<v<w<x).
__global__ void k_k4_quads(K4Src s, const uint64_t* __restrict__ k4, uint64_t first, uint64_t count,
                           uint32_t* __restrict__ quads) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t key = k4[first + i];
    const uint32_t t = static_cast<uint32_t>(key >> 32);
    const TriRows r = tri_rows(s, t);
    const RankWord* Ix = s.rows.Ain + r.ru0 + static_cast<uint64_t>(static_cast<uint32_t>(key)) * r.W;
    const uint32_t t1 = row_rank(Ix, r.jv);  // T(u,v,x)
    const uint32_t t2 = row_rank(Ix, r.jw);  // T(u,w,x)
    const uint32_t evx = static_cast<uint32_t>(s.triBC[t1]);
    const uint32_t t3 = row_rank(s.rows.Ain + r.rv0 + static_cast<uint64_t>(evx - r.ov) * r.Wv, r.c - r.ov);
    quads[4 * i + 0] = t1;
    quads[4 * i + 1] = t;
    quads[4 * i + 2] = t2;
    quads[4 * i + 3] = t3;
  }
}

// ---- peel ------------------------------------------------------------------

constexpr uint32_t kAlive = 0, kCur = 1, kNext = 2, kGone = 3;
constexpr uint32_t kNoMin = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t tri_state(const uint32_t* st, uint32_t t) {
  return (st[t >> 4] >> ((t & 15u) * 2u)) & 3u;
}

// Low bit of every kAlive field of w.
__device__ __forceinline__ uint32_t alive_fields(uint32_t w) { return ~(w | (w >> 1)) & 0x55555555u; }

// The 16 supports of state word i (dd is padded to whole words).
__device__ __forceinline__ void load_word_dd(const int32_t* dd, uint64_t i, int32_t d[16]) {
  const int4* d4 = reinterpret_cast<const int4*>(dd) + 4 * i;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int4 x = d4[k];
    d[4 * k] = x.x;
    d[4 * k + 1] = x.y;
    d[4 * k + 2] = x.z;
    d[4 * k + 3] = x.w;
  }
}

// Level in force for the round controlled by c: when the frontier ran empty,
// max(level, min alive support).
__device__ __forceinline__ int32_t round_level(const PeelCtl* c) {
  if (!c->need_level) return c->level;
  const uint32_t m = c->minval;
  return (m != kNoMin && static_cast<int32_t>(m) > c->level) ? static_cast<int32_t>(m) : c->level;
}

// Workgroup min of mn into *out, one atomic per workgroup. Every thread of the
// workgroup calls it.
__device__ __forceinline__ void block_min_to(uint32_t mn, uint32_t* out) {
  __shared__ uint32_t wmin[kWPB];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mn = min(mn, static_cast<uint32_t>(__shfl_xor(mn, off)));
  if (lane_id() == 0) wmin[threadIdx.x / kWave] = mn;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kWPB; ++i) mn = min(mn, wmin[i]);
    if (mn != kNoMin) atomic_min_u32(out, mn);
  }
}

__global__ void k_peel_ctl_init(PeelCtl* ctl) {
  for (int s = 0; s < 2; ++s) {
    PeelCtl& c = ctl[s];
    c.ncur = c.nspec = c.nnext = c.sel = 0;
    c.need_level = s == 0 ? 1u : 0u;  // round 0 collects the first level
    c.minval = kNoMin;
    c.level = 0;
    c.round = 0;
    c.seq = static_cast<uint32_t>(s);
    c.done = 0;
    c.error = 0;
    c.peeled = 0;
  }
}

// dd = support, kv = -1, every triangle alive (padding fields kGone), and the
// first level's minimum into ctl[0]. Consecutive threads take consecutive
// triangles (and consecutive state words), so every access is coalesced.
__global__ __launch_bounds__(kBlock) void k_peel_init(PeelBufs p, const int32_t* __restrict__ sup) {
  const uint64_t W = (p.T + 15) / 16;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  uint32_t mn = kNoMin;
  for (uint64_t t = gtid; t < 16 * W; t += stride) {
    if (t < p.T) {
      const int32_t d = sup[t];
      p.dd[t] = d;
      p.kv[t] = -1;
      mn = min(mn, static_cast<uint32_t>(d < 0 ? 0 : d));
    } else {
      p.dd[t] = 0;  // padding of the last state word
    }
  }
  for (uint64_t i = gtid; i < W; i += stride) {
    uint32_t w = 0;  // every field kAlive (0) but the padding past T
    for (uint32_t f = 0; f < 16; ++f)
      if (16 * i + f >= p.T) w |= kGone << (2 * f);
    p.st[i] = w;
  }
  block_min_to(mn, &p.ctl[0].minval);
}

__device__ __forceinline__ void read_states(const PeelBufs& p, const uint32_t q[3], uint32_t sq[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) sq[k] = tri_state(p.st, q[k]);
}

// Where a debit puts the survivors it pushes. A workgroup reserves slots of
// its own region with a counter in LDS: no global counter is shared by the
// grid, since atomics on one address serialize at memory (about 90 M/s on
// MI355X, against some 84 M wave appends in one R-MAT s20 + 64 K30 peel). A
// reservation that does not fit goes to the overflow list (global counter).
struct PushSink {
  uint32_t* region;   // the workgroup's slots
  uint32_t cap;
  uint32_t* cnt;      // LDS: slots reserved so far
  uint32_t* fit;      // LDS: end of the written slots (cap until a reservation misses)
  uint32_t* ovf;      // overflow list and its count
  uint32_t* ovf_cnt;
};

// Appends q[k] for every bit k of pushes, one reservation per wave.
__device__ __forceinline__ void push3(const PushSink& k, const uint32_t q[3], uint32_t pushes) {
  const uint32_t n = static_cast<uint32_t>(__popc(pushes));
  const uint64_t b0 = __ballot(n & 1u);
  const uint64_t b1 = __ballot((n >> 1) & 1u);
  const uint32_t total = static_cast<uint32_t>(__popcll(b0) + 2 * __popcll(b1));
  if (!total) return;
  const unsigned leader = static_cast<unsigned>(__builtin_ctzll(b0 | b1));
  uint32_t base = 0, over = 0;
  if (lane_id() == leader) {
    base = atomicAdd(k.cnt, total);
    if (base + total > k.cap) {
      atomicMin(k.fit, base);
      base = atomic_add_u32(k.ovf_cnt, total);
      over = 1;
    }
  }
  base = __shfl(base, static_cast<int>(leader));
  over = __shfl(over, static_cast<int>(leader));
  uint32_t* dst = (over ? k.ovf : k.region) + base + lanes_below(b0) + 2 * lanes_below(b1);
#pragma unroll
  for (int j = 0; j < 3; ++j)
    if ((pushes >> j) & 1u) *dst++ = q[j];
}

// Owner test and debit for one K4 of dying triangle t with partners q[0..2]
// in states sq[0..2]. A partner that is alive, or was pushed this round (dies
// next round), counts as surviving for ownership; an earlier death means the
// K4 is gone; a same-round death with a lower index owns it. A survivor
// falling from level+1 to level is pushed: its k is fixed now and it dies next
// round. kCur and kGone were set before this kernel started; a push racing with
// the state read can only be seen as kAlive, which gives the same owner, and
// the extra debit of a pushed triangle cannot push it again (its support is
// already at most level).
__device__ __forceinline__ void debit_k4(uint32_t t, const uint32_t q[3], const uint32_t sq[3], const PeelBufs& p,
                                         int32_t level, const PushSink& sink) {
  bool own = true;
#pragma unroll
  for (int k = 0; k < 3; ++k) own = own && sq[k] != kGone && !(sq[k] == kCur && q[k] < t);
  // the three decrements are all issued before any of their results is used
  bool live[3];
  int32_t old[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    live[k] = own && sq[k] == kAlive;
    if (live[k]) old[k] = atomic_sub_i32(&p.dd[q[k]], 1);
  }
  uint32_t pushes = 0;  // bit k: q[k] fell from level+1 to level
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (live[k] && old[k] == level + 1) {
      pushes |= 1u << k;
      atomic_or_u32(&p.st[q[k] >> 4], kNext << ((q[k] & 15u) * 2u));
      p.kv[q[k]] = level;
    }
  }
  push3(sink, q, pushes);
}

// On-demand variant: the K4 came from t's live rows, so q[0] and q[1] are not
// gone, the dying planes say whether each dies in this round (`die` bits 0, 1),
// and the word-level check has already dropped the K4 if one of them owns it.
// q[2]'s planes came with its rank (`die` bits 2, 3): retired means the K4 is
// gone, dying with a lower index means q[2] owns it. A partner that is not
// dying is debited: it is alive or was pushed earlier in this round, and a
// pushed triangle's k is fixed, so its extra debit changes nothing (it cannot
// be pushed again). The planes change only between rounds, so no read here
// races with a write.
__device__ __forceinline__ void destroy_k4_live(uint32_t t, const uint32_t q[3], uint32_t die, const PeelBufs& p,
                                                int32_t level, const PushSink& sink) {
  const bool live2 = (die >> 2) & 1u, dying2 = (die >> 3) & 1u;
  if (!live2 || (dying2 && q[2] < t)) return;
  bool live[3];
  live[0] = !(die & 1u);
  live[1] = !(die & 2u);
  live[2] = !dying2;
  int32_t old[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (live[k]) old[k] = atomic_sub_i32(&p.dd[q[k]], 1);
  uint32_t pushes = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (live[k] && old[k] == level + 1) {
      pushes |= 1u << k;
      atomic_or_u32(&p.st[q[k] >> 4], kNext << ((q[k] & 15u) * 2u));
      p.kv[q[k]] = level;
    }
  }
  push3(sink, q, pushes);
}

// Dying triangles per wave for a frontier of n: spread over every wave of the
// grid first, then up to 64 per wave so small triangles share a wave.
__device__ __forceinline__ uint32_t tile_width(uint32_t n, uint32_t waves) {
  const uint32_t g = n / waves;
  return g < 1 ? 1u : (g > static_cast<uint32_t>(kWave) ? static_cast<uint32_t>(kWave) : g);
}

// The live-row words holding triangle t's three bits, and the bits: t is bit
// jv of a's row, bit jw of b's row (both over N+(u)) and bit inpos(b) of c's
// c-role row (over N-(v)).
struct TriSlots {
  uint64_t w[3];
  uint32_t bit[3];
};

__device__ __forceinline__ TriSlots tri_slots(const K4Src& s, uint32_t t) {
  const uint32_t a = s.triA[t];
  const uint64_t bc = s.triBC[t];
  const uint32_t b = static_cast<uint32_t>(bc >> 32), c = static_cast<uint32_t>(bc);
  const uint32_t ou = s.O[s.src[b]];
  const uint32_t jv = b - ou, jw = a - ou, ip = s.inpos[b];
  TriSlots x;
  x.w[0] = s.rows.rowA[a] + (jv >> 5);
  x.bit[0] = jv & 31u;
  x.w[1] = s.rows.rowA[b] + (jw >> 5);
  x.bit[1] = jw & 31u;
  x.w[2] = s.rows.rowC[c] + (ip >> 5);
  x.bit[2] = ip & 31u;
  return x;
}

// t joins the current frontier: its bits in the dying planes.
__device__ __forceinline__ void mark_dying(const K4Src& s, uint32_t t) {
  const TriSlots x = tri_slots(s, t);
  atomic_or_u32(&s.PA[x.w[0]].dying, 1u << x.bit[0]);
  atomic_or_u32(&s.PA[x.w[1]].dying, 1u << x.bit[1]);
  atomic_or_u32(&s.PC[x.w[2]].dying, 1u << x.bit[2]);
}

__device__ __forceinline__ void clear_plane_bit(uint32_t* live, uint32_t bit) {
  const uint64_t b = 1ull << bit;
  atomic_and_u64(reinterpret_cast<uint64_t*>(live), ~(b | (b << 32)));
}

// t is retired: its bits leave both planes (one 64-bit atomic per word).
__device__ __forceinline__ void unlink_tri(const K4Src& s, uint32_t t) {
  const TriSlots x = tri_slots(s, t);
  clear_plane_bit(&s.PA[x.w[0]].live, x.bit[0]);
  clear_plane_bit(&s.PA[x.w[1]].live, x.bit[1]);
  clear_plane_bit(&s.PC[x.w[2]].live, x.bit[2]);
}

// State words per thread and pass of a level scan.
constexpr int kScanWords = 4;

// Exclusive scan of v over the workgroup (kBlock threads) into *excl; returns
// the total. Every thread calls it; sw holds kWPB words of LDS.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sw, uint32_t* excl) {
  const uint32_t inc = wave_inclusive_sum(v);
  const unsigned wv = threadIdx.x / kWave;
  if (lane_id() == kWave - 1) sw[wv] = inc;
  __syncthreads();
  uint32_t before = 0, total = 0;
#pragma unroll
  for (unsigned i = 0; i < static_cast<unsigned>(kWPB); ++i) {
    before += i < wv ? sw[i] : 0u;
    total += sw[i];
  }
  *excl = before + inc - v;
  __syncthreads();
  return total;
}

// A triangle pushed this round joins the next frontier: the dying planes (on
// demand) or state 10 -> 01.
template <bool UNLINK>
__device__ __forceinline__ void promote(const PeelBufs& p, const K4Src& s, uint32_t t) {
  if (UNLINK) mark_dying(s, t);
  else atomic_xor_u32(&p.st[t >> 4], 3u << ((t & 15u) * 2u));
}

template <bool UNLINK>
__global__ __launch_bounds__(kBlock) void k_peel_finish(PeelBufs p, PeelCtl* in, PeelCtl* out, K4Src s) {
  __shared__ uint32_t wd[kScanWords][kWPB], base;
  __shared__ uint32_t s_sw[kWPB], s_gpre[kPushGroups], s_lpre[kBlock], s_base[kBlock];
  static_assert(kPushGroups <= static_cast<uint32_t>(kBlock), "one group total per thread");
  if (in->done) return;
  const uint32_t F = in->ncur + in->nspec;
  const bool collect_round = F == 0;
  const unsigned wv = threadIdx.x / kWave;
  // The round's pushes: the overflow list nxt[0, nnext), then the debit
  // workgroups' regions, group after group, each at its offset in its group.
  const uint32_t par = static_cast<uint32_t>(in - p.ctl);
  uint32_t* const gc = p.gcnt + par * kPushGroups * kGroupStride;
  const uint32_t ngroups = (p.nwg + p.gsize - 1) / p.gsize;
  uint32_t gpre;
  const uint32_t gtot =
      block_exclusive_scan(threadIdx.x < ngroups ? gc[threadIdx.x * kGroupStride] : 0u, s_sw, &gpre);
  if (threadIdx.x < kPushGroups) s_gpre[threadIdx.x] = gpre;
  // the next round's counters (the other parity) were read by the round
  // before this one; they start the next round at zero
  if (blockIdx.x == 0 && threadIdx.x < ngroups) p.gcnt[((par ^ 1u) * kPushGroups + threadIdx.x) * kGroupStride] = 0;
  __syncthreads();
  const uint32_t novf = in->nnext;
  const uint32_t nn = novf + gtot;
  const bool need = !collect_round && nn == 0;
  const int32_t L = round_level(in);
  const uint32_t* cur = in->sel ? p.buf1 : p.buf0;
  uint32_t* nxt = in->sel ? p.buf0 : p.buf1;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t gsize = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  // Retire the frontier and promote the pushed triangles. The on-demand peel
  // (UNLINK) keeps both in the live and dying planes, which are all its debit
  // reads; its level scans need only "not alive", which states 01 and 10
  // already say, so it leaves the state words alone (two device atomics per
  // triangle fewer). The direct debit reads the states, so it updates them.
  for (uint64_t i = gtid; i < F; i += gsize) {
    const uint32_t t = cur[i];
    if (UNLINK) unlink_tri(s, t);
    else atomic_or_u32(&p.st[t >> 4], 2u << ((t & 15u) * 2u));  // 01 -> 11
  }
  for (uint64_t i = gtid; i < novf; i += gsize) promote<UNLINK>(p, s, nxt[i]);
  if (gtot) {
    // this workgroup's regions [r0, r1), kBlock at a time, their pushes
    // flattened over the threads
    const uint32_t r0 = static_cast<uint32_t>(static_cast<uint64_t>(blockIdx.x) * p.nwg / gridDim.x);
    const uint32_t r1 = static_cast<uint32_t>(static_cast<uint64_t>(blockIdx.x + 1) * p.nwg / gridDim.x);
    for (uint32_t c0 = r0; c0 < r1; c0 += kBlock) {
      const uint32_t r = c0 + threadIdx.x;
      const bool in_range = r < r1;
      const uint32_t c = in_range ? p.wcnt[r] : 0u;
      const uint32_t b = in_range && c ? novf + s_gpre[r / p.gsize] + p.woff[r] : 0u;
      uint32_t lp;
      const uint32_t sum = block_exclusive_scan(c, s_sw, &lp);
      s_lpre[threadIdx.x] = lp;
      s_base[threadIdx.x] = b;
      __syncthreads();
      for (uint32_t e = threadIdx.x; e < sum; e += kBlock) {
        uint32_t lo = 0, hi = kBlock;  // the last region starting at or before e
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (s_lpre[mid] <= e) lo = mid;
          else hi = mid;
        }
        const uint32_t j = e - s_lpre[lo];
        const uint32_t t = p.wslots[static_cast<uint64_t>(c0 + lo) * p.wcap + j];
        nxt[s_base[lo] + j] = t;
        promote<UNLINK>(p, s, t);
      }
      __syncthreads();
    }
  }
  if (need || collect_round) {
    const int32_t take = collect_round ? L : L + 1;  // also the k of the collected triangles
    const uint64_t W = (p.T + 15) / 16;
    uint32_t mn = kNoMin;
    // kScanWords state words per thread and pass (words b0 + k * kBlock +
    // thread), their loads issued together; one block prefix and one
    // collection-count atomic per pass, so the collected triangles stay in
    // index order within the block's kScanWords * kBlock words
    for (uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * kBlock * kScanWords; b0 < W;
         b0 += static_cast<uint64_t>(gridDim.x) * kBlock * kScanWords) {
      uint32_t al[kScanWords], die[kScanWords], nd[kScanWords], id[kScanWords];
#pragma unroll
      for (int k = 0; k < kScanWords; ++k) {
        const uint64_t i = b0 + static_cast<uint64_t>(k) * kBlock + threadIdx.x;
        al[k] = i < W ? alive_fields(p.st[i]) : 0u;
      }
#pragma unroll
      for (int k = 0; k < kScanWords; ++k) {
        die[k] = 0;
        if (al[k]) {
          int32_t d[16];
          load_word_dd(p.dd, b0 + static_cast<uint64_t>(k) * kBlock + threadIdx.x, d);
#pragma unroll
          for (int f = 0; f < 16; ++f) {
            if (!((al[k] >> (2 * f)) & 1u)) continue;
            mn = min(mn, static_cast<uint32_t>(d[f] < 0 ? 0 : d[f]));
            if (d[f] <= take) die[k] |= 1u << f;
          }
        }
        nd[k] = static_cast<uint32_t>(__popc(die[k]));
        id[k] = wave_inclusive_sum(nd[k]);
        if (lane_id() == kWave - 1) wd[k][wv] = id[k];
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        uint32_t td = 0;
        for (int k = 0; k < kScanWords; ++k)
          for (int q = 0; q < kWPB; ++q) {
            const uint32_t x = wd[k][q];
            wd[k][q] = td;
            td += x;
          }
        base = td ? atomic_add_u32(&out->nspec, td) : 0;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kScanWords; ++k) {
        if (!die[k]) continue;
        const uint64_t i = b0 + static_cast<uint64_t>(k) * kBlock + threadIdx.x;
        uint32_t od = base + wd[k][wv] + id[k] - nd[k];
        uint32_t set = 0;
        for (int f = 0; f < 16; ++f) {
          if (!((die[k] >> f) & 1u)) continue;
          const uint32_t t = static_cast<uint32_t>(16 * i + f);
          p.kv[t] = take;
          nxt[od++] = t;
          set |= kCur << (2 * f);
        }
        atomic_or_u32(&p.st[i], set);  // 00 -> 01
        if (UNLINK) {
          for (uint32_t d = die[k]; d; d &= d - 1u) mark_dying(s, static_cast<uint32_t>(16 * i + __builtin_ctz(d)));
        }
      }
      __syncthreads();
    }
    if (need) block_min_to(mn, &out->minval);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const uint64_t peeled = in->peeled + F;
    uint32_t error = 0, done = 0;
    if (collect_round && (!in->need_level || in->minval == kNoMin)) {
      error = in->minval == kNoMin ? 1u : 2u;  // nothing alive / collection found nothing
      done = 1;
    }
    if (peeled >= p.T) done = 1;
    out->ncur = nn;
    out->nnext = 0;
    out->sel = in->sel ^ 1u;
    out->need_level = need ? 1u : 0u;
    out->level = L;
    out->round = in->round + (F ? 1u : 0u);
    out->seq = in->seq + 1;
    out->error = error;
    out->peeled = peeled;
    out->done = done;
    // `in` is read by every workgroup of this kernel, and a late one may
    // start after this one has ended: it is left alone here (the next debit
    // kernel prepares it for reuse)
  }
}

// The debit workgroup's push sink: its region of the round's push slots,
// counters in LDS (s_cnt, s_fit), overflow to ovf / *ovf_cnt. Every thread of
// the workgroup calls it.
__device__ __forceinline__ PushSink open_sink(const PeelBufs& p, uint32_t* ovf, uint32_t* ovf_cnt, uint32_t* s_cnt,
                                              uint32_t* s_fit) {
  if (threadIdx.x == 0) {
    *s_cnt = 0;
    *s_fit = p.wcap;
  }
  __syncthreads();
  return PushSink{p.wslots + static_cast<uint64_t>(blockIdx.x) * p.wcap, p.wcap, s_cnt, s_fit, ovf, ovf_cnt};
}

// Publishes the workgroup's push count (every workgroup, every round, so the
// finish never reads a stale one) and, with pushes, takes its region's offset
// in its group from the counter of the round's parity. Every thread of the
// workgroup calls it.
__device__ __forceinline__ void close_sink(const PeelBufs& p, const PushSink& k, const PeelCtl* in) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t c = min(*k.cnt, *k.fit);
    const uint32_t par = static_cast<uint32_t>(in - p.ctl);
    p.wcnt[blockIdx.x] = c;
    p.woff[blockIdx.x] =
        c ? atomic_add_u32(p.gcnt + (par * kPushGroups + blockIdx.x / p.gsize) * kGroupStride, c) : 0u;
  }
}

// Prepares the slot this round's finish writes: its collection count and its
// min reduction start empty, and a finished peel marks it done too, so that
// rounds still queued in the captured launch do nothing whichever slot they
// read. Only the debit kernel touches `out` (its finish reads `out`'s fields
// only after it), and only at its very end, after the kernel's last read of
// its own control: any store through `out` before that (it may alias `in` as
// far as the compiler knows) turns the round's control reads into vector
// loads.
__device__ __forceinline__ void prepare_out_slot(PeelCtl* out, bool done) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    out->nspec = 0;
    out->minval = kNoMin;
    if (done) out->done = 1;
  }
}

// Per-wave LDS of the direct debit: the list extents of the tile's triangles.
struct ListTile {
  uint64_t pre[kWave + 1];  // first entry of each triangle in the tile's flattened entries
  uint64_t first[kWave];    // koff of each triangle
  uint32_t tri[kWave];
};

// Largest index i < 64 with pre[i] <= key (pre non-decreasing, pre[0] = 0):
// the triangle holding flattened entry key (empty triangles share their
// successor's start and are skipped).
__device__ __forceinline__ uint32_t find_slot_u64(const uint64_t* pre, uint64_t key) {
  uint32_t lo = 0, hi = kWave;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pre[mid] <= key) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Direct debit: a wave takes tile_width() dying triangles and flattens their
// K4 list entries over its lanes, two entries per lane and pass, so a short
// list does not leave most of a wave idle. Both entries' loads, then both
// entries' state reads, are issued before either is waited on.
// The direct debit of the n dying triangles cur[0, n): the waves numbered
// wave, wave + nwaves, ... take G of them at a time.
__device__ __forceinline__ void debit_lists(const PeelBufs& p, const uint32_t* __restrict__ cur, uint32_t n,
                                            uint64_t wave, uint64_t nwaves, uint32_t G, ListTile& L,
                                            const uint64_t* __restrict__ koff, const uint32_t* __restrict__ lists,
                                            int32_t level, const PushSink& sink) {
  const unsigned lane = lane_id();
  for (uint64_t tile = wave; tile * G < n; tile += nwaves) {
    const uint64_t i = tile * G + lane;
    uint32_t t = 0;
    uint64_t b = 0, len = 0;
    if (lane < G && i < n) {
      t = cur[i];
      b = koff[t];
      len = koff[t + 1] - b;
    }
    const uint64_t inc = wave_inclusive_sum_u64(len);
    const uint64_t total = __shfl(inc, kWave - 1);
    L.pre[lane] = inc - len;
    if (lane == kWave - 1) L.pre[kWave] = inc;
    L.first[lane] = b;
    L.tri[lane] = t;
    wave_sync();
    for (uint64_t e = lane; e < total; e += 2 * kWave) {
      const bool two = e + kWave < total;
      const uint32_t sa_ = find_slot_u64(L.pre, e);
      const uint32_t ta = L.tri[sa_];
      const uint64_t ja = L.first[sa_] + (e - L.pre[sa_]);
      uint32_t tb = 0;
      uint64_t jb = 0;
      if (two) {
        const uint32_t sb_ = find_slot_u64(L.pre, e + kWave);
        tb = L.tri[sb_];
        jb = L.first[sb_] + (e + kWave - L.pre[sb_]);
      }
      uint32_t qa[3], qb[3] = {0u, 0u, 0u}, sa[3], sb[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) qa[k] = __builtin_nontemporal_load(lists + 3 * ja + k);
      if (two) {
#pragma unroll
        for (int k = 0; k < 3; ++k) qb[k] = __builtin_nontemporal_load(lists + 3 * jb + k);
      }
      read_states(p, qa, sa);
      if (two) read_states(p, qb, sb);
      debit_k4(ta, qa, sa, p, level, sink);
      if (two) debit_k4(tb, qb, sb, p, level, sink);
    }
    wave_sync();
  }
}

// The on-demand debit of the n dying triangles cur[0, n): the waves numbered
// wave, wave + nwaves, ... take G of them at a time and enumerate their K4s
// from the live bit rows, flattened over the lanes.
__device__ __forceinline__ void debit_bits(const PeelBufs& p, const K4Src& s, const uint32_t* __restrict__ cur,
                                           uint32_t n, uint64_t wave, uint64_t nwaves, uint32_t G, TileLds& L,
                                           int32_t level, const PushSink& sink) {
  const unsigned lane = lane_id();
  for (uint64_t tile = wave; tile * G < n; tile += nwaves) {
    const uint64_t i = tile * G + lane;
    const bool valid = lane < G && i < n;
    k4_tile<false, true>(s, valid ? cur[i] : 0u, valid, 0, L,
                         [&](uint32_t tt, uint64_t, uint32_t, const uint32_t q[3], uint32_t die) {
                           destroy_k4_live(tt, q, die, p, level, sink);
                         });
  }
}

// Direct debit: a wave takes tile_width() dying triangles and flattens their
// K4 list entries over its lanes, two entries per lane and pass, so a short
// list does not leave most of a wave idle. Both entries' loads, then both
// entries' state reads, are issued before either is waited on.
__global__ __launch_bounds__(kBlock) void k_peel_debit_lists(PeelBufs p, PeelCtl* in, PeelCtl* out,
                                                             const uint64_t* __restrict__ koff,
                                                             const uint32_t* __restrict__ lists, uint32_t tile_cap) {
  __shared__ ListTile lds[kWPB];
  const bool done = in->done;
  const int32_t level = round_level(in);
  const uint32_t* cur = in->sel ? p.buf1 : p.buf0;
  uint32_t* nxt = in->sel ? p.buf0 : p.buf1;
  const uint32_t n = done ? 0u : in->ncur + in->nspec;
  const unsigned w = threadIdx.x / kWave;
  const uint64_t nw = static_cast<uint64_t>(gridDim.x) * kWPB;
  const uint32_t G = min(tile_width(n, static_cast<uint32_t>(nw)), tile_cap);
  __shared__ uint32_t s_cnt, s_fit;
  const PushSink sink = open_sink(p, nxt, &in->nnext, &s_cnt, &s_fit);
  debit_lists(p, cur, n, static_cast<uint64_t>(blockIdx.x) * kWPB + w, nw, G, lds[w], koff, lists, level, sink);
  close_sink(p, sink, in);
  prepare_out_slot(out, done);
}

// On-demand debit: a wave takes tile_width() dying triangles; their K4s are
// enumerated from the live bit rows, flattened over the lanes. Its waves wait
// on memory 84% of their cycles (R-MAT s20 + 64 K30), so it runs six per SIMD
// (80 VGPRs, no spills): 1357 -> 1323 ms there; seven spill and are slower.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_peel_debit_bits(PeelBufs p, PeelCtl* in, PeelCtl* out, K4Src s,
                                                            uint32_t tile_cap) {
  __shared__ TileLds lds[kWPB];
  const bool done = in->done;
  const int32_t level = round_level(in);
  const uint32_t* cur = in->sel ? p.buf1 : p.buf0;
  uint32_t* nxt = in->sel ? p.buf0 : p.buf1;
  const uint32_t n = done ? 0u : in->ncur + in->nspec;
  const unsigned w = threadIdx.x / kWave;
  const uint64_t nw = static_cast<uint64_t>(gridDim.x) * kWPB;
  const uint32_t G = min(tile_width(n, static_cast<uint32_t>(nw)), tile_cap);
  __shared__ uint32_t s_cnt, s_fit;
  const PushSink sink = open_sink(p, nxt, &in->nnext, &s_cnt, &s_fit);
  debit_bits(p, s, cur, n, static_cast<uint64_t>(blockIdx.x) * kWPB + w, nw, G, lds[w], level, sink);
  close_sink(p, sink, in);
  prepare_out_slot(out, done);
}

// ---- one workgroup per graph ---------------------------------------------------
//
// The graphs of a batch share no triangle and no K4, so each one can be peeled
// by itself. k_peel_graphs peels a whole batch in one launch: a workgroup
// takes graphs from a queue, largest first, and runs one graph's rounds with
// workgroup barriers between the debit and the state advance. No graph waits
// for another's rounds, and nothing returns to the host between rounds. The
// rounds are those of the graph decomposed alone, with the same level rule,
// collection and ownership as k_peel_debit_* + k_peel_finish:
// - when the frontier is empty, level = max(level, min alive support) and the
//   alive triangles with support <= level form the frontier (their k);
// - a survivor whose support falls from level+1 to level is pushed to the next
//   round (its k = level).
// A graph with no K4 gets k = 0 and no round, as its own decomposition does.
// Every round retires at least one triangle, so a graph ends after at most
// its triangle count of rounds.

// The 2-bit fields of state word i that hold triangles of [t0, t1) (low bit
// of each field; the word overlaps the range).
__device__ __forceinline__ uint32_t field_mask(uint64_t i, uint64_t t0, uint64_t t1) {
  const uint64_t lo = 16 * i;
  uint32_t m = 0x55555555u;
  if (t0 > lo) m &= 0x55555555u << (2 * (t0 - lo));
  if (t1 < lo + 16) m &= (1u << (2 * (t1 - lo))) - 1u;
  return m;
}

// Workgroup of the per-graph peel: 16 waves, so that the largest graphs of a
// batch (the launch's critical path) get a CU's worth of waves.
constexpr int kGraphBlock = 1024;
constexpr int kGraphWaves = kGraphBlock / kWave;

template <bool LIVE>
__global__ __launch_bounds__(kGraphBlock) void k_peel_graphs(PeelBufs p, K4Src s, const uint64_t* __restrict__ koff,
                                                             const uint32_t* __restrict__ lists, GraphQueue q,
                                                             uint32_t tile_cap) {
  __shared__ ListTile lt[LIVE ? 1 : kGraphWaves];
  __shared__ TileLds tl[LIVE ? kGraphWaves : 1];
  __shared__ uint32_t s_g, s_n, s_nn, s_fit, s_min, s_max;
  const unsigned w = threadIdx.x / kWave;
  const unsigned lane = lane_id();
  for (;;) {
    if (threadIdx.x == 0) s_g = atomic_add_u32(q.next, 1u);
    __syncthreads();
    const uint32_t gi = s_g;
    if (gi >= q.count) return;
    const uint32_t g = q.order[gi];
    const uint64_t t0 = q.tri_first[g], t1 = q.tri_first[g + 1];
    const uint64_t w0 = t0 / 16, w1 = (t1 + 15) / 16;
    uint32_t* cur = p.buf0 + t0;
    uint32_t* nxt = p.buf1 + t0;
    int32_t level = 0;
    uint32_t rounds = 0, n = 0;
    bool first = true;
    for (;;) {
      if (n == 0) {
        // min alive support of the graph (and the first time the max: no K4 at all?)
        if (threadIdx.x == 0) {
          s_min = kNoMin;
          s_max = 0;
        }
        __syncthreads();
        uint32_t mn = kNoMin, mx = 0;
        for (uint64_t i = w0 + threadIdx.x; i < w1; i += kGraphBlock) {
          const uint32_t al = alive_fields(p.st[i]) & field_mask(i, t0, t1);
          if (!al) continue;
          int32_t d[16];
          load_word_dd(p.dd, i, d);
#pragma unroll
          for (int f = 0; f < 16; ++f) {
            if (!((al >> (2 * f)) & 1u)) continue;
            const uint32_t v = static_cast<uint32_t>(d[f] < 0 ? 0 : d[f]);
            mn = min(mn, v);
            mx = max(mx, v);
          }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          mn = min(mn, static_cast<uint32_t>(__shfl_xor(mn, off)));
          mx = max(mx, static_cast<uint32_t>(__shfl_xor(mx, off)));
        }
        if (lane == 0) {
          atomicMin(&s_min, mn);
          atomicMax(&s_max, mx);
        }
        __syncthreads();
        const uint32_t m = s_min;
        if (m == kNoMin) break;  // every triangle of the graph has its k
        if (first && s_max == 0) {
          for (uint64_t t = t0 + threadIdx.x; t < t1; t += kGraphBlock) {
            p.kv[t] = 0;
            atomic_or_u32(&p.st[t >> 4], kGone << ((t & 15u) * 2u));
          }
          break;
        }
        first = false;
        level = max(level, static_cast<int32_t>(m));
        if (threadIdx.x == 0) s_n = 0;
        __syncthreads();
        // the alive triangles with support <= level; every lane runs every
        // pass (wave prefix sums), one LDS append per wave and pass
        for (uint64_t b = w0; b < w1; b += kGraphBlock) {
          const uint64_t i = b + threadIdx.x;
          uint32_t die = 0;
          if (i < w1) {
            const uint32_t al = alive_fields(p.st[i]) & field_mask(i, t0, t1);
            if (al) {
              int32_t d[16];
              load_word_dd(p.dd, i, d);
#pragma unroll
              for (int f = 0; f < 16; ++f)
                if (((al >> (2 * f)) & 1u) && d[f] <= level) die |= 1u << f;
            }
          }
          const uint32_t nd = static_cast<uint32_t>(__popc(die));
          const uint32_t inc = wave_inclusive_sum(nd);
          const uint32_t tot = __shfl(inc, kWave - 1);
          uint32_t base = 0;
          if (lane == kWave - 1 && tot) base = atomicAdd(&s_n, tot);
          base = __shfl(base, kWave - 1);
          if (die) {
            uint32_t o = base + inc - nd, set = 0;
            for (uint32_t d = die; d; d &= d - 1u) {
              const uint32_t f = static_cast<uint32_t>(__builtin_ctz(d));
              const uint32_t t = static_cast<uint32_t>(16 * i + f);
              p.kv[t] = level;
              cur[o++] = t;
              set |= kCur << (2 * f);
              if constexpr (LIVE) mark_dying(s, t);
            }
            atomic_or_u32(&p.st[i], set);  // 00 -> 01
          }
        }
        __syncthreads();
        n = s_n;
        if (n == 0) {  // cannot happen: the min alive support is at most level
          if (threadIdx.x == 0) atomic_or_u32(q.error, 1u);
          break;
        }
      }
      if (++rounds > t1 - t0) {  // every round retires a triangle
        if (threadIdx.x == 0) atomic_or_u32(q.error, 2u);
        break;
      }
      if (threadIdx.x == 0) {
        s_nn = 0;
        s_fit = 0xFFFFFFFFu;
      }
      __syncthreads();
      const uint32_t G = min(tile_width(n, kGraphWaves), tile_cap);
      // the graph's own next buffer holds every push (at most its triangles)
      const PushSink sink{nxt, 0xFFFFFFFFu, &s_nn, &s_fit, nxt, &s_nn};
      if constexpr (LIVE) debit_bits(p, s, cur, n, w, kGraphWaves, G, tl[w], level, sink);
      else debit_lists(p, cur, n, w, kGraphWaves, G, lt[w], koff, lists, level, sink);
      __syncthreads();
      const uint32_t nn = s_nn;
      // as in k_peel_finish: the on-demand peel keeps its states in the planes
      for (uint32_t i = threadIdx.x; i < n; i += kGraphBlock) {
        const uint32_t t = cur[i];
        if constexpr (LIVE) unlink_tri(s, t);
        else atomic_or_u32(&p.st[t >> 4], 2u << ((t & 15u) * 2u));  // 01 -> 11
      }
      for (uint32_t i = threadIdx.x; i < nn; i += kGraphBlock) {
        const uint32_t t = nxt[i];
        if constexpr (LIVE) mark_dying(s, t);
        else atomic_xor_u32(&p.st[t >> 4], 3u << ((t & 15u) * 2u));  // 10 -> 01
      }
      __syncthreads();
      uint32_t* const tmp = cur;
      cur = nxt;
      nxt = tmp;
      n = nn;
    }
    if (threadIdx.x == 0) {
      q.rounds[g] = rounds;
      q.level[g] = level;
    }
    __syncthreads();  // s_g is rewritten next
  }
}

// Sum of the residual supports dd[0..T) into *total (which starts at zero).
// Every debit lowers one dd by one, so the peel's debits are 4Q minus this sum.
__global__ __launch_bounds__(kBlock) void k_residual_sum(const int32_t* __restrict__ dd, uint64_t T, int64_t* total) {
  __shared__ int64_t wsum[kWPB];
  int64_t acc = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < T;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    acc += dd[i];
  acc = static_cast<int64_t>(wave_sum_u64(static_cast<uint64_t>(acc)));
  if (lane_id() == 0) wsum[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int i = 0; i < kWPB; ++i) t += wsum[i];
    if (t) atomic_add_u64(reinterpret_cast<uint64_t*>(total), static_cast<uint64_t>(t));
  }
}

// Peel rows: PA holds the Aout bits with their peel ranks, the live plane over
// Aout | Ain and an empty dying plane; PC the Cm bits as the live plane.
__global__ void k_live_rows(K4Src s, uint64_t WA, uint64_t WC, const uint32_t* __restrict__ IO,
                            const uint32_t* __restrict__ IE, uint32_t n, uint32_t* __restrict__ icut) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t tid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (uint64_t i = tid; i < WA; i += stride) {
    const uint32_t o = s.rows.Aout[i];
    s.PA[i] = PeelWord{o, s.rows.AoutR[i], o | s.rows.Ain[i].bits, 0u};
  }
  for (uint64_t i = tid; i < WC; i += stride) s.PC[i] = PlaneWord{s.rows.Cm[i].bits, 0u};
  for (uint64_t u = tid; u < n; u += stride)  // in-lists are sorted by source id
    icut[u] = static_cast<uint32_t>(lower_bound_dev<uint32_t>(IE, IO[u], IO[u + 1], static_cast<uint32_t>(u)) - IO[u]);
}

}  // namespace

void launch_residual_sum(const PeelBufs& p, int64_t* total, hipStream_t s) {
  if (!p.T) return;
  hipLaunchKernelGGL(k_residual_sum, dim3(grid_for(p.T, kBlock, 2048)), dim3(kBlock), 0, s, p.dd, p.T, total);
  NUC34_LAUNCH_CHECK();
}

void launch_support_bits(const K4Src& s, uint64_t WA, uint64_t WC, uint2* ab, uint32_t* cb, int32_t* sup,
                         uint32_t* own, uint64_t* total, hipStream_t st) {
  if (!s.T) return;
  const uint64_t most = WA > WC ? WA : WC;
  hipLaunchKernelGGL(k_support_rows, dim3(grid_for(most, 256, 1u << 14)), dim3(256), 0, st, s.rows, WA, WC, ab, cb);
  const uint64_t tiles = (s.T + kWave - 1) / kWave;
  hipLaunchKernelGGL(k_support_bits, dim3(grid_for(tiles, kWPB, 1u << 14)), dim3(kBlock), 0, st, s, ab, cb, sup, own,
                     total);
  NUC34_LAUNCH_CHECK();
}

void launch_k4_lists(const K4Src& s, const uint64_t* koff, uint32_t* lists, unsigned grid_cap, bool live,
                     hipStream_t st) {
  if (!s.T) return;
  const uint64_t tiles = (s.T + kWave - 1) / kWave;
  const dim3 grid(grid_for_xcd(tiles, kWPB, grid_cap ? grid_cap : 1u << 15));
  if (live) hipLaunchKernelGGL(k_k4_lists<true>, grid, dim3(kBlock), 0, st, s, koff, lists);
  else hipLaunchKernelGGL(k_k4_lists<false>, grid, dim3(kBlock), 0, st, s, koff, lists);
  NUC34_LAUNCH_CHECK();
}

void launch_k4_owned_bounds(const uint64_t* k4off, uint64_t T, uint64_t first, uint64_t count, uint64_t* out,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_k4_owned_bounds, dim3(1), dim3(64), 0, st, k4off, T, first, count, out);
  NUC34_LAUNCH_CHECK();
}

void launch_k4_owned(const K4Src& s, const uint64_t* k4off, uint64_t t_begin, uint64_t t_end, uint64_t first,
                     uint64_t count, uint64_t* k4, hipStream_t st) {
  if (t_end <= t_begin || !count) return;
  hipLaunchKernelGGL(k_k4_owned, dim3(grid_for(t_end - t_begin, kBlock, 1u << 14)), dim3(kBlock), 0, st, s, k4off,
                     t_begin, t_end, first, count, k4);
  NUC34_LAUNCH_CHECK();
}

void launch_k4_quads(const K4Src& s, const uint64_t* k4owned, uint64_t first, uint64_t count, uint32_t* quads,
                     hipStream_t st) {
  if (!count) return;
  hipLaunchKernelGGL(k_k4_quads, dim3(grid_for(count, 256, 1u << 14)), dim3(256), 0, st, s, k4owned, first, count,
                     quads);
  NUC34_LAUNCH_CHECK();
}

void launch_live_rows(const K4Src& src, uint64_t WA, uint64_t WC, const uint32_t* IO, const uint32_t* IE, uint32_t n,
                      uint32_t* icut, hipStream_t s) {
  const uint64_t most = WA > WC ? (WA > n ? WA : n) : (WC > n ? WC : n);
  if (!most) return;
  hipLaunchKernelGGL(k_live_rows, dim3(grid_for(most, 256, 1u << 14)), dim3(256), 0, s, src, WA, WC, IO, IE, n, icut);
  NUC34_LAUNCH_CHECK();
}

void launch_peel_init(const PeelBufs& p, const int32_t* sup, hipStream_t s) {
  hipLaunchKernelGGL(k_peel_ctl_init, dim3(1), dim3(1), 0, s, p.ctl);
  hipLaunchKernelGGL(k_peel_init, dim3(grid_for(p.T, kBlock, 4096)), dim3(kBlock), 0, s, p, sup);
  NUC34_LAUNCH_CHECK();
}

void launch_peel_round_direct(const PeelBufs& p, int slot, const uint64_t* koff, const uint32_t* lists,
                              const PeelShape& shape, hipStream_t s) {
  PeelCtl* in = p.ctl + (slot & 1);
  PeelCtl* out = p.ctl + ((slot + 1) & 1);
  hipLaunchKernelGGL(k_peel_debit_lists, dim3(shape.debit_grid), dim3(kBlock), 0, s, p, in, out, koff, lists,
                     static_cast<uint32_t>(shape.tile_cap ? shape.tile_cap : 1));
  hipLaunchKernelGGL(k_peel_finish<false>, dim3(shape.finish_grid), dim3(kBlock), 0, s, p, in, out, K4Src{});
  NUC34_LAUNCH_CHECK();
}

void launch_peel_graphs(const PeelBufs& p, const K4Src& src, const uint64_t* koff, const uint32_t* lists,
                        const GraphQueue& q, const PeelShape& shape, bool live, hipStream_t s) {
  if (!q.count) return;
  const unsigned grid = q.count < shape.debit_grid ? q.count : shape.debit_grid;
  const uint32_t cap = static_cast<uint32_t>(shape.tile_cap ? shape.tile_cap : 1);
  if (live)
    hipLaunchKernelGGL(k_peel_graphs<true>, dim3(grid), dim3(kGraphBlock), 0, s, p, src, koff, lists, q, cap);
  else
    hipLaunchKernelGGL(k_peel_graphs<false>, dim3(grid), dim3(kGraphBlock), 0, s, p, src, koff, lists, q, cap);
  NUC34_LAUNCH_CHECK();
}

void launch_peel_round_ondemand(const PeelBufs& p, int slot, const K4Src& src, const PeelShape& shape, hipStream_t s) {
  PeelCtl* in = p.ctl + (slot & 1);
  PeelCtl* out = p.ctl + ((slot + 1) & 1);
  hipLaunchKernelGGL(k_peel_debit_bits, dim3(shape.debit_grid), dim3(kBlock), 0, s, p, in, out, src,
                     static_cast<uint32_t>(shape.tile_cap ? shape.tile_cap : 1));
  hipLaunchKernelGGL(k_peel_finish<true>, dim3(shape.finish_grid), dim3(kBlock), 0, s, p, in, out, src);
  NUC34_LAUNCH_CHECK();
}

}  // namespace nuc34
