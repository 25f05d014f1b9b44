// (c) four-clique support / enumeration, (d) on-demand re-enumeration and (e)
// the level-synchronous peel.
//
// A four-clique on triangle t (vertices u<v<w, edges a=(u,w), b=(u,v), c=(v,w))
// is a vertex x that appears as third vertex in the incidence lists of all of
// b, a and c; the matching entries carry the three partner triangles
// (for_each_clique_of, nucleus/cuNucleus34_ondemand.cu:136-170). Lists whose
// combined length is at most kSmallL are intersected by one lane (3-way
// zipper); longer ones by a whole wave, 64 entries of each list at a time: the
// lanes hold the chunk of b's list, the chunks of a's and c's lists sit in LDS,
// every entry not above the smallest chunk maximum is resolved by binary search
// and each list advances past the resolved entries.
//
// Peel (cuNucleus34_ondemand.cu:203-305, 599-690; cuNucleus34.cu:107-141):
// level = max(level, min alive support); the triangles with support <= level
// die this round (kv = level, round stamp). A K4 is destroyed by its first
// member to die, same-round ties going to the lowest triangle index, and its
// owner debits the surviving members. A survivor whose support falls from
// level+1 to level is pushed onto the next round's frontier, so rounds after
// the first of a level touch only the triangles that crossed.
#include "cliques.hpp"

namespace nuc34 {

namespace {

constexpr int kWPB = kBlock / kWave;
constexpr uint32_t kNone = 0xFFFFFFFFu;

struct TriEdges {
  uint32_t a, b, c;
};

__device__ __forceinline__ TriEdges tri_edges(const TriView& v, uint64_t t) {
  const uint64_t bc = v.triBC[t];
  return {v.triA[t], static_cast<uint32_t>(bc >> 32), static_cast<uint32_t>(bc)};
}

__device__ __forceinline__ uint64_t tri_len(const TriView& v, const TriEdges& e) {
  return (v.incO[e.a + 1] - v.incO[e.a]) + (v.incO[e.b + 1] - v.incO[e.b]) + (v.incO[e.c + 1] - v.incO[e.c]);
}

// x is the highest vertex of the partner from c's list exactly when c is that
// partner's b-edge: then t holds the K4's lowest three vertices.
__device__ __forceinline__ bool owns_k4(const TriView& v, uint32_t partner_c, uint32_t c) {
  return static_cast<uint32_t>(v.triBC[partner_c] >> 32) == c;
}

// One lane: fn(partner via b, partner via a, partner via c) per K4, ascending x.
template <typename F>
__device__ __forceinline__ void k4_thread(const TriView& v, const TriEdges& e, F&& fn) {
  uint64_t i = v.incO[e.b], iE = v.incO[e.b + 1];
  uint64_t j = v.incO[e.a], jE = v.incO[e.a + 1];
  uint64_t k = v.incO[e.c], kE = v.incO[e.c + 1];
  while (i < iE && j < jE && k < kE) {
    const uint32_t x1 = v.third[i], x2 = v.third[j], x3 = v.third[k];
    if (x1 == x2 && x2 == x3) {
      fn(v.tri[i], v.tri[j], v.tri[k]);
      ++i;
      ++j;
      ++k;
    } else {
      const uint32_t mx = max(x1, max(x2, x3));
      i += (x1 < mx);
      j += (x2 < mx);
      k += (x3 < mx);
    }
  }
}

__device__ __forceinline__ uint32_t lds_lower_bound(const uint32_t* s, uint32_t n, uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (s[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Whole wave on one triangle: fn(match, index in b's list, in a's, in c's) is
// called by every lane once per chunk step, in ascending x order across steps
// and lanes.
template <typename F>
__device__ __forceinline__ void k4_wave(const TriView& v, const TriEdges& e, uint32_t* s1, uint32_t* s2, F&& fn) {
  const unsigned lane = lane_id();
  uint64_t p0 = v.incO[e.b], e0 = v.incO[e.b + 1];
  uint64_t p1 = v.incO[e.a], e1 = v.incO[e.a + 1];
  uint64_t p2 = v.incO[e.c], e2 = v.incO[e.c + 1];
  while (p0 < e0 && p1 < e1 && p2 < e2) {
    const uint32_t n0 = static_cast<uint32_t>(e0 - p0 < kWave ? e0 - p0 : kWave);
    const uint32_t n1 = static_cast<uint32_t>(e1 - p1 < kWave ? e1 - p1 : kWave);
    const uint32_t n2 = static_cast<uint32_t>(e2 - p2 < kWave ? e2 - p2 : kWave);
    const uint32_t x0 = lane < n0 ? v.third[p0 + lane] : kNone;
    const uint32_t x1 = lane < n1 ? v.third[p1 + lane] : kNone;
    const uint32_t x2 = lane < n2 ? v.third[p2 + lane] : kNone;
    const uint32_t m0 = __shfl(x0, static_cast<int>(n0 - 1));
    const uint32_t m1 = __shfl(x1, static_cast<int>(n1 - 1));
    const uint32_t m2 = __shfl(x2, static_cast<int>(n2 - 1));
    const uint32_t mm = min(m0, min(m1, m2));
    s1[lane] = x1;
    s2[lane] = x2;
    wave_sync();
    const bool take0 = lane < n0 && x0 <= mm;
    bool match = false;
    uint32_t j1 = 0, j2 = 0;
    if (take0) {
      j1 = lds_lower_bound(s1, n1, x0);
      j2 = lds_lower_bound(s2, n2, x0);
      match = j1 < n1 && s1[j1] == x0 && j2 < n2 && s2[j2] == x0;
    }
    fn(match, p0 + lane, p1 + j1, p2 + j2);
    p0 += static_cast<uint64_t>(__popcll(__ballot(take0)));
    p1 += static_cast<uint64_t>(__popcll(__ballot(lane < n1 && x1 <= mm)));
    p2 += static_cast<uint64_t>(__popcll(__ballot(lane < n2 && x2 <= mm)));
    wave_sync();
  }
}

// ---- support count ----------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_support_small(TriView v, int32_t* __restrict__ sup,
                                                          uint32_t* __restrict__ own, uint64_t* __restrict__ total,
                                                          uint32_t* __restrict__ bigq,
                                                          uint32_t* __restrict__ bigcount) {
  uint64_t acc = 0;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < v.T; t += stride) {
    const TriEdges e = tri_edges(v, t);
    const bool big = tri_len(v, e) > kSmallL;
    bool got;
    const uint32_t slot = wave_append(bigcount, big, &got);
    if (big) {
      bigq[slot] = static_cast<uint32_t>(t);
      continue;
    }
    int32_t cnt = 0;
    uint32_t ow = 0;
    k4_thread(v, e, [&](uint32_t, uint32_t, uint32_t pc) {
      ++cnt;
      ow += owns_k4(v, pc, e.c) ? 1u : 0u;
    });
    sup[t] = cnt;
    own[t] = ow;
    acc += static_cast<uint64_t>(cnt);
  }
  acc = wave_sum_u64(acc);
  if (lane_id() == 0 && acc) atomic_add_u64(total, acc);
}

__global__ __launch_bounds__(kBlock) void k_support_big(TriView v, const uint32_t* __restrict__ bigq,
                                                        const uint32_t* __restrict__ bigcount,
                                                        int32_t* __restrict__ sup, uint32_t* __restrict__ own,
                                                        uint64_t* __restrict__ total) {
  __shared__ uint32_t s1[kWPB][kWave], s2[kWPB][kWave];
  const unsigned w = threadIdx.x / kWave;
  const uint32_t n = *bigcount;
  const uint32_t nw = gridDim.x * kWPB;
  for (uint32_t q = blockIdx.x * kWPB + w; q < n; q += nw) {
    const uint32_t t = bigq[q];
    const TriEdges e = tri_edges(v, t);
    uint32_t cnt = 0, ow = 0;
    k4_wave(v, e, s1[w], s2[w], [&](bool match, uint64_t, uint64_t, uint64_t i2) {
      cnt += static_cast<uint32_t>(__popcll(__ballot(match)));
      const bool o = match && owns_k4(v, v.tri[i2], e.c);
      ow += static_cast<uint32_t>(__popcll(__ballot(o)));
    });
    if (lane_id() == 0) {
      sup[t] = static_cast<int32_t>(cnt);
      own[t] = ow;
      if (cnt) atomic_add_u64(total, cnt);
    }
  }
}

// ---- support and K4 lists from the bit rows --------------------------------

__global__ __launch_bounds__(kBlock) void k_support_bits(const uint32_t* __restrict__ triA,
                                                         const uint64_t* __restrict__ triBC, uint64_t T, BitRows r,
                                                         int32_t* __restrict__ sup, uint32_t* __restrict__ sup_up,
                                                         uint32_t* __restrict__ own, uint64_t* __restrict__ total) {
  uint64_t acc = 0;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < T; t += stride) {
    const uint32_t a = triA[t];
    const uint32_t b = static_cast<uint32_t>(triBC[t] >> 32);
    const uint64_t ra = r.rowA[a], rb = r.rowA[b];
    const uint32_t W = static_cast<uint32_t>(r.rowA[a + 1] - ra);
    uint32_t up = 0, ow = 0, down = 0;
    for (uint32_t i = 0; i < W; ++i) {
      const uint32_t oa = r.Aout[ra + i], ob = r.Aout[rb + i];
      up += static_cast<uint32_t>(__popc((oa | r.Ain[ra + i]) & (ob | r.Ain[rb + i])));
      ow += static_cast<uint32_t>(__popc(oa & ob));
    }
    const uint64_t ca = r.rowC[a], cb = r.rowC[b];
    const uint32_t WC = static_cast<uint32_t>(r.rowC[a + 1] - ca);
    for (uint32_t i = 0; i < WC; ++i) down += static_cast<uint32_t>(__popc(r.Cm[ca + i] & r.Cm[cb + i]));
    sup[t] = static_cast<int32_t>(up + down);
    sup_up[t] = up;
    own[t] = ow;
    acc += up + down;
  }
  acc = wave_sum_u64(acc);
  if (lane_id() == 0 && acc) atomic_add_u64(total, acc);
}

__device__ __forceinline__ void put3(uint32_t* __restrict__ lists, uint64_t slot, uint32_t x, uint32_t y, uint32_t z) {
  uint32_t* q = lists + 3 * slot;
  q[0] = x;
  q[1] = y;
  q[2] = z;
}

// Direct K4 lists, step 1: each triangle lists the K4s it owns (fourth vertex
// above all three of its vertices) as (t << 32 | position of x in N+(u)).
__global__ __launch_bounds__(kBlock) void k_k4_owned(const uint32_t* __restrict__ triA,
                                                     const uint64_t* __restrict__ triBC, uint64_t T, BitRows r,
                                                     const uint64_t* __restrict__ k4off, uint64_t* __restrict__ k4) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < T; t += stride) {
    const uint32_t a = triA[t];
    const uint32_t b = static_cast<uint32_t>(triBC[t] >> 32);
    const uint32_t* Aa = r.Aout + r.rowA[a];
    const uint32_t* Ab = r.Aout + r.rowA[b];
    const uint32_t W = static_cast<uint32_t>(r.rowA[a + 1] - r.rowA[a]);
    uint64_t pos = k4off[t];
    for (uint32_t wi = 0; wi < W; ++wi) {
      uint32_t bits = Aa[wi] & Ab[wi];
      while (bits) {
        const uint32_t jx = wi * 32 + static_cast<uint32_t>(__builtin_ctz(bits));
        bits &= bits - 1;
        k4[pos++] = (t << 32) | jx;
      }
    }
  }
}

// The four triangles of the K4 {u < v < w < x} owned by t = T(u,v,w), x at
// position jx of N+(u): T(u,v,x) and T(u,w,x) from the a-role rows of (u,x),
// T(v,w,x) from the a-role row of (v,x) = c-edge of T(u,v,x).
struct K4Ids {
  uint32_t t, t1, t2, t3;  // T(u,v,w), T(u,v,x), T(u,w,x), T(v,w,x)
  uint32_t a, b, c, c1;    // edges (u,w), (u,v), (v,w), (v,x)
  uint32_t ou, jv, jw, ex;
};

__device__ __forceinline__ K4Ids k4_ids(uint64_t key, const uint32_t* __restrict__ triA,
                                        const uint64_t* __restrict__ triBC, const uint32_t* __restrict__ O,
                                        const uint32_t* __restrict__ E, const uint32_t* __restrict__ src,
                                        const BitRows& r, const uint64_t* __restrict__ offA) {
  K4Ids k;
  k.t = static_cast<uint32_t>(key >> 32);
  const uint32_t jx = static_cast<uint32_t>(key);
  k.a = triA[k.t];
  const uint64_t bc = triBC[k.t];
  k.b = static_cast<uint32_t>(bc >> 32);
  k.c = static_cast<uint32_t>(bc);
  k.ou = O[src[k.b]];
  k.jv = k.b - k.ou;
  k.jw = k.a - k.ou;
  k.ex = k.ou + jx;
  const uint32_t* Ix = r.Ain + r.rowA[k.ex];
  k.t1 = static_cast<uint32_t>(offA[k.ex] + bits_below(Ix, k.jv));
  k.t2 = static_cast<uint32_t>(offA[k.ex] + bits_below(Ix, k.jw));
  k.c1 = static_cast<uint32_t>(triBC[k.t1]);
  const uint32_t ov = O[E[k.b]];
  k.t3 = static_cast<uint32_t>(offA[k.c1] + bits_below(r.Ain + r.rowA[k.c1], k.c - ov));
  return k;
}

// Direct K4 lists, step 2: one thread per K4; an entry (the other three
// triangle ids) in the list of each of its four triangles, at a position fixed
// by prefix counts: up-K4s (fourth vertex above the triangle's lowest vertex)
// first, in fourth-vertex order, then down-K4s in order of the lowest vertex's
// in-list position.
__global__ __launch_bounds__(kBlock) void k_k4_place(const uint64_t* __restrict__ k4, uint64_t Q,
                                                     const uint32_t* __restrict__ triA,
                                                     const uint64_t* __restrict__ triBC,
                                                     const uint32_t* __restrict__ O, const uint32_t* __restrict__ E,
                                                     const uint32_t* __restrict__ src,
                                                     const uint32_t* __restrict__ inpos, BitRows r,
                                                     const uint64_t* __restrict__ offA,
                                                     const uint64_t* __restrict__ koff,
                                                     const uint32_t* __restrict__ sup_up,
                                                     uint32_t* __restrict__ lists) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t q = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < Q; q += stride) {
    const K4Ids k = k4_ids(k4[q], triA, triBC, O, E, src, r, offA);
    const uint32_t* Aa = r.Aout + r.rowA[k.a];
    const uint32_t* Ab = r.Aout + r.rowA[k.b];
    const uint32_t* Ia = r.Ain + r.rowA[k.a];
    const uint32_t* Ib = r.Ain + r.rowA[k.b];
    const uint32_t* Ax = r.Aout + r.rowA[k.ex];
    const uint32_t* Ix = r.Ain + r.rowA[k.ex];
    const uint32_t jx = k.ex - k.ou;
    put3(lists, koff[k.t] + sym_bits_below(Aa, Ia, Ab, Ib, jx), k.t1, k.t2, k.t3);
    put3(lists, koff[k.t1] + sym_bits_below(Ab, Ib, Ax, Ix, k.jw), k.t, k.t2, k.t3);
    put3(lists, koff[k.t2] + sym_bits_below(Aa, Ia, Ax, Ix, k.jv), k.t, k.t1, k.t3);
    put3(lists,
         koff[k.t3] + sup_up[k.t3] + and_bits_below(r.Cm + r.rowC[k.c], r.Cm + r.rowC[k.c1], inpos[k.b]), k.t,
         k.t1, k.t2);
  }
}

__global__ void k_k4_quads(const uint64_t* __restrict__ k4, uint64_t first, uint64_t count,
                           const uint32_t* __restrict__ triA, const uint64_t* __restrict__ triBC,
                           const uint32_t* __restrict__ O, const uint32_t* __restrict__ E,
                           const uint32_t* __restrict__ src, BitRows r, const uint64_t* __restrict__ offA,
                           uint32_t* __restrict__ quads) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const K4Ids k = k4_ids(k4[first + i], triA, triBC, O, E, src, r, offA);
    quads[4 * i + 0] = k.t1;
    quads[4 * i + 1] = k.t;
    quads[4 * i + 2] = k.t2;
    quads[4 * i + 3] = k.t3;
  }
}

// ---- peel ------------------------------------------------------------------

constexpr int32_t kDead = 0x7FFFFFFF;
constexpr uint32_t kAliveRound = 0xFFFFFFFFu;

// Level in force for this round: when the frontier ran empty, the new level is
// max(level, min alive support) from k_peel_level (committed by k_peel_finish).
__device__ __forceinline__ int32_t round_level(const PeelCtl* c) {
  if (!c->need_level) return c->level;
  const uint32_t m = c->minval;
  return (m != 0xFFFFFFFFu && static_cast<int32_t>(m) > c->level) ? static_cast<int32_t>(m) : c->level;
}

__global__ __launch_bounds__(kBlock) void k_peel_init(PeelBufs p, const int32_t* __restrict__ sup) {
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < p.T;
       t += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    p.dd[t] = sup[t];
    p.rnd[t] = kAliveRound;
    p.kv[t] = -1;
    p.alive0[t] = static_cast<uint32_t>(t);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    PeelCtl* c = p.ctl;
    c->count[0] = c->count[1] = 0;
    c->parity = 0;
    c->need_level = 1;
    c->minval = 0xFFFFFFFFu;
    c->level = 0;
    c->round = 0;
    c->done = 0;
    c->peeled = 0;
    c->big = 0;
    c->error = 0;
    c->nalive[0] = static_cast<uint32_t>(p.T);
    c->nalive[1] = 0;
    c->apar = 0;
  }
}

// Level selection, step 1 (only when the frontier ran empty): min support over
// the alive list; one atomic per workgroup.
__global__ __launch_bounds__(kBlock) void k_peel_level(PeelBufs p) {
  const PeelCtl* c = p.ctl;
  if (c->done || !c->need_level) return;
  __shared__ uint32_t wmin[kWPB];
  uint32_t mn = 0xFFFFFFFFu;
  const uint32_t* alive = c->apar ? p.alive1 : p.alive0;
  const uint32_t na = c->nalive[c->apar];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < na; i += gridDim.x * blockDim.x) {
    const int32_t d = p.dd[alive[i]];
    if (d != kDead) mn = min(mn, static_cast<uint32_t>(d < 0 ? 0 : d));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mn = min(mn, static_cast<uint32_t>(__shfl_xor(mn, off)));
  if (lane_id() == 0) wmin[threadIdx.x / kWave] = mn;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kWPB; ++i) mn = min(mn, wmin[i]);
    if (mn != 0xFFFFFFFFu) atomic_min_u32(&p.ctl->minval, mn);
  }
}

// Level selection, step 2: every alive triangle with support <= level dies this
// round (k = level); the others move to the other alive buffer. Each thread
// takes kCollectItems list entries per tile, so one pair of atomics reserves
// the output slots of a whole 4096-entry tile.
constexpr int kCollectItems = 16;

__global__ __launch_bounds__(kBlock) void k_peel_collect(PeelBufs p) {
  PeelCtl* c = p.ctl;
  if (c->done || !c->need_level) return;
  __shared__ uint32_t wd[kWPB], wk[kWPB], base[2];
  const int32_t level = round_level(c);
  const uint32_t round = c->round;
  const uint32_t par = c->parity;
  const uint32_t ap = c->apar;
  const uint32_t* alive = ap ? p.alive1 : p.alive0;
  uint32_t* keep = ap ? p.alive0 : p.alive1;
  const uint32_t na = c->nalive[ap];
  uint32_t* cur = par ? p.buf1 : p.buf0;
  const unsigned w = threadIdx.x / kWave;
  const unsigned lane = lane_id();
  constexpr uint32_t tile = kBlock * kCollectItems;
  for (uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * tile; b0 < na; b0 += static_cast<uint64_t>(gridDim.x) * tile) {
    uint32_t tv[kCollectItems];
    uint32_t die = 0, stay = 0;  // bit j: item j
#pragma unroll
    for (int j = 0; j < kCollectItems; ++j) {
      const uint64_t i = b0 + static_cast<uint64_t>(j) * kBlock + threadIdx.x;
      tv[j] = 0;
      if (i < na) {
        const uint32_t t = alive[i];
        const int32_t d = p.dd[t];
        tv[j] = t;
        if (d <= level) die |= 1u << j;
        else if (d != kDead) stay |= 1u << j;
      }
    }
    const uint32_t nd = static_cast<uint32_t>(__popc(die)), nk = static_cast<uint32_t>(__popc(stay));
    const uint32_t id = wave_inclusive_sum(nd), ik = wave_inclusive_sum(nk);
    if (lane == kWave - 1) {
      wd[w] = id;
      wk[w] = ik;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t td = 0, tk = 0;
      for (int q = 0; q < kWPB; ++q) {
        const uint32_t x = wd[q], y = wk[q];
        wd[q] = td;
        wk[q] = tk;
        td += x;
        tk += y;
      }
      base[0] = td ? atomic_add_u32(&c->count[par], td) : 0;
      base[1] = tk ? atomic_add_u32(&c->nalive[ap ^ 1u], tk) : 0;
    }
    __syncthreads();
    uint32_t od = base[0] + wd[w] + id - nd, ok = base[1] + wk[w] + ik - nk;
#pragma unroll
    for (int j = 0; j < kCollectItems; ++j) {
      if (die & (1u << j)) {
        p.rnd[tv[j]] = round;
        p.kv[tv[j]] = level;
        p.dd[tv[j]] = kDead;
        cur[od++] = tv[j];
      } else if (stay & (1u << j)) {
        keep[ok++] = tv[j];
      }
    }
    __syncthreads();
  }
}

// Owner test and debit for one K4 of dying triangle t with partners q[0..2].
// A partner that is alive, or was pushed this round (dies next round), counts
// as surviving for ownership; an earlier death means the K4 is gone; a
// same-round death with a lower index owns it. A survivor falling from
// level+1 to level is pushed: its k is fixed now and it dies next round.
__device__ __forceinline__ void destroy_k4(uint32_t t, const uint32_t q[3], const PeelBufs& p, int32_t level,
                                           uint32_t round, uint32_t* __restrict__ nxt, uint32_t* __restrict__ ncount) {
  uint32_t rq[3];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    rq[s] = p.rnd[q[s]];
    if (rq[s] < round) return;
    if (rq[s] == round && q[s] < t) return;
  }
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    if (rq[s] != kAliveRound) continue;
    const int32_t old = atomic_sub_i32(&p.dd[q[s]], 1);
    if (old == level + 1) {
      p.rnd[q[s]] = round + 1;
      p.kv[q[s]] = level;
      __hip_atomic_store(&p.dd[q[s]], kDead, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      nxt[atomic_add_u32(ncount, 1u)] = q[s];
    }
  }
}

// End of a round, one workgroup: commit the level, flip the buffers, count the
// fixed triangles; an empty next frontier triggers level selection.
__global__ void k_peel_finish(PeelBufs p) {
  if (threadIdx.x != 0) return;
  PeelCtl* c = p.ctl;
  if (c->done) return;
  if (c->need_level) {
    c->level = round_level(c);
    if (c->minval == 0xFFFFFFFFu || c->count[c->parity] == 0) {
      c->error = c->minval == 0xFFFFFFFFu ? 1u : 2u;  // nothing alive / nothing collected
      c->done = 1;
      return;
    }
    c->minval = 0xFFFFFFFFu;
    c->nalive[c->apar] = 0;
    c->apar ^= 1u;
  }
  const uint32_t par = c->parity;
  c->peeled += c->count[par];
  c->count[par] = 0;
  c->parity = par ^ 1u;
  c->round += 1;
  c->big = 0;
  c->need_level = c->count[par ^ 1u] == 0 ? 1u : 0u;
  if (c->peeled >= p.T) c->done = 1;
}

__global__ __launch_bounds__(kBlock) void k_peel_debit_lists(PeelBufs p, const uint64_t* __restrict__ koff,
                                                             const uint32_t* __restrict__ lists) {
  const PeelCtl* c = p.ctl;
  if (c->done) return;
  const int32_t level = round_level(c);
  const uint32_t round = c->round;
  const uint32_t par = c->parity;
  const uint32_t* cur = par ? p.buf1 : p.buf0;
  uint32_t* nxt = par ? p.buf0 : p.buf1;
  uint32_t* ncount = &p.ctl->count[par ^ 1u];
  const uint32_t n = c->count[par];
  const unsigned w = threadIdx.x / kWave;
  const unsigned lane = lane_id();
  const uint32_t nw = gridDim.x * kWPB;
  for (uint32_t i = blockIdx.x * kWPB + w; i < n; i += nw) {
    const uint32_t t = cur[i];
    const uint64_t b = koff[t], e = koff[t + 1];
    for (uint64_t j = b + lane; j < e; j += kWave) {
      const uint32_t* q = lists + 3 * j;
      const uint32_t qq[3] = {q[0], q[1], q[2]};
      destroy_k4(t, qq, p, level, round, nxt, ncount);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_peel_debit_od_small(PeelBufs p, TriView v) {
  PeelCtl* c = p.ctl;
  if (c->done) return;
  const int32_t level = round_level(c);
  const uint32_t round = c->round;
  const uint32_t par = c->parity;
  const uint32_t* cur = par ? p.buf1 : p.buf0;
  uint32_t* nxt = par ? p.buf0 : p.buf1;
  uint32_t* ncount = &c->count[par ^ 1u];
  const uint32_t n = c->count[par];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t t = cur[i];
    const TriEdges e = tri_edges(v, t);
    const bool big = tri_len(v, e) > kSmallL;
    bool got;
    const uint32_t slot = wave_append(&c->big, big, &got);
    if (big) {
      p.bigq[slot] = t;
      continue;
    }
    k4_thread(v, e, [&](uint32_t pb, uint32_t pa, uint32_t pcc) {
      const uint32_t qq[3] = {pb, pa, pcc};
      destroy_k4(t, qq, p, level, round, nxt, ncount);
    });
  }
}

__global__ __launch_bounds__(kBlock) void k_peel_debit_od_big(PeelBufs p, TriView v) {
  __shared__ uint32_t s1[kWPB][kWave], s2[kWPB][kWave];
  PeelCtl* c = p.ctl;
  if (c->done) return;
  const int32_t level = round_level(c);
  const uint32_t round = c->round;
  const uint32_t par = c->parity;
  uint32_t* nxt = par ? p.buf0 : p.buf1;
  uint32_t* ncount = &c->count[par ^ 1u];
  const uint32_t n = c->big;
  const unsigned w = threadIdx.x / kWave;
  const uint32_t nw = gridDim.x * kWPB;
  for (uint32_t q = blockIdx.x * kWPB + w; q < n; q += nw) {
    const uint32_t t = p.bigq[q];
    const TriEdges e = tri_edges(v, t);
    k4_wave(v, e, s1[w], s2[w], [&](bool match, uint64_t i0, uint64_t i1, uint64_t i2) {
      if (match) {
        const uint32_t qq[3] = {v.tri[i0], v.tri[i1], v.tri[i2]};
        destroy_k4(t, qq, p, level, round, nxt, ncount);
      }
    });
  }
}

constexpr unsigned kRoundGrid = 2048;

}  // namespace

void launch_support(const TriView& v, int32_t* support, uint32_t* owned, uint64_t* total, uint32_t* bigq,
                    uint32_t* bigcount, hipStream_t s) {
  if (!v.T) return;
  hipLaunchKernelGGL(k_support_small, dim3(grid_for(v.T, kBlock, 1u << 14)), dim3(kBlock), 0, s, v, support, owned,
                     total, bigq, bigcount);
  NUC34_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_support_big, dim3(8192), dim3(kBlock), 0, s, v, bigq, bigcount, support, owned, total);
  NUC34_LAUNCH_CHECK();
}

void launch_support_bits(const uint32_t* triA, const uint64_t* triBC, uint64_t T, const BitRows& rows, int32_t* sup,
                         uint32_t* sup_up, uint32_t* own, uint64_t* total, hipStream_t s) {
  if (!T) return;
  hipLaunchKernelGGL(k_support_bits, dim3(grid_for(T, kBlock, 1u << 14)), dim3(kBlock), 0, s, triA, triBC, T, rows,
                     sup, sup_up, own, total);
  NUC34_LAUNCH_CHECK();
}

void launch_k4_owned(const uint32_t* triA, const uint64_t* triBC, uint64_t T, const BitRows& rows,
                     const uint64_t* k4off, uint64_t* k4, hipStream_t s) {
  if (!T) return;
  hipLaunchKernelGGL(k_k4_owned, dim3(grid_for(T, kBlock, 1u << 14)), dim3(kBlock), 0, s, triA, triBC, T, rows, k4off,
                     k4);
  NUC34_LAUNCH_CHECK();
}

void launch_k4_place(const uint64_t* k4, uint64_t Q, const uint32_t* triA, const uint64_t* triBC, const uint32_t* O,
                     const uint32_t* E, const uint32_t* src, const uint32_t* inpos, const BitRows& rows,
                     const uint64_t* offA, const uint64_t* koff, const uint32_t* sup_up, uint32_t* lists,
                     hipStream_t s) {
  if (!Q) return;
  hipLaunchKernelGGL(k_k4_place, dim3(grid_for(Q, kBlock, 1u << 16)), dim3(kBlock), 0, s, k4, Q, triA, triBC, O, E,
                     src, inpos, rows, offA, koff, sup_up, lists);
  NUC34_LAUNCH_CHECK();
}

void launch_k4_quads(const uint64_t* k4owned, uint64_t first, uint64_t count, const uint32_t* triA,
                     const uint64_t* triBC, const uint32_t* O, const uint32_t* E, const uint32_t* src,
                     const BitRows& rows, const uint64_t* offA, uint32_t* quads, hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(k_k4_quads, dim3(grid_for(count, 256, 1u << 14)), dim3(256), 0, s, k4owned, first, count, triA,
                     triBC, O, E, src, rows, offA, quads);
  NUC34_LAUNCH_CHECK();
}

constexpr unsigned kScanGrid = 512;

void launch_peel_init(const PeelBufs& p, const int32_t* sup, hipStream_t s) {
  hipLaunchKernelGGL(k_peel_init, dim3(grid_for(p.T ? p.T : 1, kBlock, 4096)), dim3(kBlock), 0, s, p, sup);
  NUC34_LAUNCH_CHECK();
}

void launch_peel_round_direct(const PeelBufs& p, const uint64_t* koff, const uint32_t* lists, hipStream_t s) {
  hipLaunchKernelGGL(k_peel_level, dim3(kScanGrid), dim3(kBlock), 0, s, p);
  hipLaunchKernelGGL(k_peel_collect, dim3(kScanGrid), dim3(kBlock), 0, s, p);
  hipLaunchKernelGGL(k_peel_debit_lists, dim3(kRoundGrid), dim3(kBlock), 0, s, p, koff, lists);
  hipLaunchKernelGGL(k_peel_finish, dim3(1), dim3(kWave), 0, s, p);
  NUC34_LAUNCH_CHECK();
}

void launch_peel_round_ondemand(const PeelBufs& p, const TriView& v, hipStream_t s) {
  hipLaunchKernelGGL(k_peel_level, dim3(kScanGrid), dim3(kBlock), 0, s, p);
  hipLaunchKernelGGL(k_peel_collect, dim3(kScanGrid), dim3(kBlock), 0, s, p);
  hipLaunchKernelGGL(k_peel_debit_od_small, dim3(kRoundGrid), dim3(kBlock), 0, s, p, v);
  hipLaunchKernelGGL(k_peel_debit_od_big, dim3(kRoundGrid), dim3(kBlock), 0, s, p, v);
  hipLaunchKernelGGL(k_peel_finish, dim3(1), dim3(kWave), 0, s, p);
  NUC34_LAUNCH_CHECK();
}

}  // namespace nuc34
