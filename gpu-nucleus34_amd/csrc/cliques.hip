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

constexpr uint32_t kAlive = 0, kCur = 1, kNext = 2, kGone = 3;
constexpr uint32_t kNoMin = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t tri_state(const uint32_t* st, uint32_t t) {
  return (st[t >> 4] >> ((t & 15u) * 2u)) & 3u;
}

// Start of a round for 16 fields at once: kCur -> kGone, kNext -> kCur, kAlive
// and kGone unchanged (new high bit = old low bit, new low bit = low | high).
__device__ __forceinline__ uint32_t advance_states(uint32_t w) {
  const uint32_t lo = w & 0x55555555u, hi = (w >> 1) & 0x55555555u;
  return (lo << 1) | lo | hi;
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
    c.ncur = c.nnext = c.sel = 0;
    c.need_level = s == 0 ? 1u : 0u;
    c.minval = kNoMin;
    c.level = 0;
    c.round = 0;
    c.done = 0;
    c.big = 0;
    c.error = 0;
    c.peeled = 0;
  }
}

// dd = support, kv = -1, every triangle alive (padding fields kGone), and the
// first level's minimum into ctl[0].
__global__ __launch_bounds__(kBlock) void k_peel_init(PeelBufs p, const int32_t* __restrict__ sup) {
  const uint64_t W = (p.T + 15) / 16;
  uint32_t mn = kNoMin;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < W;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t w = 0;
#pragma unroll
    for (uint32_t f = 0; f < 16; ++f) {
      const uint64_t t = 16 * i + f;
      if (t < p.T) {
        const int32_t d = sup[t];
        p.dd[t] = d;
        p.kv[t] = -1;
        mn = min(mn, static_cast<uint32_t>(d < 0 ? 0 : d));
      } else {
        p.dd[t] = 0;
        w |= kGone << (2 * f);
      }
    }
    p.st[i] = w;
  }
  block_min_to(mn, &p.ctl[0].minval);
}

// Level change (the frontier ran empty): every alive triangle with support
// <= level dies this round (k = level). One state word (16 triangles) per
// thread; one atomic per workgroup tile reserves the frontier slots.
__global__ __launch_bounds__(kBlock) void k_peel_collect(PeelBufs p, PeelCtl* in) {
  if (in->done || !in->need_level) return;
  __shared__ uint32_t wd[kWPB], base;
  const int32_t level = round_level(in);
  uint32_t* cur = in->sel ? p.buf1 : p.buf0;
  const uint64_t W = (p.T + 15) / 16;
  const unsigned wv = threadIdx.x / kWave;
  for (uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * kBlock; b0 < W; b0 += static_cast<uint64_t>(gridDim.x) * kBlock) {
    const uint64_t i = b0 + threadIdx.x;
    uint32_t w = 0, die = 0;
    if (i < W) {
      w = p.st[i];
      const uint32_t al = alive_fields(w);
      if (al) {
        int32_t d[16];
        load_word_dd(p.dd, i, d);
#pragma unroll
        for (int f = 0; f < 16; ++f)
          if (((al >> (2 * f)) & 1u) && d[f] <= level) die |= 1u << f;
      }
    }
    const uint32_t nd = static_cast<uint32_t>(__popc(die));
    const uint32_t id = wave_inclusive_sum(nd);
    if (lane_id() == kWave - 1) wd[wv] = id;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t td = 0;
      for (int q = 0; q < kWPB; ++q) {
        const uint32_t x = wd[q];
        wd[q] = td;
        td += x;
      }
      base = td ? atomic_add_u32(&in->ncur, td) : 0;
    }
    __syncthreads();
    if (die) {
      uint32_t od = base + wd[wv] + id - nd;
      for (int f = 0; f < 16; ++f) {
        if (!((die >> f) & 1u)) continue;
        const uint32_t t = static_cast<uint32_t>(16 * i + f);
        p.kv[t] = level;
        cur[od++] = t;
        w |= kCur << (2 * f);
      }
      p.st[i] = w;
    }
    __syncthreads();
  }
}

// Owner test and debit for one K4 of dying triangle t with partners q[0..2].
// A partner that is alive, or was pushed this round (dies next round), counts
// as surviving for ownership; an earlier death means the K4 is gone; a
// same-round death with a lower index owns it. A survivor falling from
// level+1 to level is pushed: its k is fixed now and it dies next round.
// kCur and kGone were set before this kernel started; a push racing with this
// read can only be seen as kAlive, which gives the same owner, and the extra
// debit of a pushed triangle cannot push it again (its support is already at
// most level).
__device__ __forceinline__ void destroy_k4(uint32_t t, const uint32_t q[3], const PeelBufs& p, int32_t level,
                                           uint32_t* __restrict__ nxt, uint32_t* __restrict__ ncount) {
  uint32_t sq[3];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    sq[s] = tri_state(p.st, q[s]);
    if (sq[s] == kGone) return;
    if (sq[s] == kCur && q[s] < t) return;
  }
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    if (sq[s] != kAlive) continue;
    const int32_t old = atomic_sub_i32(&p.dd[q[s]], 1);
    if (old == level + 1) {
      atomic_or_u32(&p.st[q[s] >> 4], kNext << ((q[s] & 15u) * 2u));
      p.kv[q[s]] = level;
      nxt[atomic_add_u32(ncount, 1u)] = q[s];
    }
  }
}

// End of a round: advance every state word (this round's dead become kGone,
// the pushed become the next frontier); when nothing was pushed, the min alive
// support for the next level goes into out->minval. Workgroup 0 writes the
// next round's control into `out`; no other workgroup reads `out`.
__global__ __launch_bounds__(kBlock) void k_peel_finish(PeelBufs p, PeelCtl* in, PeelCtl* out) {
  if (in->done) return;
  const bool need = in->nnext == 0;
  const uint64_t W = (p.T + 15) / 16;
  uint32_t mn = kNoMin;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < W;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t w = advance_states(p.st[i]);
    p.st[i] = w;
    if (need) {
      const uint32_t al = alive_fields(w);
      if (al) {
        int32_t d[16];
        load_word_dd(p.dd, i, d);
#pragma unroll
        for (int f = 0; f < 16; ++f)
          if ((al >> (2 * f)) & 1u) mn = min(mn, static_cast<uint32_t>(d[f] < 0 ? 0 : d[f]));
      }
    }
  }
  if (need) block_min_to(mn, &out->minval);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const uint64_t peeled = in->peeled + in->ncur;
    uint32_t error = 0, done = 0;
    if (in->need_level && (in->minval == kNoMin || in->ncur == 0)) {
      error = in->minval == kNoMin ? 1u : 2u;  // nothing alive / nothing collected
      done = 1;
    }
    if (peeled >= p.T) done = 1;
    out->ncur = in->nnext;
    out->nnext = 0;
    out->sel = in->sel ^ 1u;
    out->need_level = need ? 1u : 0u;
    out->level = round_level(in);
    out->round = in->round + 1;
    out->big = 0;
    out->error = error;
    out->peeled = peeled;
    out->done = done;
    in->minval = kNoMin;  // the next round's finish reduces into this slot
    if (done) in->done = 1;
  }
}

__global__ __launch_bounds__(kBlock) void k_peel_debit_lists(PeelBufs p, PeelCtl* in,
                                                             const uint64_t* __restrict__ koff,
                                                             const uint32_t* __restrict__ lists) {
  if (in->done) return;
  const int32_t level = round_level(in);
  const uint32_t* cur = in->sel ? p.buf1 : p.buf0;
  uint32_t* nxt = in->sel ? p.buf0 : p.buf1;
  uint32_t* ncount = &in->nnext;
  const uint32_t n = in->ncur;
  const unsigned w = threadIdx.x / kWave;
  const unsigned lane = lane_id();
  const uint32_t nw = gridDim.x * kWPB;
  for (uint32_t i = blockIdx.x * kWPB + w; i < n; i += nw) {
    const uint32_t t = cur[i];
    const uint64_t b = koff[t], e = koff[t + 1];
    for (uint64_t j = b + lane; j < e; j += kWave) {
      const uint32_t* q = lists + 3 * j;
      const uint32_t qq[3] = {__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1),
                              __builtin_nontemporal_load(q + 2)};
      destroy_k4(t, qq, p, level, nxt, ncount);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_peel_debit_od_small(PeelBufs p, PeelCtl* in, TriView v) {
  if (in->done) return;
  const int32_t level = round_level(in);
  const uint32_t* cur = in->sel ? p.buf1 : p.buf0;
  uint32_t* nxt = in->sel ? p.buf0 : p.buf1;
  uint32_t* ncount = &in->nnext;
  const uint32_t n = in->ncur;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t t = cur[i];
    const TriEdges e = tri_edges(v, t);
    const bool big = tri_len(v, e) > kSmallL;
    bool got;
    const uint32_t slot = wave_append(&in->big, big, &got);
    if (big) {
      p.bigq[slot] = t;
      continue;
    }
    k4_thread(v, e, [&](uint32_t pb, uint32_t pa, uint32_t pcc) {
      const uint32_t qq[3] = {pb, pa, pcc};
      destroy_k4(t, qq, p, level, nxt, ncount);
    });
  }
}

__global__ __launch_bounds__(kBlock) void k_peel_debit_od_big(PeelBufs p, PeelCtl* in, TriView v) {
  __shared__ uint32_t s1[kWPB][kWave], s2[kWPB][kWave];
  if (in->done) return;
  const int32_t level = round_level(in);
  uint32_t* nxt = in->sel ? p.buf0 : p.buf1;
  uint32_t* ncount = &in->nnext;
  const uint32_t n = in->big;
  const unsigned w = threadIdx.x / kWave;
  const uint32_t nw = gridDim.x * kWPB;
  for (uint32_t q = blockIdx.x * kWPB + w; q < n; q += nw) {
    const uint32_t t = p.bigq[q];
    const TriEdges e = tri_edges(v, t);
    k4_wave(v, e, s1[w], s2[w], [&](bool match, uint64_t i0, uint64_t i1, uint64_t i2) {
      if (match) {
        const uint32_t qq[3] = {v.tri[i0], v.tri[i1], v.tri[i2]};
        destroy_k4(t, qq, p, level, nxt, ncount);
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
  hipLaunchKernelGGL(k_peel_ctl_init, dim3(1), dim3(1), 0, s, p.ctl);
  hipLaunchKernelGGL(k_peel_init, dim3(grid_for((p.T + 15) / 16, kBlock, 4096)), dim3(kBlock), 0, s, p, sup);
  NUC34_LAUNCH_CHECK();
}

void launch_peel_round_direct(const PeelBufs& p, int slot, const uint64_t* koff, const uint32_t* lists,
                              hipStream_t s) {
  PeelCtl* in = p.ctl + (slot & 1);
  PeelCtl* out = p.ctl + ((slot + 1) & 1);
  hipLaunchKernelGGL(k_peel_collect, dim3(kScanGrid), dim3(kBlock), 0, s, p, in);
  hipLaunchKernelGGL(k_peel_debit_lists, dim3(kRoundGrid), dim3(kBlock), 0, s, p, in, koff, lists);
  hipLaunchKernelGGL(k_peel_finish, dim3(kScanGrid), dim3(kBlock), 0, s, p, in, out);
  NUC34_LAUNCH_CHECK();
}

void launch_peel_round_ondemand(const PeelBufs& p, int slot, const TriView& v, hipStream_t s) {
  PeelCtl* in = p.ctl + (slot & 1);
  PeelCtl* out = p.ctl + ((slot + 1) & 1);
  hipLaunchKernelGGL(k_peel_collect, dim3(kScanGrid), dim3(kBlock), 0, s, p, in);
  hipLaunchKernelGGL(k_peel_debit_od_small, dim3(kRoundGrid), dim3(kBlock), 0, s, p, in, v);
  hipLaunchKernelGGL(k_peel_debit_od_big, dim3(kRoundGrid), dim3(kBlock), 0, s, p, in, v);
  hipLaunchKernelGGL(k_peel_finish, dim3(kScanGrid), dim3(kBlock), 0, s, p, in, out);
  NUC34_LAUNCH_CHECK();
}

}  // namespace nuc34
