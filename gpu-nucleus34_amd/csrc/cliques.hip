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

// ---- K4 list (direct) -------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_k4_emit_small(TriView v, const uint32_t* __restrict__ own,
                                                          const uint64_t* __restrict__ k4off,
                                                          uint32_t* __restrict__ qud) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < v.T; t += stride) {
    if (!own[t]) continue;
    const TriEdges e = tri_edges(v, t);
    if (tri_len(v, e) > kSmallL) continue;
    uint64_t pos = k4off[t];
    k4_thread(v, e, [&](uint32_t pb, uint32_t pa, uint32_t pc) {
      if (!owns_k4(v, pc, e.c)) return;
      uint32_t* q = qud + 4 * pos;
      q[0] = pb;
      q[1] = static_cast<uint32_t>(t);
      q[2] = pa;
      q[3] = pc;
      ++pos;
    });
  }
}

__global__ __launch_bounds__(kBlock) void k_k4_emit_big(TriView v, const uint32_t* __restrict__ bigq,
                                                        const uint32_t* __restrict__ bigcount,
                                                        const uint32_t* __restrict__ own,
                                                        const uint64_t* __restrict__ k4off,
                                                        uint32_t* __restrict__ qud) {
  __shared__ uint32_t s1[kWPB][kWave], s2[kWPB][kWave];
  const unsigned w = threadIdx.x / kWave;
  const uint32_t n = *bigcount;
  const uint32_t nw = gridDim.x * kWPB;
  for (uint32_t q = blockIdx.x * kWPB + w; q < n; q += nw) {
    const uint32_t t = bigq[q];
    if (!own[t]) continue;
    const TriEdges e = tri_edges(v, t);
    uint64_t pos = k4off[t];
    k4_wave(v, e, s1[w], s2[w], [&](bool match, uint64_t i0, uint64_t i1, uint64_t i2) {
      uint32_t pc = 0;
      bool o = false;
      if (match) {
        pc = v.tri[i2];
        o = owns_k4(v, pc, e.c);
      }
      const uint64_t mask = __ballot(o);
      if (o) {
        uint32_t* out = qud + 4 * (pos + lanes_below(mask));
        out[0] = v.tri[i0];
        out[1] = t;
        out[2] = v.tri[i1];
        out[3] = pc;
      }
      pos += static_cast<uint64_t>(__popcll(mask));
    });
  }
}

__global__ void k_k4_incidence(const uint32_t* __restrict__ qud, uint64_t Q, const uint64_t* __restrict__ koff,
                               uint32_t* __restrict__ cur, uint32_t* __restrict__ incK) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t q = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < Q; q += stride) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const uint32_t t = qud[4 * q + s];
      incK[koff[t] + atomic_add_u32(&cur[t], 1u)] = static_cast<uint32_t>(q);
    }
  }
}

// ---- peel -------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_alive_min(const uint32_t* __restrict__ alive, uint32_t nalive,
                                                      const int32_t* __restrict__ kv, const int32_t* __restrict__ dd,
                                                      PeelCounters* __restrict__ pc) {
  uint32_t mn = kNone;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nalive; i += gridDim.x * blockDim.x) {
    const uint32_t t = alive[i];
    if (kv[t] == -1) {
      const int32_t d = dd[t];
      mn = min(mn, static_cast<uint32_t>(d < 0 ? 0 : d));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mn = min(mn, static_cast<uint32_t>(__shfl_xor(mn, off)));
  if (lane_id() == 0 && mn != kNone) atomic_min_u32(&pc->minval, mn);
}

__global__ __launch_bounds__(kBlock) void k_collect(const uint32_t* __restrict__ alive, uint32_t nalive,
                                                    int32_t* __restrict__ kv, const int32_t* __restrict__ dd,
                                                    uint32_t* __restrict__ rnd, int32_t level, uint32_t round,
                                                    uint32_t* __restrict__ front, PeelCounters* __restrict__ pc) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nalive; i += gridDim.x * blockDim.x) {
    const uint32_t t = alive[i];
    const bool die = kv[t] == -1 && dd[t] <= level;
    bool got;
    const uint32_t slot = wave_append(&pc->front, die, &got);
    if (die) {
      kv[t] = level;
      rnd[t] = round;
      front[slot] = t;
    }
  }
}

__global__ void k_mark(const uint32_t* __restrict__ front, const PeelCounters* __restrict__ pc,
                       int32_t* __restrict__ kv, uint32_t* __restrict__ rnd, int32_t level, uint32_t round) {
  const uint32_t n = pc->front;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t t = front[i];
    kv[t] = level;
    rnd[t] = round;
  }
}

// Owner test and debit for one K4 of dying triangle t with partners p[0..2].
__device__ __forceinline__ void destroy_k4(uint32_t t, const uint32_t p[3], const int32_t* __restrict__ kv,
                                           const uint32_t* __restrict__ rnd, int32_t* __restrict__ dd,
                                           int32_t level, uint32_t round, uint32_t* __restrict__ next,
                                           PeelCounters* __restrict__ pc) {
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    if (kv[p[s]] == -1) continue;     // still alive
    if (rnd[p[s]] < round) return;    // destroyed in an earlier round
    if (p[s] < t) return;             // same round: the lowest index owns it
  }
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    if (kv[p[s]] != -1) continue;
    const int32_t old = atomic_sub_i32(&dd[p[s]], 1);
    if (old == level + 1) next[atomic_add_u32(&pc->next, 1u)] = p[s];
  }
}

__global__ __launch_bounds__(kBlock) void k_debit_od_small(TriView v, const uint32_t* __restrict__ front,
                                                           PeelCounters* __restrict__ pc,
                                                           const int32_t* __restrict__ kv,
                                                           const uint32_t* __restrict__ rnd, int32_t* __restrict__ dd,
                                                           int32_t level, uint32_t round, uint32_t* __restrict__ next,
                                                           uint32_t* __restrict__ bigq) {
  const uint32_t n = pc->front;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t t = front[i];
    const TriEdges e = tri_edges(v, t);
    const bool big = tri_len(v, e) > kSmallL;
    bool got;
    const uint32_t slot = wave_append(&pc->big, big, &got);
    if (big) {
      bigq[slot] = t;
      continue;
    }
    k4_thread(v, e, [&](uint32_t pb, uint32_t pa, uint32_t pcc) {
      const uint32_t p[3] = {pb, pa, pcc};
      destroy_k4(t, p, kv, rnd, dd, level, round, next, pc);
    });
  }
}

__global__ __launch_bounds__(kBlock) void k_debit_od_big(TriView v, const uint32_t* __restrict__ bigq,
                                                         PeelCounters* __restrict__ pc,
                                                         const int32_t* __restrict__ kv,
                                                         const uint32_t* __restrict__ rnd, int32_t* __restrict__ dd,
                                                         int32_t level, uint32_t round,
                                                         uint32_t* __restrict__ next) {
  __shared__ uint32_t s1[kWPB][kWave], s2[kWPB][kWave];
  const unsigned w = threadIdx.x / kWave;
  const uint32_t n = pc->big;
  const uint32_t nw = gridDim.x * kWPB;
  for (uint32_t q = blockIdx.x * kWPB + w; q < n; q += nw) {
    const uint32_t t = bigq[q];
    const TriEdges e = tri_edges(v, t);
    k4_wave(v, e, s1[w], s2[w], [&](bool match, uint64_t i0, uint64_t i1, uint64_t i2) {
      if (match) {
        const uint32_t p[3] = {v.tri[i0], v.tri[i1], v.tri[i2]};
        destroy_k4(t, p, kv, rnd, dd, level, round, next, pc);
      }
    });
  }
}

__global__ __launch_bounds__(kBlock) void k_debit_direct(const uint64_t* __restrict__ koff,
                                                         const uint32_t* __restrict__ incK,
                                                         const uint32_t* __restrict__ qud,
                                                         const uint32_t* __restrict__ front,
                                                         PeelCounters* __restrict__ pc,
                                                         const int32_t* __restrict__ kv,
                                                         const uint32_t* __restrict__ rnd, int32_t* __restrict__ dd,
                                                         int32_t level, uint32_t round,
                                                         uint32_t* __restrict__ next) {
  const unsigned w = threadIdx.x / kWave;
  const unsigned lane = lane_id();
  const uint32_t n = pc->front;
  const uint32_t nw = gridDim.x * kWPB;
  for (uint32_t i = blockIdx.x * kWPB + w; i < n; i += nw) {
    const uint32_t t = front[i];
    const uint64_t b = koff[t], e = koff[t + 1];
    for (uint64_t j = b + lane; j < e; j += kWave) {
      const uint32_t q = incK[j];
      const uint4 m = *reinterpret_cast<const uint4*>(qud + 4ull * q);
      uint32_t p[3];
      int k = 0;
      if (m.x != t) p[k++] = m.x;
      if (m.y != t) p[k++] = m.y;
      if (m.z != t) p[k++] = m.z;
      if (m.w != t && k < 3) p[k++] = m.w;
      destroy_k4(t, p, kv, rnd, dd, level, round, next, pc);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_compact(const uint32_t* __restrict__ alive, uint32_t nalive,
                                                    const int32_t* __restrict__ kv, uint32_t* __restrict__ out,
                                                    PeelCounters* __restrict__ pc) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nalive; i += gridDim.x * blockDim.x) {
    const uint32_t t = alive[i];
    const bool keep = kv[t] == -1;
    bool got;
    const uint32_t slot = wave_append(&pc->alive, keep, &got);
    if (keep) out[slot] = t;
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

void launch_k4_emit(const TriView& v, const uint32_t* owned, const uint64_t* k4off, uint32_t* qud,
                    const uint32_t* bigq, const uint32_t* bigcount, hipStream_t s) {
  if (!v.T) return;
  hipLaunchKernelGGL(k_k4_emit_small, dim3(grid_for(v.T, kBlock, 1u << 14)), dim3(kBlock), 0, s, v, owned, k4off,
                     qud);
  NUC34_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_k4_emit_big, dim3(8192), dim3(kBlock), 0, s, v, bigq, bigcount, owned, k4off, qud);
  NUC34_LAUNCH_CHECK();
}

void launch_k4_incidence(const uint32_t* qud, uint64_t Q, const uint64_t* koff, uint32_t* cur, uint32_t* incK,
                         hipStream_t s) {
  if (!Q) return;
  hipLaunchKernelGGL(k_k4_incidence, dim3(grid_for(Q, 256, 1u << 16)), dim3(256), 0, s, qud, Q, koff, cur, incK);
  NUC34_LAUNCH_CHECK();
}

void launch_alive_min(const uint32_t* alive, uint32_t nalive, const int32_t* kv, const int32_t* dd,
                      PeelCounters* pc, hipStream_t s) {
  if (!nalive) return;
  hipLaunchKernelGGL(k_alive_min, dim3(grid_for(nalive, kBlock, 4096)), dim3(kBlock), 0, s, alive, nalive, kv, dd,
                     pc);
  NUC34_LAUNCH_CHECK();
}

void launch_collect(const uint32_t* alive, uint32_t nalive, int32_t* kv, const int32_t* dd, uint32_t* rnd,
                    int32_t level, uint32_t round, uint32_t* front, PeelCounters* pc, hipStream_t s) {
  if (!nalive) return;
  hipLaunchKernelGGL(k_collect, dim3(grid_for(nalive, kBlock, 4096)), dim3(kBlock), 0, s, alive, nalive, kv, dd,
                     rnd, level, round, front, pc);
  NUC34_LAUNCH_CHECK();
}

void launch_mark(const uint32_t* front, const PeelCounters* pc, int32_t* kv, uint32_t* rnd, int32_t level,
                 uint32_t round, hipStream_t s) {
  hipLaunchKernelGGL(k_mark, dim3(kRoundGrid), dim3(kBlock), 0, s, front, pc, kv, rnd, level, round);
  NUC34_LAUNCH_CHECK();
}

void launch_debit_ondemand(const TriView& v, const uint32_t* front, PeelCounters* pc, const int32_t* kv,
                           const uint32_t* rnd, int32_t* dd, int32_t level, uint32_t round, uint32_t* next,
                           uint32_t* bigq, hipStream_t s) {
  hipLaunchKernelGGL(k_debit_od_small, dim3(kRoundGrid), dim3(kBlock), 0, s, v, front, pc, kv, rnd, dd, level,
                     round, next, bigq);
  NUC34_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_debit_od_big, dim3(kRoundGrid), dim3(kBlock), 0, s, v, bigq, pc, kv, rnd, dd, level, round,
                     next);
  NUC34_LAUNCH_CHECK();
}

void launch_debit_direct(const uint64_t* koff, const uint32_t* incK, const uint32_t* qud, const uint32_t* front,
                         PeelCounters* pc, const int32_t* kv, const uint32_t* rnd, int32_t* dd, int32_t level,
                         uint32_t round, uint32_t* next, hipStream_t s) {
  hipLaunchKernelGGL(k_debit_direct, dim3(kRoundGrid), dim3(kBlock), 0, s, koff, incK, qud, front, pc, kv, rnd, dd,
                     level, round, next);
  NUC34_LAUNCH_CHECK();
}

void launch_compact_alive(const uint32_t* alive, uint32_t nalive, const int32_t* kv, uint32_t* out,
                          PeelCounters* pc, hipStream_t s) {
  if (!nalive) return;
  hipLaunchKernelGGL(k_compact, dim3(grid_for(nalive, kBlock, 4096)), dim3(kBlock), 0, s, alive, nalive, kv, out,
                     pc);
  NUC34_LAUNCH_CHECK();
}

}  // namespace nuc34
