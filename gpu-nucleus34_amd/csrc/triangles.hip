// (b) Triangle enumeration by sorted-adjacency intersection.
//
// Reference: cu_cliques_tris + intersect_lists (nucleus/utils.cuh:259-343) and
// cliques_tri_par (nucleus/cliques.hpp:98-226): for each oriented edge b=(u,v),
// every w in N+(u) ∩ N+(v) gives the triangle {a = idx(u,w), b, c = idx(v,w)}.
//
// Here one wave takes a tile of G consecutive oriented edges (G = 16, 32 or 64,
// a launch-shape choice) and flattens the work of the tile (sum of d+(v) over
// its edges) across its 64 lanes, so lane utilisation does not depend on list
// lengths; smaller tiles stage shorter owner ranges. The tile's owner vertices are
// consecutive, so their out-lists are one contiguous range of E, staged in LDS
// when it fits; each candidate w = E[c] is then found in N+(u) by binary search
// in LDS (global memory otherwise).
//   bits pass: every triangle sets one bit in each of three per-edge rows (the
//              bit-packed edge->triangle incidence, bitrows.hpp): the b-role row
//              of b, the a-role row of a and the c-role row of c;
//   emit pass: a triangle's canonical index is the rank of v's position in the
//              a-role row of a (offA[a] plus the a-role bits below it), i.e.
//              exactly its place in the reference's (a, b)-sorted triangle list
//              (tri_t::operator<, utils.cuh:93-102), with no sort and no
//              atomics. The triangle is written at its peel id instead, the
//              rank of w's position in the b-role row of b (by b-edge, then
//              a-edge: the order the peel's decrements share cache lines in,
//              and the order this pass enumerates in, so the writes coalesce),
//              with its canonical index beside it; its peel id and its a-edge's
//              in-list position are also stored at its c-slot (cSlot), the rank
//              of the lowest vertex's in-list position in the c-role row of c.
//
// Degree-binned schedule (tri_schedule 1, or 2 from out-degree 9): the edges
// of an owner u with out-degree d <= 640 are enumerated a vertex at a time
// instead, by a wave (d <= 64), a 256-thread workgroup (d <= 256) or a
// 1024-thread workgroup (d <= 640). The group stages N+(u) and the candidate counts in LDS and
// flattens all of u's candidates over its lanes. Every triangle whose lowest
// vertex is u sets its b-role and a-role bits in rows of u's own out-edges,
// so the bits pass builds those rows in LDS (LDS atomics) and writes them out
// whole, leaving one global atomic per triangle (the c-role bit); the emit
// pass reads u's ranked a-role rows from LDS. Owners above 640 (and with
// tri_schedule 2 those up to 8, whose few candidates leave a wave idle) stay
// on the edge tiles.
#include "kernels.hpp"

namespace nuc34 {

namespace {

constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kListCap = 1536;  // staged owner-list entries per wave

template <typename P>
__device__ __forceinline__ uint32_t search_u32(P list, uint32_t lo, uint32_t hi, uint32_t key) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (list[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Emit pass: triangle (a, b, c) with w at position jw of N+(u) (b = (u, v))
// goes to its peel id, the rank of its bit in b's b-role row (word wq of the
// rows), with its canonical index p and its c-slot.
__device__ __forceinline__ void emit_tri(const BitRows& rows, const TriOut& out, uint32_t a, uint32_t b, uint32_t c,
                                         uint32_t jw, uint64_t wq, uint32_t p, const uint32_t* __restrict__ inpos) {
  const uint32_t q = rows.AoutR[wq] + static_cast<uint32_t>(__popc(rows.Aout[wq] & ((1u << (jw & 31u)) - 1u)));
  out.triA[q] = a;
  out.triBC[q] = (static_cast<uint64_t>(b) << 32) | c;
  out.canon[q] = p;
  out.cSlot[row_rank(rows.Cm + rows.rowC[c], inpos[b])] = uint2{q, inpos[a]};
}

template <bool EMIT>
__global__ __launch_bounds__(kBlock) void k_tri_enum(const uint32_t* __restrict__ O, const uint32_t* __restrict__ E,
                                                     const uint32_t* __restrict__ src, uint64_t e_begin,
                                                     uint64_t e_end, BitRows rows,
                                                     const uint32_t* __restrict__ inpos, TriOut out,
                                                     uint32_t G, uint32_t skip_lo, uint32_t skip_hi) {
  __shared__ uint32_t s_pre[kWavesPerBlock][kWave + 1];
  __shared__ uint32_t s_l0[kWavesPerBlock][kWave];
  __shared__ uint32_t s_len[kWavesPerBlock][kWave];
  __shared__ uint32_t s_vb[kWavesPerBlock][kWave];
  __shared__ uint32_t s_ob[kWavesPerBlock][kWave];
  __shared__ uint32_t s_list[kWavesPerBlock][kListCap];
  const unsigned w = threadIdx.x / kWave;
  const unsigned lane = lane_id();
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * kWavesPerBlock;
  for (uint64_t tile = static_cast<uint64_t>(blockIdx.x) * kWavesPerBlock + w; e_begin + tile * G < e_end;
       tile += nwaves) {
    const uint64_t e0 = e_begin + tile * G;
    const uint64_t e = e0 + lane;
    const bool valid = lane < G && e < e_end;
    uint32_t work = 0, ob = 0, oe = 0, vb = 0;
    if (valid) {
      const uint32_t u = src[e];
      const uint32_t v = E[e];
      vb = O[v];
      ob = O[u];
      oe = O[u + 1];
      // owners of out-degree in (skip_lo, skip_hi] are enumerated a vertex at
      // a time (k_tri_vertex)
      work = (oe - ob > skip_lo && oe - ob <= skip_hi) ? 0u : O[v + 1] - vb;
    }
    const uint32_t inc = wave_inclusive_sum(work);
    const uint32_t total = __shfl(inc, kWave - 1);
    const uint64_t last_lane = (e_end - e0 < static_cast<uint64_t>(G) ? e_end - e0 : G) - 1;
    const uint32_t lo = __shfl(ob, 0);
    const uint32_t hi = __shfl(oe, static_cast<int>(last_lane));
    const bool staged = (hi - lo) <= kListCap;
    if (staged)
      for (uint32_t i = lane; i < hi - lo; i += kWave) s_list[w][i] = E[lo + i];
    s_pre[w][lane] = inc - work;
    if (lane == kWave - 1) s_pre[w][kWave] = inc;
    s_l0[w][lane] = staged ? ob - lo : ob;
    s_len[w][lane] = oe - ob;
    s_vb[w][lane] = vb;
    s_ob[w][lane] = ob;
    wave_sync();
    for (uint32_t base = 0; base < total; base += kWave) {
      const uint32_t item = base + lane;
      if (item < total) {
        uint32_t slo = 0, shi = kWave;
        while (shi - slo > 1) {
          const uint32_t mid = (slo + shi) >> 1;
          if (s_pre[w][mid] <= item) slo = mid;
          else shi = mid;
        }
        const uint32_t i = slo;
        const uint32_t c = s_vb[w][i] + (item - s_pre[w][i]);
        const uint32_t x = E[c];
        const uint32_t l0 = s_l0[w][i];
        const uint32_t l1 = l0 + s_len[w][i];
        uint32_t pos;
        bool found;
        if (staged) {
          pos = search_u32(s_list[w], l0, l1, x);
          found = pos < l1 && s_list[w][pos] == x;
          pos += lo;
        } else {
          pos = search_u32(E, l0, l1, x);
          found = pos < l1 && E[pos] == x;
        }
        if (found) {
          const uint32_t a = pos;
          const uint32_t b = static_cast<uint32_t>(e0 + i);
          const uint32_t ob_u = s_ob[w][i];
          const uint32_t jv = b - ob_u, jw = a - ob_u;
          if constexpr (!EMIT) {
            // b-role: w seen from (u,v); a-role: v seen from (u,w); c-role: u's
            // position among the in-neighbours of v, seen from (v,w).
            atomicOr(&rows.Aout[rows.rowA[b] + (jw >> 5)], 1u << (jw & 31));
            atomicOr(&rows.Ain[rows.rowA[a] + (jv >> 5)].bits, 1u << (jv & 31));
            const uint32_t iu = inpos[b];
            atomicOr(&rows.Cm[rows.rowC[c] + (iu >> 5)].bits, 1u << (iu & 31));
          } else {
            const uint32_t p = row_rank(rows.Ain + rows.rowA[a], jv);
            emit_tri(rows, out, a, b, c, jw, rows.rowA[b] + (jw >> 5), p, inpos);
          }
        }
      }
    }
    wave_sync();
  }
}

// ---- a vertex at a time ------------------------------------------------------

template <int NT>
__device__ __forceinline__ void group_sync() {
  if constexpr (NT == kWave) wave_sync();
  else __syncthreads();
}

// Exclusive prefix of v over the NT threads of a group (thread order); *total
// gets the sum. tmp: NT / 64 words of LDS. Every thread of the group calls it.
template <int NT>
__device__ __forceinline__ uint32_t group_exclusive_scan(uint32_t v, uint32_t* tmp, uint32_t* total, unsigned tid) {
  const uint32_t inc = wave_inclusive_sum(v);
  if constexpr (NT == kWave) {
    *total = __shfl(inc, kWave - 1);
    return inc - v;
  } else {
    const unsigned wv = tid / kWave;
    if (lane_id() == kWave - 1) tmp[wv] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int i = 0; i < NT / kWave; ++i) {
      const uint32_t x = tmp[i];
      before += static_cast<unsigned>(i) < wv ? x : 0u;
      all += x;
    }
    __syncthreads();  // tmp is reused by the next call
    *total = all;
    return before + inc - v;
  }
}

// Per-group LDS of the vertex schedule for out-degrees up to D.
template <int NT, int D>
struct VertexLds {
  // bits pass: the b-role rows (R) and a-role rows (C) of u's out-edges,
  // d * W words each; emit pass: u's ranked a-role rows (d * W words of 8
  // bytes, so first and 8-byte aligned)
  alignas(8) uint32_t rows[2 * D * ((D + 31) / 32)];
  uint32_t list[D];      // N+(u)
  uint32_t vb[D];        // O[v] of each v in N+(u)
  uint32_t pre[D + 1];   // first candidate of each v
  uint32_t tmp[NT / kWave];
};

// Triangles with lowest vertex u (out-degree 1 <= d <= D), by the NT threads
// of a group (tid = thread index in the group).
template <bool EMIT, int NT, int D>
__device__ __forceinline__ void tri_vertex(uint32_t u, const uint32_t* __restrict__ O, const uint32_t* __restrict__ E,
                                           const BitRows& rows, const uint32_t* __restrict__ inpos, const TriOut& out,
                                           VertexLds<NT, D>& L, unsigned tid) {
  const uint32_t ob = O[u];
  const uint32_t d = O[u + 1] - ob;
  const uint32_t W = (d + 31) >> 5;
  const uint64_t r0 = rows.rowA[ob];  // rows of u's out-edges: d rows of W words from here
  uint32_t cnt = 0;
  if (tid < d) {
    const uint32_t v = E[ob + tid];
    const uint32_t vb = O[v];
    L.list[tid] = v;
    L.vb[tid] = vb;
    cnt = O[v + 1] - vb;
  }
  uint32_t total = 0;
  const uint32_t excl = group_exclusive_scan<NT>(cnt, L.tmp, &total, tid);
  if (tid < d) L.pre[tid] = excl;
  if (tid == 0) L.pre[d] = total;
  if constexpr (EMIT) {
    RankWord* a_rows = reinterpret_cast<RankWord*>(L.rows);
    for (uint32_t i = tid; i < d * W; i += NT) a_rows[i] = rows.Ain[r0 + i];
  } else {
    for (uint32_t i = tid; i < 2 * d * W; i += NT) L.rows[i] = 0u;
  }
  group_sync<NT>();
  uint32_t* const R = L.rows;          // bits pass: b-role rows
  uint32_t* const Cc = L.rows + d * W;  // bits pass: a-role rows
  for (uint32_t base = 0; base < total; base += NT) {
    const uint32_t item = base + tid;
    if (item >= total) continue;
    uint32_t lo = 0, hi = d;  // the largest j with pre[j] <= item
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (L.pre[mid] <= item) lo = mid;
      else hi = mid;
    }
    const uint32_t j = lo;
    const uint32_t c = L.vb[j] + (item - L.pre[j]);
    const uint32_t x = E[c];
    const uint32_t jw = search_u32(L.list, 0u, d, x);
    if (jw >= d || L.list[jw] != x) continue;
    const uint32_t b = ob + j, a = ob + jw;
    if constexpr (!EMIT) {
      atomicOr(&R[j * W + (jw >> 5)], 1u << (jw & 31));
      atomicOr(&Cc[jw * W + (j >> 5)], 1u << (j & 31));
      const uint32_t iu = inpos[b];
      atomicOr(&rows.Cm[rows.rowC[c] + (iu >> 5)].bits, 1u << (iu & 31));
    } else {
      const RankWord* a_rows = reinterpret_cast<const RankWord*>(L.rows);
      const uint32_t p = row_rank(a_rows + jw * W, j);
      emit_tri(rows, out, a, b, c, jw, r0 + static_cast<uint64_t>(j) * W + (jw >> 5), p, inpos);
    }
  }
  group_sync<NT>();
  if constexpr (!EMIT) {
    for (uint32_t i = tid; i < d * W; i += NT) {
      rows.Aout[r0 + i] = R[i];
      rows.Ain[r0 + i].bits = Cc[i];
    }
    group_sync<NT>();  // the LDS is reused by the next vertex
  }
}

// Vertex classes of the binned schedule (out-degree limits) and their groups.
constexpr uint32_t kVertexD0 = 64, kVertexD1 = 256, kVertexD2 = 640;

// Vertices of out-degree min_deg+1..640 into three lists by class (lists[k * n
// ..], counts[k]); the order within a list does not matter.
__global__ void k_vertex_bins(const uint32_t* __restrict__ O, uint32_t n, uint32_t min_deg,
                              uint32_t* __restrict__ lists, uint32_t* __restrict__ counts) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t b = static_cast<uint64_t>(blockIdx.x) * blockDim.x; b < n; b += stride) {
    const uint64_t u = b + threadIdx.x;
    uint32_t d = 0;
    if (u < n) d = O[u + 1] - O[u];
    const bool c0 = d > min_deg && d <= kVertexD0, c1 = d > kVertexD0 && d <= kVertexD1,
               c2 = d > kVertexD1 && d <= kVertexD2;
    bool got = false;
    const uint32_t s0 = wave_append(&counts[0], c0, &got);
    if (got) lists[s0] = static_cast<uint32_t>(u);
    const uint32_t s1 = wave_append(&counts[1], c1, &got);
    if (got) lists[static_cast<uint64_t>(n) + s1] = static_cast<uint32_t>(u);
    const uint32_t s2 = wave_append(&counts[2], c2, &got);
    if (got) lists[2 * static_cast<uint64_t>(n) + s2] = static_cast<uint32_t>(u);
  }
}

// Class 0: a wave per vertex (kBlock / kWave vertices per workgroup at a time).
template <bool EMIT>
__global__ __launch_bounds__(kBlock) void k_tri_vertex_wave(const uint32_t* __restrict__ O,
                                                            const uint32_t* __restrict__ E, BitRows rows,
                                                            const uint32_t* __restrict__ inpos, TriOut out,
                                                            const uint32_t* __restrict__ list,
                                                            const uint32_t* __restrict__ count, uint64_t e_begin,
                                                            uint64_t e_end) {
  __shared__ VertexLds<kWave, kVertexD0> lds[kWavesPerBlock];
  const unsigned w = threadIdx.x / kWave;
  const uint32_t nv = *count;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kWavesPerBlock + w; i < nv;
       i += static_cast<uint64_t>(gridDim.x) * kWavesPerBlock) {
    const uint32_t u = list[i];
    if (O[u] < e_begin || O[u] >= e_end) continue;  // another segment's vertex
    tri_vertex<EMIT, kWave, kVertexD0>(u, O, E, rows, inpos, out, lds[w], lane_id());
  }
}

// Classes 1 and 2: a workgroup of NT threads per vertex.
template <bool EMIT, int NT, int D>
__global__ __launch_bounds__(NT) void k_tri_vertex_group(const uint32_t* __restrict__ O,
                                                         const uint32_t* __restrict__ E, BitRows rows,
                                                         const uint32_t* __restrict__ inpos, TriOut out,
                                                         const uint32_t* __restrict__ list,
                                                         const uint32_t* __restrict__ count, uint64_t e_begin,
                                                         uint64_t e_end) {
  __shared__ VertexLds<NT, D> lds;
  const uint32_t nv = *count;
  for (uint64_t i = blockIdx.x; i < nv; i += gridDim.x) {
    const uint32_t u = list[i];
    if (O[u] < e_begin || O[u] >= e_end) continue;
    tri_vertex<EMIT, NT, D>(u, O, E, rows, inpos, out, lds, threadIdx.x);
  }
}

__global__ void k_tri_abc(const uint32_t* __restrict__ triA, const uint64_t* __restrict__ triBC, uint64_t count,
                          uint32_t* __restrict__ abc) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t bc = triBC[i];
    abc[3 * i] = triA[i];
    abc[3 * i + 1] = static_cast<uint32_t>(bc >> 32);
    abc[3 * i + 2] = static_cast<uint32_t>(bc);
  }
}

__global__ void k_canon_tris(const uint32_t* __restrict__ triA, const uint64_t* __restrict__ triBC,
                             const uint32_t* __restrict__ canon, uint64_t T, uint32_t* __restrict__ ctriA,
                             uint64_t* __restrict__ ctriBC) {
  for (uint64_t q = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < T;
       q += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t p = canon[q];
    ctriA[p] = triA[q];
    ctriBC[p] = triBC[q];
  }
}

__global__ void k_canon_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ canon, uint64_t T,
                            uint32_t* __restrict__ dst) {
  for (uint64_t q = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < T;
       q += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    dst[canon[q]] = src[q];
}

__global__ void k_tri_abc_canon(const uint32_t* __restrict__ triA, const uint64_t* __restrict__ triBC,
                                const uint32_t* __restrict__ canon, uint64_t T, uint32_t* __restrict__ abc) {
  for (uint64_t q = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < T;
       q += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t p = canon[q];
    const uint64_t bc = triBC[q];
    abc[3 * p] = triA[q];
    abc[3 * p + 1] = static_cast<uint32_t>(bc >> 32);
    abc[3 * p + 2] = static_cast<uint32_t>(bc);
  }
}

unsigned tri_grid(uint64_t edges, uint32_t G) {
  const uint64_t tiles = (edges + G - 1) / G;
  return grid_for(tiles, kWavesPerBlock, 1u << 15);
}

uint32_t tile_edges(uint32_t G) {
  if (G != 16 && G != 32 && G != 64) throw std::logic_error("triangle tile width must be 16, 32 or 64");
  return G;
}

}  // namespace

void launch_vertex_bins(const uint32_t* O, uint32_t n, uint32_t min_deg, uint32_t* lists, uint32_t* counts,
                        hipStream_t s) {
  if (!n) return;
  if (min_deg >= kVertexD0) throw std::logic_error("vertex classes start below out-degree 64");
  hipLaunchKernelGGL(k_vertex_bins, dim3(grid_for(n, 256, 1u << 14)), dim3(256), 0, s, O, n, min_deg, lists,
                     counts);
  NUC34_LAUNCH_CHECK();
}

namespace {

template <bool EMIT>
void launch_tri(const uint32_t* O, const uint32_t* E, const uint32_t* src, uint64_t e_begin, uint64_t e_end,
                const BitRows& rows, const uint32_t* inpos, const TriOut& out, uint32_t tile,
                const VertexBins* bins, hipStream_t s) {
  if (e_end <= e_begin) return;
  const uint32_t G = tile_edges(tile);
  // with bins, the tiles keep only the owners outside the vertex classes
  const uint32_t skip_lo = bins ? bins->min_deg : 0u, skip_hi = bins ? kVertexD2 : 0u;
  hipLaunchKernelGGL(k_tri_enum<EMIT>, dim3(tri_grid(e_end - e_begin, G)), dim3(kBlock), 0, s, O, E, src, e_begin,
                     e_end, rows, inpos, out, G, skip_lo, skip_hi);
  if (bins) {
    const uint64_t n = bins->n;
    hipLaunchKernelGGL(k_tri_vertex_wave<EMIT>, dim3(grid_for(n, kWavesPerBlock, 8192)), dim3(kBlock), 0, s, O, E,
                       rows, inpos, out, bins->lists, bins->counts, e_begin, e_end);
    hipLaunchKernelGGL((k_tri_vertex_group<EMIT, 256, kVertexD1>), dim3(grid_for(n, 1, 4096)), dim3(256), 0, s, O, E,
                       rows, inpos, out, bins->lists + n, bins->counts + 1, e_begin, e_end);
    hipLaunchKernelGGL((k_tri_vertex_group<EMIT, 1024, kVertexD2>), dim3(grid_for(n, 1, 512)), dim3(1024), 0, s, O,
                       E, rows, inpos, out, bins->lists + 2 * n, bins->counts + 2, e_begin, e_end);
  }
  NUC34_LAUNCH_CHECK();
}

}  // namespace

void launch_tri_bits(const uint32_t* O, const uint32_t* E, const uint32_t* src, uint64_t e_begin, uint64_t e_end,
                     const BitRows& rows, const uint32_t* inpos, uint32_t tile, const VertexBins* bins,
                     hipStream_t s) {
  launch_tri<false>(O, E, src, e_begin, e_end, rows, inpos, TriOut{}, tile, bins, s);
}

void launch_tri_emit(const uint32_t* O, const uint32_t* E, const uint32_t* src, uint64_t e_begin, uint64_t e_end,
                     const BitRows& rows, const uint32_t* inpos, const TriOut& out, uint32_t tile,
                     const VertexBins* bins, hipStream_t s) {
  launch_tri<true>(O, E, src, e_begin, e_end, rows, inpos, out, tile, bins, s);
}

void launch_canon_tris(const uint32_t* triA, const uint64_t* triBC, const uint32_t* canon, uint64_t T,
                       uint32_t* ctriA, uint64_t* ctriBC, hipStream_t s) {
  if (!T) return;
  hipLaunchKernelGGL(k_canon_tris, dim3(grid_for(T, 256, 1u << 16)), dim3(256), 0, s, triA, triBC, canon, T, ctriA,
                     ctriBC);
  NUC34_LAUNCH_CHECK();
}

void launch_canon_u32(const uint32_t* src, const uint32_t* canon, uint64_t T, uint32_t* dst, hipStream_t s) {
  if (!T) return;
  hipLaunchKernelGGL(k_canon_u32, dim3(grid_for(T, 256, 1u << 16)), dim3(256), 0, s, src, canon, T, dst);
  NUC34_LAUNCH_CHECK();
}

void launch_tri_abc_canon(const uint32_t* triA, const uint64_t* triBC, const uint32_t* canon, uint64_t T,
                          uint32_t* abc, hipStream_t s) {
  if (!T) return;
  hipLaunchKernelGGL(k_tri_abc_canon, dim3(grid_for(T, 256, 1u << 16)), dim3(256), 0, s, triA, triBC, canon, T, abc);
  NUC34_LAUNCH_CHECK();
}

void launch_tri_abc(const uint32_t* triA, const uint64_t* triBC, uint64_t count, uint32_t* abc, hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(k_tri_abc, dim3(grid_for(count, 256, 1u << 15)), dim3(256), 0, s, triA, triBC, count, abc);
  NUC34_LAUNCH_CHECK();
}

}  // namespace nuc34
