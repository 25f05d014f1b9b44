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
//              atomics; its id and its a-edge's in-list position are also
//              stored at its c-slot (cSlot), the rank of the lowest vertex's
//              in-list position in the c-role row of c.
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

template <bool EMIT>
__global__ __launch_bounds__(kBlock) void k_tri_enum(const uint32_t* __restrict__ O, const uint32_t* __restrict__ E,
                                                     const uint32_t* __restrict__ src, uint64_t e_begin,
                                                     uint64_t e_end, BitRows rows,
                                                     const uint32_t* __restrict__ inpos,
                                                     uint32_t* __restrict__ triA, uint64_t* __restrict__ triBC,
                                                     uint2* __restrict__ cSlot,
                                                     uint32_t G) {
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
      work = O[v + 1] - vb;
      ob = O[u];
      oe = O[u + 1];
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
            triA[p] = a;
            triBC[p] = (static_cast<uint64_t>(b) << 32) | c;
            cSlot[row_rank(rows.Cm + rows.rowC[c], inpos[b])] = uint2{p, inpos[a]};
          }
        }
      }
    }
    wave_sync();
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

unsigned tri_grid(uint64_t edges, uint32_t G) {
  const uint64_t tiles = (edges + G - 1) / G;
  return grid_for(tiles, kWavesPerBlock, 1u << 15);
}

uint32_t tile_edges(uint32_t G) {
  if (G != 16 && G != 32 && G != 64) throw std::logic_error("triangle tile width must be 16, 32 or 64");
  return G;
}

}  // namespace

void launch_tri_bits(const uint32_t* O, const uint32_t* E, const uint32_t* src, uint64_t e_begin, uint64_t e_end,
                     const BitRows& rows, const uint32_t* inpos, uint32_t tile, hipStream_t s) {
  if (e_end <= e_begin) return;
  const uint32_t G = tile_edges(tile);
  hipLaunchKernelGGL(k_tri_enum<false>, dim3(tri_grid(e_end - e_begin, G)), dim3(kBlock), 0, s, O, E, src, e_begin,
                     e_end, rows, inpos, nullptr, nullptr, nullptr, G);
  NUC34_LAUNCH_CHECK();
}

void launch_tri_emit(const uint32_t* O, const uint32_t* E, const uint32_t* src, uint64_t e_begin, uint64_t e_end,
                     const BitRows& rows, const uint32_t* inpos, uint32_t* triA, uint64_t* triBC, uint2* cSlot,
                     uint32_t tile, hipStream_t s) {
  if (e_end <= e_begin) return;
  const uint32_t G = tile_edges(tile);
  hipLaunchKernelGGL(k_tri_enum<true>, dim3(tri_grid(e_end - e_begin, G)), dim3(kBlock), 0, s, O, E, src, e_begin,
                     e_end, rows, inpos, triA, triBC, cSlot, G);
  NUC34_LAUNCH_CHECK();
}

void launch_tri_abc(const uint32_t* triA, const uint64_t* triBC, uint64_t count, uint32_t* abc, hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(k_tri_abc, dim3(grid_for(count, 256, 1u << 15)), dim3(256), 0, s, triA, triBC, count, abc);
  NUC34_LAUNCH_CHECK();
}

}  // namespace nuc34
