// Builders for the bit-packed incidence: the in-CSR of the oriented graph (the
// c-role rows are indexed by in-list position) and the per-edge row layout.
#include "bitrows.hpp"
#include "kernels.hpp"
#include "scan.hpp"
#include "segsort.hpp"

namespace nuc34 {

namespace {

__global__ void k_indeg(const uint32_t* __restrict__ E, uint64_t m, uint32_t* __restrict__ indeg) {
  for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
       e += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    atomic_add_u32(&indeg[E[e]], 1u);
}

__global__ void k_in_fill(const uint32_t* __restrict__ E, const uint32_t* __restrict__ src, uint64_t m,
                          const uint32_t* __restrict__ IO, uint32_t* __restrict__ cur, uint64_t* __restrict__ keys) {
  for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
       e += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t v = E[e];
    keys[IO[v] + atomic_add_u32(&cur[v], 1u)] = (static_cast<uint64_t>(src[e]) << 32) | e;
  }
}

__global__ void k_in_split(const uint64_t* __restrict__ keys, uint64_t m, const uint32_t* __restrict__ E,
                           const uint32_t* __restrict__ IO, uint32_t* __restrict__ IE, uint32_t* __restrict__ IEe,
                           uint32_t* __restrict__ inpos) {
  for (uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < m;
       p += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t k = keys[p];
    const uint32_t e = static_cast<uint32_t>(k);
    IE[p] = static_cast<uint32_t>(k >> 32);
    IEe[p] = e;
    inpos[e] = static_cast<uint32_t>(p - IO[E[e]]);
  }
}

__global__ void k_row_words(const uint32_t* __restrict__ O, const uint32_t* __restrict__ IO,
                            const uint32_t* __restrict__ src, uint64_t m, uint32_t* __restrict__ wa,
                            uint32_t* __restrict__ wc) {
  for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
       e += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t u = src[e];
    wa[e] = (O[u + 1] - O[u] + 31) >> 5;
    wc[e] = (IO[u + 1] - IO[u] + 31) >> 5;
  }
}

__global__ void k_edge_triangles(BitRows r, uint64_t m, uint32_t* __restrict__ cnt) {
  for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
       e += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t c = 0;
    for (uint64_t i = r.rowA[e]; i < r.rowA[e + 1]; ++i)
      c += static_cast<uint32_t>(__popc(r.Aout[i]) + __popc(r.Ain[i].bits));
    for (uint64_t i = r.rowC[e]; i < r.rowC[e + 1]; ++i) c += static_cast<uint32_t>(__popc(r.Cm[i].bits));
    cnt[e] = c;
  }
}

// rank[i] = P[i]: the exclusive scan of the words' set bits over the whole
// array (rows are laid out edge after edge, so the scan at a row's first word
// is that edge's first triangle id, or first c-slot).
__global__ void k_set_ranks(RankWord* __restrict__ w, const uint64_t* __restrict__ P, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    w[i].rank = static_cast<uint32_t>(P[i]);
}

__global__ void k_set_ranks32(uint32_t* __restrict__ r, const uint64_t* __restrict__ P, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    r[i] = static_cast<uint32_t>(P[i]);
}

}  // namespace

void build_in_csr(const uint32_t* O, const uint32_t* E, const uint32_t* src, uint32_t n, uint64_t m, uint32_t* IO,
                  uint32_t* IE, uint32_t* IEe, uint32_t* inpos, Allocator& alloc, hipStream_t s) {
  (void)O;
  DevBuf indeg = alloc.get((static_cast<uint64_t>(n) + 1) * 4, "in-degrees");
  NUC34_HIP(hipMemsetAsync(indeg.as<void>(), 0, indeg.bytes(), s));
  if (m) {
    hipLaunchKernelGGL(k_indeg, dim3(grid_for(m, 256, 1u << 16)), dim3(256), 0, s, E, m, indeg.as<uint32_t>());
    NUC34_LAUNCH_CHECK();
  }
  {
    DevBuf scratch = alloc.get(scan_scratch_elems(n) * 4, "scan scratch");
    exclusive_scan<uint32_t, uint32_t>(indeg.as<uint32_t>(), IO, n, scratch.as<uint32_t>(), s);
  }
  if (!m) return;
  DevBuf keys = alloc.get(m * 8, "in-list keys");
  NUC34_HIP(hipMemsetAsync(indeg.as<void>(), 0, indeg.bytes(), s));
  hipLaunchKernelGGL(k_in_fill, dim3(grid_for(m, 256, 1u << 16)), dim3(256), 0, s, E, src, m, IO,
                     indeg.as<uint32_t>(), keys.as<uint64_t>());
  NUC34_LAUNCH_CHECK();
  {
    // segment offsets as u64 for the segmented sort; k_in_fill's per-vertex
    // cursors have counted every in-edge again, so they are the in-degrees
    DevBuf off64 = alloc.get((static_cast<uint64_t>(n) + 1) * 8, "in-list offsets");
    DevBuf scratch = alloc.get(scan_scratch_elems(n) * 8, "scan scratch");
    exclusive_scan<uint32_t, uint64_t>(indeg.as<uint32_t>(), off64.as<uint64_t>(), n, scratch.as<uint64_t>(), s);
    segmented_sort_u64(keys.as<uint64_t>(), off64.as<uint64_t>(), n, alloc, s);
  }
  hipLaunchKernelGGL(k_in_split, dim3(grid_for(m, 256, 1u << 16)), dim3(256), 0, s, keys.as<uint64_t>(), m, E, IO,
                     IE, IEe, inpos);
  NUC34_LAUNCH_CHECK();
}

void launch_row_words(const uint32_t* O, const uint32_t* IO, const uint32_t* src, uint64_t m, uint32_t* wa,
                      uint32_t* wc, hipStream_t s) {
  if (!m) return;
  hipLaunchKernelGGL(k_row_words, dim3(grid_for(m, 256, 1u << 16)), dim3(256), 0, s, O, IO, src, m, wa, wc);
  NUC34_LAUNCH_CHECK();
}

void launch_edge_triangles(const BitRows& r, uint64_t m, uint32_t* cnt, hipStream_t s) {
  if (!m) return;
  hipLaunchKernelGGL(k_edge_triangles, dim3(grid_for(m, 256, 1u << 16)), dim3(256), 0, s, r, m, cnt);
  NUC34_LAUNCH_CHECK();
}

void launch_set_ranks32(uint32_t* r, const uint64_t* P, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_set_ranks32, dim3(grid_for(n, 256, 1u << 16)), dim3(256), 0, s, r, P, n);
  NUC34_LAUNCH_CHECK();
}

void launch_set_ranks(RankWord* w, const uint64_t* P, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_set_ranks, dim3(grid_for(n, 256, 1u << 16)), dim3(256), 0, s, w, P, n);
  NUC34_LAUNCH_CHECK();
}

}  // namespace nuc34
