// (d) Edge -> triangle incidence keyed by third vertex.
//
// Reference: build_incidence_entries + sort_by_key + split_keys + lower_bound
// (nucleus/cuNucleus34_ondemand.cu:92-128, 502-555). Each triangle t with
// vertices u<v<w (rank) and edges a=(u,w), b=(u,v), c=(v,w) is listed under b
// with third vertex w, under a with v and under c with u; each edge's list is
// sorted by third vertex. Offsets come from the per-edge role counts of the
// triangle count pass (no global sort): entries are scattered into their edge
// segment and each segment is sorted on its own.
#include "kernels.hpp"

namespace nuc34 {

namespace {

__global__ void k_inc_len(const uint32_t* __restrict__ cntA, const uint32_t* __restrict__ cntB,
                          const uint32_t* __restrict__ cntC, uint64_t m, uint32_t* __restrict__ len) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    len[i] = cntA[i] + cntB[i] + cntC[i];
}

__global__ void k_inc_fill(const uint32_t* __restrict__ triA, const uint64_t* __restrict__ triBC, uint64_t T,
                           const uint32_t* __restrict__ E, const uint32_t* __restrict__ src,
                           const uint64_t* __restrict__ incO, uint32_t* __restrict__ cur,
                           uint64_t* __restrict__ inc) {
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < T;
       t += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t a = triA[t];
    const uint64_t bc = triBC[t];
    const uint32_t b = static_cast<uint32_t>(bc >> 32);
    const uint32_t c = static_cast<uint32_t>(bc);
    const uint32_t u = src[b], v = E[b], w = E[a];
    const uint64_t id = t;
    inc[incO[b] + atomic_add_u32(&cur[b], 1u)] = (static_cast<uint64_t>(w) << 32) | id;
    inc[incO[a] + atomic_add_u32(&cur[a], 1u)] = (static_cast<uint64_t>(v) << 32) | id;
    inc[incO[c] + atomic_add_u32(&cur[c], 1u)] = (static_cast<uint64_t>(u) << 32) | id;
  }
}

__global__ void k_inc_split(const uint64_t* __restrict__ inc, uint64_t n, uint32_t* __restrict__ third,
                            uint32_t* __restrict__ tri) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t x = inc[i];
    third[i] = static_cast<uint32_t>(x >> 32);
    tri[i] = static_cast<uint32_t>(x);
  }
}

}  // namespace

void launch_inc_len(const uint32_t* cntA, const uint32_t* cntB, const uint32_t* cntC, uint64_t m, uint32_t* len,
                    hipStream_t s) {
  if (!m) return;
  hipLaunchKernelGGL(k_inc_len, dim3(grid_for(m, 256, 1u << 16)), dim3(256), 0, s, cntA, cntB, cntC, m, len);
  NUC34_LAUNCH_CHECK();
}

void launch_inc_fill(const uint32_t* triA, const uint64_t* triBC, uint64_t T, const uint32_t* E,
                     const uint32_t* src, const uint64_t* incO, uint32_t* cur, uint64_t* inc, hipStream_t s) {
  if (!T) return;
  hipLaunchKernelGGL(k_inc_fill, dim3(grid_for(T, 256, 1u << 16)), dim3(256), 0, s, triA, triBC, T, E, src, incO,
                     cur, inc);
  NUC34_LAUNCH_CHECK();
}

void launch_inc_split(const uint64_t* inc, uint64_t n, uint32_t* third, uint32_t* tri, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_inc_split, dim3(grid_for(n, 256, 1u << 16)), dim3(256), 0, s, inc, n, third, tri);
  NUC34_LAUNCH_CHECK();
}

}  // namespace nuc34
