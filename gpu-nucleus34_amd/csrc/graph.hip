// (a) Degree orientation on the device: keep u->v iff (deg u, u) < (deg v, v)
// with the undirected degree, list order preserved, so edge indices equal those
// of the reference's host make_oriented (nucleus/graph/graph.hpp:176-218,
// vertex_less at nucleus/graph/graph_api.hpp:124-127).
// One wave per vertex; keep flags are ballots, positions come from mbcnt.
#include "kernels.hpp"

namespace nuc34 {

namespace {

__device__ __forceinline__ bool keep_edge(uint32_t u, uint32_t du, uint32_t v, uint32_t dv) {
  return du < dv || (du == dv && u < v);
}

__global__ __launch_bounds__(kBlock) void k_orient_count(const uint32_t* __restrict__ Ou,
                                                         const uint32_t* __restrict__ Eu, uint32_t n,
                                                         uint32_t* __restrict__ outdeg) {
  const unsigned lane = lane_id();
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kBlock / kWave);
  for (uint64_t u = static_cast<uint64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; u < n;
       u += waves) {
    const uint32_t b = Ou[u], e = Ou[u + 1], du = e - b;
    uint32_t cnt = 0;
    for (uint32_t base = 0; base < du; base += kWave) {
      const uint32_t i = base + lane;
      bool keep = false;
      if (i < du) {
        const uint32_t v = Eu[b + i];
        keep = keep_edge(static_cast<uint32_t>(u), du, v, Ou[v + 1] - Ou[v]);
      }
      cnt += static_cast<uint32_t>(__popcll(__ballot(keep)));
    }
    if (lane == 0) outdeg[u] = cnt;
  }
}

__global__ __launch_bounds__(kBlock) void k_orient_fill(const uint32_t* __restrict__ Ou,
                                                        const uint32_t* __restrict__ Eu, uint32_t n,
                                                        const uint32_t* __restrict__ O, uint32_t* __restrict__ E,
                                                        uint32_t* __restrict__ src) {
  const unsigned lane = lane_id();
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kBlock / kWave);
  for (uint64_t u = static_cast<uint64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; u < n;
       u += waves) {
    const uint32_t b = Ou[u], e = Ou[u + 1], du = e - b;
    uint32_t out = O[u];
    for (uint32_t base = 0; base < du; base += kWave) {
      const uint32_t i = base + lane;
      bool keep = false;
      uint32_t v = 0;
      if (i < du) {
        v = Eu[b + i];
        keep = keep_edge(static_cast<uint32_t>(u), du, v, Ou[v + 1] - Ou[v]);
      }
      const uint64_t mask = __ballot(keep);
      if (keep) {
        const uint32_t pos = out + lanes_below(mask);
        E[pos] = v;
        src[pos] = static_cast<uint32_t>(u);
      }
      out += static_cast<uint32_t>(__popcll(mask));
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_fill_src(const uint32_t* __restrict__ O, uint32_t n,
                                                     uint32_t* __restrict__ src) {
  const unsigned lane = lane_id();
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kBlock / kWave);
  for (uint64_t u = static_cast<uint64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; u < n;
       u += waves) {
    for (uint32_t i = O[u] + lane; i < O[u + 1]; i += kWave) src[i] = static_cast<uint32_t>(u);
  }
}

__global__ void k_iota(uint32_t* __restrict__ a, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    a[i] = static_cast<uint32_t>(i);
}

template <typename T>
__global__ void k_fill(T* __restrict__ a, uint64_t n, T v) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    a[i] = v;
}

// Undirected CSR check, wave per vertex (coalesced over the list, so a hub
// vertex does not serialize one thread): offsets ordered and within m, each
// list strictly increasing with ids < n and no self-loop; *bad = lowest
// offending vertex. A list is read only after its offsets passed, so every
// read stays inside O[0..n] and E[0..m).
__global__ __launch_bounds__(kBlock) void k_check_csr(const uint32_t* __restrict__ O, const uint32_t* __restrict__ E,
                                                      uint32_t n, uint64_t m, uint32_t* __restrict__ bad) {
  const unsigned lane = lane_id();
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kBlock / kWave);
  for (uint64_t u = static_cast<uint64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; u < n;
       u += waves) {
    const uint32_t o0 = O[u], o1 = O[u + 1];
    bool ok = o0 <= o1 && o1 <= m;
    if (ok) {
      for (uint32_t i = o0 + lane; i < o1; i += kWave) {
        const uint32_t x = E[i];
        ok = ok && x < n && x != u && (i == o0 || E[i - 1] < x);
      }
    }
    if (__ballot(!ok) && lane == 0) atomic_min_u32(bad, static_cast<uint32_t>(u));
  }
}

// Disjoint union of a batch, thread per union vertex v (graph g, local id u):
// graph g's offsets sit at O[vfirst[g] + g ..] and its edges at E[efirst[g] ..]
// with local ids; Ou[v] = efirst[g] + O_g[u] and every id gains vfirst[g].
// The same checks as k_check_csr, per graph (O_g[0] = 0 and O_g[n_g] = m_g are
// checked on the host); *bad = lowest offending union vertex.
__global__ void k_batch_union(const uint32_t* __restrict__ O, const uint32_t* __restrict__ E,
                              const uint64_t* __restrict__ vfirst, const uint64_t* __restrict__ efirst, uint32_t count,
                              uint64_t nv, uint32_t* __restrict__ Ou, uint32_t* __restrict__ Eu,
                              uint32_t* __restrict__ bad) {
  for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < nv;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint32_t lo = 0, hi = count;  // largest g with vfirst[g] <= v
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (vfirst[mid] <= v) lo = mid;
      else hi = mid;
    }
    const uint32_t g = lo;
    const uint64_t u = v - vfirst[g];
    const uint64_t ng = vfirst[g + 1] - vfirst[g], mg = efirst[g + 1] - efirst[g];
    const uint32_t* Og = O + vfirst[g] + g;
    const uint32_t o0 = Og[u], o1 = Og[u + 1];
    bool ok = o0 <= o1 && o1 <= mg;
    Ou[v] = static_cast<uint32_t>(efirst[g] + (ok ? o0 : 0));
    const uint32_t* Eg = E + efirst[g];
    uint32_t* Eo = Eu + efirst[g];
    for (uint32_t i = o0; ok && i < o1; ++i) {
      const uint32_t x = Eg[i];
      ok = x < ng && x != u && (i == o0 || Eg[i - 1] < x);
      Eo[i] = static_cast<uint32_t>(x + vfirst[g]);
    }
    if (!ok) atomic_min_u32(bad, static_cast<uint32_t>(v));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) Ou[nv] = static_cast<uint32_t>(efirst[count]);
}

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint64_t* __restrict__ idx,
                             uint64_t* __restrict__ out, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    out[i] = src[idx[i]];
}

__global__ void k_gather_u64(const uint64_t* __restrict__ src, const uint64_t* __restrict__ idx,
                             uint64_t* __restrict__ out, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    out[i] = src[idx[i]];
}

}  // namespace

void launch_check_csr(const uint32_t* O, const uint32_t* E, uint32_t n, uint64_t m, uint32_t* bad, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_check_csr, dim3(grid_for(n, kBlock / kWave, 1u << 16)), dim3(kBlock), 0, s, O, E, n, m, bad);
  NUC34_LAUNCH_CHECK();
}

void launch_batch_union(const uint32_t* O, const uint32_t* E, const uint64_t* vfirst, const uint64_t* efirst,
                        uint32_t count, uint64_t nv, uint32_t* Ou, uint32_t* Eu, uint32_t* bad, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_union, dim3(grid_for(nv ? nv : 1, 256, 1u << 16)), dim3(256), 0, s, O, E, vfirst, efirst,
                     count, nv, Ou, Eu, bad);
  NUC34_LAUNCH_CHECK();
}

void launch_gather_u32(const uint32_t* src, const uint64_t* idx, uint64_t* out, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_gather_u32, dim3(grid_for(n, 256, 1u << 14)), dim3(256), 0, s, src, idx, out, n);
  NUC34_LAUNCH_CHECK();
}

void launch_gather_u64(const uint64_t* src, const uint64_t* idx, uint64_t* out, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_gather_u64, dim3(grid_for(n, 256, 1u << 14)), dim3(256), 0, s, src, idx, out, n);
  NUC34_LAUNCH_CHECK();
}

void launch_orient_count(const uint32_t* Ou, const uint32_t* Eu, uint32_t n, uint32_t* outdeg, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_orient_count, dim3(grid_for(n, kBlock / kWave, 1u << 16)), dim3(kBlock), 0, s, Ou, Eu, n,
                     outdeg);
  NUC34_LAUNCH_CHECK();
}

void launch_orient_fill(const uint32_t* Ou, const uint32_t* Eu, uint32_t n, const uint32_t* O, uint32_t* E,
                        uint32_t* src, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_orient_fill, dim3(grid_for(n, kBlock / kWave, 1u << 16)), dim3(kBlock), 0, s, Ou, Eu, n, O,
                     E, src);
  NUC34_LAUNCH_CHECK();
}

void launch_fill_src(const uint32_t* O, uint32_t n, uint32_t* src, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_src, dim3(grid_for(n, kBlock / kWave, 1u << 16)), dim3(kBlock), 0, s, O, n, src);
  NUC34_LAUNCH_CHECK();
}

void launch_iota(uint32_t* a, uint64_t n, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 256, 1u << 16)), dim3(256), 0, s, a, n);
  NUC34_LAUNCH_CHECK();
}

void launch_fill_i32(int32_t* a, uint64_t n, int32_t v, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill<int32_t>, dim3(grid_for(n, 256, 1u << 16)), dim3(256), 0, s, a, n, v);
  NUC34_LAUNCH_CHECK();
}

void launch_fill_u32(uint32_t* a, uint64_t n, uint32_t v, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill<uint32_t>, dim3(grid_for(n, 256, 1u << 16)), dim3(256), 0, s, a, n, v);
  NUC34_LAUNCH_CHECK();
}

}  // namespace nuc34
