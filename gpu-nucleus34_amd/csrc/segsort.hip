#include <vector>

#include "segsort.hpp"

namespace nuc34 {

namespace {

constexpr uint32_t kChunk = 4096;  // keys per workgroup LDS sort (32 KiB)

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  const uint32_t lo = __shfl_xor(static_cast<uint32_t>(v), m);
  const uint32_t hi = __shfl_xor(static_cast<uint32_t>(v >> 32), m);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// One wave per segment; segments longer than 64 are queued for the workgroup pass.
__global__ __launch_bounds__(kBlock) void k_segsort_wave(uint64_t* __restrict__ keys,
                                                         const uint64_t* __restrict__ off, uint64_t nseg,
                                                         uint32_t* __restrict__ big, uint32_t* __restrict__ nbig,
                                                         uint32_t* __restrict__ lng, uint32_t* __restrict__ nlng) {
  const unsigned lane = lane_id();
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (kBlock / kWave);
  for (uint64_t s = static_cast<uint64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave; s < nseg;
       s += waves) {
    const uint64_t b = off[s];
    const uint64_t len = off[s + 1] - b;
    if (len <= 1) continue;
    if (len > kWave) {
      if (lane == 0) {
        if (len > kChunk) lng[atomic_add_u32(nlng, 1)] = static_cast<uint32_t>(s);
        else big[atomic_add_u32(nbig, 1)] = static_cast<uint32_t>(s);
      }
      continue;
    }
    uint64_t x = lane < len ? keys[b + lane] : ~0ull;
#pragma unroll
    for (unsigned k = 2; k <= kWave; k <<= 1) {
#pragma unroll
      for (unsigned j = k >> 1; j > 0; j >>= 1) {
        const uint64_t y = shfl_xor_u64(x, static_cast<int>(j));
        const bool up = (lane & k) == 0;
        const bool lower = (lane & j) == 0;
        x = (lower == up) ? (x < y ? x : y) : (x < y ? y : x);
      }
    }
    if (lane < len) keys[b + lane] = x;
  }
}

// Bitonic sort of n (power of two, <= kChunk) keys in LDS by one workgroup.
__device__ void lds_bitonic(uint64_t* sh, unsigned n) {
  for (unsigned k = 2; k <= n; k <<= 1) {
    for (unsigned j = k >> 1; j > 0; j >>= 1) {
      for (unsigned i = threadIdx.x; i < n; i += kBlock) {
        const unsigned ixj = i ^ j;
        if (ixj > i) {
          const uint64_t a = sh[i], c = sh[ixj];
          const bool up = (i & k) == 0;
          if ((a > c) == up) {
            sh[i] = c;
            sh[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

// One workgroup per queued segment: sorts it (or each of its 4096-key chunks).
__global__ __launch_bounds__(kBlock) void k_segsort_block(uint64_t* __restrict__ keys,
                                                          const uint64_t* __restrict__ off,
                                                          const uint32_t* __restrict__ big, uint32_t nbig) {
  __shared__ uint64_t sh[kChunk];
  for (uint32_t q = blockIdx.x; q < nbig; q += gridDim.x) {
    const uint32_t s = big[q];
    const uint64_t b = off[s];
    const uint64_t len = off[s + 1] - b;
    for (uint64_t c0 = 0; c0 < len; c0 += kChunk) {
      const unsigned cl = static_cast<unsigned>(len - c0 < kChunk ? len - c0 : kChunk);
      unsigned n = 64;
      while (n < cl) n <<= 1;
      for (unsigned i = threadIdx.x; i < n; i += kBlock) sh[i] = i < cl ? keys[b + c0 + i] : ~0ull;
      __syncthreads();
      lds_bitonic(sh, n);
      for (unsigned i = threadIdx.x; i < cl; i += kBlock) keys[b + c0 + i] = sh[i];
      __syncthreads();
    }
  }
}

// Rank merge of the sorted chunks of one long segment: key at chunk position p
// goes to p + (number of smaller keys in every other chunk). One workgroup per
// (segment, chunk) work item.
__global__ __launch_bounds__(kBlock) void k_segsort_rank(const uint64_t* __restrict__ keys,
                                                         const uint64_t* __restrict__ off,
                                                         const uint32_t* __restrict__ items_seg,
                                                         const uint32_t* __restrict__ items_chunk,
                                                         const uint64_t* __restrict__ items_tmp,
                                                         uint32_t nitems, uint64_t* __restrict__ tmp) {
  for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const uint32_t s = items_seg[it];
    const uint32_t c = items_chunk[it];
    const uint64_t b = off[s];
    const uint64_t len = off[s + 1] - b;
    const uint64_t nch = (len + kChunk - 1) / kChunk;
    const uint64_t c0 = static_cast<uint64_t>(c) * kChunk;
    const uint64_t cl = len - c0 < kChunk ? len - c0 : kChunk;
    for (uint64_t p = threadIdx.x; p < cl; p += kBlock) {
      const uint64_t x = keys[b + c0 + p];
      uint64_t rank = p;
      for (uint64_t o = 0; o < nch; ++o) {
        if (o == c) continue;
        const uint64_t ob = b + o * kChunk;
        const uint64_t ol = len - o * kChunk < kChunk ? len - o * kChunk : kChunk;
        rank += lower_bound_dev<uint64_t>(keys, ob, ob + ol, x) - ob;
      }
      tmp[items_tmp[it] + rank] = x;
    }
  }
}

__global__ void k_gather_bounds(const uint64_t* __restrict__ off, const uint32_t* __restrict__ segs,
                                uint32_t nsegs, uint64_t* __restrict__ out) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nsegs) {
    out[2 * q] = off[segs[q]];
    out[2 * q + 1] = off[segs[q] + 1];
  }
}

__global__ void k_copy_back(uint64_t* __restrict__ keys, const uint64_t* __restrict__ off,
                            const uint32_t* __restrict__ segs, const uint64_t* __restrict__ tmp_off,
                            uint32_t nsegs, const uint64_t* __restrict__ tmp) {
  for (uint32_t q = blockIdx.x; q < nsegs; q += gridDim.x) {
    const uint32_t s = segs[q];
    const uint64_t b = off[s];
    const uint64_t len = off[s + 1] - b;
    for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) keys[b + i] = tmp[tmp_off[q] + i];
  }
}

}  // namespace

void segmented_sort_u64(uint64_t* keys, const uint64_t* off, uint64_t nseg, Allocator& alloc, hipStream_t s) {
  if (!nseg) return;
  if (nseg > 0xFFFFFFFFull) throw Overflow("segmented sort: more than 2^32 segments");
  DevBuf big = alloc.get(nseg * sizeof(uint32_t), "segsort queue");
  DevBuf lng = alloc.get(nseg * sizeof(uint32_t), "segsort long queue");
  DevBuf cnt = alloc.get(2 * sizeof(uint32_t), "segsort counts");
  NUC34_HIP(hipMemsetAsync(cnt.as<void>(), 0, 2 * sizeof(uint32_t), s));
  hipLaunchKernelGGL(k_segsort_wave, dim3(grid_for(nseg, kBlock / kWave, 65536)), dim3(kBlock), 0, s, keys, off,
                     nseg, big.as<uint32_t>(), cnt.as<uint32_t>(), lng.as<uint32_t>(), cnt.as<uint32_t>() + 1);
  NUC34_LAUNCH_CHECK();
  uint32_t h_cnt[2] = {0, 0};
  NUC34_HIP(hipMemcpyAsync(h_cnt, cnt.as<void>(), sizeof(h_cnt), hipMemcpyDeviceToHost, s));
  NUC34_HIP(hipStreamSynchronize(s));
  if (h_cnt[0]) {
    hipLaunchKernelGGL(k_segsort_block, dim3(std::min<uint32_t>(h_cnt[0], 65536)), dim3(kBlock), 0, s, keys,
                       off, big.as<uint32_t>(), h_cnt[0]);
    NUC34_LAUNCH_CHECK();
  }
  const uint32_t nl = h_cnt[1];
  if (!nl) return;
  // Segments longer than one chunk: sort each chunk, then merge chunks by rank.
  hipLaunchKernelGGL(k_segsort_block, dim3(std::min<uint32_t>(nl, 65536)), dim3(kBlock), 0, s, keys, off,
                     lng.as<uint32_t>(), nl);
  NUC34_LAUNCH_CHECK();
  std::vector<uint32_t> segs(nl);
  std::vector<uint64_t> qo(2ull * nl);
  DevBuf qoff = alloc.get(2ull * nl * sizeof(uint64_t), "segsort bounds");
  hipLaunchKernelGGL(k_gather_bounds, dim3((nl + 255) / 256), dim3(256), 0, s, off, lng.as<uint32_t>(), nl,
                     qoff.as<uint64_t>());
  NUC34_LAUNCH_CHECK();
  NUC34_HIP(hipMemcpyAsync(segs.data(), lng.as<void>(), nl * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  NUC34_HIP(hipMemcpyAsync(qo.data(), qoff.as<void>(), qo.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  NUC34_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> it_seg, it_chunk;
  std::vector<uint64_t> it_tmp, long_tmp_off;
  uint64_t tmp_total = 0;
  for (uint32_t q = 0; q < nl; ++q) {
    const uint64_t len = qo[2 * q + 1] - qo[2 * q];
    long_tmp_off.push_back(tmp_total);
    const uint64_t nch = (len + kChunk - 1) / kChunk;
    for (uint64_t c = 0; c < nch; ++c) {
      it_seg.push_back(segs[q]);
      it_chunk.push_back(static_cast<uint32_t>(c));
      it_tmp.push_back(tmp_total);
    }
    tmp_total += len;
  }
  const uint32_t nit = static_cast<uint32_t>(it_seg.size());
  DevBuf d_itseg = alloc.get(nit * 4ull, "segsort items");
  DevBuf d_itch = alloc.get(nit * 4ull, "segsort items");
  DevBuf d_ittmp = alloc.get(nit * 8ull, "segsort items");
  DevBuf d_ltmp = alloc.get(nl * 8ull, "segsort long");
  DevBuf tmp = alloc.get(tmp_total * 8ull, "segsort merge buffer");
  NUC34_HIP(hipMemcpyAsync(d_itseg.as<void>(), it_seg.data(), nit * 4ull, hipMemcpyHostToDevice, s));
  NUC34_HIP(hipMemcpyAsync(d_itch.as<void>(), it_chunk.data(), nit * 4ull, hipMemcpyHostToDevice, s));
  NUC34_HIP(hipMemcpyAsync(d_ittmp.as<void>(), it_tmp.data(), nit * 8ull, hipMemcpyHostToDevice, s));
  NUC34_HIP(hipMemcpyAsync(d_ltmp.as<void>(), long_tmp_off.data(), nl * 8ull, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_segsort_rank, dim3(std::min<uint32_t>(nit, 65536)), dim3(kBlock), 0, s, keys, off,
                     d_itseg.as<uint32_t>(), d_itch.as<uint32_t>(), d_ittmp.as<uint64_t>(), nit,
                     tmp.as<uint64_t>());
  NUC34_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_copy_back, dim3(std::min<uint32_t>(nl, 65536)), dim3(kBlock), 0, s, keys, off,
                     lng.as<uint32_t>(), d_ltmp.as<uint64_t>(), nl, tmp.as<uint64_t>());
  NUC34_LAUNCH_CHECK();
  NUC34_HIP(hipStreamSynchronize(s));
}

}  // namespace nuc34
