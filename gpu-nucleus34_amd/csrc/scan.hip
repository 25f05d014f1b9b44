#include "bitrows.hpp"
#include "scan.hpp"

namespace nuc34 {

namespace {

// The value a scan input contributes: itself, or for a ranked row word the
// number of its set bits (the scan then gives every word its running rank).
template <typename T>
__device__ __forceinline__ T scan_value(T x) {
  return x;
}
__device__ __forceinline__ uint32_t scan_value(RankWord w) { return static_cast<uint32_t>(__popc(w.bits)); }
__device__ __forceinline__ uint32_t scan_value(BitWord w) { return static_cast<uint32_t>(__popc(w.bits)); }

// Block-wide inclusive scan of one value per thread; returns this thread's
// inclusive prefix and writes the block total to *total.
template <typename T>
__device__ __forceinline__ T block_inclusive(T v, T* lds, T* total) {
  const unsigned lane = threadIdx.x & (kWave - 1);
  const unsigned wid = threadIdx.x / kWave;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const T t = __shfl_up(v, static_cast<unsigned>(off));
    if (lane >= static_cast<unsigned>(off)) v += t;
  }
  if (lane == kWave - 1) lds[wid] = v;
  __syncthreads();
  T before = 0, all = 0;
#pragma unroll
  for (unsigned w = 0; w < kBlock / kWave; ++w) {
    const T x = lds[w];
    if (w < wid) before += x;
    all += x;
  }
  __syncthreads();
  *total = all;
  return v + before;
}

template <typename In, typename Out>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const In* __restrict__ in, uint64_t n,
                                                        Out* __restrict__ partial) {
  __shared__ Out lds[kBlock / kWave];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile;
  Out acc = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const uint64_t i = base + static_cast<uint64_t>(j) * kBlock + threadIdx.x;
    if (i < n) acc += static_cast<Out>(scan_value(in[i]));
  }
  Out total;
  block_inclusive<Out>(acc, lds, &total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Exclusive scan of the tile totals in place (single workgroup), total at [tiles].
template <typename Out>
__global__ __launch_bounds__(kBlock) void k_scan_partials(Out* __restrict__ partial, uint64_t tiles) {
  __shared__ Out lds[kBlock / kWave];
  Out carry = 0;
  for (uint64_t b = 0; b < tiles; b += kBlock) {
    const uint64_t i = b + threadIdx.x;
    const Out v = i < tiles ? partial[i] : Out(0);
    Out total;
    const Out inc = block_inclusive<Out>(v, lds, &total);
    if (i < tiles) partial[i] = carry + inc - v;
    carry += total;
  }
  if (threadIdx.x == 0) partial[tiles] = carry;
}

template <typename In, typename Out>
__global__ __launch_bounds__(kBlock) void k_scan_tiles(const In* __restrict__ in, uint64_t n,
                                                       const Out* __restrict__ partial,
                                                       Out* __restrict__ out) {
  __shared__ Out lds[kBlock / kWave];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile;
  Out carry = partial[blockIdx.x];
  for (int j = 0; j < kScanItems; ++j) {
    const uint64_t i = base + static_cast<uint64_t>(j) * kBlock + threadIdx.x;
    const Out v = i < n ? static_cast<Out>(scan_value(in[i])) : Out(0);
    Out total;
    const Out inc = block_inclusive<Out>(v, lds, &total);
    if (i < n) out[i] = carry + inc - v;
    carry += total;
  }
}

template <typename Out>
__global__ void k_scan_total(const Out* __restrict__ partial, uint64_t tiles, Out* __restrict__ out,
                             uint64_t n) {
  out[n] = partial[tiles];
}

}  // namespace

template <typename In, typename Out>
void exclusive_scan(const In* in, Out* out, uint64_t n, Out* scratch, hipStream_t s) {
  const uint64_t tiles = scan_tiles(n);
  if (tiles) {
    hipLaunchKernelGGL((k_scan_reduce<In, Out>), dim3(static_cast<unsigned>(tiles)), dim3(kBlock), 0, s,
                       in, n, scratch);
    NUC34_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL((k_scan_partials<Out>), dim3(1), dim3(kBlock), 0, s, scratch, tiles);
  NUC34_LAUNCH_CHECK();
  if (tiles) {
    hipLaunchKernelGGL((k_scan_tiles<In, Out>), dim3(static_cast<unsigned>(tiles)), dim3(kBlock), 0, s,
                       in, n, scratch, out);
    NUC34_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL((k_scan_total<Out>), dim3(1), dim3(1), 0, s, scratch, tiles, out, n);
  NUC34_LAUNCH_CHECK();
}

template void exclusive_scan<uint32_t, uint32_t>(const uint32_t*, uint32_t*, uint64_t, uint32_t*, hipStream_t);
template void exclusive_scan<uint32_t, uint64_t>(const uint32_t*, uint64_t*, uint64_t, uint64_t*, hipStream_t);
template void exclusive_scan<int32_t, uint64_t>(const int32_t*, uint64_t*, uint64_t, uint64_t*, hipStream_t);
template void exclusive_scan<uint64_t, uint64_t>(const uint64_t*, uint64_t*, uint64_t, uint64_t*, hipStream_t);
template void exclusive_scan<RankWord, uint64_t>(const RankWord*, uint64_t*, uint64_t, uint64_t*, hipStream_t);
template void exclusive_scan<BitWord, uint64_t>(const BitWord*, uint64_t*, uint64_t, uint64_t*, hipStream_t);

}  // namespace nuc34
