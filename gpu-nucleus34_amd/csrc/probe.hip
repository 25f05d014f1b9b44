// Measured device limits that bound this path's kernels: streaming copy
// bandwidth, random 4-byte atomic throughput and random 4-byte load throughput
// (L2-resident and HBM-sized working sets). The peel's debits are random
// atomics and its ownership tests random loads, so these rates, not the HBM
// peak alone, are its roofline.
#include "../../include/nuc34.h"
#include "common.hpp"

namespace nuc34 {

namespace {

__device__ __forceinline__ uint32_t mix32(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return static_cast<uint32_t>(x);
}

__global__ void k_random_atomic(int32_t* a, uint32_t n, uint64_t ops) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < ops;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    atomic_sub_i32(&a[mix32(i) % n], 1);
}

__global__ void k_random_load(const uint32_t* a, uint32_t n, uint64_t ops, uint32_t* sink) {
  uint32_t acc = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < ops;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    acc += a[mix32(i) % n];
  if (acc == 0x9E3779B9u) *sink = acc;  // keeps the loads
}

}  // namespace

}  // namespace nuc34

using namespace nuc34;

namespace {

// Buffers and events of one measurement, released on every way out.
struct ProbeRes {
  void* a = nullptr;
  void* b = nullptr;
  void* sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~ProbeRes() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (sink) (void)hipFree(sink);
  }
};

}  // namespace

extern "C" int nuc34_measure_device(int device, nuc34_device_rates* out) {
  if (!out) return NUC34_E_INVALID;
  try {
    NUC34_HIP(hipSetDevice(device));
    const size_t bytes = size_t(1) << 30;
    ProbeRes r;
    for (void** p : {&r.a, &r.b}) {
      const hipError_t e = hipMalloc(p, bytes);
      if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        return NUC34_E_OOM;
      }
      NUC34_HIP(e);
    }
    NUC34_HIP(hipMalloc(&r.sink, 4));
    NUC34_HIP(hipEventCreate(&r.e0));
    NUC34_HIP(hipEventCreate(&r.e1));
    NUC34_HIP(hipMemset(r.a, 0, bytes));
    auto best_ms = [&](auto&& launch) {
      float best = 1e30f;
      for (int rep = 0; rep < 3; ++rep) {
        NUC34_HIP(hipEventRecord(r.e0, nullptr));
        launch();
        NUC34_HIP(hipEventRecord(r.e1, nullptr));
        NUC34_HIP(hipEventSynchronize(r.e1));
        float ms = 0;
        NUC34_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
        best = std::min(best, ms);
      }
      return static_cast<double>(best);
    };
    const double copy_ms =
        best_ms([&] { NUC34_HIP(hipMemcpyAsync(r.b, r.a, bytes, hipMemcpyDeviceToDevice, nullptr)); });
    const uint64_t ops = uint64_t(1) << 28;
    const uint32_t n_atomic = 16u << 20, n_l2 = 1u << 20, n_hbm = 64u << 20;  // 64 MiB, 4 MiB, 256 MiB
    const double atomic_ms = best_ms([&] {
      hipLaunchKernelGGL(k_random_atomic, dim3(8192), dim3(256), 0, nullptr, static_cast<int32_t*>(r.a), n_atomic,
                         ops);
      NUC34_LAUNCH_CHECK();
    });
    const double l2_ms = best_ms([&] {
      hipLaunchKernelGGL(k_random_load, dim3(8192), dim3(256), 0, nullptr, static_cast<const uint32_t*>(r.a), n_l2,
                         ops, static_cast<uint32_t*>(r.sink));
      NUC34_LAUNCH_CHECK();
    });
    const double hbm_ms = best_ms([&] {
      hipLaunchKernelGGL(k_random_load, dim3(8192), dim3(256), 0, nullptr, static_cast<const uint32_t*>(r.a), n_hbm,
                         ops, static_cast<uint32_t*>(r.sink));
      NUC34_LAUNCH_CHECK();
    });
    out->copy_gbps = 2.0 * static_cast<double>(bytes) / (copy_ms * 1e6);
    out->atomic_gops = static_cast<double>(ops) / (atomic_ms * 1e6);
    out->load_l2_gops = static_cast<double>(ops) / (l2_ms * 1e6);
    out->load_hbm_gops = static_cast<double>(ops) / (hbm_ms * 1e6);
    return NUC34_OK;
  } catch (const std::exception&) {
    return NUC34_E_HIP;
  }
}
