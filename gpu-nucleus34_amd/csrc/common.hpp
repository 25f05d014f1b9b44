// Shared definitions for the gfx950 kernels of the (3,4)-nucleus path:
// wave64 helpers, HIP error handling and device buffers owned by a context.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>

namespace nuc34 {

constexpr int kWave = 64;
constexpr int kBlock = 256;  // 4 waves per workgroup

struct HipError : std::runtime_error {
  hipError_t code;
  HipError(hipError_t c, const std::string& where)
      : std::runtime_error(where + ": " + hipGetErrorString(c)), code(c) {}
};

struct Overflow : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct OutOfMemory : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// The HIP call in progress on this thread, named in error messages when the
// runtime (or a tool attached to it) throws instead of returning an error.
inline thread_local const char* tls_hip_call = "";

#define NUC34_HIP(expr)                                                                   \
  do {                                                                                    \
    ::nuc34::tls_hip_call = #expr;                                                        \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) throw ::nuc34::HipError(e_, std::string(#expr) + " at " __FILE__); \
  } while (0)

#define NUC34_LAUNCH_CHECK() NUC34_HIP(hipGetLastError())

inline unsigned grid_for(uint64_t items, unsigned per_block, unsigned cap = 1u << 20) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g == 0) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

// grid_for rounded to a multiple of the 8 XCDs (up, or down at the cap), for
// kernels that walk their tiles in xcd_block() order with a grid-strided loop.
inline unsigned grid_for_xcd(uint64_t items, unsigned per_block, unsigned cap = 1u << 20) {
  const unsigned g = grid_for(items, per_block, cap);
  if (g < 8) return g;
  const unsigned up = (g + 7u) & ~7u;
  return up <= cap ? up : g & ~7u;  // never above the cap
}

// ---- device-side helpers -------------------------------------------------

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// Orders this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup number renumbered so that each of the 8 XCDs gets a contiguous run:
// the hardware deals workgroups to the XCDs in turn, and each XCD has its own
// L2, so neighbouring tiles that read the same rows then share one L2. It is a
// permutation of [0, gridDim.x) (the identity when the grid is not a multiple
// of 8): results never depend on it.
__device__ __forceinline__ unsigned xcd_block() {
  constexpr unsigned kXcds = 8;
  if (gridDim.x % kXcds) return blockIdx.x;
  return (blockIdx.x % kXcds) * (gridDim.x / kXcds) + blockIdx.x / kXcds;
}

// Number of set bits of `mask` below this lane (mbcnt on the 64-bit ballot).
__device__ __forceinline__ unsigned lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint64_t atomic_add_u64(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_add_u32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t atomic_sub_i32(int32_t* p, int32_t v) {
  return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_min_u32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_or_u32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_and_u32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t atomic_and_u64(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_xor_u32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_xor(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-aggregated append: every active lane with `pred` gets a unique slot of
// *counter; one atomic per wave.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred, bool* got) {
  const uint64_t mask = __ballot(pred);
  *got = pred;
  if (!mask) return 0;
  const unsigned leader = static_cast<unsigned>(__builtin_ctzll(mask));
  uint32_t base = 0;
  if (lane_id() == leader) base = atomic_add_u32(counter, static_cast<uint32_t>(__popcll(mask)));
  base = __shfl(base, static_cast<int>(leader));
  return base + lanes_below(mask);
}

// Wave-aggregated append of n in [0, 3] slots per active lane: returns the
// lane's first slot of *counter; one atomic per wave.
__device__ __forceinline__ uint32_t wave_append3(uint32_t* counter, uint32_t n) {
  const uint64_t b0 = __ballot(n & 1u);
  const uint64_t b1 = __ballot((n >> 1) & 1u);
  const uint32_t total = static_cast<uint32_t>(__popcll(b0) + 2 * __popcll(b1));
  if (!total) return 0;
  const unsigned leader = static_cast<unsigned>(__builtin_ctzll(b0 | b1));
  uint32_t base = 0;
  if (lane_id() == leader) base = atomic_add_u32(counter, total);
  base = __shfl(base, static_cast<int>(leader));
  return base + lanes_below(b0) + 2 * lanes_below(b1);
}

// Wave-wide inclusive prefix sum of a 32-bit value.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
  const unsigned lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t t = __shfl_up(v, static_cast<unsigned>(off));
    if (lane >= static_cast<unsigned>(off)) v += t;
  }
  return v;
}

// Wave-wide inclusive prefix sum of a 64-bit value.
__device__ __forceinline__ uint64_t wave_inclusive_sum_u64(uint64_t v) {
  const unsigned lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint64_t t = __shfl_up(v, static_cast<unsigned>(off));
    if (lane >= static_cast<unsigned>(off)) v += t;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// First index in sorted a[lo, hi) whose value is >= key.
template <typename T>
__device__ __forceinline__ uint64_t lower_bound_dev(const T* a, uint64_t lo, uint64_t hi, T key) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

}  // namespace nuc34

namespace nuc34 {

// Device allocations charged against a byte budget (0 = only the device limit).
class Allocator;

class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept { swap(o); }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      swap(o);
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release();
  template <typename T>
  T* as() const { return static_cast<T*>(p_); }
  size_t bytes() const { return bytes_; }
  explicit operator bool() const { return p_ != nullptr; }

 private:
  friend class Allocator;
  void swap(DevBuf& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(bytes_, o.bytes_);
    std::swap(owner_, o.owner_);
  }
  void* p_ = nullptr;
  size_t bytes_ = 0;
  Allocator* owner_ = nullptr;
};

class Allocator {
 public:
  uint64_t budget = 0;  // bytes; 0 = no limit beyond the device
  uint64_t used = 0;    // bytes held by live buffers
  uint64_t peak = 0;
  uint64_t cached = 0;  // bytes held in the reuse cache
  bool caching = true;
  ~Allocator() { trim(); }
  // Arena mode: every buffer is carved from [base, base + size) of device
  // memory the caller owns, and the device allocator is never called. Blocks
  // given back on top of the arena are reclaimed (buffers are released in
  // reverse order of their requests, mostly); the rest when the arena's user
  // is done with it.
  void use_arena(void* base, size_t size) {
    arena_ = static_cast<char*>(base);
    arena_size_ = size;
    arena_top_ = 0;
  }
  DevBuf get(size_t bytes, const char* what) {
    DevBuf b;
    if (!bytes) return b;
    bytes = (bytes + 255) & ~static_cast<size_t>(255);
    if (arena_) {
      if (bytes > arena_size_ - arena_top_)
        throw OutOfMemory(std::string("scratch array exhausted allocating ") + what + " (" + std::to_string(bytes) +
                          " bytes; " + std::to_string(arena_top_) + " of " + std::to_string(arena_size_) +
                          " in use)");
      b.p_ = arena_ + arena_top_;
      b.bytes_ = bytes;
      b.owner_ = this;
      arena_top_ += bytes;
      used += bytes;
      if (used > peak) peak = used;
      return b;
    }
    // reuse a cached block of at most twice the request
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first <= 2 * bytes) {
      b.p_ = it->second;
      b.bytes_ = it->first;
      b.owner_ = this;
      cached -= it->first;
      free_.erase(it);
      used += b.bytes_;
      if (used > peak) peak = used;
      return b;
    }
    if (budget && used + cached + bytes > budget) trim();
    if (budget && used + bytes > budget)
      throw OutOfMemory(std::string("HBM budget exceeded allocating ") + what + " (" +
                        std::to_string(bytes) + " bytes; in use " + std::to_string(used) + " of " +
                        std::to_string(budget) + ")");
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if ((e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) && cached) {
      (void)hipGetLastError();
      trim();
      e = hipMalloc(&p, bytes);
    }
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
      (void)hipGetLastError();
      throw OutOfMemory(std::string("device out of memory allocating ") + what + " (" +
                        std::to_string(bytes) + " bytes)");
    }
    if (e != hipSuccess) throw HipError(e, std::string("hipMalloc ") + what);
    b.p_ = p;
    b.bytes_ = bytes;
    b.owner_ = this;
    used += bytes;
    if (used > peak) peak = used;
    return b;
  }
  // Called by DevBuf::release. Work using the block is ordered before any later
  // user on the single stream of the owning context.
  void give_back(void* p, size_t bytes) {
    used -= bytes;
    if (arena_) {
      if (static_cast<char*>(p) + bytes == arena_ + arena_top_) arena_top_ -= bytes;
      return;
    }
    if (caching) {
      free_.emplace(bytes, p);
      cached += bytes;
    } else {
      (void)hipFree(p);
    }
  }
  void trim() {
    for (auto& kv : free_) (void)hipFree(kv.second);
    free_.clear();
    cached = 0;
  }

 private:
  std::multimap<size_t, void*> free_;
  char* arena_ = nullptr;
  size_t arena_size_ = 0, arena_top_ = 0;
};

inline void DevBuf::release() {
  if (p_) {
    if (owner_) owner_->give_back(p_, bytes_);
    else (void)hipFree(p_);
  }
  p_ = nullptr;
  bytes_ = 0;
  owner_ = nullptr;
}

}  // namespace nuc34
