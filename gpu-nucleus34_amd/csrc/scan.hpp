// Device-wide exclusive prefix sum (reduce-then-scan over 4096-element tiles),
// out[0..N] with out[N] = total. Used for CSR offsets: oriented offsets,
// triangles per edge, incidence offsets, four-cliques per triangle; over
// ranked row words (RankWord) it sums their set bits.
#pragma once
#include "common.hpp"

namespace nuc34 {

constexpr int kScanItems = 16;
constexpr uint64_t kScanTile = static_cast<uint64_t>(kBlock) * kScanItems;

inline uint64_t scan_tiles(uint64_t n) { return (n + kScanTile - 1) / kScanTile; }

// Scratch elements (of the output type) that exclusive_scan needs for n inputs.
inline uint64_t scan_scratch_elems(uint64_t n) { return scan_tiles(n) + 1; }

template <typename In, typename Out>
void exclusive_scan(const In* in, Out* out, uint64_t n, Out* scratch, hipStream_t s);

}  // namespace nuc34
