// Segmented sort of unique 64-bit keys: every segment [off[s], off[s+1]) of
// `keys` is sorted ascending in place. Segments of up to 64 keys are sorted in
// registers by one wave (bitonic over shuffles), up to 4096 keys by one
// workgroup in LDS, longer ones as LDS-sorted 4096-key chunks merged by rank.
#pragma once
#include "common.hpp"

namespace nuc34 {

void segmented_sort_u64(uint64_t* keys, const uint64_t* off, uint64_t nseg, Allocator& alloc,
                        hipStream_t s);

}  // namespace nuc34
