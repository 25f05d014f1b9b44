// Launch wrappers for the gfx950 kernels of the (3,4)-nucleus path. All device
// pointers; every wrapper enqueues on stream `s` and checks the launch.
#pragma once
#include "bitrows.hpp"
#include "common.hpp"

namespace nuc34 {

// (a) degree orientation of an undirected CSR (Ou[n+1], Eu[2m]).
void launch_orient_count(const uint32_t* Ou, const uint32_t* Eu, uint32_t n, uint32_t* outdeg, hipStream_t s);
void launch_orient_fill(const uint32_t* Ou, const uint32_t* Eu, uint32_t n, const uint32_t* O, uint32_t* E,
                        uint32_t* src, hipStream_t s);
void launch_fill_src(const uint32_t* O, uint32_t n, uint32_t* src, hipStream_t s);
// Checks of an uploaded undirected CSR (O[0] = 0 and O[n] = m checked by the
// caller): *bad (preset to UINT32_MAX) becomes the lowest malformed vertex.
void launch_check_csr(const uint32_t* O, const uint32_t* E, uint32_t n, uint64_t m, uint32_t* bad, hipStream_t s);
// Disjoint union of `count` undirected graphs uploaded back to back (offsets
// at O[vfirst[g] + g ..], edges at E[efirst[g] ..], local ids) into Ou[nv+1],
// Eu, with the same checks per graph into *bad.
void launch_batch_union(const uint32_t* O, const uint32_t* E, const uint64_t* vfirst, const uint64_t* efirst,
                        uint32_t count, uint64_t nv, uint32_t* Ou, uint32_t* Eu, uint32_t* bad, hipStream_t s);

// (b) triangle enumeration over oriented edges [e_begin, e_end) (whole
// vertices): the bits pass fills the incidence rows, the emit pass (after the
// row ranks) writes the triangles at their peel ids (the rank of their bit in
// their b-edge's b-role row: by b-edge, then a-edge) with their canonical
// (a, b) indices, and at their c-slots their peel ids with the in-list
// position of their a-edge.
// `tile` = oriented edges per wave tile: 16, 32 or 64 (std::logic_error
// otherwise). With `bins` (the degree-binned schedule), owners of out-degree
// min_deg+1..640 are enumerated a vertex at a time by a wave or a workgroup
// and the tiles take the rest.
struct VertexBins {
  const uint32_t* lists;   // 3 * n entries: vertices of each class
  const uint32_t* counts;  // 3 counts (on the device)
  uint32_t n;
  uint32_t min_deg;        // owners of out-degree up to this stay on the tiles
};
// Sorts the vertices of out-degree min_deg+1..640 into the three classes
// (counts zeroed by the caller); min_deg < 64.
void launch_vertex_bins(const uint32_t* O, uint32_t n, uint32_t min_deg, uint32_t* lists, uint32_t* counts,
                        hipStream_t s);
void launch_tri_bits(const uint32_t* O, const uint32_t* E, const uint32_t* src, uint64_t e_begin, uint64_t e_end,
                     const BitRows& rows, const uint32_t* inpos, uint32_t tile, const VertexBins* bins,
                     hipStream_t s);
struct TriOut {
  uint32_t* triA = nullptr;   // a-edge, by peel id
  uint64_t* triBC = nullptr;  // (b << 32) | c, by peel id
  uint32_t* canon = nullptr;  // canonical (a, b) index, by peel id
  uint2* cSlot = nullptr;     // per c-slot: peel id, in-list position of the a-edge
};
void launch_tri_emit(const uint32_t* O, const uint32_t* E, const uint32_t* src, uint64_t e_begin, uint64_t e_end,
                     const BitRows& rows, const uint32_t* inpos, const TriOut& out, uint32_t tile,
                     const VertexBins* bins, hipStream_t s);
// Canonical order from peel order: dst[canon[q]] = src[q] for q < T.
void launch_canon_tris(const uint32_t* triA, const uint64_t* triBC, const uint32_t* canon, uint64_t T,
                       uint32_t* ctriA, uint64_t* ctriBC, hipStream_t s);
void launch_canon_u32(const uint32_t* src, const uint32_t* canon, uint64_t T, uint32_t* dst, hipStream_t s);

// abc[3i..3i+2] = (a, b, c) of triangles [0, count) of the given arrays.
void launch_tri_abc(const uint32_t* triA, const uint64_t* triBC, uint64_t count, uint32_t* abc, hipStream_t s);
// The same for all T triangles of the peel-ordered arrays, each at its
// canonical index canon[q].
void launch_tri_abc_canon(const uint32_t* triA, const uint64_t* triBC, const uint32_t* canon, uint64_t T,
                          uint32_t* abc, hipStream_t s);
// out[i] = src[idx[i]] for i < n (idx in bounds of src, checked by the caller).
void launch_gather_u32(const uint32_t* src, const uint64_t* idx, uint64_t* out, uint64_t n, hipStream_t s);
void launch_gather_u64(const uint64_t* src, const uint64_t* idx, uint64_t* out, uint64_t n, hipStream_t s);
void launch_iota(uint32_t* a, uint64_t n, hipStream_t s);
void launch_fill_i32(int32_t* a, uint64_t n, int32_t v, hipStream_t s);
void launch_fill_u32(uint32_t* a, uint64_t n, uint32_t v, hipStream_t s);

}  // namespace nuc34
