// C ABI of the (3,4)-nucleus path: a device context that owns every device
// buffer (no global device symbols, so contexts are reentrant), and the stage
// orchestration: orientation -> triangles + bit-packed incidence -> support ->
// [per-triangle K4 lists] -> peel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nuc34.h"
#include "../host/gen.hpp"
#include "../host/graph_io.hpp"
#include "../host/presets.hpp"
#include "cliques.hpp"
#include "kernels.hpp"
#include "scan.hpp"
#include "segsort.hpp"

using namespace nuc34;

struct nuc34_graph {
  LoadedGraph lg;
};

struct nuc34_ctx {
  int device = 0;
  // no-allocation mode (nuc34_decompose_noalloc): buffers come from the
  // device's scratch array, nothing is timed, peel rounds are launched eagerly
  // and their control is read back with synchronous copies
  bool noalloc = false;
  hipStream_t s = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Allocator alloc;
  std::string err;
  // oriented graph
  bool has_graph = false;
  uint32_t n = 0;
  uint64_t m = 0;
  DevBuf O, E, src;
  // batch: vertex and oriented-edge boundaries of the graphs of a union
  std::vector<uint64_t> batch_v, batch_e;
  // per graph: its rounds, when the last peel ran one workgroup per graph
  std::vector<uint32_t> batch_rounds;
  // triangles and the bit-packed incidence
  bool has_tris = false;
  uint64_t T = 0;
  DevBuf IO, IE, IEe, inpos;        // in-CSR of the oriented graph
  DevBuf rowA, rowC, Aout, Ain, Cm;  // bit rows
  uint64_t WA = 0, WC = 0;           // words of the Aout/Ain rows and of the Cm rows
  DevBuf offA;
  DevBuf AoutR;                      // running rank of the Aout bits per word (peel ids)
  DevBuf triA, triBC;                // by peel id (b-edge, then a-edge)
  DevBuf canon;                      // canonical (a, b) index per peel id
  DevBuf cSlot;                      // per c-slot: peel id, in-list position of its a-edge
  DevBuf ctriA, ctriBC;              // canonical order, built when first asked for
  DevBuf vbins;                      // degree-binned schedule: vertex classes (during the triangle passes)
  // support
  bool has_sup = false;
  uint64_t Q = 0;
  DevBuf sup, own, small;  // small: counters
  // direct: per-triangle K4 partner lists
  bool has_k4 = false;
  DevBuf koff, lists;
  // peel result
  bool has_k = false;
  DevBuf kv;
  nuc34_stats st{};
  nuc34_tuning tune{};  // 0 = defaults
  // peel polling: pinned host copies of the two control slots after each
  // captured launch, and the events that mark them written
  PeelCtl* ctl_host = nullptr;  // 2 launches x 2 slots
  hipEvent_t poll_ev[2] = {nullptr, nullptr};
  // results back to the host: two pinned staging halves and their events
  char* stage = nullptr;
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  // per-kernel event timers
  struct KTimer {
    const char* name = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    uint64_t launches = 0;
    double ms = 0;
    bool pending = false;
  };
  KTimer kt[NUC34_MAX_KERNEL_STATS];
};

namespace {

int fail(nuc34_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

template <typename F>
int guard(nuc34_ctx* c, F&& f) {
  try {
    NUC34_HIP(hipSetDevice(c->device));
    return f();
  } catch (const OutOfMemory& e) {
    return fail(c, NUC34_E_OOM, e.what());
  } catch (const Overflow& e) {
    return fail(c, NUC34_E_OVERFLOW, e.what());
  } catch (const HipError& e) {
    return fail(c, NUC34_E_HIP, e.what());
  } catch (const std::invalid_argument& e) {
    return fail(c, NUC34_E_INVALID, e.what());
  } catch (const std::logic_error& e) {
    return fail(c, NUC34_E_INVALID, e.what());
  } catch (const std::exception& e) {
    // thrown from inside the runtime: name the call
    return fail(c, NUC34_E_HIP, std::string(e.what()) + " (in " + tls_hip_call + ")");
  }
}

// Milliseconds between two events recorded around `f` on the context stream.
template <typename F>
double timed(nuc34_ctx* c, F&& f) {
  if (c->noalloc) {  // no events: nothing may be created
    f();
    return 0.0;
  }
  NUC34_HIP(hipEventRecord(c->ev0, c->s));
  f();
  NUC34_HIP(hipEventRecord(c->ev1, c->s));
  NUC34_HIP(hipEventSynchronize(c->ev1));
  float ms = 0;
  NUC34_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  return ms;
}

constexpr size_t kStageBytes = size_t(32) << 20;  // each of the two pinned halves

// memcpy with up to 8 host threads (large blocks only).
void par_memcpy(char* dst, const char* src, size_t bytes) {
  const size_t kMin = size_t(4) << 20;
  const unsigned nt = bytes < 2 * kMin ? 1u : static_cast<unsigned>(std::min<size_t>(8, bytes / kMin));
  if (nt == 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t part = (bytes + nt - 1) / nt;
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) {
    const size_t lo = t * part, n = lo < bytes ? std::min(part, bytes - lo) : 0;
    if (n) th.emplace_back([=] { std::memcpy(dst + lo, src + lo, n); });
  }
  std::memcpy(dst, src, std::min(part, bytes));
  for (auto& x : th) x.join();
}

// Device to host through the two pinned halves: the device fills one while
// host threads move the other into the caller's buffer (pageable memory would
// otherwise be copied at a fraction of the link's rate). Work already queued
// on the context stream runs first.
void to_host(nuc34_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  if (bytes <= (size_t(1) << 20) || c->noalloc) {
    NUC34_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->s));
    NUC34_HIP(hipStreamSynchronize(c->s));
    return;
  }
  if (!c->stage) {
    NUC34_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->stage), 2 * kStageBytes, hipHostMallocDefault));
    for (hipEvent_t& e : c->stage_ev) NUC34_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  const size_t nchunks = (bytes + kStageBytes - 1) / kStageBytes;
  auto enqueue = [&](size_t i) {
    const size_t off = i * kStageBytes, n = std::min(kStageBytes, bytes - off);
    NUC34_HIP(hipMemcpyAsync(c->stage + (i & 1) * kStageBytes, static_cast<const char*>(src) + off, n,
                             hipMemcpyDeviceToHost, c->s));
    NUC34_HIP(hipEventRecord(c->stage_ev[i & 1], c->s));
  };
  enqueue(0);
  for (size_t i = 0; i < nchunks; ++i) {
    if (i + 1 < nchunks) enqueue(i + 1);  // into the half the host finished with last step
    NUC34_HIP(hipEventSynchronize(c->stage_ev[i & 1]));
    const size_t off = i * kStageBytes, n = std::min(kStageBytes, bytes - off);
    par_memcpy(static_cast<char*>(dst) + off, c->stage + (i & 1) * kStageBytes, n);
  }
}

// Host to device for one contiguous device range fed from several host
// segments (graphs' arrays back to back), through the same two pinned halves:
// host threads fill one half while the device copies the other. Returns with
// the copies queued on the context stream; a half is rewritten only after its
// previous copy has finished.
void to_device(nuc34_ctx* c, void* dst, const std::vector<std::pair<const void*, size_t>>& segs) {
  size_t total = 0;
  for (const auto& sg : segs) total += sg.second;
  if (!total) return;
  if (total <= (size_t(1) << 20) || c->noalloc) {
    size_t off = 0;
    for (const auto& sg : segs) {
      if (sg.second)
        NUC34_HIP(hipMemcpyAsync(static_cast<char*>(dst) + off, sg.first, sg.second, hipMemcpyHostToDevice, c->s));
      off += sg.second;
    }
    return;
  }
  if (!c->stage) {
    NUC34_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->stage), 2 * kStageBytes, hipHostMallocDefault));
    for (hipEvent_t& e : c->stage_ev) NUC34_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  unsigned half = 0;
  size_t fill = 0, off = 0;
  NUC34_HIP(hipEventSynchronize(c->stage_ev[half]));
  auto flush = [&] {
    NUC34_HIP(hipMemcpyAsync(static_cast<char*>(dst) + off, c->stage + half * kStageBytes, fill,
                             hipMemcpyHostToDevice, c->s));
    NUC34_HIP(hipEventRecord(c->stage_ev[half], c->s));
    off += fill;
    fill = 0;
    half ^= 1u;
    NUC34_HIP(hipEventSynchronize(c->stage_ev[half]));
  };
  for (const auto& sg : segs) {
    const char* p = static_cast<const char*>(sg.first);
    size_t left = sg.second;
    while (left) {
      const size_t n = std::min(left, kStageBytes - fill);
      par_memcpy(c->stage + half * kStageBytes + fill, p, n);
      fill += n;
      p += n;
      left -= n;
      if (fill == kStageBytes) flush();
    }
  }
  if (fill) flush();
}

enum KernelId { K_TRI_BITS, K_TRI_EMIT, K_SUPPORT, K_K4_LISTS, K_PEEL, K_COUNT };
const char* const kKernelNames[K_COUNT] = {"tri_enum_bits", "tri_enum_emit", "support_bits", "k4_lists",
                                          "peel_rounds"};

// Records events around `f` on the context stream; read back by ktime_collect.
template <typename F>
void ktime(nuc34_ctx* c, int id, uint64_t launches, F&& f) {
  if (c->noalloc) {
    f();
    return;
  }
  auto& k = c->kt[id];
  if (k.pending) {
    NUC34_HIP(hipEventSynchronize(k.b));
    float ms = 0;
    NUC34_HIP(hipEventElapsedTime(&ms, k.a, k.b));
    k.ms += ms;
    k.pending = false;
  }
  if (!k.a) {
    NUC34_HIP(hipEventCreate(&k.a));
    NUC34_HIP(hipEventCreate(&k.b));
  }
  k.name = kKernelNames[id];
  NUC34_HIP(hipEventRecord(k.a, c->s));
  f();
  NUC34_HIP(hipEventRecord(k.b, c->s));
  k.launches += launches;
  k.pending = true;
}

void ktime_reset(nuc34_ctx* c) {
  for (auto& k : c->kt) {
    k.launches = 0;
    k.ms = 0;
    k.pending = false;
  }
}

void ktime_collect(nuc34_ctx* c) {
  c->st.n_kernels = 0;
  for (int id = 0; id < K_COUNT; ++id) {
    auto& k = c->kt[id];
    if (k.pending) {
      NUC34_HIP(hipEventSynchronize(k.b));
      float ms = 0;
      NUC34_HIP(hipEventElapsedTime(&ms, k.a, k.b));
      k.ms += ms;
      k.pending = false;
    }
    if (!k.launches) continue;
    nuc34_kernel_stat& o = c->st.kernels[c->st.n_kernels++];
    std::memset(o.name, 0, sizeof(o.name));
    std::strncpy(o.name, kKernelNames[id], sizeof(o.name) - 1);
    o.ms = k.ms;
    o.launches = k.launches;
  }
}

template <typename T>
T read_scalar(nuc34_ctx* c, const void* dev) {
  T v{};
  NUC34_HIP(hipMemcpyAsync(&v, dev, sizeof(T), hipMemcpyDeviceToHost, c->s));
  NUC34_HIP(hipStreamSynchronize(c->s));
  return v;
}

void zero(nuc34_ctx* c, DevBuf& b) {
  if (b) NUC34_HIP(hipMemsetAsync(b.as<void>(), 0, b.bytes(), c->s));
}

void reset_stages(nuc34_ctx* c, int from) {
  // from: 0 graph, 1 triangles, 2 support, 3 k4 lists, 4 k
  if (from <= 4) {
    c->kv.release();
    c->has_k = false;
    c->batch_rounds.clear();
  }
  if (from <= 3) {
    for (DevBuf* b : {&c->koff, &c->lists}) b->release();
    c->has_k4 = false;
  }
  if (from <= 2) {
    for (DevBuf* b : {&c->sup, &c->own, &c->small}) b->release();
    c->has_sup = false;
    c->Q = 0;
  }
  if (from <= 1) {
    for (DevBuf* b : {&c->IO, &c->IE, &c->IEe, &c->inpos, &c->rowA, &c->rowC, &c->Aout, &c->Ain, &c->Cm, &c->offA,
                      &c->AoutR, &c->triA, &c->triBC, &c->canon, &c->cSlot, &c->ctriA, &c->ctriBC, &c->vbins})
      b->release();
    c->has_tris = false;
    c->T = 0;
  }
  if (from <= 0) {
    c->O.release();
    c->E.release();
    c->src.release();
    c->has_graph = false;
  }
}

// Edge ranges of `parts` vertex segments (make_split_parts on the oriented CSR).
std::vector<std::pair<uint64_t, uint64_t>> edge_parts(nuc34_ctx* c, uint32_t parts) {
  std::vector<std::pair<uint64_t, uint64_t>> r;
  if (parts <= 1) {
    r.emplace_back(0, c->m);
    return r;
  }
  Csr h;
  h.n = c->n;
  h.O.resize(static_cast<size_t>(c->n) + 1);
  NUC34_HIP(hipMemcpyAsync(h.O.data(), c->O.as<void>(), h.O.size() * 4, hipMemcpyDeviceToHost, c->s));
  NUC34_HIP(hipStreamSynchronize(c->s));
  const auto sp = split_parts(h, parts);
  for (size_t i = 0; i < sp.size(); ++i) {
    const uint64_t vb = sp[i];
    const uint64_t ve = i + 1 < sp.size() ? sp[i + 1] : c->n;
    if (ve <= vb) continue;
    r.emplace_back(h.O[vb], h.O[ve]);
  }
  return r;
}

uint32_t tri_tile(const nuc34_ctx* c) { return c->tune.tri_tile_edges ? c->tune.tri_tile_edges : 64u; }
// Degree-binned intersection: owners up to this out-degree stay on the edge tiles.
uint32_t tri_bins_min_degree(const nuc34_ctx* c) { return c->tune.tri_schedule == 2 ? 8u : 0u; }

BitRows bitrows(nuc34_ctx* c) {
  return BitRows{c->rowA.as<uint64_t>(), c->rowC.as<uint64_t>(), c->Aout.as<uint32_t>(), c->Ain.as<RankWord>(),
                 c->Cm.as<RankWord>(),       c->AoutR.as<uint32_t>()};
}

K4Src k4src(nuc34_ctx* c) {
  return K4Src{c->O.as<uint32_t>(),    c->E.as<uint32_t>(),     c->src.as<uint32_t>(), c->inpos.as<uint32_t>(),
               bitrows(c),             c->triA.as<uint32_t>(),  c->triBC.as<uint64_t>(), c->cSlot.as<uint2>(),
               c->T,                  nullptr,                 nullptr,
               nullptr};
}

void run_triangles(nuc34_ctx* c, uint32_t parts) {
  const uint64_t m = c->m;
  const uint32_t n = c->n;
  const auto ranges = edge_parts(c, parts);
  // enumeration: in-CSR, row layout, the bits pass
  const double ms_enum = timed(c, [&] {
    // in-CSR: the c-role rows are indexed by position in the source's in-list
    c->IO = c->alloc.get((static_cast<uint64_t>(n) + 1) * 4, "in-offsets");
    c->IE = c->alloc.get(m * 4 + 4, "in-lists");
    c->IEe = c->alloc.get(m * 4 + 4, "in-list edges");
    c->inpos = c->alloc.get(m * 4 + 4, "in-list positions");
    build_in_csr(c->O.as<uint32_t>(), c->E.as<uint32_t>(), c->src.as<uint32_t>(), n, m, c->IO.as<uint32_t>(),
                 c->IE.as<uint32_t>(), c->IEe.as<uint32_t>(), c->inpos.as<uint32_t>(), c->alloc, c->s);
    // row layout
    c->rowA = c->alloc.get((m + 1) * 8, "row offsets");
    c->rowC = c->alloc.get((m + 1) * 8, "row offsets");
    {
      DevBuf wa = c->alloc.get(m * 4 + 4, "row words");
      DevBuf wc = c->alloc.get(m * 4 + 4, "row words");
      launch_row_words(c->O.as<uint32_t>(), c->IO.as<uint32_t>(), c->src.as<uint32_t>(), m, wa.as<uint32_t>(),
                       wc.as<uint32_t>(), c->s);
      DevBuf scratch = c->alloc.get(scan_scratch_elems(m) * 8, "scan scratch");
      exclusive_scan<uint32_t, uint64_t>(wa.as<uint32_t>(), c->rowA.as<uint64_t>(), m, scratch.as<uint64_t>(), c->s);
      exclusive_scan<uint32_t, uint64_t>(wc.as<uint32_t>(), c->rowC.as<uint64_t>(), m, scratch.as<uint64_t>(), c->s);
    }
    const uint64_t WA = c->WA = read_scalar<uint64_t>(c, c->rowA.as<uint64_t>() + m);
    const uint64_t WC = c->WC = read_scalar<uint64_t>(c, c->rowC.as<uint64_t>() + m);
    c->Aout = c->alloc.get(WA * 4 + 4, "b-role rows");
    c->Ain = c->alloc.get(WA * 8 + 8, "a-role rows");
    c->Cm = c->alloc.get(WC * 8 + 8, "c-role rows");
    zero(c, c->Aout);
    zero(c, c->Ain);
    zero(c, c->Cm);
    const uint32_t min_deg = tri_bins_min_degree(c);
    if (c->tune.tri_schedule && n) {
      // degree-binned schedule: the vertices of each class, for both passes
      c->vbins = c->alloc.get(3 * static_cast<uint64_t>(n) * 4 + 16, "vertex classes");
      uint32_t* const counts = c->vbins.as<uint32_t>() + 3 * static_cast<uint64_t>(n);
      NUC34_HIP(hipMemsetAsync(counts, 0, 16, c->s));
      launch_vertex_bins(c->O.as<uint32_t>(), n, min_deg, c->vbins.as<uint32_t>(), counts, c->s);
    }
    const VertexBins bins{c->vbins.as<uint32_t>(), c->vbins.as<uint32_t>() + 3 * static_cast<uint64_t>(n), n,
                          min_deg};
    const VertexBins* pb = c->vbins ? &bins : nullptr;
    const BitRows rows = bitrows(c);
    ktime(c, K_TRI_BITS, ranges.size(), [&] {
      for (const auto& r : ranges)
        launch_tri_bits(c->O.as<uint32_t>(), c->E.as<uint32_t>(), c->src.as<uint32_t>(), r.first, r.second, rows,
                        c->inpos.as<uint32_t>(), tri_tile(c), pb, c->s);
    });
  });
  // ids: role counts and ranks, then the records at their peel ids
  c->st.ms_tri_order = timed(c, [&] {
    c->offA = c->alloc.get((m + 1) * 8, "triangle offsets");
    {
      // one exclusive scan of the set bits over all a-role words gives every
      // word its running rank (the triangle id of its first bit), and at each
      // row's first word the edge's first triangle id; likewise the c-role
      // words and the c-slots
      const uint64_t WA = c->WA, WC = c->WC;
      DevBuf pA = c->alloc.get((WA + 1) * 8, "a-role ranks");
      DevBuf pC = c->alloc.get((WC + 1) * 8, "c-role ranks");
      DevBuf scratch = c->alloc.get(scan_scratch_elems(std::max(WA, WC)) * 8, "scan scratch");
      exclusive_scan<RankWord, uint64_t>(c->Ain.as<RankWord>(), pA.as<uint64_t>(), WA, scratch.as<uint64_t>(), c->s);
      exclusive_scan<RankWord, uint64_t>(c->Cm.as<RankWord>(), pC.as<uint64_t>(), WC, scratch.as<uint64_t>(), c->s);
      c->T = read_scalar<uint64_t>(c, pA.as<uint64_t>() + WA);
      if (c->T >= 0xFFFFFFFFull)
        throw Overflow("triangle count " + std::to_string(c->T) + " exceeds the 32-bit triangle id range");
      if (read_scalar<uint64_t>(c, pC.as<uint64_t>() + WC) != c->T)
        throw std::logic_error("c-role bits disagree with the triangle count");
      launch_set_ranks(c->Ain.as<RankWord>(), pA.as<uint64_t>(), WA, c->s);
      launch_set_ranks(c->Cm.as<RankWord>(), pC.as<uint64_t>(), WC, c->s);
      launch_gather_u64(pA.as<uint64_t>(), c->rowA.as<uint64_t>(), c->offA.as<uint64_t>(), m + 1, c->s);
      // the peel ids: the same scan over the b-role words
      exclusive_scan<BitWord, uint64_t>(reinterpret_cast<const BitWord*>(c->Aout.as<uint32_t>()), pA.as<uint64_t>(),
                                        WA, scratch.as<uint64_t>(), c->s);
      if (read_scalar<uint64_t>(c, pA.as<uint64_t>() + WA) != c->T)
        throw std::logic_error("b-role bits disagree with the triangle count");
      c->AoutR = c->alloc.get(WA * 4 + 4, "b-role ranks");
      launch_set_ranks32(c->AoutR.as<uint32_t>(), pA.as<uint64_t>(), WA, c->s);
    }
    const BitRows rows = bitrows(c);
    c->triA = c->alloc.get(c->T * 4 + 4, "triangle a-edges");
    c->triBC = c->alloc.get(c->T * 8 + 8, "triangle b,c edges");
    c->canon = c->alloc.get(c->T * 4 + 4, "canonical triangle order");
    c->cSlot = c->alloc.get(c->T * 8 + 8, "triangles by c-slot");
    TriOut out;
    out.triA = c->triA.as<uint32_t>();
    out.triBC = c->triBC.as<uint64_t>();
    out.canon = c->canon.as<uint32_t>();
    out.cSlot = c->cSlot.as<uint2>();
    const VertexBins bins{c->vbins.as<uint32_t>(), c->vbins.as<uint32_t>() + 3 * static_cast<uint64_t>(n), n,
                          tri_bins_min_degree(c)};
    ktime(c, K_TRI_EMIT, ranges.size(), [&] {
      for (const auto& r : ranges)
        launch_tri_emit(c->O.as<uint32_t>(), c->E.as<uint32_t>(), c->src.as<uint32_t>(), r.first, r.second, rows,
                        c->inpos.as<uint32_t>(), out, tri_tile(c), c->vbins ? &bins : nullptr, c->s);
    });
    c->vbins.release();
  });
  c->st.ms_triangles = ms_enum + c->st.ms_tri_order;
  c->has_tris = true;
  c->st.triangles = c->T;
  c->st.incidence_entries = 3 * c->T;
}

void run_support(nuc34_ctx* c) {
  const uint64_t T = c->T;
  c->st.ms_support = timed(c, [&] {
    c->sup = c->alloc.get(T * 4 + 4, "support");
    c->own = c->alloc.get(T * 4 + 4, "owned K4 counts");
    c->small = c->alloc.get(64, "counters");
    zero(c, c->small);
    DevBuf ab = c->alloc.get(c->WA * 8 + 8, "support rows");
    DevBuf cb = c->alloc.get(c->WC * 4 + 4, "support rows");
    ktime(c, K_SUPPORT, 1, [&] {
      launch_support_bits(k4src(c), c->WA, c->WC, ab.as<uint2>(), cb.as<uint32_t>(), c->sup.as<int32_t>(),
                          c->own.as<uint32_t>(), reinterpret_cast<uint64_t*>(c->small.as<char>() + 8), c->s);
    });
    const uint64_t total = read_scalar<uint64_t>(c, c->small.as<char>() + 8);
    if (total % 4) throw std::logic_error("support sum is not a multiple of 4");
    c->Q = total / 4;
  });
  c->has_sup = true;
  c->st.four_cliques = c->Q;
}

// The peel rows: the Aout bits with their peel ranks beside a live plane (over
// Aout | Ain) and a dying plane, the c-role rows' planes, and the in-list cuts.
// o.v enumerates K4s from them (fresh planes: every K4) and names partners by
// peel id, the order the triangles are numbered in.
struct PeelRows {
  DevBuf liveA, liveC, icut;
  K4Src v{};
};

void build_peel_rows(nuc34_ctx* c, PeelRows& o) {
  o.v = k4src(c);
  o.liveA = c->alloc.get(c->WA * sizeof(PeelWord) + sizeof(PeelWord), "peel a/b rows");
  o.liveC = c->alloc.get(c->WC * sizeof(PlaneWord) + sizeof(PlaneWord), "peel c planes");
  o.icut = c->alloc.get((static_cast<uint64_t>(c->n) + 1) * 4, "in-list cuts");
  o.v.PA = o.liveA.as<PeelWord>();
  o.v.PC = o.liveC.as<PlaneWord>();
  o.v.icut = o.icut.as<uint32_t>();
  launch_live_rows(o.v, c->WA, c->WC, c->IO.as<uint32_t>(), c->IE.as<uint32_t>(), c->n, o.icut.as<uint32_t>(), c->s);
}

// The triangle arrays in canonical (a, b) order, for the callers that copy
// triangles out (built once per triangle listing).
void ensure_canonical_tris(nuc34_ctx* c) {
  if (c->ctriA || !c->T) return;
  c->ctriA = c->alloc.get(c->T * 4 + 4, "triangles in canonical order");
  c->ctriBC = c->alloc.get(c->T * 8 + 8, "triangles in canonical order");
  launch_canon_tris(c->triA.as<uint32_t>(), c->triBC.as<uint64_t>(), c->canon.as<uint32_t>(), c->T,
                    c->ctriA.as<uint32_t>(), c->ctriBC.as<uint64_t>(), c->s);
}

// Direct variant: every triangle's K4 list (three partners per K4, by peel id),
// from the peel rows with fresh planes.
void run_k4_lists(nuc34_ctx* c) {
  const uint64_t T = c->T, Q = c->Q;
  c->st.ms_k4_list = timed(c, [&] {
    PeelRows o;
    build_peel_rows(c, o);
    c->koff = c->alloc.get((T + 1) * 8, "K4 list offsets");
    {
      DevBuf scratch = c->alloc.get(scan_scratch_elems(T) * 8, "scan scratch");
      exclusive_scan<int32_t, uint64_t>(c->sup.as<int32_t>(), c->koff.as<uint64_t>(), T, scratch.as<uint64_t>(),
                                        c->s);
    }
    if (read_scalar<uint64_t>(c, c->koff.as<uint64_t>() + T) != 4 * Q)
      throw std::logic_error("K4 list sizes disagree with the support sum");
    c->lists = c->alloc.get(4 * Q * 12 + 12, "K4 partner lists");
    ktime(c, K_K4_LISTS, 1, [&] {
      launch_k4_lists(o.v, c->koff.as<uint64_t>(), c->lists.as<uint32_t>(), c->tune.k4_list_grid_cap, true, c->s);
    });
  });
  c->has_k4 = true;
}

// Rounds enqueued per host check of the done flag (captured once as a graph).
constexpr int kRoundsPerLaunch = 32;
// Most triangles of a batch's graph peeled by one workgroup (by default).
constexpr uint32_t kGraphPeelMaxTris = 1u << 18;

PeelShape peel_shape(const nuc34_ctx* c) {
  PeelShape sh;
  if (c->tune.peel_debit_grid) sh.debit_grid = c->tune.peel_debit_grid;
  if (c->tune.peel_tile_cap) sh.tile_cap = c->tune.peel_tile_cap;
  if (c->tune.peel_finish_grid) sh.finish_grid = c->tune.peel_finish_grid;
  return sh;
}

// The graphs' triangle slices of a loaded batch (count + 1 offsets, from offA at
// the graphs' first edges).
std::vector<uint64_t> batch_tri_first(nuc34_ctx* c) {
  const uint64_t k = c->batch_e.size();
  std::vector<uint64_t> tf(k);
  DevBuf idx = c->alloc.get(k * 8, "batch boundaries");
  DevBuf out = c->alloc.get(k * 8, "batch triangle offsets");
  NUC34_HIP(hipMemcpyAsync(idx.as<void>(), c->batch_e.data(), k * 8, hipMemcpyHostToDevice, c->s));
  launch_gather_u64(c->offA.as<uint64_t>(), idx.as<uint64_t>(), out.as<uint64_t>(), k, c->s);
  NUC34_HIP(hipMemcpyAsync(tf.data(), out.as<void>(), k * 8, hipMemcpyDeviceToHost, c->s));
  NUC34_HIP(hipStreamSynchronize(c->s));
  if (tf.front() != 0 || tf.back() != c->T) throw std::logic_error("batch triangle offsets inconsistent");
  for (uint64_t g = 0; g + 1 < k; ++g)
    if (tf[g + 1] < tf[g]) throw std::logic_error("batch triangle offsets not monotone");
  return tf;
}

// Every graph of the batch small enough for one workgroup's peel.
bool graphs_fit_one_workgroup(const nuc34_ctx* c, const std::vector<uint64_t>& tf) {
  const uint64_t most = c->tune.graph_peel_max_tris ? c->tune.graph_peel_max_tris : kGraphPeelMaxTris;
  for (size_t g = 0; g + 1 < tf.size(); ++g)
    if (tf[g + 1] - tf[g] > most) return false;
  return true;
}

void run_peel(nuc34_ctx* c, int variant) {
  const uint64_t T = c->T;
  c->kv = c->alloc.get(T * 4 + 4, "k values");
  c->has_k = true;
  c->st.rounds = 0;
  c->st.k_max = 0;
  c->st.unpeeled = 0;
  c->st.ms_peel = c->st.ms_peel_init = 0;
  if (!T) return;
  if (!c->Q) {
    zero(c, c->kv);
    NUC34_HIP(hipStreamSynchronize(c->s));
    return;
  }
  const uint64_t words = (T + 15) / 16;
  DevBuf dd = c->alloc.get(words * 64, "peel supports");
  DevBuf stw = c->alloc.get(words * 4, "peel states");
  DevBuf b0 = c->alloc.get(T * 4, "frontier");
  DevBuf b1 = c->alloc.get(T * 4, "frontier");
  DevBuf ctlb = c->alloc.get(2 * sizeof(PeelCtl), "peel control");
  PeelBufs pb{dd.as<int32_t>(),  stw.as<uint32_t>(), c->kv.as<int32_t>(),
              b0.as<uint32_t>(), b1.as<uint32_t>(),  ctlb.as<PeelCtl>(), T};
  // Both variants peel by peel id (supports, states, frontiers, k); k goes
  // back to the canonical order at the end. On demand the peel rows are built
  // here, fresh planes for every peel.
  K4Src v{};
  PeelRows o;
  DevBuf pkv = c->alloc.get(T * 4 + 4, "k values in peel order");
  pb.kv = pkv.as<int32_t>();
  // A loaded batch whose graphs are all small: one workgroup per graph, the
  // whole peel one launch (k_peel_graphs). The graphs' triangle slices come
  // from offA at the graphs' first edges; workgroups take the largest first.
  const uint64_t ngraphs = c->batch_e.empty() ? 0 : c->batch_e.size() - 1;
  std::vector<uint64_t> tfirst;
  DevBuf gtf, gord, gout;
  GraphQueue gq{};
  bool by_graph = ngraphs > 0 && c->tune.batch_peel != 1 && !c->noalloc;
  if (by_graph) {
    tfirst = batch_tri_first(c);
    by_graph = graphs_fit_one_workgroup(c, tfirst);
  }
  if (by_graph) {
    gtf = c->alloc.get((ngraphs + 1) * 8, "batch triangle offsets");
    NUC34_HIP(hipMemcpyAsync(gtf.as<void>(), tfirst.data(), (ngraphs + 1) * 8, hipMemcpyHostToDevice, c->s));
    std::vector<uint32_t> order(ngraphs);
    for (uint32_t g = 0; g < ngraphs; ++g) order[g] = g;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
      return tfirst[x + 1] - tfirst[x] > tfirst[y + 1] - tfirst[y];
    });
    gord = c->alloc.get(ngraphs * 4, "batch peel order");
    gout = c->alloc.get(ngraphs * 8 + 8, "batch peel rounds and levels");
    NUC34_HIP(hipMemcpyAsync(gord.as<void>(), order.data(), ngraphs * 4, hipMemcpyHostToDevice, c->s));
    NUC34_HIP(hipMemsetAsync(gout.as<void>(), 0, ngraphs * 8 + 8, c->s));
    gq.tri_first = gtf.as<uint64_t>();
    gq.order = gord.as<uint32_t>();
    gq.count = static_cast<uint32_t>(ngraphs);
    gq.rounds = gout.as<uint32_t>();
    gq.level = reinterpret_cast<int32_t*>(gout.as<uint32_t>() + ngraphs);
    gq.next = gout.as<uint32_t>() + 2 * ngraphs;
    gq.error = gq.next + 1;
  }
  // round parity must equal slot parity: groups of even size
  const int per_launch = c->tune.peel_rounds_per_launch ? static_cast<int>(c->tune.peel_rounds_per_launch & ~1u)
                                                        : kRoundsPerLaunch;
  const PeelShape shape = peel_shape(c);
  // push regions of the debit workgroups: room for twice their even share of
  // the triangles, 32 to 1024 slots each (a fuller workgroup overflows to the
  // shared list, which stays correct)
  DevBuf wslots, wcnt, woff, gcnt;
  if (!by_graph) {
    const uint64_t nwg = shape.debit_grid;
    const uint64_t share = (2 * T + nwg - 1) / nwg;
    const uint32_t cap = static_cast<uint32_t>(std::min<uint64_t>(1024, std::max<uint64_t>(32, (share + 31) & ~31ull)));
    wslots = c->alloc.get(nwg * cap * 4, "peel push regions");
    wcnt = c->alloc.get(nwg * 4, "peel push counts");
    woff = c->alloc.get(nwg * 4, "peel push offsets");
    gcnt = c->alloc.get(2 * kPushGroups * kGroupStride * 4, "peel push group counters");
    NUC34_HIP(hipMemsetAsync(gcnt.as<void>(), 0, gcnt.bytes(), c->s));
    pb.wslots = wslots.as<uint32_t>();
    pb.wcnt = wcnt.as<uint32_t>();
    pb.woff = woff.as<uint32_t>();
    pb.gcnt = gcnt.as<uint32_t>();
    pb.wcap = cap;
    pb.nwg = static_cast<uint32_t>(nwg);
    pb.gsize = static_cast<uint32_t>((nwg + kPushGroups - 1) / kPushGroups);
  }
  auto enqueue_rounds = [&](int n) {
    for (int r = 0; r < n; ++r) {
      if (variant == NUC34_DIRECT)
        launch_peel_round_direct(pb, r, c->koff.as<uint64_t>(), c->lists.as<uint32_t>(), shape, c->s);
      else
        launch_peel_round_ondemand(pb, r, v, shape, c->s);
    }
  };
  PeelCtl hs[2] = {};
  PeelCtl h{};
  c->st.ms_peel_init = timed(c, [&] {
    if (variant != NUC34_DIRECT) {
      build_peel_rows(c, o);
      v = o.v;
    }
    launch_peel_init(pb, c->sup.as<int32_t>(), c->s);
  });
  if (by_graph) {
    c->st.ms_peel = c->st.ms_peel_init + timed(c, [&] {
      ktime(c, K_PEEL, 1, [&] {
        launch_peel_graphs(pb, v, c->koff.as<uint64_t>(), c->lists.as<uint32_t>(), gq, shape, variant != NUC34_DIRECT,
                           c->s);
      });
    });
    std::vector<uint32_t> out(2 * ngraphs + 2);
    NUC34_HIP(hipMemcpyAsync(out.data(), gout.as<void>(), out.size() * 4, hipMemcpyDeviceToHost, c->s));
    NUC34_HIP(hipStreamSynchronize(c->s));
    if (out[2 * ngraphs + 1]) throw std::logic_error("peel: inconsistent state (code " + std::to_string(out[2 * ngraphs + 1]) + ")");
    c->batch_rounds.assign(out.begin(), out.begin() + ngraphs);
    uint32_t rmax = 0;
    int32_t lmax = 0;
    for (uint64_t g = 0; g < ngraphs; ++g) {
      rmax = std::max(rmax, out[g]);
      lmax = std::max(lmax, static_cast<int32_t>(out[ngraphs + g]));
    }
    h.round = rmax;
    h.level = lmax;
    h.peeled = T;
  } else {
  c->st.ms_peel = c->st.ms_peel_init + timed(c, [&] {
    struct Captured {  // destroyed on every way out, errors included
      hipStream_t s = nullptr;
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      ~Captured() {
        if (exec) {
          (void)hipStreamSynchronize(s);  // no launch of it may still be queued
          (void)hipGraphExecDestroy(exec);
        }
        if (graph) (void)hipGraphDestroy(graph);
      }
    } cap;
    cap.s = c->s;
    const uint64_t max_launches = 2 * T / per_launch + 16;  // every second round fixes >= 1 triangle
    if (c->noalloc) {
      // rounds launched eagerly (instantiating a graph allocates); both
      // control slots read back after every group
      PeelCtl got[2];
      for (uint64_t l = 0;; ++l) {
        enqueue_rounds(per_launch);
        NUC34_HIP(hipMemcpyAsync(got, ctlb.as<void>(), sizeof(got), hipMemcpyDeviceToHost, c->s));
        NUC34_HIP(hipStreamSynchronize(c->s));
        h = got[0].seq >= got[1].seq ? got[0] : got[1];
        if (h.done) break;
        if (l > max_launches) throw std::logic_error("peel did not finish");
      }
      return;
    }
    if (!std::getenv("NUC34_NO_GRAPH") && hipStreamBeginCapture(c->s, hipStreamCaptureModeRelaxed) == hipSuccess) {
      try {
        enqueue_rounds(per_launch);
      } catch (...) {  // leave the stream out of capture mode before reporting
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(c->s, &partial);
        if (partial) (void)hipGraphDestroy(partial);
        throw;
      }
      if (hipStreamEndCapture(c->s, &cap.graph) != hipSuccess ||
          hipGraphInstantiate(&cap.exec, cap.graph, nullptr, nullptr, 0) != hipSuccess) {
        cap.exec = nullptr;
        (void)hipGetLastError();
      }
    } else {
      (void)hipGetLastError();
    }
    const hipGraphExec_t exec = cap.exec;
    if (!c->ctl_host) {
      NUC34_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->ctl_host), 4 * sizeof(PeelCtl), hipHostMallocDefault));
      for (hipEvent_t& e : c->poll_ev) NUC34_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    // One launch stays queued behind the one being checked, so the device never
    // waits for the host between launches; rounds after the last are no-ops.
    // Each launch is followed by a copy of both control slots into its own
    // pinned pair and an event the host waits on.
    uint64_t launches = 0;
    auto launch = [&] {
      if (exec) NUC34_HIP(hipGraphLaunch(exec, c->s));
      else enqueue_rounds(per_launch);
      PeelCtl* dst = c->ctl_host + 2 * (launches & 1);
      NUC34_HIP(hipMemcpyAsync(dst, ctlb.as<void>(), 2 * sizeof(PeelCtl), hipMemcpyDeviceToHost, c->s));
      NUC34_HIP(hipEventRecord(c->poll_ev[launches & 1], c->s));
      ++launches;
    };
    ktime(c, K_PEEL, 0, [&] {
      launch();
      for (uint64_t l = 0;; ++l) {
        launch();
        NUC34_HIP(hipEventSynchronize(c->poll_ev[l & 1]));
        const PeelCtl* got = c->ctl_host + 2 * (l & 1);
        hs[0] = got[0];
        hs[1] = got[1];
        // the last round written (the higher sequence number) holds the final
        // counts; the round that finishes marks both slots done
        h = hs[0].seq >= hs[1].seq ? hs[0] : hs[1];
        if (h.done) break;
        if (l > max_launches) throw std::logic_error("peel did not finish");
      }
      // the launch still queued must end before its graph is destroyed
      NUC34_HIP(hipStreamSynchronize(c->s));
    });
    c->kt[K_PEEL].launches += launches * static_cast<uint64_t>(per_launch);
  });
  }
  if (h.error) throw std::logic_error("peel: inconsistent state (code " + std::to_string(h.error) + ")");
  c->st.ms_peel += timed(c, [&] {
    launch_canon_u32(pkv.as<uint32_t>(), c->canon.as<uint32_t>(), T, c->kv.as<uint32_t>(), c->s);
  });
  c->st.rounds = h.round;
  c->st.k_max = h.level;
  c->st.unpeeled = T - h.peeled;
  {
    // every debit lowers one residual support by one
    DevBuf rsum = c->alloc.get(8, "residual support sum");
    NUC34_HIP(hipMemsetAsync(rsum.as<void>(), 0, 8, c->s));
    launch_residual_sum(pb, rsum.as<int64_t>(), c->s);
    c->st.debits = static_cast<uint64_t>(4 * static_cast<int64_t>(c->st.four_cliques) - read_scalar<int64_t>(c, rsum.as<void>()));
  }
}

// Host-side checks that keep every device index in bounds: monotone offsets,
// ids < n, strictly increasing lists without self-loops, and for oriented
// input an acyclic orientation (Kahn's algorithm).
std::string validate_csr(const uint32_t* O, const uint32_t* E, uint32_t n, bool oriented) {
  for (uint32_t u = 0; u < n; ++u) {
    if (O[u] > O[u + 1]) return "offsets not monotone at vertex " + std::to_string(u);
    for (uint32_t i = O[u]; i < O[u + 1]; ++i) {
      if (E[i] >= n) return "vertex id " + std::to_string(E[i]) + " out of range";
      if (E[i] == u) return "self-loop at vertex " + std::to_string(u);
      if (i > O[u] && E[i - 1] >= E[i]) return "adjacency of vertex " + std::to_string(u) + " not strictly increasing";
    }
  }
  if (oriented) {
    std::vector<uint32_t> indeg(n, 0), queue;
    for (uint64_t i = 0; i < O[n]; ++i) indeg[E[i]]++;
    queue.reserve(n);
    for (uint32_t u = 0; u < n; ++u)
      if (!indeg[u]) queue.push_back(u);
    for (size_t h = 0; h < queue.size(); ++h) {
      const uint32_t u = queue[h];
      for (uint32_t i = O[u]; i < O[u + 1]; ++i)
        if (--indeg[E[i]] == 0) queue.push_back(E[i]);
    }
    if (queue.size() != n) return "oriented graph has a directed cycle";
  }
  return std::string();
}

// The message of this thread's last failed nuc34_ctx_create (there is no
// context to hold it).
thread_local std::string tls_create_error;

}  // namespace

extern "C" {

int nuc34_graph_load(const char* path, nuc34_graph** out, char* err, size_t errlen) {
  try {
    auto* g = new nuc34_graph;
    try {
      g->lg = load_dataset(path);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
    return NUC34_OK;
  } catch (const std::exception& e) {
    if (err && errlen) {
      std::strncpy(err, e.what(), errlen - 1);
      err[errlen - 1] = 0;
    }
    return NUC34_E_IO;
  }
}

int nuc34_graph_from_pairs(const uint32_t* src, const uint32_t* dst, uint64_t count, nuc34_graph** out) {
  try {
    auto g = std::make_unique<nuc34_graph>();
    g->lg.g = csr_from_pairs(src, dst, count);
    g->lg.lines = count;
    *out = g.release();
    return NUC34_OK;
  } catch (const std::exception&) {
    return NUC34_E_INVALID;
  }
}

int nuc34_graph_info(const nuc34_graph* g, uint32_t* n, uint64_t* m, int* oriented) {
  if (!g) return NUC34_E_INVALID;
  if (n) *n = g->lg.g.n;
  if (m) *m = g->lg.g.m();
  if (oriented) *oriented = g->lg.oriented ? 1 : 0;
  return NUC34_OK;
}

int nuc34_graph_copy(const nuc34_graph* g, uint32_t* O, uint32_t* E) {
  if (!g) return NUC34_E_INVALID;
  std::memcpy(O, g->lg.g.O.data(), g->lg.g.O.size() * 4);
  if (!g->lg.g.E.empty()) std::memcpy(E, g->lg.g.E.data(), g->lg.g.E.size() * 4);
  return NUC34_OK;
}

int nuc34_graph_orient_host(nuc34_graph* g) {
  if (!g) return NUC34_E_INVALID;
  if (!g->lg.oriented) {
    g->lg.g = orient_host(g->lg.g);
    g->lg.oriented = true;
  }
  return NUC34_OK;
}

void nuc34_graph_free(nuc34_graph* g) { delete g; }

namespace {
int graph_from_generated(const gen::Edges& e, nuc34_graph** out) {
  std::vector<uint32_t> s(e.size()), d(e.size());
  for (size_t i = 0; i < e.size(); ++i) {
    s[i] = e[i].first;
    d[i] = e[i].second;
  }
  auto g = std::make_unique<nuc34_graph>();
  g->lg.g = csr_from_pairs(s.data(), d.data(), s.size());
  g->lg.lines = e.size();
  *out = g.release();
  return NUC34_OK;
}
}  // namespace

int nuc34_gen_preset(const char* name, nuc34_graph** out) {
  try {
    return graph_from_generated(preset_edges(name), out);
  } catch (const std::exception&) {
    return NUC34_E_INVALID;
  }
}

int nuc34_gen_batch_graph(uint64_t seed, uint64_t index, nuc34_graph** out) {
  try {
    return graph_from_generated(gen::batch_graph(seed, index, kBatchNLo, kBatchNHi), out);
  } catch (const std::exception&) {
    return NUC34_E_INVALID;
  }
}

int nuc34_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                   uint32_t planted_cliques, uint32_t planted_size, nuc34_graph** out) {
  try {
    auto e = gen::rmat(scale, edge_factor, a, b, c, seed);
    if (planted_cliques) add_planted(e, scale, seed, planted_cliques, planted_size);
    return graph_from_generated(e, out);
  } catch (const std::exception&) {
    return NUC34_E_INVALID;
  }
}


int nuc34_ctx_create(int device, uint64_t hbm_budget, nuc34_ctx** out) {
  if (!out) return NUC34_E_INVALID;
  auto* c = new nuc34_ctx;
  c->device = device;
  c->alloc.budget = hbm_budget;
  const int rc = guard(c, [&]() -> int {
    NUC34_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    NUC34_HIP(hipEventCreate(&c->ev0));
    NUC34_HIP(hipEventCreate(&c->ev1));
    return NUC34_OK;
  });
  if (rc != NUC34_OK) {
    tls_create_error = c->err;  // nuc34_last_error(NULL) reports it
    nuc34_ctx_destroy(c);       // releases whatever was created
    *out = nullptr;
    return rc;
  }
  *out = c;
  return NUC34_OK;
}

void nuc34_ctx_destroy(nuc34_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->s) (void)hipStreamSynchronize(c->s);
  reset_stages(c, 0);
  c->alloc.trim();
  for (auto& k : c->kt) {
    if (k.a) (void)hipEventDestroy(k.a);
    if (k.b) (void)hipEventDestroy(k.b);
  }
  for (hipEvent_t& e : c->poll_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t& e : c->stage_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stage) (void)hipHostFree(c->stage);
  if (c->ctl_host) (void)hipHostFree(c->ctl_host);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
}

const char* nuc34_last_error(const nuc34_ctx* c) {
  if (c) return c->err.c_str();
  return tls_create_error.empty() ? "null context" : tls_create_error.c_str();
}

}  // extern "C"

namespace {

// Installs an uploaded CSR (device buffers, already checked) as the context's
// graph: used as given when oriented, else oriented on the device.
int install_graph(nuc34_ctx* c, DevBuf& Ou, DevBuf& Eu, uint32_t n, uint64_t m, bool oriented) {
  c->n = n;
  if (oriented) {
    c->m = m;
    c->O = std::move(Ou);
    c->E = std::move(Eu);
    c->src = c->alloc.get(m * 4 + 4, "edge sources");
    launch_fill_src(c->O.as<uint32_t>(), n, c->src.as<uint32_t>(), c->s);
    NUC34_HIP(hipStreamSynchronize(c->s));
    c->st.ms_orient = 0;
  } else {
    if (m % 2) throw std::invalid_argument("undirected CSR must hold both directions of every edge");
    c->st.ms_orient = timed(c, [&] {
      DevBuf deg = c->alloc.get((static_cast<uint64_t>(n) + 1) * 4, "out-degrees");
      launch_orient_count(Ou.as<uint32_t>(), Eu.as<uint32_t>(), n, deg.as<uint32_t>(), c->s);
      c->O = c->alloc.get((static_cast<uint64_t>(n) + 1) * 4, "oriented offsets");
      DevBuf scratch = c->alloc.get(scan_scratch_elems(n) * 4, "scan scratch");
      exclusive_scan<uint32_t, uint32_t>(deg.as<uint32_t>(), c->O.as<uint32_t>(), n, scratch.as<uint32_t>(), c->s);
      const uint32_t mo = read_scalar<uint32_t>(c, c->O.as<uint32_t>() + n);
      if (mo != m / 2) throw std::invalid_argument("undirected CSR is not symmetric");
      c->m = mo;
      c->E = c->alloc.get(c->m * 4 + 4, "oriented edges");
      c->src = c->alloc.get(c->m * 4 + 4, "edge sources");
      launch_orient_fill(Ou.as<uint32_t>(), Eu.as<uint32_t>(), n, c->O.as<uint32_t>(), c->E.as<uint32_t>(),
                         c->src.as<uint32_t>(), c->s);
    });
  }
  c->has_graph = true;
  c->st.n = n;
  c->st.m = c->m;
  return NUC34_OK;
}

// Reads the device check flag; kNoVertex when the input passed.
constexpr uint32_t kNoVertex = 0xFFFFFFFFu;

}  // namespace

extern "C" {

int nuc34_load_csr(nuc34_ctx* c, const uint32_t* O, const uint32_t* E, uint32_t n, uint64_t m, int oriented) {
  if (c) {
    c->batch_v.clear();
    c->batch_e.clear();
  }
  if (!c || !O || (m && !E)) return fail(c, NUC34_E_INVALID, "null argument");
  if (O[0] != 0 || O[n] != m) return fail(c, NUC34_E_INVALID, "offsets do not span the edge array");
  if (m >= 0xFFFFFFFFull) return fail(c, NUC34_E_OVERFLOW, "more than 2^32-1 edges");
  if (oriented) {  // the acyclicity check runs on the host
    const std::string bad = validate_csr(O, E, n, true);
    if (!bad.empty()) return fail(c, NUC34_E_INVALID, bad);
  }
  return guard(c, [&]() -> int {
    reset_stages(c, 0);
    c->st = nuc34_stats{};
    DevBuf Ou = c->alloc.get((static_cast<uint64_t>(n) + 1) * 4, "input offsets");
    DevBuf Eu = c->alloc.get(m * 4 + 4, "input edges");
    to_device(c, Ou.as<void>(), {{O, (static_cast<uint64_t>(n) + 1) * 4}});
    if (m) to_device(c, Eu.as<void>(), {{E, m * 4}});
    if (!oriented) {
      DevBuf bad = c->alloc.get(4, "check flag");
      launch_fill_u32(bad.as<uint32_t>(), 1, kNoVertex, c->s);
      launch_check_csr(Ou.as<uint32_t>(), Eu.as<uint32_t>(), n, m, bad.as<uint32_t>(), c->s);
      const uint32_t v = read_scalar<uint32_t>(c, bad.as<uint32_t>());
      if (v != kNoVertex)
        return fail(c, NUC34_E_INVALID, "vertex " + std::to_string(v) +
                                            ": offsets out of order or range, id out of range, self-loop, or list "
                                            "not strictly increasing");
    }
    return install_graph(c, Ou, Eu, n, m, oriented != 0);
  });
}

}  // extern "C"

namespace {

// The graphs of a batch: back to back in two arrays, or one pointer pair per graph.
struct BatchSrc {
  const uint32_t* O = nullptr;
  const uint32_t* E = nullptr;
  const uint32_t* const* Op = nullptr;
  const uint32_t* const* Ep = nullptr;
  const uint32_t* og(uint32_t g, const std::vector<uint64_t>& vfirst) const { return Op ? Op[g] : O + vfirst[g] + g; }
  const uint32_t* eg(uint32_t g, const std::vector<uint64_t>& efirst) const { return Ep ? Ep[g] : E + efirst[g]; }
};

int load_batch(nuc34_ctx* c, uint32_t count, const uint32_t* n, const uint64_t* m, const BatchSrc& src, int oriented) {
  std::vector<uint64_t> vfirst(count + 1, 0), efirst(count + 1, 0);
  for (uint32_t g = 0; g < count; ++g) {
    vfirst[g + 1] = vfirst[g] + n[g];
    efirst[g + 1] = efirst[g] + m[g];
  }
  if (src.Op)
    for (uint32_t g = 0; g < count; ++g)
      if (!src.Op[g] || (m[g] && !src.Ep[g])) return fail(c, NUC34_E_INVALID, "null graph " + std::to_string(g));
  if (count == 1) {  // the union of one graph is the graph
    const uint32_t* O0 = src.og(0, vfirst);
    const int rc = nuc34_load_csr(c, O0, m[0] ? src.eg(0, efirst) : O0, n[0], m[0], oriented);
    if (rc != NUC34_OK) return rc;
    c->batch_v = {0, n[0]};
    c->batch_e = {0, c->m};
    return NUC34_OK;
  }
  const uint64_t nv = vfirst[count], ne = efirst[count];
  if (nv >= 0xFFFFFFFFull) return fail(c, NUC34_E_OVERFLOW, "batch has more than 2^32-1 vertices");
  if (ne >= 0xFFFFFFFFull) return fail(c, NUC34_E_OVERFLOW, "batch has more than 2^32-1 edges");
  for (uint32_t g = 0; g < count; ++g) {
    const uint32_t* Og = src.og(g, vfirst);
    if (Og[0] != 0 || Og[n[g]] != m[g])
      return fail(c, NUC34_E_INVALID, "graph " + std::to_string(g) + ": offsets do not span its edges");
  }
  if (oriented) {  // rare: build the union on the host, where the acyclicity check runs
    std::vector<uint32_t> Ou(nv + 1), Eu(ne);
    for (uint32_t g = 0; g < count; ++g) {
      const uint32_t* Og = src.og(g, vfirst);
      const uint32_t* Eg = src.eg(g, efirst);
      for (uint32_t v = 0; v < n[g]; ++v) Ou[vfirst[g] + v] = static_cast<uint32_t>(efirst[g] + Og[v]);
      for (uint64_t i = 0; i < m[g]; ++i) Eu[efirst[g] + i] = Eg[i] + static_cast<uint32_t>(vfirst[g]);
    }
    Ou[nv] = static_cast<uint32_t>(ne);
    const int rc = nuc34_load_csr(c, Ou.data(), Eu.empty() ? Ou.data() : Eu.data(), static_cast<uint32_t>(nv), ne, 1);
    if (rc != NUC34_OK) return rc;
  } else {
    const int rc = guard(c, [&]() -> int {
      c->batch_v.clear();
      c->batch_e.clear();
      reset_stages(c, 0);
      c->st = nuc34_stats{};
      DevBuf Oin = c->alloc.get((nv + count) * 4, "batch offsets");
      DevBuf Ein = c->alloc.get(ne * 4 + 4, "batch edges");
      DevBuf firsts = c->alloc.get((count + 1) * 16, "batch boundaries");
      // the graphs' arrays land back to back in Oin and Ein, each taken from
      // where it lies (no host-side concatenation)
      std::vector<std::pair<const void*, size_t>> oseg(count), eseg(count);
      for (uint32_t g = 0; g < count; ++g) {
        oseg[g] = {src.og(g, vfirst), (n[g] + 1ull) * 4};
        eseg[g] = {m[g] ? src.eg(g, efirst) : nullptr, m[g] * 4};
      }
      to_device(c, Oin.as<void>(), oseg);
      to_device(c, Ein.as<void>(), eseg);
      NUC34_HIP(hipMemcpyAsync(firsts.as<void>(), vfirst.data(), (count + 1) * 8, hipMemcpyHostToDevice, c->s));
      NUC34_HIP(hipMemcpyAsync(firsts.as<uint64_t>() + count + 1, efirst.data(), (count + 1) * 8,
                               hipMemcpyHostToDevice, c->s));
      DevBuf Ou = c->alloc.get((nv + 1) * 4, "input offsets");
      DevBuf Eu = c->alloc.get(ne * 4 + 4, "input edges");
      DevBuf bad = c->alloc.get(4, "check flag");
      launch_fill_u32(bad.as<uint32_t>(), 1, kNoVertex, c->s);
      launch_batch_union(Oin.as<uint32_t>(), Ein.as<uint32_t>(), firsts.as<uint64_t>(),
                         firsts.as<uint64_t>() + count + 1, count, nv, Ou.as<uint32_t>(), Eu.as<uint32_t>(),
                         bad.as<uint32_t>(), c->s);
      const uint32_t v = read_scalar<uint32_t>(c, bad.as<uint32_t>());
      if (v != kNoVertex) {
        const uint32_t g = static_cast<uint32_t>(std::upper_bound(vfirst.begin(), vfirst.end(), v) - vfirst.begin()) - 1;
        return fail(c, NUC34_E_INVALID, "graph " + std::to_string(g) + ", vertex " + std::to_string(v - vfirst[g]) +
                                            ": offsets out of order or range, id out of range, self-loop, or list "
                                            "not strictly increasing");
      }
      Oin.release();
      Ein.release();
      return install_graph(c, Ou, Eu, static_cast<uint32_t>(nv), ne, false);
    });
    if (rc != NUC34_OK) return rc;
  }
  return guard(c, [&]() -> int {
    // oriented edge boundaries: the union's oriented offsets at each graph's first vertex
    DevBuf idx = c->alloc.get((count + 1) * 8, "batch boundaries");
    DevBuf out = c->alloc.get((count + 1) * 8, "batch edge boundaries");
    NUC34_HIP(hipMemcpyAsync(idx.as<void>(), vfirst.data(), (count + 1) * 8, hipMemcpyHostToDevice, c->s));
    launch_gather_u32(c->O.as<uint32_t>(), idx.as<uint64_t>(), out.as<uint64_t>(), count + 1, c->s);
    c->batch_e.resize(count + 1);
    NUC34_HIP(hipMemcpyAsync(c->batch_e.data(), out.as<void>(), (count + 1) * 8, hipMemcpyDeviceToHost, c->s));
    NUC34_HIP(hipStreamSynchronize(c->s));
    c->batch_v = vfirst;
    return NUC34_OK;
  });
}

}  // namespace

extern "C" {

int nuc34_load_batch(nuc34_ctx* c, uint32_t count, const uint32_t* n, const uint64_t* m, const uint32_t* O,
                     const uint32_t* E, int oriented) {
  if (!c || !n || !m || !O || (count && !E)) return fail(c, NUC34_E_INVALID, "null argument");
  BatchSrc src;
  src.O = O;
  src.E = E;
  return load_batch(c, count, n, m, src, oriented);
}

int nuc34_load_batch_ptrs(nuc34_ctx* c, uint32_t count, const uint32_t* n, const uint64_t* m,
                          const uint32_t* const* O, const uint32_t* const* E, int oriented) {
  if (!c || !n || !m || !O || (count && !E)) return fail(c, NUC34_E_INVALID, "null argument");
  BatchSrc src;
  src.Op = O;
  src.Ep = E;
  return load_batch(c, count, n, m, src, oriented);
}

int nuc34_batch_graph(nuc34_ctx* c, uint32_t g, uint64_t* vertex_first, uint64_t* edge_first, uint64_t* tri_first,
                      uint64_t* tri_count) {
  if (!c || c->batch_v.empty()) return fail(c, NUC34_E_STATE, "no batch loaded");
  if (static_cast<uint64_t>(g) + 1 >= c->batch_v.size()) return fail(c, NUC34_E_INVALID, "graph index out of range");
  if (vertex_first) *vertex_first = c->batch_v[g];
  if (edge_first) *edge_first = c->batch_e[g];
  if (!tri_first && !tri_count) return NUC34_OK;
  if (!c->has_tris) return fail(c, NUC34_E_STATE, "triangles not computed");
  return guard(c, [&]() -> int {
    uint64_t lo = 0, hi = 0;
    NUC34_HIP(hipMemcpyAsync(&lo, c->offA.as<uint64_t>() + c->batch_e[g], 8, hipMemcpyDeviceToHost, c->s));
    NUC34_HIP(hipMemcpyAsync(&hi, c->offA.as<uint64_t>() + c->batch_e[g + 1], 8, hipMemcpyDeviceToHost, c->s));
    NUC34_HIP(hipStreamSynchronize(c->s));
    if (tri_first) *tri_first = lo;
    if (tri_count) *tri_count = hi - lo;
    return NUC34_OK;
  });
}

int nuc34_batch_slices(nuc34_ctx* c, uint64_t* edge_first, uint64_t* tri_first) {
  if (!c || c->batch_v.empty()) return fail(c, NUC34_E_STATE, "no batch loaded");
  const uint64_t k = c->batch_e.size();  // graphs + 1
  if (edge_first) std::copy(c->batch_e.begin(), c->batch_e.end(), edge_first);
  if (!tri_first) return NUC34_OK;
  if (!c->has_tris) return fail(c, NUC34_E_STATE, "triangles not computed");
  return guard(c, [&]() -> int {
    for (uint64_t g = 0; g < k; ++g)
      if (c->batch_e[g] > c->m) throw std::logic_error("batch edge boundary out of range");
    DevBuf idx = c->alloc.get(k * 8, "batch boundaries");
    DevBuf out = c->alloc.get(k * 8, "batch triangle offsets");
    NUC34_HIP(hipMemcpyAsync(idx.as<void>(), c->batch_e.data(), k * 8, hipMemcpyHostToDevice, c->s));
    launch_gather_u64(c->offA.as<uint64_t>(), idx.as<uint64_t>(), out.as<uint64_t>(), k, c->s);
    NUC34_HIP(hipMemcpyAsync(tri_first, out.as<void>(), k * 8, hipMemcpyDeviceToHost, c->s));
    NUC34_HIP(hipStreamSynchronize(c->s));
    return NUC34_OK;
  });
}

int nuc34_batch_rounds(nuc34_ctx* c, uint32_t* rounds) {
  if (!c || !rounds) return fail(c, NUC34_E_INVALID, "null argument");
  if (c->batch_v.empty()) return fail(c, NUC34_E_STATE, "no batch loaded");
  if (!c->has_k || c->batch_rounds.size() + 1 != c->batch_e.size())
    return fail(c, NUC34_E_STATE, "the last peel did not run one workgroup per graph");
  std::copy(c->batch_rounds.begin(), c->batch_rounds.end(), rounds);
  return NUC34_OK;
}

int nuc34_graph_dims(const nuc34_ctx* c, uint32_t* n, uint64_t* m) {
  if (!c || !c->has_graph) return NUC34_E_STATE;
  if (n) *n = c->n;
  if (m) *m = c->m;
  return NUC34_OK;
}

int nuc34_copy_oriented(nuc34_ctx* c, uint32_t* O, uint32_t* E) {
  if (!c || !c->has_graph) return fail(c, NUC34_E_STATE, "no graph loaded");
  return guard(c, [&]() -> int {
    NUC34_HIP(hipMemcpyAsync(O, c->O.as<void>(), (static_cast<uint64_t>(c->n) + 1) * 4, hipMemcpyDeviceToHost, c->s));
    if (c->m) NUC34_HIP(hipMemcpyAsync(E, c->E.as<void>(), c->m * 4, hipMemcpyDeviceToHost, c->s));
    NUC34_HIP(hipStreamSynchronize(c->s));
    return NUC34_OK;
  });
}

int nuc34_triangles(nuc34_ctx* c, uint32_t parts, uint64_t* T) {
  if (!c || !c->has_graph) return fail(c, NUC34_E_STATE, "no graph loaded");
  return guard(c, [&]() -> int {
    if (!c->has_tris) {
      reset_stages(c, 1);
      run_triangles(c, parts);
    }
    if (T) *T = c->T;
    return NUC34_OK;
  });
}

int nuc34_copy_triangles(nuc34_ctx* c, uint64_t first, uint64_t count, uint32_t* abc) {
  if (!c || !c->has_tris) return fail(c, NUC34_E_STATE, "triangles not computed");
  if (first > c->T || count > c->T - first) return fail(c, NUC34_E_INVALID, "triangle range out of bounds");
  return guard(c, [&]() -> int {
    if (!count) return NUC34_OK;
    ensure_canonical_tris(c);
    // formatted as (a, b, c) triples on the device, copied back in chunks
    const uint64_t chunk = std::min<uint64_t>(count, 16ull << 20);
    DevBuf buf = c->alloc.get(chunk * 12, "triangle copy");
    for (uint64_t done = 0; done < count; done += chunk) {
      const uint64_t n = std::min(chunk, count - done);
      launch_tri_abc(c->ctriA.as<uint32_t>() + first + done, c->ctriBC.as<uint64_t>() + first + done, n,
                     buf.as<uint32_t>(), c->s);
      to_host(c, abc + 3 * done, buf.as<void>(), n * 12);  // the buffer is reused only after this returns
    }
    return NUC34_OK;
  });
}

int nuc34_copy_edge_triangles(nuc34_ctx* c, uint32_t* counts) {
  if (!c || !c->has_tris) return fail(c, NUC34_E_STATE, "triangles not computed");
  if (!counts) return fail(c, NUC34_E_INVALID, "null argument");
  return guard(c, [&]() -> int {
    if (!c->m) return NUC34_OK;
    DevBuf cnt = c->alloc.get(c->m * 4, "edge triangle counts");
    launch_edge_triangles(bitrows(c), c->m, cnt.as<uint32_t>(), c->s);
    to_host(c, counts, cnt.as<void>(), c->m * 4);
    return NUC34_OK;
  });
}

int nuc34_support(nuc34_ctx* c, uint64_t* Q) {
  if (!c || !c->has_graph) return fail(c, NUC34_E_STATE, "no graph loaded");
  return guard(c, [&]() -> int {
    if (!c->has_tris) run_triangles(c, 1);
    if (!c->has_sup) {
      reset_stages(c, 2);
      if (c->T) run_support(c);
      else c->has_sup = true;
    }
    if (Q) *Q = c->Q;
    return NUC34_OK;
  });
}

int nuc34_copy_support(nuc34_ctx* c, int32_t* support) {
  if (!c || !c->has_sup) return fail(c, NUC34_E_STATE, "support not computed");
  return guard(c, [&]() -> int {
    if (!c->T) return NUC34_OK;
    DevBuf cs = c->alloc.get(c->T * 4, "support in canonical order");
    launch_canon_u32(c->sup.as<uint32_t>(), c->canon.as<uint32_t>(), c->T, cs.as<uint32_t>(), c->s);
    to_host(c, support, cs.as<void>(), c->T * 4);
    return NUC34_OK;
  });
}

int nuc34_decompose(nuc34_ctx* c, int variant, uint32_t parts, nuc34_stats* stats) {
  if (!c || !c->has_graph) return fail(c, NUC34_E_STATE, "no graph loaded");
  if (variant != NUC34_DIRECT && variant != NUC34_ONDEMAND && variant != NUC34_AUTO && variant != NUC34_AUTO_FIT)
    return fail(c, NUC34_E_INVALID, "unknown variant");
  return guard(c, [&]() -> int {
    if (!c->has_tris) {
      ktime_reset(c);
      reset_stages(c, 1);
      run_triangles(c, parts ? parts : 1);
    }
    if (!c->has_sup) {
      reset_stages(c, 2);
      if (c->T) run_support(c);
      else c->has_sup = true;
    }
    // AUTO for one graph: the on-demand peel enumerates each dying triangle's
    // K4s from the ranked rows at O(1) per K4 and skips K4s destroyed in
    // earlier rounds, so it needs no 48 bytes per K4 and ran faster than the
    // direct variant on every measured single graph (10^4 to 10^10 K4s). AUTO
    // for a batch of small graphs (each within one workgroup's peel) is
    // AUTO_FIT: their lists are cheap to build, and the direct peel beat the
    // on-demand one (512-graph mix: 22.6 against 26.1 ms). A batch of large
    // graphs stays on demand (two R-MAT s18: 1174 ms direct, against 362 ms
    // per graph on demand for four).
    int v = variant == NUC34_AUTO ? NUC34_ONDEMAND : variant;
    const bool small_batch =
        variant == NUC34_AUTO && c->batch_e.size() > 2 && graphs_fit_one_workgroup(c, batch_tri_first(c));
    if (variant == NUC34_AUTO_FIT || small_batch) {
      // the direct variant's extra memory: the lists (48 bytes per K4) and
      // their offsets, plus the peel's O(T) state both variants hold
      const uint64_t need = 48 * c->Q + 8 * (c->T + 1) + 20 * c->T;
      uint64_t room = 0;
      if (c->alloc.budget) {
        room = c->alloc.budget > c->alloc.used ? c->alloc.budget - c->alloc.used : 0;
      } else {
        size_t free_b = 0, total_b = 0;
        NUC34_HIP(hipMemGetInfo(&free_b, &total_b));
        room = free_b + c->alloc.cached;
      }
      v = need + need / 4 <= room ? NUC34_DIRECT : NUC34_ONDEMAND;
    }
    reset_stages(c, 4);
    // the on-demand variant needs nothing beyond the bit rows: each dying
    // triangle's K4s are enumerated from them in the peel
    c->st.ms_incidence = 0;
    if (v != NUC34_DIRECT) c->st.ms_k4_list = 0;
    if (c->Q && v == NUC34_DIRECT && !c->has_k4) run_k4_lists(c);
    run_peel(c, v);
    c->st.variant = v;
    c->st.peak_bytes = c->alloc.peak;
    ktime_collect(c);
    c->st.ms_total = c->st.ms_triangles + c->st.ms_support + (v == NUC34_DIRECT ? c->st.ms_k4_list : 0.0) +
                     c->st.ms_peel;
    if (stats) *stats = c->st;
    return NUC34_OK;
  });
}

int nuc34_copy_k(nuc34_ctx* c, int32_t* k) {
  if (!c || !c->has_k) return fail(c, NUC34_E_STATE, "decomposition not run");
  return guard(c, [&]() -> int {
    if (c->T) to_host(c, k, c->kv.as<void>(), c->T * 4);
    return NUC34_OK;
  });
}

int nuc34_copy_fourcliques(nuc34_ctx* c, uint64_t first, uint64_t count, uint32_t* quds) {
  if (!c || !c->has_k4) return fail(c, NUC34_E_STATE, "K4 list not built (direct variant only)");
  if (first > c->Q || count > c->Q - first) return fail(c, NUC34_E_INVALID, "K4 range out of bounds");
  return guard(c, [&]() -> int {
    if (!count) return NUC34_OK;
    const uint64_t T = c->T;
    // the listing walks the triangles in canonical order, with canonical ids
    // (the a-role ranks)
    ensure_canonical_tris(c);
    K4Src kc = k4src(c);
    kc.triA = c->ctriA.as<uint32_t>();
    kc.triBC = c->ctriBC.as<uint64_t>();
    kc.cSlot = nullptr;
    DevBuf k4off = c->alloc.get((T + 1) * 8, "owned K4 offsets");
    {
      DevBuf own = c->alloc.get(T * 4, "owned K4 counts in canonical order");
      launch_canon_u32(c->own.as<uint32_t>(), c->canon.as<uint32_t>(), T, own.as<uint32_t>(), c->s);
      DevBuf scratch = c->alloc.get(scan_scratch_elems(T) * 8, "scan scratch");
      exclusive_scan<uint32_t, uint64_t>(own.as<uint32_t>(), k4off.as<uint64_t>(), T, scratch.as<uint64_t>(), c->s);
    }
    // only the triangles whose owned K4s reach [first, first + count) are walked
    DevBuf bounds = c->alloc.get(16, "K4 copy bounds");
    launch_k4_owned_bounds(k4off.as<uint64_t>(), T, first, count, bounds.as<uint64_t>(), c->s);
    uint64_t tb[2] = {0, 0};
    NUC34_HIP(hipMemcpyAsync(tb, bounds.as<void>(), 16, hipMemcpyDeviceToHost, c->s));
    NUC34_HIP(hipStreamSynchronize(c->s));
    if (tb[0] > tb[1] || tb[1] >= T) throw std::logic_error("owned K4 offsets inconsistent");
    DevBuf k4 = c->alloc.get(count * 8, "owned K4s");
    launch_k4_owned(kc, k4off.as<uint64_t>(), tb[0], tb[1] + 1, first, count, k4.as<uint64_t>(), c->s);
    DevBuf out = c->alloc.get(count * 16, "K4 copy");
    launch_k4_quads(kc, k4.as<uint64_t>(), 0, count, out.as<uint32_t>(), c->s);
    NUC34_HIP(hipMemcpyAsync(quds, out.as<void>(), count * 16, hipMemcpyDeviceToHost, c->s));
    NUC34_HIP(hipStreamSynchronize(c->s));
    return NUC34_OK;
  });
}

int nuc34_reset(nuc34_ctx* c) {
  if (!c) return NUC34_E_INVALID;
  return guard(c, [&]() -> int {
    NUC34_HIP(hipStreamSynchronize(c->s));
    reset_stages(c, 1);
    const nuc34_stats keep = c->st;
    c->st = nuc34_stats{};
    c->st.n = keep.n;
    c->st.m = keep.m;
    c->st.ms_orient = keep.ms_orient;
    ktime_reset(c);
    return NUC34_OK;
  });
}

int nuc34_get_stats(const nuc34_ctx* c, nuc34_stats* stats) {
  if (!c || !stats) return NUC34_E_INVALID;
  *stats = c->st;
  return NUC34_OK;
}

int nuc34_set_tuning(nuc34_ctx* c, const nuc34_tuning* t) {
  if (!c || !t) return fail(c, NUC34_E_INVALID, "null argument");
  if (t->peel_tile_cap > 64) return fail(c, NUC34_E_INVALID, "peel_tile_cap above 64");
  if (t->peel_rounds_per_launch == 1) return fail(c, NUC34_E_INVALID, "peel_rounds_per_launch must be even");
  if (t->peel_debit_grid > (1u << 20) || t->k4_list_grid_cap > (1u << 20) || t->peel_finish_grid > (1u << 20))
    return fail(c, NUC34_E_INVALID, "grid above 2^20 workgroups");
  if (t->tri_tile_edges && t->tri_tile_edges != 16 && t->tri_tile_edges != 32 && t->tri_tile_edges != 64)
    return fail(c, NUC34_E_INVALID, "tri_tile_edges must be 16, 32 or 64");
  if (t->batch_peel > 2) return fail(c, NUC34_E_INVALID, "batch_peel must be 0, 1 or 2");
  if (t->tri_schedule > 2) return fail(c, NUC34_E_INVALID, "tri_schedule must be 0, 1 or 2");
  c->tune = *t;
  c->tune.peel_rounds_per_launch &= ~1u;
  return NUC34_OK;
}

int nuc34_get_tuning(const nuc34_ctx* c, nuc34_tuning* t) {
  if (!c || !t) return NUC34_E_INVALID;
  *t = c->tune;
  if (!t->peel_debit_grid) t->peel_debit_grid = PeelShape{}.debit_grid;
  if (!t->peel_tile_cap) t->peel_tile_cap = PeelShape{}.tile_cap;
  if (!t->peel_rounds_per_launch) t->peel_rounds_per_launch = kRoundsPerLaunch;
  if (!t->k4_list_grid_cap) t->k4_list_grid_cap = 1u << 15;
  if (!t->tri_tile_edges) t->tri_tile_edges = 64;
  if (!t->peel_finish_grid) t->peel_finish_grid = PeelShape{}.finish_grid;
  if (!t->batch_peel) t->batch_peel = 2;
  if (!t->graph_peel_max_tris) t->graph_peel_max_tris = kGraphPeelMaxTris;
  return NUC34_OK;
}

int nuc34_autotune(nuc34_ctx* c, int variant, nuc34_tuning* best, double* best_ms) {
  if (!c || !c->has_graph) return fail(c, NUC34_E_STATE, "no graph loaded");
  if (variant != NUC34_DIRECT && variant != NUC34_ONDEMAND && variant != NUC34_AUTO && variant != NUC34_AUTO_FIT)
    return fail(c, NUC34_E_INVALID, "unknown variant");
  nuc34_tuning cur;
  nuc34_get_tuning(c, &cur);
  int v = variant;  // the variant that runs (AUTO and AUTO_FIT resolve in the first decomposition)
  // a candidate's time: the median of three whole decompositions
  auto time_of = [&](const nuc34_tuning& t, double* ms) -> int {
    c->tune = t;
    double runs[3];
    for (double& r_ms : runs) {
      int rc = nuc34_reset(c);
      nuc34_stats st{};
      if (rc == NUC34_OK) rc = nuc34_decompose(c, v, 1, &st);
      if (rc != NUC34_OK) return rc;
      r_ms = st.ms_total;
      v = st.variant;
    }
    std::sort(runs, runs + 3);
    *ms = runs[1];
    return NUC34_OK;
  };
  double cur_ms = 0;
  int rc = time_of(cur, &cur_ms);
  if (rc != NUC34_OK) return rc;
  struct Knob {
    uint32_t nuc34_tuning::*field;
    std::vector<uint32_t> values;
    bool direct_only, ondemand_only, batch_only;
  };
  // the peel schedule of a batch first (rounds shared by the union, or one
  // workgroup per graph), then the launch shapes under it
  const Knob knobs[] = {
      {&nuc34_tuning::batch_peel, {1, 2}, false, false, true},
      {&nuc34_tuning::peel_debit_grid, {1024, 2048, 4096, 8192, 16384}, false, false, false},
      {&nuc34_tuning::peel_tile_cap, {8, 16, 32, 64}, false, false, false},
      {&nuc34_tuning::peel_rounds_per_launch, {16, 32, 64}, false, false, false},
      {&nuc34_tuning::k4_list_grid_cap, {8192, 32768}, true, false, false},
      {&nuc34_tuning::tri_schedule, {0, 1, 2}, false, false, false},
      {&nuc34_tuning::tri_tile_edges, {16, 32, 64}, false, false, false},
      {&nuc34_tuning::peel_finish_grid, {256, 512, 1024}, false, false, false},
  };
  // coordinate descent, repeated until a pass changes nothing (at most three
  // passes); a candidate replaces the current shape only when it is faster by
  // more than 1%, so near-equal shapes do not flip the choice between runs
  for (int pass = 0; pass < 3; ++pass) {
    bool changed = false;
    for (const Knob& k : knobs) {
      if ((k.direct_only && v != NUC34_DIRECT) || (k.ondemand_only && v != NUC34_ONDEMAND) ||
          (k.batch_only && c->batch_v.empty()))
        continue;
      for (uint32_t val : k.values) {
        if (cur.*k.field == val) continue;
        nuc34_tuning t = cur;
        t.*k.field = val;
        double ms = 0;
        rc = time_of(t, &ms);
        if (rc != NUC34_OK) return rc;
        if (ms < 0.99 * cur_ms) {
          cur = t;
          cur_ms = ms;
          changed = true;
        }
      }
    }
    if (!changed) break;
  }
  c->tune = cur;
  if (best) *best = cur;
  if (best_ms) *best_ms = cur_ms;
  return NUC34_OK;
}

int nuc34_device_info(int device, char* buf, size_t len) {
  hipDeviceProp_t p{};
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return NUC34_E_HIP;
  int rt = 0, drv = 0;
  (void)hipRuntimeGetVersion(&rt);
  (void)hipDriverGetVersion(&drv);
  std::string s;
  s += "name=" + std::string(p.name) + "\n";
  s += "arch=" + std::string(p.gcnArchName) + "\n";
  s += "global_memory_bytes=" + std::to_string(p.totalGlobalMem) + "\n";
  s += "compute_units=" + std::to_string(p.multiProcessorCount) + "\n";
  s += "clock_khz=" + std::to_string(p.clockRate) + "\n";
  s += "memory_clock_khz=" + std::to_string(p.memoryClockRate) + "\n";
  s += "memory_bus_width=" + std::to_string(p.memoryBusWidth) + "\n";
  s += "l2_cache_bytes=" + std::to_string(p.l2CacheSize) + "\n";
  s += "lds_per_block_bytes=" + std::to_string(p.sharedMemPerBlock) + "\n";
  s += "warp_size=" + std::to_string(p.warpSize) + "\n";
  s += "pci_bus_id=" + std::to_string(p.pciBusID) + "\n";
  s += "hip_runtime=" + std::to_string(rt) + "\n";
  s += "hip_driver=" + std::to_string(drv) + "\n";
  int cur = 0;
  size_t fr = 0, tot = 0;
  if (hipGetDevice(&cur) == hipSuccess && hipSetDevice(device) == hipSuccess && hipMemGetInfo(&fr, &tot) == hipSuccess)
    s += "free_memory_bytes=" + std::to_string(fr) + "\n";
  (void)hipSetDevice(cur);
  if (!buf || !len) return NUC34_E_INVALID;
  std::strncpy(buf, s.c_str(), len - 1);
  buf[len - 1] = 0;
  return NUC34_OK;
}

}  // extern "C"

// ---- decomposition without device allocation --------------------------------
//
// Every device buffer of nuc34_decompose_noalloc comes from `scratch`, a
// file-scope device array: each device the library runs on holds its own copy
// for the life of the process. Host code reaches a device's copy only through
// the address hipGetSymbolAddress returns with that device current (the
// array's name in host code is a host-side stand-in, not a device address).
namespace {

static __device__ __align__(256) unsigned char scratch[NUC34_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;

struct ScratchSlot {
  std::mutex mu;  // one call at a time per device: the calls share the array
  void* ptr = nullptr;
};
ScratchSlot g_scratch[kMaxDevices];

}  // namespace

extern "C" int nuc34_decompose_noalloc(int device, const uint32_t* O, const uint32_t* E, uint32_t n, uint64_t m,
                                       int variant, uint64_t tri_capacity, uint32_t* abc, int32_t* k,
                                       nuc34_stats* stats) {
  auto reject = [](int code, const std::string& msg) {
    tls_create_error = msg;
    return code;
  };
  if (device < 0 || device >= kMaxDevices) return reject(NUC34_E_INVALID, "device index out of range");
  if (!O || (m && !E) || (tri_capacity && (!abc || !k))) return reject(NUC34_E_INVALID, "null argument");
  if (variant != NUC34_DIRECT && variant != NUC34_ONDEMAND && variant != NUC34_AUTO)
    return reject(NUC34_E_INVALID, "variant must be NUC34_DIRECT, NUC34_ONDEMAND or NUC34_AUTO");
  if (O[0] != 0 || O[n] != m) return reject(NUC34_E_INVALID, "offsets do not span the edge array");
  if (m % 2) return reject(NUC34_E_INVALID, "undirected CSR must hold both directions of every edge");
  // the input alone (uploaded and oriented) takes 12 bytes per vertex and 10
  // per edge entry; larger graphs cannot fit whatever their triangles
  const uint64_t input_bytes = 12 * (static_cast<uint64_t>(n) + 1) + 10 * m;
  if (input_bytes > NUC34_NOALLOC_SCRATCH_BYTES / 4)
    return reject(NUC34_E_INVALID, "graph too large for the no-allocation scratch array (" +
                                       std::to_string(input_bytes) + " bytes of input state, at most " +
                                       std::to_string(NUC34_NOALLOC_SCRATCH_BYTES / 4) + ")");
  ScratchSlot& slot = g_scratch[device];
  std::lock_guard<std::mutex> lk(slot.mu);
  if (hipSetDevice(device) != hipSuccess) {
    (void)hipGetLastError();
    return reject(NUC34_E_HIP, "hipSetDevice failed");
  }
  if (!slot.ptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return reject(NUC34_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    slot.ptr = p;
  }
  nuc34_ctx c;  // no stream (the null stream), no events, buffers from the array
  c.device = device;
  c.noalloc = true;
  c.alloc.use_arena(slot.ptr, NUC34_NOALLOC_SCRATCH_BYTES);
  const int rc = guard(&c, [&]() -> int {
    {
      DevBuf Ou = c.alloc.get((static_cast<uint64_t>(n) + 1) * 4, "input offsets");
      DevBuf Eu = c.alloc.get(m * 4 + 4, "input edges");
      to_device(&c, Ou.as<void>(), {{O, (static_cast<uint64_t>(n) + 1) * 4}});
      if (m) to_device(&c, Eu.as<void>(), {{E, m * 4}});
      DevBuf bad = c.alloc.get(4, "check flag");
      launch_fill_u32(bad.as<uint32_t>(), 1, kNoVertex, c.s);
      launch_check_csr(Ou.as<uint32_t>(), Eu.as<uint32_t>(), n, m, bad.as<uint32_t>(), c.s);
      const uint32_t v = read_scalar<uint32_t>(&c, bad.as<uint32_t>());
      if (v != kNoVertex)
        return fail(&c, NUC34_E_INVALID, "vertex " + std::to_string(v) +
                                             ": offsets out of order or range, id out of range, self-loop, or "
                                             "list not strictly increasing");
      bad.release();
      install_graph(&c, Ou, Eu, n, m, false);
    }
    run_triangles(&c, 1);
    if (c.T > tri_capacity) {
      c.st.triangles = c.T;
      if (stats) *stats = c.st;
      return fail(&c, NUC34_E_INVALID, "triangle count " + std::to_string(c.T) + " exceeds tri_capacity " +
                                           std::to_string(tri_capacity));
    }
    if (c.T) run_support(&c);
    c.has_sup = true;
    const int v = variant == NUC34_AUTO ? NUC34_ONDEMAND : variant;
    if (c.Q && v == NUC34_DIRECT) run_k4_lists(&c);
    run_peel(&c, v);
    c.st.variant = v;
    c.st.peak_bytes = c.alloc.peak;
    if (c.T) {
      DevBuf buf = c.alloc.get(c.T * 12, "triangle copy");
      launch_tri_abc_canon(c.triA.as<uint32_t>(), c.triBC.as<uint64_t>(), c.canon.as<uint32_t>(), c.T,
                           buf.as<uint32_t>(), c.s);
      to_host(&c, abc, buf.as<void>(), c.T * 12);
      to_host(&c, k, c.kv.as<void>(), c.T * 4);
    }
    NUC34_HIP(hipStreamSynchronize(c.s));
    if (stats) *stats = c.st;
    return NUC34_OK;
  });
  if (rc != NUC34_OK) {
    tls_create_error = c.err;
    (void)hipStreamSynchronize(nullptr);  // nothing of this call may still run on the array
  }
  reset_stages(&c, 0);
  return rc;
}

