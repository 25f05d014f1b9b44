/*
 * nuc34 — C ABI of the MI355X (gfx950) (3,4)-nucleus decomposition.
 *
 * Plain pointers and sizes only; device memory never crosses the boundary.
 * Every entry point returns a status code (NUC34_OK on success); the message
 * of the last failure on a context is available from nuc34_last_error().
 *
 * Each entry point replaces a piece of the reference's operator surface
 * (paths relative to the reference tree):
 *   nuc34_graph_load        graph_t::from_edges / deserialize + extension
 *                           dispatch   (nucleus/graph/graph.hpp:286-411, 491-536;
 *                           nucleus/cuNucleus34.cu:235-256)
 *   nuc34_load_csr          host make_oriented + H2D of O, D, E
 *                           (graph.hpp:176-218; cuNucleus34.cu:290-293)
 *   nuc34_triangles         cu::compute_triangles / _segmented / _parted +
 *                           thrust::sort(tri_t)  (nucleus/utils.cuh:1282, 1411, 1880,
 *                           1897-1912; cuNucleus34_ondemand.cu:497-500)
 *   nuc34_support           edge->triangle incidence + count_support
 *                           (cuNucleus34_ondemand.cu:502-584)
 *   nuc34_decompose         direct: cu::compute_fourcliques + incidence CSR +
 *                           peel_fourcliques loop (utils.cuh:2173; cuNucleus34.cu:511-714);
 *                           on-demand: mark_dying/debit_partners loop
 *                           (cuNucleus34_ondemand.cu:599-690)
 *   nuc34_copy_*            the D2H copies of d_tris, d_kv, d_quds and the
 *                           cudaMemcpyFromSymbol counters (cuNucleus34.cu:391, 556, 716-721)
 */
#ifndef NUC34_H_
#define NUC34_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  NUC34_OK = 0,
  NUC34_E_OOM = 1,      /* device memory or HBM budget exhausted */
  NUC34_E_OVERFLOW = 2, /* a count exceeds the 32-bit id range it indexes */
  NUC34_E_HIP = 3,      /* HIP runtime error */
  NUC34_E_INVALID = 4,  /* bad argument or malformed input */
  NUC34_E_IO = 5,       /* dataset unreadable or unknown format */
  NUC34_E_STATE = 6     /* call out of order (e.g. no graph loaded) */
};

enum { NUC34_DIRECT = 0, NUC34_ONDEMAND = 1, NUC34_AUTO = 2, NUC34_AUTO_FIT = 3 };

typedef struct nuc34_ctx nuc34_ctx;
typedef struct nuc34_graph nuc34_graph;

#define NUC34_MAX_KERNEL_STATS 8

/* Device time of one kernel (or kernel group) from HIP events recorded on the
 * context stream around its launches during the last decomposition. */
typedef struct nuc34_kernel_stat {
  char name[32];
  double ms;         /* summed over launches */
  uint64_t launches;
} nuc34_kernel_stat;

typedef struct nuc34_stats {
  uint64_t n;                 /* vertices */
  uint64_t m;                 /* oriented edges */
  uint64_t triangles;         /* T */
  uint64_t four_cliques;      /* Q */
  uint64_t incidence_entries; /* 3T edge->triangle entries */
  uint64_t rounds;            /* peel rounds */
  uint64_t unpeeled;          /* triangles left without k (0 on success) */
  uint64_t peak_bytes;        /* peak device bytes held by the context */
  int32_t k_max;
  int32_t variant;            /* NUC34_DIRECT or NUC34_ONDEMAND: what ran */
  double ms_orient;           /* device times from HIP events, milliseconds */
  double ms_triangles;
  double ms_incidence;
  double ms_support;
  double ms_k4_list;          /* direct: K4 list + triangle->K4 incidence */
  double ms_peel;
  double ms_total;            /* triangles .. peel (orientation excluded) */
  uint32_t n_kernels;
  nuc34_kernel_stat kernels[NUC34_MAX_KERNEL_STATS];
  uint64_t debits;            /* support decrements the peel applied (device atomics) */
  double ms_tri_order;        /* part of ms_triangles: canonical (a,b) ids, ranked incidence rows and
                                 the triangle records (the rest is the enumeration into the bit rows) */
  double ms_peel_init;        /* part of ms_peel: supports, states and live rows set up before round 0 */
} nuc34_stats;

/* ---- host-side datasets (reference loaders, exact semantics) ------------ */

/* Loads .mtx / .edges / .txt (undirected result of from_edges) or .grh (stored,
 * already oriented) by extension. */
int nuc34_graph_load(const char* path, nuc34_graph** out, char* err, size_t errlen);
/* Undirected graph from raw (src[i], dst[i]) pairs, as from_edges builds it
 * after its header line. */
int nuc34_graph_from_pairs(const uint32_t* src, const uint32_t* dst, uint64_t count, nuc34_graph** out);
int nuc34_graph_info(const nuc34_graph* g, uint32_t* n, uint64_t* m, int* oriented);
/* O: n+1 offsets, E: m vertex ids. */
int nuc34_graph_copy(const nuc34_graph* g, uint32_t* O, uint32_t* E);
/* Host orientation (make_oriented) of an undirected dataset, in place. */
int nuc34_graph_orient_host(nuc34_graph* g);
void nuc34_graph_free(nuc34_graph* g);

/* Seeded synthetic datasets (undirected, as if loaded from the generator's
 * .edges file). No counterpart in the reference, which downloads datasets:
 * named workloads "er1000", "rmat16", "rmat18", "rmat20k30", and graph
 * `index` of the mixed many-graphs batch drawn from `seed` (n log-uniform in
 * [200, 50000]; G(n,M) with mean degree 8-64, Barabasi-Albert m 4-16, or
 * G(n,M) degree 8 with planted 10-40 cliques, by index mod 3). */
int nuc34_gen_preset(const char* name, nuc34_graph** out);
int nuc34_gen_batch_graph(uint64_t seed, uint64_t index, nuc34_graph** out);
int nuc34_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                   uint32_t planted_cliques, uint32_t planted_size, nuc34_graph** out);

/* ---- device context ------------------------------------------------------ */

/* hbm_budget: bytes the context may hold on the device (0 = no limit). */
int nuc34_ctx_create(int device, uint64_t hbm_budget, nuc34_ctx** out);
void nuc34_ctx_destroy(nuc34_ctx* ctx);
/* The message of the context's last failure; with ctx NULL, that of this
 * thread's last failed nuc34_ctx_create (for example no device) or
 * nuc34_decompose_noalloc. */
const char* nuc34_last_error(const nuc34_ctx* ctx);

/* Uploads a CSR. oriented == 0: undirected (both directions, lists sorted by
 * id); the device orients it with the degree order. oriented != 0: used as
 * given (lists sorted, acyclic). Resets every later stage. */
int nuc34_load_csr(nuc34_ctx* ctx, const uint32_t* O, const uint32_t* E, uint32_t n, uint64_t m, int oriented);
int nuc34_graph_dims(const nuc34_ctx* ctx, uint32_t* n, uint64_t* m_oriented);

/* Many independent graphs as one batch: graph g has n[g] vertices and m[g]
 * edges; O holds the graphs' offset arrays back to back (n[g]+1 local offsets
 * each, starting at 0) and E their edge arrays (local vertex ids). They are
 * loaded as one disjoint union, so every later stage and every peel round
 * covers all graphs at once; results are per-graph slices. No counterpart in
 * the reference (which runs one graph per process): the batched mode of the
 * many-graphs workload. */
int nuc34_load_batch(nuc34_ctx* ctx, uint32_t count, const uint32_t* n, const uint64_t* m, const uint32_t* O,
                     const uint32_t* E, int oriented);
/* The same union from one (offsets, edges) pointer pair per graph (E[g] may be
 * NULL when m[g] = 0): each graph is uploaded straight into its place, with no
 * host-side concatenation. No counterpart in the reference. */
int nuc34_load_batch_ptrs(nuc34_ctx* ctx, uint32_t count, const uint32_t* n, const uint64_t* m,
                          const uint32_t* const* O, const uint32_t* const* E, int oriented);
/* Graph g of the loaded batch: its first vertex and first oriented edge in the
 * union and, once triangles are computed, its triangle slice [tri_first,
 * tri_first + tri_count) of the union's (a,b)-ordered list. The union's edge
 * indices of graph g minus edge_first are the graph's own. */
int nuc34_batch_graph(nuc34_ctx* ctx, uint32_t g, uint64_t* vertex_first, uint64_t* edge_first, uint64_t* tri_first,
                      uint64_t* tri_count);
/* All graphs' boundaries at once: edge_first and tri_first receive count + 1
 * entries each (the last one is the union's total); either may be NULL. */
int nuc34_batch_slices(nuc34_ctx* ctx, uint64_t* edge_first, uint64_t* tri_first);
/* The peel rounds of every graph of the batch (count entries) from the last
 * decomposition, when it peeled the batch one workgroup per graph (tuning
 * batch_peel); each equals the rounds of that graph decomposed alone, and the
 * union's stats.rounds is their maximum. NUC34_E_STATE when the last peel
 * shared its rounds across the union. No counterpart in the reference, which
 * decomposes one graph per process (nucleus/cuNucleus34.cu:144-175). */
int nuc34_batch_rounds(nuc34_ctx* ctx, uint32_t* rounds);
int nuc34_copy_oriented(nuc34_ctx* ctx, uint32_t* O, uint32_t* E);

/* Triangles in the reference's canonical (a,b) order; `parts` >= 1 splits the
 * enumeration into that many vertex ranges (same result for every value). */
int nuc34_triangles(nuc34_ctx* ctx, uint32_t parts, uint64_t* T);
/* abc: 3 * count entries (a, b, c) for triangles [first, first + count). */
int nuc34_copy_triangles(nuc34_ctx* ctx, uint64_t first, uint64_t count, uint32_t* abc);

/* counts: m entries, the number of triangles on each oriented edge (any
 * role). It replaces the per-edge extents of the on-demand edge->triangle
 * incidence CSR, incO[e+1] - incO[e] (nucleus/cuNucleus34_ondemand.cu:502-555). */
int nuc34_copy_edge_triangles(nuc34_ctx* ctx, uint32_t* counts);

/* Edge->triangle incidence and per-triangle K4 support; Q = sum / 4. */
int nuc34_support(nuc34_ctx* ctx, uint64_t* Q);
int nuc34_copy_support(nuc34_ctx* ctx, int32_t* support);

/* Runs the missing stages and the peel. variant: NUC34_DIRECT (per-triangle
 * K4 lists in HBM, 48 bytes per K4), NUC34_ONDEMAND (K4s re-enumerated from the
 * bit-packed incidence as triangles die, O(T) memory), NUC34_AUTO (one graph:
 * on demand, never slower and never larger on the measured single graphs; a
 * batch of two or more graphs each within tuning graph_peel_max_tris
 * triangles: as NUC34_AUTO_FIT) or NUC34_AUTO_FIT
 * (direct when, with Q known from the support count, its lists fit what is
 * left of the context's budget, or of free device memory without one, with a
 * quarter to spare; on demand otherwise: the reference's own choice between
 * one-shot and segmented work from free memory, nucleus/utils.cuh:1329-1340). */
int nuc34_decompose(nuc34_ctx* ctx, int variant, uint32_t parts, nuc34_stats* stats);
/* k for every triangle in (a,b) order (0 for triangles in no K4). */
int nuc34_copy_k(nuc34_ctx* ctx, int32_t* k);
/* Direct variant only: K4s [first, first+count) as 4 triangle ids each. */
int nuc34_copy_fourcliques(nuc34_ctx* ctx, uint64_t first, uint64_t count, uint32_t* quds);
int nuc34_get_stats(const nuc34_ctx* ctx, nuc34_stats* stats);
/* Drops every stage computed from the loaded graph (triangles onwards) and
 * keeps the graph, so the next call recomputes them; device buffers return to
 * the context's reuse cache. No counterpart in the reference: a context utility
 * for repeated runs (benchmarks, serving). */
int nuc34_reset(nuc34_ctx* ctx);

/* Launch shapes a context uses; 0 in any field selects its default. */
typedef struct nuc34_tuning {
  uint32_t peel_debit_grid;        /* workgroups of a peel round's debit kernel (2048) */
  uint32_t peel_tile_cap;          /* most dying triangles one wave of the debit kernel takes at a time, 1..64 (64) */
  uint32_t peel_rounds_per_launch; /* peel rounds per captured graph launch, even (32) */
  uint32_t k4_list_grid_cap;       /* most workgroups of the direct K4 list build (32768) */
  uint32_t tri_tile_edges;         /* oriented edges per wave tile of the triangle intersection: 16, 32, 64 (64) */
  uint32_t peel_finish_grid;       /* workgroups of a peel round's finish kernel (512) */
  uint32_t batch_peel;             /* a loaded batch: 1 = peel rounds shared by the whole union; 2 = one
                                      workgroup per graph, the whole peel one launch, when no graph has
                                      more than graph_peel_max_tris triangles; 0 = 2 (0) */
  uint32_t graph_peel_max_tris;    /* most triangles of a graph peeled by one workgroup (1 << 18) */
  uint32_t tri_schedule;           /* triangle intersection: 0 = wave tiles of tri_tile_edges oriented edges;
                                      1 = by the owner's out-degree d: a wave per vertex (d <= 64), a
                                      256-thread workgroup (d <= 256) or a 1024-thread one (d <= 640), with
                                      the vertex's rows built in LDS, and the tiles above 640; 2 = as 1 with
                                      the tiles also for d <= 8 (0) */
} nuc34_tuning;

int nuc34_set_tuning(nuc34_ctx* ctx, const nuc34_tuning* tuning);
int nuc34_get_tuning(const nuc34_ctx* ctx, nuc34_tuning* tuning);
/* Times candidate schedules and launch shapes on the loaded graph, one
 * parameter at a time from the current setting (for a batch, the peel schedule
 * first: shared rounds or one workgroup per graph), in passes until a pass
 * changes nothing (at most three; each candidate: the median of three whole
 * decompositions in `variant`, adopted when more than 1% faster), keeps the
 * fastest in the context and returns it with its decomposition time in
 * milliseconds. The reference launches with fixed shapes (MAX_THRD_BLK,
 * nucleus/utils.cuh:29); no counterpart there. */
int nuc34_autotune(nuc34_ctx* ctx, int variant, nuc34_tuning* best, double* best_ms);

/* Device description as "key=value\n" lines (name, arch, memory, CUs, ...). */

/* Measured limits of a device (about 1 GiB of device memory, well under a
 * second): streaming device-to-device copy in (read + write) GB/s, and 10^9
 * random 4-byte operations per second for device-scope atomic subtractions
 * (64 MiB array) and for loads from a 4 MiB (L2-resident) and a 256 MiB array.
 * The peel's debits are random atomics and its ownership tests random loads,
 * so these rates bound it together with the HBM peak. No counterpart in the
 * reference. */
typedef struct nuc34_device_rates {
  double copy_gbps;
  double atomic_gops;
  double load_l2_gops;
  double load_hbm_gops;
} nuc34_device_rates;
int nuc34_measure_device(int device, nuc34_device_rates* out);
int nuc34_device_info(int device, char* buf, size_t len);

/* ---- decomposition without device allocation ----------------------------- */

/* Bytes of the per-device scratch array behind nuc34_decompose_noalloc. */
#define NUC34_NOALLOC_SCRATCH_BYTES (256ull << 20)

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid it): the work of nuc34_load_csr (oriented = 0), nuc34_decompose and
 * the copies of the triangles and k, for one undirected CSR, with every
 * device buffer carved from a file-scope device array of
 * NUC34_NOALLOC_SCRATCH_BYTES that each device the library runs on holds until
 * the process ends. No hipMalloc, hipFree, stream, event, graph or pinned host
 * buffer is created; work goes to the null stream and the call returns when
 * it is done. abc receives 3 * T and k receives T entries, (a,b) order, when
 * T <= tri_capacity; otherwise NUC34_E_INVALID, with stats->triangles = T.
 * variant: NUC34_DIRECT, NUC34_ONDEMAND or NUC34_AUTO (on demand). Graphs whose
 * input alone passes a quarter of the array are rejected with
 * NUC34_E_INVALID; a graph whose intermediate state does not fit gets
 * NUC34_E_OOM. Calls on one device share its array and are serialized by a
 * per-device lock in this process (the array keeps the last call's contents;
 * every call clears what it needs). Errors: nuc34_last_error(NULL). stats
 * carries counts, rounds and the peak scratch bytes, no times. */
int nuc34_decompose_noalloc(int device, const uint32_t* O, const uint32_t* E, uint32_t n, uint64_t m, int variant,
                            uint64_t tri_capacity, uint32_t* abc, int32_t* k, nuc34_stats* stats);

/* ---- serving loop -------------------------------------------------------- */
/*
 * No counterpart in the reference (one graph per process): a decomposition
 * service for a stream of independent graphs. Any thread submits jobs; worker
 * threads, each with its own context and stream on the device, take them in
 * arrival order: `small_workers` small-job workers and `large_workers`
 * large-job workers. A free small-job worker packs every queued job that fits
 * its caps into one disjoint-union batch (one pass of every kernel and one
 * peel); with more than one, jobs that arrive while a batch is on the device
 * start at once on another worker, their work overlapping the running batch
 * (continuous batching), instead of waiting for the batch's slowest graph; a
 * batch of small graphs is peeled one workgroup per graph (tuning batch_peel).
 * The
 * large-job lane runs graphs at or above `large_edges` edge entries or
 * `large_wedges` wedges one at a time, so they never hold small jobs back.
 * Admission is by HBM footprint: each lane holds its share of the budget, a batch runs
 * in the configured variant (NUC34_AUTO: direct when its lists fit for a batch
 * of several small graphs, on demand otherwise), and a batch
 * that runs out of its lane's budget is split in halves and retried.
 */
typedef struct nuc34_server nuc34_server;

typedef struct nuc34_serve_config {
  int device;
  uint64_t hbm_budget;       /* bytes for all workers (0 = 90% of the free device memory) */
  double large_fraction;     /* share of the budget for the large-job workers (0 = 0.75) */
  uint64_t large_edges;      /* undirected edge entries (2 per edge) at which a job is large (0 = 4 Mi) */
  uint64_t max_batch_edges;  /* edge entries per small-job batch (0 = 256 Mi) */
  uint32_t max_batch_graphs; /* graphs per small-job batch (0 = 4096) */
  int variant;               /* NUC34_AUTO (default), NUC34_AUTO_FIT, NUC34_DIRECT or NUC34_ONDEMAND */
  double batch_window_ms;    /* small-job lane: wait this long after its oldest queued job
                                arrived before starting a batch (0 = start at once) */
  uint32_t large_workers;    /* workers (contexts, streams) for large jobs (0 = 2, at most 64) */
  uint64_t large_wedges;     /* wedges, sum over vertices of d(d-1)/2, at which a job is large
                                whatever its edge count (0 = 2^28): a graph with hubs costs more
                                than its edge count says */
  uint32_t small_workers;    /* workers (contexts, streams) for small-job batches (0 = 2, at most 64);
                                they share the small lane's budget equally */
} nuc34_serve_config;

typedef struct nuc34_job_info {
  int32_t status;        /* NUC34_OK, or the failure code of the job's last attempt */
  int32_t variant;       /* what ran */
  uint64_t n;            /* vertices */
  uint64_t m;            /* oriented edges */
  uint64_t triangles;
  uint64_t four_cliques;
  int32_t k_max;
  uint32_t batch_graphs; /* jobs decomposed together with this one (itself included) */
  uint64_t batch_id;
  double t_submit;       /* seconds on the server clock (nuc34_serve_now) */
  double t_start;        /* its batch started */
  double t_done;         /* its results are on the host */
  char error[96];
} nuc34_job_info;

typedef struct nuc34_serve_stats {
  uint64_t submitted, completed, failed, batches, splits;
  double busy_s[2];      /* device-side wall time per lane (small, large) */
} nuc34_serve_stats;

int nuc34_serve_create(const nuc34_serve_config* cfg, nuc34_server** out);
/* Drains the queues, then stops the lanes and frees everything. */
void nuc34_serve_destroy(nuc34_server* srv);
double nuc34_serve_now(const nuc34_server* srv);
/* Queues an undirected CSR (as nuc34_load_csr with oriented = 0); the server
 * copies it before returning. */
int nuc34_serve_submit(nuc34_server* srv, const uint32_t* O, const uint32_t* E, uint32_t n, uint64_t m,
                       uint64_t* job);
/* Waits up to timeout_ms (negative: no limit) for the job; NUC34_E_STATE when
 * it is not done by then or unknown. */
int nuc34_serve_wait(nuc34_server* srv, uint64_t job, int64_t timeout_ms, nuc34_job_info* info);
/* Copies a finished job's triangles (3T: a, b, c as the graph's own oriented
 * edge ids, (a,b) order) and k (T); either may be NULL. Then forgets the job. */
int nuc34_serve_take(nuc34_server* srv, uint64_t job, uint32_t* abc, int32_t* k);
int nuc34_serve_get_stats(nuc34_server* srv, nuc34_serve_stats* st);

#ifdef __cplusplus
}
#endif

#endif /* NUC34_H_ */
