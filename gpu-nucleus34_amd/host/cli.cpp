// Command-line front ends with the reference's drop-in surface:
//   nucleus34          <dataset> [parts] [output file]   (nucleus/cuNucleus34.cu:144-839)
//   nucleus34_ondemand <dataset> [output file]           (nucleus/cuNucleus34_ondemand.cu:308-784)
// Same extension rules, "[phase]" lines on stderr, the JSON record on stdout
// with the reference's keys and order, the k file "(a,b,c)\tk" (rows in (a,b)
// order in both modes), early exits when there are no triangles or no
// four-cliques, and error JSON with "errorstep"/"error" (exit status 0, as the
// reference). Environment keys name HIP where the reference names CUDA.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nuc34.h"
#include "cli_common.hpp"
#include "graph_io.hpp"
#include "json_out.hpp"

#if defined(NUC34_CLI_DIRECT)
#define APP_NAME "hip-nucleus34-direct"
#define DIRECT 1
#else
#define APP_NAME "hip-nucleus34-ondemand"
#define DIRECT 0
#endif
#define APP_VER "1.0.0"

using nuc34::JsonOut;

namespace {

constexpr size_t kPrintCap = 5;

std::string cxx_version() {
  std::ostringstream os;
  os << __GNUC__ << "." << __GNUC_MINOR__ << "-" << __GNUC_PATCHLEVEL__;
  return os.str();
}

void check(int rc, nuc34_ctx* ctx) {
  if (rc != NUC34_OK) throw std::runtime_error(nuc34_last_error(ctx));
}

}  // namespace

int main(int argc, char** argv) {
  const std::string prog = nuc34::strip_path(argv[0]);
  if (argc < 2) {
#if DIRECT
    std::cerr << "Usage: " << prog << " <dataset file (.mtx or .grh)>"
              << " [parts (default 1; vertex-range segments of the counting passes)]"
              << " [results output file (optional)]" << std::endl;
#else
    std::cerr << "Usage: " << prog << " <dataset file (.mtx or .grh)>"
              << " [results output file (optional)]" << std::endl;
#endif
    return EXIT_FAILURE;
  }
  const std::string dataset(argv[1]);
  std::string output;
  uint32_t parts = 1;
#if DIRECT
  if (argc > 2) {
    parts = static_cast<uint32_t>(std::stoul(argv[2]));
    if (parts < 1) parts = 1;
  }
  if (argc > 3) output = argv[3];
#else
  if (argc > 2) output = argv[2];
#endif
  std::string step;
  auto set_step = [&](const char* s) {
    step = s;
    std::cerr << "[phase] " << step << std::endl;
  };

  set_step("0.1 Environment info collection");
  const auto cpu = nuc34cli::cpuinfo();
  const auto gpu = nuc34cli::device_props(0);
  const long procs = sysconf(_SC_NPROCESSORS_ONLN);
  JsonOut js;
  js.begin_object(true, "environment");
  js.field("app", APP_NAME);
  js.field("version", APP_VER);
  js.field("g++", cxx_version());
  js.field("hip", gpu.empty() ? std::string("none") : [&] {
    for (const auto& p : gpu)
      if (p.first == "hip_runtime") return p.second;
    return std::string("unknown");
  }());
#if DIRECT
  js.field("parts", parts);
  js.field("slurm_requested_gpus", nuc34cli::env_or_empty("SLURM_GPUS"));
  js.field("slurm_available_gpus", nuc34cli::env_or_empty("SLURM_VISIBLE_DEVICES"));
  js.field("slurm_job_gpus", nuc34cli::env_or_empty("SLURM_JOB_GPUS"));
  js.field("peer_gpu_count", 0);
#endif
  js.field("thrust", "none (hand-written HIP)");
  js.field("policy", "device");
  js.field("procs", procs);
  js.field("cpu_core_count", cpu.size());
#if DIRECT
  js.field("cpu_affinity", nuc34cli::cpu_affinity());
#endif
  js.field("cpu_model", cpu.empty() ? std::string("Unknown") : cpu[0].count("model name") ? cpu[0].at("model name") : std::string("Unknown"));
#if DIRECT
  js.field("cpu_physical_cores", cpu.empty() ? std::string("Unknown") : cpu[0].count("cpu cores") ? cpu[0].at("cpu cores") : std::string("Unknown"));
#endif
  for (const auto& p : gpu) {
    std::string k = "gpu_" + p.first;
    std::replace(k.begin(), k.end(), ' ', '_');
    js.field(k, p.second);
  }
  js.field("host", nuc34cli::hostname());
  js.end_object();
  js.begin_object(false, "dataset");
  js.field("file", nuc34::strip_path(dataset));

  nuc34_ctx* ctx = nullptr;
  nuc34_graph* g = nullptr;
  auto finish = [&]() {
    js.end_object();
    js.last_object();
    std::cout << js.str() << std::endl;
    if (ctx) nuc34_ctx_destroy(ctx);
    if (g) nuc34_graph_free(g);
  };
  try {
    set_step("0.2 Data loading and graph construction");
    const auto t0 = std::chrono::steady_clock::now();
    char err[512] = {0};
    if (nuc34_graph_load(dataset.c_str(), &g, err, sizeof(err)) != NUC34_OK) throw std::runtime_error(err);
    uint32_t n = 0;
    uint64_t mu = 0;
    int oriented = 0;
    nuc34_graph_info(g, &n, &mu, &oriented);
    std::vector<uint32_t> O(static_cast<size_t>(n) + 1), E(mu);
    nuc34_graph_copy(g, O.data(), E.data());
    set_step("0.3 Device initialization and upload");
    check(nuc34_ctx_create(0, 0, &ctx), ctx);
    check(nuc34_load_csr(ctx, O.data(), E.data(), n, mu, oriented), ctx);
    uint64_t m = 0;
    nuc34_graph_dims(ctx, &n, &m);
    const double t_data = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    js.field("vertex_count", n);
    js.field("edge_count", m * 2);
    {
      std::ostringstream os;
      os << std::fixed << std::setprecision(2) << t_data;
      js.raw("data_loading_time_sec", os.str());
    }
#if DIRECT
    js.field("data_memory_bytes", (static_cast<uint64_t>(n) + 1) * 4 + mu * 4);
#endif
    {
      std::ostringstream os;
      os << "vertices=" << n << " oriented_edges=" << m << " host_graph=" << (((static_cast<uint64_t>(n) + 1) * 4 + mu * 4) >> 20)
         << " MiB load=" << std::fixed << std::setprecision(1) << t_data << "s";
      nuc34cli::mem_note("graph", os.str());
    }
#if DIRECT
    {
      // no path-count scratch: the intersection writes bit-packed rows of
      // ceil(d+/32) and ceil(d-/32) words per edge
      std::ostringstream os;
      os << "parts=" << parts << " mode=bit-rows (no path scratch)";
      nuc34cli::mem_note("tri-plan", os.str());
    }
#endif
    set_step("1.1 Triangle counting");
    uint64_t T = 0;
    check(nuc34_triangles(ctx, parts, &T), ctx);
    nuc34_stats st{};
    nuc34_get_stats(ctx, &st);
    // enumeration into the bit rows; the canonical ordering is the reference's
    // triangle graph (direct) or incidence build (on demand)
    js.field("triangle_time_sec", (st.ms_triangles - st.ms_tri_order) / 1e3);
    js.field("triangle_count", T);
    {
      std::ostringstream os;
      os << "triangles=" << T << " device_tris=" << ((T * 12) >> 20) << " MiB time=" << std::fixed
         << std::setprecision(1) << st.ms_triangles / 1e3 << "s";
      nuc34cli::mem_note("tri-done", os.str());
    }
    std::vector<uint32_t> first(3 * std::min<uint64_t>(kPrintCap, T));
    check(nuc34_copy_triangles(ctx, 0, first.size() / 3, first.data()), ctx);
#if DIRECT
    {
      std::vector<std::string> rows;
      for (size_t i = 0; i < first.size() / 3; ++i) rows.push_back(nuc34cli::tuple3(&first[3 * i]));
      js.array("triangles", rows.begin(), rows.end(), [](std::ostream& os, const std::string& r) { os << '"' << r << '"'; });
    }
#endif
    if (!T) {
      finish();
      return EXIT_SUCCESS;
    }
    set_step(DIRECT ? "2.0 Building triangle incidence" : "2.1 Building edge -> triangle incidence (3 entries per triangle)");
    uint64_t Q = 0;
    check(nuc34_support(ctx, &Q), ctx);
    nuc34_get_stats(ctx, &st);
#if DIRECT
    js.field("triangle_graph_time_sec", st.ms_tri_order / 1e3);
    js.field("four_clique_time_sec", st.ms_support / 1e3);
    js.field("four_clique_count", Q);
#else
    js.field("incidence_time_sec", st.ms_tri_order / 1e3);
    js.field("incidence_entries", st.incidence_entries);
    js.field("four_clique_time_sec", st.ms_support / 1e3);
    js.field("four_clique_count", Q);
#endif
    {
      std::ostringstream os;
      os << "four_cliques=" << Q << " peel_resident=" << ((T * 24 + (DIRECT ? 48 * Q + 8 * (T + 1) : 0)) >> 20)
         << " MiB time=" << std::fixed << std::setprecision(1) << st.ms_support / 1e3 << "s";
      nuc34cli::mem_note("fc-done", os.str());
    }
    if (!Q) {
#if DIRECT
      js.array("four-cliques", first.begin(), first.begin(), [](std::ostream&, uint32_t) {});
#endif
      finish();
      return EXIT_SUCCESS;
    }
    set_step(DIRECT ? "4.0 Peeling (device resident, K4 lists)" : "4.0 Peeling (on-demand clique re-enumeration)");
    check(nuc34_decompose(ctx, DIRECT ? NUC34_DIRECT : NUC34_ONDEMAND, parts, &st), ctx);
#if DIRECT
    {
      std::vector<uint32_t> q(4 * std::min<uint64_t>(kPrintCap, Q));
      check(nuc34_copy_fourcliques(ctx, 0, q.size() / 4, q.data()), ctx);
      std::vector<std::string> rows;
      for (size_t i = 0; i < q.size() / 4; ++i) rows.push_back(nuc34cli::tuple4(&q[4 * i]));
      js.array("four-cliques", rows.begin(), rows.end(), [](std::ostream& os, const std::string& r) { os << '"' << r << '"'; });
    }
    nuc34cli::mem_note("inc-done", "triangle->four-clique lists built");
    js.field("four_clique_graph_time_sec", st.ms_k4_list / 1e3);
    js.field("peeling_mode", "resident-direct");
    js.field("peeling_preprocess_time_sec", st.ms_peel_init / 1e3);
    js.field("peeling_time_sec", (st.ms_peel - st.ms_peel_init) / 1e3);
#else
    js.field("unpeeled_count", st.unpeeled);
    if (st.unpeeled) std::cerr << "[warn] " << st.unpeeled << " triangles left unpeeled" << std::endl;
    js.field("peeling_mode", "ondemand-reenumerate");
    js.field("peeling_time_sec", st.ms_peel / 1e3);
#endif
    js.field("peeling_iterations", st.rounds);
    js.field("K_max", st.k_max);
    js.field("total_time_sec", st.ms_total / 1e3);
    set_step("4.3 Finalizing results on host");
    std::vector<int32_t> k(T);
    check(nuc34_copy_k(ctx, k.data()), ctx);
    js.field("K_count", k.size());
    js.array("K", k.begin(), k.begin() + static_cast<long>(std::min<uint64_t>(kPrintCap, T)),
             [](std::ostream& os, int32_t v) { os << v; });
    if (!output.empty()) {
      std::vector<uint32_t> abc(3 * T);
      check(nuc34_copy_triangles(ctx, 0, T, abc.data()), ctx);
      nuc34cli::write_kfile(output, abc, k);
    }
    long long sum = 0;
    for (int32_t x : k) sum += x;
    js.field("K_avg", static_cast<double>(sum) / static_cast<double>(k.size()));
    finish();
  } catch (const std::exception& e) {
    std::cerr << "Dataset: " << dataset << std::endl
              << "Step: " << step << std::endl
              << "Error: " << e.what() << std::endl
              << "ROCm SMI: " << std::endl
              << nuc34cli::smi_dump() << std::endl;
    js.field("errorstep", step);
    js.field("error", std::string(e.what()));
    finish();
  }
  return EXIT_SUCCESS;
}
