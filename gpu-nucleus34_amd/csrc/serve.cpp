// Serving loop over the C ABI: a queue of independent graphs per job class
// (small jobs, batched as disjoint unions; large jobs, one at a time) and
// worker threads with their own device contexts and streams (`small_workers`
// for small jobs, `large_workers` for large ones), so a large job never holds
// small ones back, one job's host work overlaps another's device work, and a
// small job that arrives while a batch runs starts on the next free small
// worker at once (continuous batching) rather than after that batch.
//
// The reference runs one graph per process (nucleus/cuNucleus34.cu:147-175);
// its memory plans (nucleus/utils.cuh:1329-1340, nucleus/cuNucleus34.cu:563-585)
// decide between one-shot and segmented work from free memory. Here each lane
// holds a share of the HBM budget; a batch runs in the configured variant (by
// default on demand, O(T) memory, or direct with 48 bytes per K4 once the
// support count has given Q), and a batch that exceeds the lane's budget (or
// overflows, or holds a malformed graph) is split in halves and retried, down
// to single jobs.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/nuc34.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
  uint64_t id = 0;
  uint32_t n = 0;
  uint64_t m = 0;
  std::vector<uint32_t> O, E;
  nuc34_job_info info{};
  std::vector<uint32_t> abc;
  std::vector<int32_t> k;
  bool done = false;
};

// A worker thread with its own context (and stream) serving one job class.
struct Worker {
  nuc34_ctx* ctx = nullptr;
  int cls = 0;  // 0: small jobs, batched; 1: large jobs, one at a time
  std::thread thread;
};

void set_error(nuc34_job_info& info, const char* msg) {
  std::strncpy(info.error, msg ? msg : "", sizeof(info.error) - 1);
  info.error[sizeof(info.error) - 1] = 0;
}

}  // namespace

struct nuc34_server {
  nuc34_serve_config cfg{};
  Clock::time_point t0 = Clock::now();
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  uint64_t next_id = 1;
  uint64_t next_batch = 1;
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
  std::deque<Job*> queue[2];  // per class: 0 small, 1 large
  std::vector<Worker> workers;
  nuc34_serve_stats st{};

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
  void work(Worker* w);
  void run_batch(Worker* w, std::vector<Job*>& batch);
  int decompose_into(Worker* w, std::vector<Job*>& batch, uint64_t batch_id, double t_start);
};

// One attempt at a batch: union load, decomposition, results sliced per job.
// Returns the status; on NUC34_E_OOM the caller splits the batch.
int nuc34_server::decompose_into(Worker* w, std::vector<Job*>& batch, uint64_t batch_id, double t_start) {
  nuc34_ctx* ctx = w->ctx;
  const uint32_t count = static_cast<uint32_t>(batch.size());
  std::vector<uint32_t> ns(count);
  std::vector<uint64_t> ms(count);
  std::vector<const uint32_t*> Op(count), Ep(count);
  for (uint32_t g = 0; g < count; ++g) {
    ns[g] = batch[g]->n;
    ms[g] = batch[g]->m;
    Op[g] = batch[g]->O.data();
    Ep[g] = batch[g]->m ? batch[g]->E.data() : nullptr;
  }
  // each job's arrays go straight into the union on the device
  int rc = nuc34_load_batch_ptrs(ctx, count, ns.data(), ms.data(), Op.data(), Ep.data(), 0);
  nuc34_stats s{};
  if (rc == NUC34_OK) rc = nuc34_decompose(ctx, cfg.variant, 1, &s);
  if (rc == NUC34_OK && count == 1) {  // one job: results straight into it
    Job* j = batch[0];
    j->abc.resize(3 * s.triangles);
    j->k.resize(s.triangles);
    rc = nuc34_copy_triangles(ctx, 0, s.triangles, j->abc.data());
    if (rc == NUC34_OK) rc = nuc34_copy_k(ctx, j->k.data());
    if (rc == NUC34_OK) {
      j->info.status = NUC34_OK;
      j->info.error[0] = 0;
      j->info.variant = s.variant;
      j->info.m = s.m;
      j->info.triangles = s.triangles;
      j->info.four_cliques = s.four_cliques;
      j->info.k_max = s.triangles ? s.k_max : 0;
      j->info.batch_graphs = 1;
      j->info.batch_id = batch_id;
      j->info.t_start = t_start;
      return NUC34_OK;
    }
  }
  std::vector<uint64_t> ef(count + 1), tf(count + 1);
  std::vector<uint32_t> abc;
  std::vector<int32_t> k, sup;
  if (rc == NUC34_OK) rc = nuc34_batch_slices(ctx, ef.data(), tf.data());
  if (rc == NUC34_OK) {
    abc.resize(3 * s.triangles);
    k.resize(s.triangles);
    sup.resize(s.triangles);
    rc = nuc34_copy_triangles(ctx, 0, s.triangles, abc.data());
    if (rc == NUC34_OK) rc = nuc34_copy_k(ctx, k.data());
    if (rc == NUC34_OK) rc = nuc34_copy_support(ctx, sup.data());
  }
  if (rc != NUC34_OK) {
    const char* msg = nuc34_last_error(ctx);
    for (Job* j : batch) {
      j->info.status = rc;
      set_error(j->info, msg);
    }
    return rc;
  }
  for (uint32_t g = 0; g < count; ++g) {
    Job* j = batch[g];
    const uint64_t a = tf[g], b = tf[g + 1];
    const uint32_t e0 = static_cast<uint32_t>(ef[g]);
    j->abc.resize(3 * (b - a));
    for (uint64_t i = 0; i < 3 * (b - a); ++i) j->abc[i] = abc[3 * a + i] - e0;
    j->k.assign(k.begin() + static_cast<std::ptrdiff_t>(a), k.begin() + static_cast<std::ptrdiff_t>(b));
    uint64_t q4 = 0;
    int32_t km = 0;
    for (uint64_t i = a; i < b; ++i) {
      q4 += static_cast<uint64_t>(sup[i]);
      km = std::max(km, k[i]);
    }
    j->info.status = NUC34_OK;
    j->info.error[0] = 0;
    j->info.variant = s.variant;
    j->info.m = ef[g + 1] - ef[g];
    j->info.triangles = b - a;
    j->info.four_cliques = q4 / 4;
    j->info.k_max = km;
    j->info.batch_graphs = count;
    j->info.batch_id = batch_id;
    j->info.t_start = t_start;
  }
  return NUC34_OK;
}

void nuc34_server::run_batch(Worker* w, std::vector<Job*>& batch) {
  uint64_t bid;
  {
    std::lock_guard<std::mutex> lk(mu);
    bid = next_batch++;
    st.batches++;
  }
  const double t_start = now();
  const int rc = decompose_into(w, batch, bid, t_start);
  // a batch that does not fit, whose union overflows the 32-bit ids, or that
  // holds a malformed graph is split in halves and retried, so the failure
  // ends up with the job that causes it and its batch-mates still run
  const bool splittable = rc == NUC34_E_OOM || rc == NUC34_E_OVERFLOW || rc == NUC34_E_INVALID;
  if (splittable && batch.size() > 1) {
    {
      std::lock_guard<std::mutex> lk(mu);
      st.splits++;
    }
    std::vector<Job*> lo(batch.begin(), batch.begin() + static_cast<std::ptrdiff_t>(batch.size() / 2));
    std::vector<Job*> hi(batch.begin() + static_cast<std::ptrdiff_t>(batch.size() / 2), batch.end());
    run_batch(w, lo);
    run_batch(w, hi);
    return;
  }
  const double t_done = now();
  {
    std::lock_guard<std::mutex> lk(mu);
    st.busy_s[w->cls] += t_done - t_start;
    for (Job* j : batch) {
      j->info.t_done = t_done;
      if (j->info.status == NUC34_OK) st.completed++;
      else st.failed++;
      j->done = true;
      j->O.clear();
      j->O.shrink_to_fit();
      j->E.clear();
      j->E.shrink_to_fit();
    }
  }
  cv_done.notify_all();
}

void nuc34_server::work(Worker* w) {
  std::deque<Job*>& Q = queue[w->cls];
  const bool batched = w->cls == 0;
  const uint64_t max_edges = cfg.max_batch_edges;
  const uint32_t max_graphs = cfg.max_batch_graphs;
  for (;;) {
    std::vector<Job*> batch;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || !Q.empty(); });
      if (Q.empty()) return;  // stop requested and drained
      if (batched && cfg.batch_window_ms > 0) {
        const double until = Q.front()->info.t_submit + cfg.batch_window_ms * 1e-3;
        const auto deadline = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(until));
        cv_work.wait_until(lk, deadline, [&] { return stop; });
        if (Q.empty()) continue;  // another worker of the class took them
      }
      uint64_t edges = 0;
      while (!Q.empty()) {
        Job* j = Q.front();
        if (!batch.empty() && (!batched || batch.size() >= max_graphs || edges + j->m > max_edges)) break;
        batch.push_back(j);
        edges += j->m;
        Q.pop_front();
      }
    }
    run_batch(w, batch);
  }
}

extern "C" {

int nuc34_serve_create(const nuc34_serve_config* cfg, nuc34_server** out) {
  if (!out) return NUC34_E_INVALID;
  *out = nullptr;
  auto s = std::make_unique<nuc34_server>();
  if (cfg) s->cfg = *cfg;
  nuc34_serve_config& c = s->cfg;
  if (c.variant != NUC34_DIRECT && c.variant != NUC34_ONDEMAND && c.variant != NUC34_AUTO_FIT) c.variant = NUC34_AUTO;
  if (c.large_fraction <= 0 || c.large_fraction >= 1) c.large_fraction = 0.75;
  if (!c.large_edges) c.large_edges = 4ull << 20;
  if (!c.large_wedges) c.large_wedges = 1ull << 28;
  if (!c.max_batch_edges) c.max_batch_edges = 256ull << 20;
  if (!c.max_batch_graphs) c.max_batch_graphs = 4096;
  if (!c.large_workers) c.large_workers = 2;
  if (!c.small_workers) c.small_workers = 2;
  // each worker holds a context, a stream and a thread
  if (c.large_workers > 64 || c.small_workers > 64) return NUC34_E_INVALID;
  if (!c.hbm_budget) {
    char buf[1024];
    if (nuc34_device_info(c.device, buf, sizeof(buf)) != NUC34_OK) return NUC34_E_HIP;
    const char* f = std::strstr(buf, "free_memory_bytes=");
    if (!f) return NUC34_E_HIP;
    c.hbm_budget = static_cast<uint64_t>(0.9 * static_cast<double>(std::strtoull(f + 18, nullptr, 10)));
  }
  const uint64_t big = static_cast<uint64_t>(c.large_fraction * static_cast<double>(c.hbm_budget));
  s->workers.resize(c.small_workers + c.large_workers);
  for (size_t i = 0; i < s->workers.size(); ++i) {
    Worker& w = s->workers[i];
    w.cls = i < c.small_workers ? 0 : 1;
    const uint64_t budget = w.cls == 0 ? (c.hbm_budget - big) / c.small_workers : big / c.large_workers;
    const int rc = nuc34_ctx_create(c.device, budget, &w.ctx);
    if (rc != NUC34_OK) {
      for (size_t k = 0; k < i; ++k) nuc34_ctx_destroy(s->workers[k].ctx);
      return rc;
    }
  }
  nuc34_server* raw = s.release();
  for (Worker& w : raw->workers) w.thread = std::thread([raw, wp = &w] { raw->work(wp); });
  *out = raw;
  return NUC34_OK;
}

void nuc34_serve_destroy(nuc34_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (Worker& w : s->workers) {
    if (w.thread.joinable()) w.thread.join();
    nuc34_ctx_destroy(w.ctx);
  }
  delete s;
}

double nuc34_serve_now(const nuc34_server* s) { return s ? s->now() : 0.0; }

int nuc34_serve_submit(nuc34_server* s, const uint32_t* O, const uint32_t* E, uint32_t n, uint64_t m, uint64_t* job) {
  if (!s || !O || (m && !E) || !job) return NUC34_E_INVALID;
  if (O[0] != 0 || O[n] != m) return NUC34_E_INVALID;
  auto j = std::make_unique<Job>();
  j->n = n;
  j->m = m;
  j->O.assign(O, O + n + 1);
  if (m) j->E.assign(E, E + m);
  j->info.n = n;
  // wedges through each vertex, d(d-1)/2 (one term fits 64 bits: d < 2^32),
  // summed until the threshold is met; w stays below the threshold meanwhile
  bool large = m >= s->cfg.large_edges;
  for (uint64_t v = 0, w = 0; v < n && !large; ++v) {
    const uint64_t d = O[v + 1] - O[v];
    const uint64_t t = d ? d * (d - 1) / 2 : 0;
    if (t >= s->cfg.large_wedges - w) large = true;
    else w += t;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return NUC34_E_STATE;
  j->id = s->next_id++;
  j->info.t_submit = s->now();
  *job = j->id;
  Job* raw = j.get();
  s->jobs.emplace(raw->id, std::move(j));
  s->queue[large ? 1 : 0].push_back(raw);
  s->st.submitted++;
  s->cv_work.notify_all();
  return NUC34_OK;
}

int nuc34_serve_wait(nuc34_server* s, uint64_t job, int64_t timeout_ms, nuc34_job_info* info) {
  if (!s) return NUC34_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job);
  if (it == s->jobs.end()) return NUC34_E_STATE;
  Job* j = it->second.get();
  auto ready = [&] { return j->done; };
  if (timeout_ms < 0) s->cv_done.wait(lk, ready);
  else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return NUC34_E_STATE;
  if (info) *info = j->info;
  return NUC34_OK;
}

int nuc34_serve_take(nuc34_server* s, uint64_t job, uint32_t* abc, int32_t* k) {
  if (!s) return NUC34_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job);
  if (it == s->jobs.end() || !it->second->done) return NUC34_E_STATE;
  Job* j = it->second.get();
  if (abc && !j->abc.empty()) std::memcpy(abc, j->abc.data(), j->abc.size() * 4);
  if (k && !j->k.empty()) std::memcpy(k, j->k.data(), j->k.size() * 4);
  s->jobs.erase(it);
  return NUC34_OK;
}

int nuc34_serve_get_stats(nuc34_server* s, nuc34_serve_stats* st) {
  if (!s || !st) return NUC34_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->st;
  return NUC34_OK;
}

}  // extern "C"
