// Batched and serving forms of the :gradepro estimator.
//
// Batched: a list of independent requests is spread over stream lanes
// (longest-processing-time first) and enqueued with no host synchronisation
// until the end, so small estimates run concurrently.  Each lane keeps its
// plans (device buffers) resident between calls; host data moves through a
// pinned staging buffer so asynchronous copies never block the host.
//
// Serving: a worker thread runs continuous batching over `max_inflight`
// device streams.  A request is admitted as soon as a stream slot is free and
// its device footprint fits the memory budget; admitted requests overlap on
// the GPU.  Each slot keeps its plan (grid tables, spectra, rung buffers)
// resident and reuses it for the next request of the same shape.  Results
// do not depend on co-scheduled work: binning is fixed-point (order
// independent) and every other reduction has a fixed order.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "pkde_plan.hpp"

using namespace pkde;

namespace {

struct PlanKey {
  int32_t ndims;
  int64_t n[PKDE_MAX_DIMS];
  double lo[PKDE_MAX_DIMS], hi[PKDE_MAX_DIMS], h[PKDE_MAX_DIMS];
  pkde_gradepro_params p;
  bool operator==(const PlanKey& o) const { return memcmp(this, &o, sizeof(PlanKey)) == 0; }
};

PlanKey key_of(const pkde_request& r) {
  PlanKey k;
  memset(&k, 0, sizeof(k));
  k.ndims = r.ndims;
  for (int i = 0; i < r.ndims && i < PKDE_MAX_DIMS; ++i) {
    k.n[i] = r.n_per_axis[i];
    k.lo[i] = r.lo[i];
    k.hi[i] = r.hi[i];
    k.h[i] = r.h[i];
  }
  k.p = r.params;
  return k;
}

int make_plan(pkde_ctx* ctx, const pkde_request& r, pkde_plan** out) {
  return pkde_gradepro_plan_create(ctx, r.ndims, r.n_per_axis, r.lo, r.hi, r.h, &r.params, out);
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// Estimated device work of a request: the ladder (grid points x rungs x
// bootstraps) plus binning (draws x corners).
static double request_work(const pkde_request& r) {
  double G = 1.0;
  for (int i = 0; i < r.ndims; ++i) G *= (double)r.n_per_axis[i];
  const double B = r.params.n_bootstraps > 0 ? r.params.n_bootstraps : 100;
  const double S = r.params.n_steps > 0 ? r.params.n_steps + 1 : 251;
  return G * S * B + (double)r.n * B * (double)(1 << r.ndims);
}

static std::vector<unsigned char> key_bytes(const pkde_request& r) {
  PlanKey k = key_of(r);
  const unsigned char* p = (const unsigned char*)&k;
  return std::vector<unsigned char>(p, p + sizeof(k));
}

extern "C" int pkde_gradepro_estimate_batch(pkde_ctx* ctx, int32_t count, const pkde_request* reqs,
                                            int32_t* status_out) {
  if (!ctx || count < 0 || (count > 0 && !reqs)) return fail(PKDE_INVALID_ARGUMENT, "invalid argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = check_hip(hipSetDevice(ctx->device), "hipSetDevice");
  if (rc) return rc;
  // stream lanes: requests are spread over independent streams so that small
  // estimates run concurrently; each lane keeps its plans resident
  int nl = 8;
  if (const char* e = getenv("PKDE_BATCH_STREAMS")) nl = std::max(1, std::min(16, atoi(e)));
  while ((int)ctx->lanes.size() < nl) {
    pkde_batch_lane ln;
    rc = check_hip(hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking), "stream");
    if (rc) return rc;
    ctx->lanes.push_back(ln);
  }
  // pinned staging: samples in, density / chosen / flag out
  std::vector<size_t> off_x(count), off_d(count), off_c(count), off_e(count);
  size_t total = 0;
  auto take = [&](size_t bytes) {
    size_t o = total;
    total += (bytes + 255) & ~(size_t)255;
    return o;
  };
  for (int i = 0; i < count; ++i) {
    const pkde_request& r = reqs[i];
    int64_t G = 1;
    for (int a = 0; a < r.ndims && a < PKDE_MAX_DIMS; ++a) G *= r.n_per_axis[a];
    off_x[i] = take((size_t)std::max<int64_t>(r.n, 0) * std::max(r.ndims, 0) * 8);
    off_d[i] = take((size_t)std::max<int64_t>(G, 0) * 8);
    off_c[i] = take((size_t)std::max<int64_t>(G, 0) * 2);
    off_e[i] = take(4);
  }
  if (total > ctx->stage_bytes) {
    if (ctx->stage) (void)hipHostFree(ctx->stage);
    ctx->stage = nullptr;
    ctx->stage_bytes = 0;
    rc = check_hip(hipHostMalloc((void**)&ctx->stage, total, hipHostMallocDefault), "pinned staging");
    if (rc) return rc;
    ctx->stage_bytes = total;
  }
  // longest-processing-time assignment of requests to lanes
  std::vector<int> order(count), lane_of(count, 0);
  for (int i = 0; i < count; ++i) order[i] = i;
  std::sort(order.begin(), order.end(),
            [&](int x, int y) { return request_work(reqs[x]) > request_work(reqs[y]); });
  for (auto& ln : ctx->lanes) ln.work = 0.0;
  for (int i : order) {
    int best = 0;
    for (int l = 1; l < nl; ++l)
      if (ctx->lanes[l].work < ctx->lanes[best].work) best = l;
    lane_of[i] = best;
    ctx->lanes[best].work += request_work(reqs[i]);
  }
  std::vector<int> st(count, PKDE_OK);
  std::vector<pkde_plan*> used(count, nullptr);
  int first_rc = PKDE_OK;
  for (int i : order) {
    const pkde_request& r = reqs[i];
    pkde_batch_lane& ln = ctx->lanes[lane_of[i]];
    hipStream_t s = ln.stream;
    if (!r.samples || !r.density_out || r.n < 2 || r.ndims < 1 || r.ndims > PKDE_MAX_DIMS) {
      st[i] = fail(PKDE_INVALID_ARGUMENT, "invalid request");
      continue;
    }
    pkde_plan* pl = nullptr;
    std::vector<unsigned char> k = key_bytes(r);
    for (auto& kv : ln.plans)
      if (kv.first == k) pl = kv.second;
    if (!pl) {
      st[i] = make_plan(ctx, r, &pl);
      if (st[i]) continue;
      ln.plans.push_back({k, pl});
    }
    used[i] = pl;
    memcpy(ctx->stage + off_x[i], r.samples, (size_t)r.n * r.ndims * 8);
    st[i] = plan_reserve(pl, r.n, false);
    if (!st[i])
      st[i] = check_hip(hipMemcpyAsync(pl->d.x, ctx->stage + off_x[i], (size_t)r.n * r.ndims * 8,
                                       hipMemcpyHostToDevice, s), "upload");
    if (!st[i]) st[i] = plan_enqueue(pl, pl->d.x, r.n, nullptr, pl->d.density, pl->d.chosen, s, false);
    if (!st[i])
      st[i] = check_hip(hipMemcpyAsync(ctx->stage + off_d[i], pl->d.density, (size_t)pl->g.G * 8,
                                       hipMemcpyDeviceToHost, s), "download");
    if (!st[i])
      st[i] = check_hip(hipMemcpyAsync(ctx->stage + off_c[i], pl->d.chosen, (size_t)pl->g.G * 2,
                                       hipMemcpyDeviceToHost, s), "download");
    if (!st[i])
      st[i] = check_hip(hipMemcpyAsync(ctx->stage + off_e[i], pl->d.err, 4, hipMemcpyDeviceToHost, s),
                        "flag");
  }
  for (int l = 0; l < nl; ++l) {
    int e = check_hip(hipStreamSynchronize(ctx->lanes[l].stream), "batch");
    if (e && !rc) rc = e;
  }
  for (int i = 0; i < count; ++i) {
    if (st[i] == PKDE_OK && used[i]) {
      const int64_t G = used[i]->g.G;
      int err = 0;
      memcpy(&err, ctx->stage + off_e[i], 4);
      if (err & 1)
        st[i] = fail(PKDE_OUT_OF_GRID, "Data points lie beyond grid.");
      else {
        memcpy(reqs[i].density_out, ctx->stage + off_d[i], (size_t)G * 8);
        if (reqs[i].chosen_out) memcpy(reqs[i].chosen_out, ctx->stage + off_c[i], (size_t)G * 2);
      }
    }
    if (status_out) status_out[i] = st[i];
    if (st[i] != PKDE_OK && first_rc == PKDE_OK) first_rc = st[i];
  }
  return rc ? rc : first_rc;
}

// ------------------------------------------------------------------ server
struct Job {
  uint64_t ticket = 0;
  pkde_request req{};
  std::vector<double> samples;
  int status = PKDE_PENDING;
  bool done = false;
  double t_submit = 0, t_done = 0;
  uint64_t bytes = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  pkde_plan* plan = nullptr;
  PlanKey key{};
  std::shared_ptr<Job> job;
  int* err_host = nullptr;  // pinned
  double* dens_host = nullptr;
  int16_t* chosen_host = nullptr;
  int64_t host_cap = 0;
};

struct pkde_server {
  pkde_ctx* ctx = nullptr;
  pkde_server_cfg cfg{};
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv;
  std::condition_variable done_cv;
  std::deque<std::shared_ptr<Job>> queue;
  std::map<uint64_t, std::shared_ptr<Job>> jobs;
  std::vector<Slot> slots;
  uint64_t next_ticket = 1;
  uint64_t completed = 0, rejected = 0, bytes_in_use = 0;
  std::vector<double> lat;
  bool stop = false;
};

static void finish_job(pkde_server* sv, Slot& sl) {
  auto job = sl.job;
  const int64_t G = sl.plan->g.G;
  int st = PKDE_OK;
  if (*sl.err_host & 1) st = PKDE_OUT_OF_GRID;
  if (st == PKDE_OK) {
    memcpy(job->req.density_out, sl.dens_host, (size_t)G * 8);
    if (job->req.chosen_out) memcpy(job->req.chosen_out, sl.chosen_host, (size_t)G * 2);
  }
  std::lock_guard<std::mutex> lk(sv->mu);
  job->status = st;
  job->done = true;
  job->t_done = now_ms();
  sv->lat.push_back(job->t_done - job->t_submit);
  sv->completed++;
  sv->bytes_in_use -= job->bytes;
  sl.job.reset();
  sv->done_cv.notify_all();
}

static int start_job(pkde_server* sv, Slot& sl, std::shared_ptr<Job> job) {
  const pkde_request& r = job->req;
  PlanKey k = key_of(r);
  int st = PKDE_OK;
  if (!sl.plan || !(sl.key == k)) {
    if (sl.plan) pkde_gradepro_plan_destroy(sl.plan);
    sl.plan = nullptr;
    st = make_plan(sv->ctx, r, &sl.plan);
    if (st) return st;
    sl.key = k;
  }
  pkde_plan* pl = sl.plan;
  const int64_t G = pl->g.G;
  if (G > sl.host_cap) {
    if (sl.dens_host) (void)hipHostFree(sl.dens_host);
    if (sl.chosen_host) (void)hipHostFree(sl.chosen_host);
    sl.dens_host = nullptr;
    sl.chosen_host = nullptr;
    if ((st = check_hip(hipHostMalloc((void**)&sl.dens_host, (size_t)G * 8, hipHostMallocDefault), "pinned")))
      return st;
    if ((st = check_hip(hipHostMalloc((void**)&sl.chosen_host, (size_t)G * 2, hipHostMallocDefault), "pinned")))
      return st;
    sl.host_cap = G;
  }
  if ((st = plan_reserve(pl, r.n, false))) return st;
  hipStream_t s = sl.stream;
  if ((st = check_hip(hipMemcpyAsync(pl->d.x, job->samples.data(), (size_t)r.n * r.ndims * 8,
                                     hipMemcpyHostToDevice, s), "upload")))
    return st;
  if ((st = plan_enqueue(pl, pl->d.x, r.n, nullptr, pl->d.density, pl->d.chosen, s, false))) return st;
  if ((st = check_hip(hipMemcpyAsync(sl.dens_host, pl->d.density, (size_t)G * 8, hipMemcpyDeviceToHost, s), "dl")))
    return st;
  if ((st = check_hip(hipMemcpyAsync(sl.chosen_host, pl->d.chosen, (size_t)G * 2, hipMemcpyDeviceToHost, s), "dl")))
    return st;
  if ((st = check_hip(hipMemcpyAsync(sl.err_host, pl->d.err, 4, hipMemcpyDeviceToHost, s), "dl"))) return st;
  if ((st = check_hip(hipEventRecord(sl.done, s), "event"))) return st;
  sl.job = job;
  return PKDE_OK;
}

static void server_loop(pkde_server* sv) {
  (void)hipSetDevice(sv->ctx->device);
  for (;;) {
    bool progressed = false;
    // completions
    for (auto& sl : sv->slots)
      if (sl.job && hipEventQuery(sl.done) == hipSuccess) {
        finish_job(sv, sl);
        progressed = true;
      }
    // admissions (continuous batching)
    for (auto& sl : sv->slots) {
      if (sl.job) continue;
      std::shared_ptr<Job> job;
      {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->queue.empty()) break;
        auto& front = sv->queue.front();
        if (front->bytes > sv->cfg.memory_budget) {
          front->status = PKDE_OUT_OF_MEMORY;
          front->done = true;
          front->t_done = now_ms();
          sv->rejected++;
          sv->queue.pop_front();
          sv->done_cv.notify_all();
          progressed = true;
          continue;
        }
        if (sv->bytes_in_use + front->bytes > sv->cfg.memory_budget) break;
        job = front;
        sv->queue.pop_front();
        sv->bytes_in_use += job->bytes;
      }
      int st = start_job(sv, sl, job);
      if (st != PKDE_OK) {
        std::lock_guard<std::mutex> lk(sv->mu);
        job->status = st;
        job->done = true;
        job->t_done = now_ms();
        sv->bytes_in_use -= job->bytes;
        sv->done_cv.notify_all();
      }
      progressed = true;
    }
    std::unique_lock<std::mutex> lk(sv->mu);
    bool busy = false;
    for (auto& sl : sv->slots) busy = busy || (bool)sl.job;
    if (sv->stop && !busy && sv->queue.empty()) break;
    if (!progressed) {
      if (busy)
        sv->cv.wait_for(lk, std::chrono::microseconds(50));
      else
        sv->cv.wait_for(lk, std::chrono::milliseconds(5), [&] { return sv->stop || !sv->queue.empty(); });
    }
  }
}

extern "C" {

int pkde_server_create(pkde_ctx* ctx, const pkde_server_cfg* cfg, pkde_server** out) {
  if (!ctx || !out) return fail(PKDE_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  int rc = check_hip(hipSetDevice(ctx->device), "hipSetDevice");
  if (rc) return rc;
  pkde_server* sv = new pkde_server();
  sv->ctx = ctx;
  sv->cfg.max_inflight = (cfg && cfg->max_inflight > 0) ? cfg->max_inflight : 4;
  if (sv->cfg.max_inflight > 16) sv->cfg.max_inflight = 16;
  size_t fr = 0, tot = 0;
  (void)hipMemGetInfo(&fr, &tot);
  sv->cfg.memory_budget = (cfg && cfg->memory_budget > 0) ? cfg->memory_budget : (uint64_t)(fr * 0.8);
  sv->slots.resize(sv->cfg.max_inflight);
  for (auto& sl : sv->slots) {
    if (!rc) rc = check_hip(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "stream");
    if (!rc) rc = check_hip(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "event");
    if (!rc) rc = check_hip(hipHostMalloc((void**)&sl.err_host, 64, hipHostMallocDefault), "pinned");
  }
  if (rc) {
    pkde_server_destroy(sv);
    return rc;
  }
  sv->worker = std::thread(server_loop, sv);
  *out = sv;
  return PKDE_OK;
}

int pkde_server_submit(pkde_server* sv, const pkde_request* req, uint64_t* ticket) {
  if (!sv || !req || !ticket || !req->samples || !req->density_out || req->n < 2)
    return fail(PKDE_INVALID_ARGUMENT, "invalid request");
  auto job = std::make_shared<Job>();
  job->req = *req;
  job->samples.assign(req->samples, req->samples + (size_t)req->n * req->ndims);
  job->req.samples = job->samples.data();
  // footprint from a host-only plan description
  pkde_plan* probe = nullptr;
  int st = make_plan(sv->ctx, *req, &probe);
  if (st) return st;
  job->bytes = plan_bytes(probe, req->n);
  pkde_gradepro_plan_destroy(probe);
  std::lock_guard<std::mutex> lk(sv->mu);
  job->ticket = sv->next_ticket++;
  job->t_submit = now_ms();
  sv->jobs[job->ticket] = job;
  sv->queue.push_back(job);
  *ticket = job->ticket;
  sv->cv.notify_all();
  return PKDE_OK;
}

int pkde_server_poll(pkde_server* sv, uint64_t ticket, int* done) {
  if (!sv || !done) return fail(PKDE_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->jobs.find(ticket);
  if (it == sv->jobs.end()) return fail(PKDE_INVALID_ARGUMENT, "unknown ticket");
  *done = it->second->done ? 1 : 0;
  if (!it->second->done) return PKDE_PENDING;
  const int st = it->second->status;
  sv->jobs.erase(it);
  return st;
}

int pkde_server_wait(pkde_server* sv, uint64_t ticket) {
  if (!sv) return fail(PKDE_INVALID_ARGUMENT, "null server");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->jobs.find(ticket);
  if (it == sv->jobs.end()) return fail(PKDE_INVALID_ARGUMENT, "unknown ticket");
  auto job = it->second;
  sv->done_cv.wait(lk, [&] { return job->done; });
  sv->jobs.erase(ticket);
  return job->status;
}

int pkde_server_stats_get(pkde_server* sv, pkde_server_stats* out) {
  if (!sv || !out) return fail(PKDE_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  memset(out, 0, sizeof(*out));
  out->completed = sv->completed;
  out->rejected = sv->rejected;
  out->queued = sv->queue.size();
  uint64_t infl = 0;
  for (auto& sl : sv->slots) infl += sl.job ? 1 : 0;
  out->inflight = infl;
  out->bytes_in_use = sv->bytes_in_use;
  if (!sv->lat.empty()) {
    std::vector<double> l = sv->lat;
    std::sort(l.begin(), l.end());
    auto q = [&](double f) {
      size_t i = (size_t)std::min<double>((double)l.size() - 1, std::max(0.0, f * (double)l.size() - 1e-9));
      return l[i];
    };
    out->p50_ms = q(0.50);
    out->p99_ms = q(0.99);
  }
  return PKDE_OK;
}

void pkde_server_destroy(pkde_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
    sv->cv.notify_all();
  }
  if (sv->worker.joinable()) sv->worker.join();
  (void)hipSetDevice(sv->ctx->device);
  for (auto& sl : sv->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.plan) pkde_gradepro_plan_destroy(sl.plan);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.err_host) (void)hipHostFree(sl.err_host);
    if (sl.dens_host) (void)hipHostFree(sl.dens_host);
    if (sl.chosen_host) (void)hipHostFree(sl.chosen_host);
  }
  delete sv;
}

}  // extern "C"
