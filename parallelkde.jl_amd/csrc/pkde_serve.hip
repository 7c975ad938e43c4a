// Batched and serving forms of the :gradepro estimator.
//
// Batched: a list of independent requests is spread over stream lanes
// (longest-processing-time first) and enqueued with no host synchronisation
// until the end, so small estimates run concurrently.  Each lane keeps one
// plan per grid shape and keywords (the bootstrap seed is not part of the
// key: it is a launch argument) resident between calls, and one rung-chunk
// buffer that its plans borrow in turn (estimates on one stream run one after
// another).  Host data moves through a pinned staging buffer so asynchronous
// copies never block the host.
//
// Serving: a worker thread runs continuous batching over `max_inflight`
// device streams (slots).  A request's device footprint is computed on the
// host when it is submitted; a request whose footprint alone exceeds the
// memory budget is rejected with PKDE_OUT_OF_MEMORY and the others go on.
// Short requests go first (least estimated work among the first queued ones,
// with a bound on how often one request can be passed over).  A request is
// admitted as soon as a slot is free and the memory it still
// needs (a new plan for its shape, a larger chunk buffer for the slot) fits
// the budget next to everything resident; idle plans are evicted, least
// recently used first, to make room.  Plans live in a pool keyed by shape:
// a request takes an idle plan of its shape, returns it when done, and a
// plan is never freed while a stream uses it.  Admitted requests overlap on
// the GPU.  Results do not depend on what else is in flight: binning is
// fixed-point (order independent) and every other reduction has a fixed
// order.  The memory model extends get_necessary_memory / check_memory
// (GradeProEstimator.jl:1091-1112) from one estimate to the resident set of
// a serving loop.
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

// Grid and keywords; the Philox seed of the bootstrap draws is left out (it
// is passed to each estimate), so requests that differ only in their seed
// share plans.
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
  k.p.rng_seed = 0;
  return k;
}

std::vector<unsigned char> key_bytes(const pkde_request& r) {
  PlanKey k = key_of(r);
  const unsigned char* p = (const unsigned char*)&k;
  return std::vector<unsigned char>(p, p + sizeof(k));
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

int64_t pow2_ceil(int64_t v) {
  int64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Host-side footprint of a request (no device call): the plan's buffers for
// n samples and the chunk buffer its stream must lend it under the lane
// budget.
struct Footprint {
  int status = PKDE_OK;
  uint64_t plan = 0;   // plan buffers, chunk buffer excluded
  uint64_t chunk = 0;  // the lent chunk buffer
  uint64_t chunk_own = 0;  // the chunk buffer a plan of its own would take
  int64_t G = 0;
  double work = 0.0;   // estimated device work (ladder + binning)
  double ladder_work = 0.0;  // G S B: grid-point rung bootstraps of the ladder
};

Footprint footprint_of(pkde_ctx* ctx, const pkde_request& r, int64_t n_cap,
                       uint64_t chunk_budget) {
  Footprint f;
  pkde_plan hp;
  f.status = plan_setup_host(ctx, r.ndims, r.n_per_axis, r.lo, r.hi, r.h, &r.params, &hp);
  if (f.status) return f;
  hp.chunk_budget = 0;
  f.chunk_own = plan_chunk_bytes(&hp);
  hp.chunk_budget = chunk_budget;
  f.plan = plan_bytes(&hp, n_cap, false);
  f.chunk = plan_chunk_bytes(&hp);
  f.G = hp.g.G;
  f.ladder_work = (double)hp.g.G * hp.S * hp.B;
  f.work = f.ladder_work + (double)r.n * hp.B * (double)(1 << r.ndims);
  return f;
}

int check_request(const pkde_request& r) {
  if (!r.samples || !r.density_out || r.n < 2 || r.ndims < 1 || r.ndims > PKDE_MAX_DIMS)
    return fail(PKDE_INVALID_ARGUMENT, "invalid request");
  return PKDE_OK;
}

// A plan for a stream lane or slot: the lane chunk budget, the lane's chunk
// buffer lent to it.
int make_lane_plan(pkde_ctx* ctx, const pkde_request& r, pkde_plan** out,
                   uint64_t chunk_budget) {
  const int rc =
      pkde_gradepro_plan_create(ctx, r.ndims, r.n_per_axis, r.lo, r.hi, r.h, &r.params, out);
  if (rc == PKDE_OK) (*out)->chunk_budget = chunk_budget;
  return rc;
}

// Grows a stream's chunk buffer to `need` bytes (the stream is idle).
int grow_chunk(void*& Y, uint64_t& Y_bytes, uint64_t need, const char* what) {
  if (need <= Y_bytes) return PKDE_OK;
  if (Y) (void)hipFree(Y);
  Y = nullptr;
  Y_bytes = 0;
  const hipError_t e = hipMalloc(&Y, (size_t)need);
  if (e != hipSuccess) {
    Y = nullptr;
    (void)hipGetLastError();
    return check_hip(e, what);
  }
  Y_bytes = need;
  return PKDE_OK;
}

}  // namespace

// ------------------------------------------------------------------- batch
// device_io: samples, density_out and chosen_out are device pointers (the
// samples already in device memory, the outputs written there); otherwise
// host pointers staged through pinned memory.
static int estimate_batch_impl(pkde_ctx* ctx, int32_t count, const pkde_request* reqs,
                               int32_t* status_out, bool device_io) {
  if (!ctx || count < 0 || (count > 0 && !reqs)) return fail(PKDE_INVALID_ARGUMENT, "invalid argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = check_hip(hipSetDevice(ctx->device), "hipSetDevice");
  if (rc) return rc;
  (void)hipGetLastError();
  int nl = 8;
  if (const char* e = getenv("PKDE_BATCH_STREAMS")) nl = std::max(1, std::min(16, atoi(e)));
  while ((int)ctx->lanes.size() < nl) {
    pkde_batch_lane ln;
    rc = check_hip(hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking), "stream");
    if (rc) return rc;
    ctx->lanes.push_back(ln);
  }
  std::vector<int> st(count, PKDE_OK);
  std::vector<Footprint> fp(count);
  for (int i = 0; i < count; ++i) {
    st[i] = check_request(reqs[i]);
    if (st[i] == PKDE_OK) {
      fp[i] = footprint_of(ctx, reqs[i], reqs[i].n, ctx->lane_chunk_bytes);
      st[i] = fp[i].status;
    }
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
    const int64_t G = st[i] == PKDE_OK && !device_io ? fp[i].G : 0;
    const int64_t n = st[i] == PKDE_OK && !device_io ? r.n : 0;
    off_x[i] = take((size_t)n * r.ndims * 8);
    off_d[i] = take((size_t)G * 8);
    off_c[i] = take((size_t)G * 2);
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
  std::vector<int> order;
  for (int i = 0; i < count; ++i)
    if (st[i] == PKDE_OK) order.push_back(i);
  std::sort(order.begin(), order.end(), [&](int x, int y) { return fp[x].work > fp[y].work; });
  std::vector<int> lane_of(count, 0);
  for (auto& ln : ctx->lanes) ln.work = 0.0;
  for (int i : order) {
    int best = 0;
    for (int l = 1; l < nl; ++l)
      if (ctx->lanes[l].work < ctx->lanes[best].work) best = l;
    lane_of[i] = best;
    ctx->lanes[best].work += fp[i].work;
  }
  // Every lane is idle here (the previous call synchronised them).  When
  // the plans this batch still has to create do not fit the free memory,
  // the lanes drop the plans this batch does not use.
  std::vector<std::vector<unsigned char>> keys(count);
  std::vector<pkde_plan*> used(count, nullptr);
  {
    uint64_t need = 0;
    std::vector<uint64_t> lane_chunk(nl, 0);
    for (int i : order) {
      keys[i] = key_bytes(reqs[i]);
      bool have = false;
      for (auto& kv : ctx->lanes[lane_of[i]].plans) have = have || kv.first == keys[i];
      if (!have) need += fp[i].plan;
      lane_chunk[lane_of[i]] = std::max(lane_chunk[lane_of[i]], fp[i].chunk);
    }
    for (int l = 0; l < nl; ++l)
      if (lane_chunk[l] > ctx->lanes[l].Y_bytes) need += lane_chunk[l] - ctx->lanes[l].Y_bytes;
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    if (need + (1ull << 30) > fr) {
      for (int l = 0; l < nl; ++l) {
        auto& pls = ctx->lanes[l].plans;
        for (size_t j = 0; j < pls.size();) {
          bool in_batch = false;
          for (int i : order) in_batch = in_batch || (lane_of[i] == l && keys[i] == pls[j].first);
          if (in_batch) {
            ++j;
            continue;
          }
          pkde_gradepro_plan_destroy(pls[j].second);
          pls.erase(pls.begin() + j);
        }
      }
    }
    for (int l = 0; l < nl; ++l) {
      pkde_batch_lane& ln = ctx->lanes[l];
      const int e = grow_chunk(ln.Y, ln.Y_bytes, lane_chunk[l], "lane chunk buffer");
      if (e)
        for (int i : order)
          if (lane_of[i] == l) st[i] = e;
    }
  }
  // plans, reserved for the largest request of their shape in this batch
  for (int i : order) {
    if (st[i]) continue;
    pkde_batch_lane& ln = ctx->lanes[lane_of[i]];
    pkde_plan* pl = nullptr;
    for (auto& kv : ln.plans)
      if (kv.first == keys[i]) pl = kv.second;
    if (!pl) {
      st[i] = make_lane_plan(ctx, reqs[i], &pl, ctx->lane_chunk_bytes);
      if (st[i]) continue;
      ln.plans.push_back({keys[i], pl});
    }
    if (pl->g.N >= 2) plan_borrow_chunk(pl, ln.Y, ln.Y_bytes);
    used[i] = pl;
  }
  for (int i : order) {
    if (st[i]) continue;
    int64_t nmax = reqs[i].n;
    for (int j : order)
      if (used[j] == used[i]) nmax = std::max<int64_t>(nmax, reqs[j].n);
    st[i] = plan_reserve(used[i], nmax, false);
  }
  for (int i : order) {
    if (st[i]) continue;
    const pkde_request& r = reqs[i];
    pkde_plan* pl = used[i];
    hipStream_t s = ctx->lanes[lane_of[i]].stream;
    pl->p.rng_seed = r.params.rng_seed;
    if (device_io) {
      st[i] = plan_enqueue(pl, r.samples, r.n, nullptr, r.density_out,
                           r.chosen_out ? r.chosen_out : pl->d.chosen, s, false);
      if (!st[i])
        st[i] = check_hip(hipMemcpyAsync(ctx->stage + off_e[i], pl->d.err, 4, hipMemcpyDeviceToHost, s),
                          "flag");
      continue;
    }
    memcpy(ctx->stage + off_x[i], r.samples, (size_t)r.n * r.ndims * 8);
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
  int first_rc = PKDE_OK;
  for (int i = 0; i < count; ++i) {
    if (st[i] == PKDE_OK && used[i]) {
      const int64_t G = used[i]->g.G;
      int err = 0;
      memcpy(&err, ctx->stage + off_e[i], 4);
      if (err & 1)
        st[i] = fail(PKDE_OUT_OF_GRID, "Data points lie beyond grid.");
      else if (!device_io) {
        memcpy(reqs[i].density_out, ctx->stage + off_d[i], (size_t)G * 8);
        if (reqs[i].chosen_out) memcpy(reqs[i].chosen_out, ctx->stage + off_c[i], (size_t)G * 2);
      }
    }
    if (status_out) status_out[i] = st[i];
    if (st[i] != PKDE_OK && first_rc == PKDE_OK) first_rc = st[i];
  }
  return rc ? rc : first_rc;
}

extern "C" int pkde_gradepro_estimate_batch(pkde_ctx* ctx, int32_t count, const pkde_request* reqs,
                                            int32_t* status_out) {
  return estimate_batch_impl(ctx, count, reqs, status_out, false);
}

extern "C" int pkde_gradepro_estimate_batch_device(pkde_ctx* ctx, int32_t count,
                                                   const pkde_request* reqs, int32_t* status_out) {
  return estimate_batch_impl(ctx, count, reqs, status_out, true);
}

extern "C" int pkde_gradepro_request_footprint(pkde_ctx* ctx, const pkde_request* req,
                                               uint64_t* bytes_out) {
  if (!ctx || !req || !bytes_out) return fail(PKDE_INVALID_ARGUMENT, "null argument");
  const int rc = check_request(*req);
  if (rc) return rc;
  const Footprint f = footprint_of(ctx, *req, req->n, ctx->lane_chunk_bytes);
  if (f.status) return f.status;
  *bytes_out = f.plan + f.chunk;
  return PKDE_OK;
}

// ------------------------------------------------------------------ server
struct Job {
  uint64_t ticket = 0;
  pkde_request req{};
  PlanKey key{};
  std::vector<double> samples;
  int status = PKDE_PENDING;
  bool done = false;
  double t_submit = 0, t_done = 0;
  Footprint fp;
  int64_t n_cap = 0;  // samples the plan is reserved for (a power of two)
  int bypassed = 0;   // times a later, smaller request was admitted first
};

// Admission order: among the first kLookahead queued requests, the one with
// the least estimated device work goes first (short requests do not wait
// behind long ones), unless a request ahead of it has already been passed
// over kMaxBypass times -- then that one goes, so no request starves.
static constexpr int kLookahead = 16;
static constexpr int kMaxBypass = 16;
// A request whose ladder alone fills the GPU (G S B at least this: 256^2
// grids in 2-D, 64^3 and up in 3-D) runs with no other such request beside
// it; smaller requests fill the other slots.  Two GPU-filling ladders side by
// side only share the bandwidth and their chunk buffers (60-request stream:
// 0.986x the plain path's throughput with 4 slots of any request, 0.80x with
// 8, 1.008x with 1).
static constexpr double kFillWork = 1e9;

static bool fills_gpu(const Job& j) { return j.fp.ladder_work >= kFillWork; }

struct PoolPlan {
  pkde_plan* pl = nullptr;
  PlanKey key{};
  uint64_t bytes = 0;  // resident device bytes (chunk buffer excluded)
  int64_t n_cap = 0;
  bool busy = false;
  double last_used = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  void* Y = nullptr;  // the slot's chunk buffer, lent to the plan it runs
  uint64_t Y_bytes = 0;
  PoolPlan* pp = nullptr;
  std::shared_ptr<Job> job;
  int* err_host = nullptr;  // pinned
  double* dens_host = nullptr;
  int16_t* chosen_host = nullptr;
  int64_t host_cap = 0;
  double* x_host = nullptr;  // pinned samples staging
  int64_t x_cap = 0;
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
  std::vector<std::unique_ptr<PoolPlan>> pool;  // worker thread only
  uint64_t next_ticket = 1;
  uint64_t completed = 0, rejected = 0;
  uint64_t bytes_in_use = 0;  // pool plans + slot chunk buffers (read under mu)
  uint64_t slot_chunk = 0;    // chunk budget of a slot's plans
  std::vector<double> lat;
  bool stop = false;
};

// Pooled plans are reserved for at least this many samples, and the slots'
// pinned buffers are sized for this many grid points and samples up front:
// growing a device buffer frees the old one, and a free synchronises the
// whole device, which would stall every slot in flight.
static constexpr int64_t kServeMinSamples = 1 << 20;
static constexpr int64_t kServeMinPoints = 1 << 21;

static uint64_t resident_bytes(const pkde_server* sv) {
  uint64_t b = 0;
  for (const auto& p : sv->pool) b += p->bytes;
  for (const auto& sl : sv->slots) b += sl.Y_bytes;
  return b;
}

static void publish_bytes(pkde_server* sv) {
  const uint64_t b = resident_bytes(sv);
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->bytes_in_use = b;
}

static void complete(pkde_server* sv, const std::shared_ptr<Job>& job, int status, bool rejected) {
  std::lock_guard<std::mutex> lk(sv->mu);
  job->status = status;
  job->done = true;
  job->t_done = now_ms();
  if (status == PKDE_OK) {
    sv->lat.push_back(job->t_done - job->t_submit);
    sv->completed++;
  }
  if (rejected) sv->rejected++;
  sv->done_cv.notify_all();
}

static void finish_job(pkde_server* sv, Slot& sl) {
  auto job = sl.job;
  const int64_t G = sl.pp->pl->g.G;
  int st = PKDE_OK;
  if (*sl.err_host & 1) st = fail(PKDE_OUT_OF_GRID, "Data points lie beyond grid.");
  if (st == PKDE_OK) {
    memcpy(job->req.density_out, sl.dens_host, (size_t)G * 8);
    if (job->req.chosen_out) memcpy(job->req.chosen_out, sl.chosen_host, (size_t)G * 2);
  }
  sl.pp->busy = false;
  sl.pp->last_used = now_ms();
  sl.pp = nullptr;
  sl.job.reset();
  complete(sv, job, st, false);
}

// Frees idle pool plans, least recently used first (never `keep`), then the
// chunk buffers of idle slots (never `keep_slot`'s; the plans that borrowed
// one borrow their next slot's buffer when they start), until `need` more
// bytes fit the budget.  True when they fit.
static bool make_room(pkde_server* sv, uint64_t need, const PoolPlan* keep,
                      const Slot* keep_slot) {
  for (;;) {
    if (resident_bytes(sv) + need <= sv->cfg.memory_budget) return true;
    int victim = -1;
    for (int i = 0; i < (int)sv->pool.size(); ++i) {
      const PoolPlan* p = sv->pool[i].get();
      if (p->busy || p == keep) continue;
      if (victim < 0 || p->last_used < sv->pool[victim]->last_used) victim = i;
    }
    if (victim >= 0) {
      pkde_gradepro_plan_destroy(sv->pool[victim]->pl);
      sv->pool.erase(sv->pool.begin() + victim);
      continue;
    }
    Slot* idle = nullptr;
    for (auto& o : sv->slots)
      if (&o != keep_slot && !o.job && o.Y_bytes > 0 && (!idle || o.Y_bytes > idle->Y_bytes))
        idle = &o;
    if (!idle) return false;
    (void)hipFree(idle->Y);
    idle->Y = nullptr;
    idle->Y_bytes = 0;
  }
}

static int grow_host(Slot& sl, int64_t G, int64_t nx) {
  int st = PKDE_OK;
  if (G > sl.host_cap) {
    if (sl.dens_host) (void)hipHostFree(sl.dens_host);
    if (sl.chosen_host) (void)hipHostFree(sl.chosen_host);
    sl.dens_host = nullptr;
    sl.chosen_host = nullptr;
    sl.host_cap = 0;
    if ((st = check_hip(hipHostMalloc((void**)&sl.dens_host, (size_t)G * 8, hipHostMallocDefault), "pinned")))
      return st;
    if ((st = check_hip(hipHostMalloc((void**)&sl.chosen_host, (size_t)G * 2, hipHostMallocDefault), "pinned")))
      return st;
    sl.host_cap = G;
  }
  if (nx > sl.x_cap) {
    if (sl.x_host) (void)hipHostFree(sl.x_host);
    sl.x_host = nullptr;
    sl.x_cap = 0;
    if ((st = check_hip(hipHostMalloc((void**)&sl.x_host, (size_t)nx * 8, hipHostMallocDefault), "pinned")))
      return st;
    sl.x_cap = nx;
  }
  return PKDE_OK;
}

// Tries to start the job on the free slot.  Returns 1 when started, 0 when it
// must wait for memory, and -status when it failed.
static int start_job(pkde_server* sv, Slot& sl, const std::shared_ptr<Job>& job) {
  const pkde_request& r = job->req;
  // an idle plan of the request's shape, the one reserved for the most
  // samples first
  PoolPlan* pp = nullptr;
  for (auto& p : sv->pool)
    if (!p->busy && p->key == job->key && (!pp || p->n_cap > pp->n_cap)) pp = p.get();
  uint64_t need = sl.Y_bytes < job->fp.chunk ? job->fp.chunk - sl.Y_bytes : 0;
  if (!pp)
    need += job->fp.plan;
  else if (pp->n_cap < job->n_cap)
    need += job->fp.plan > pp->bytes ? job->fp.plan - pp->bytes : 0;
  if (!make_room(sv, need, pp, &sl)) return 0;
  int st = PKDE_OK;
  const uint64_t plan_need = need - (sl.Y_bytes < job->fp.chunk ? job->fp.chunk - sl.Y_bytes : 0);
  const uint64_t pref = std::max(job->fp.chunk_own, job->fp.chunk);
  if (sl.Y_bytes < pref &&
      resident_bytes(sv) + (pref - sl.Y_bytes) + plan_need <= sv->cfg.memory_budget) {
    // the larger of the slot budget's chunk buffer and the one a plan of its
    // own would take (rungs of 128^3: 16 to a chunk rather than 8), when it
    // fits beside the resident set as it is
    if ((st = grow_chunk(sl.Y, sl.Y_bytes, pref, "slot chunk buffer"))) return -st;
  } else if (sl.Y_bytes < job->fp.chunk) {
    // grow to the largest chunk buffer any slot holds (up to the slot
    // budget's size), so that the slots reach their final size in few growths
    uint64_t want = job->fp.chunk;
    for (const auto& o : sv->slots) want = std::max(want, std::min(o.Y_bytes, sv->slot_chunk));
    if (!make_room(sv, want - sl.Y_bytes + plan_need, pp, &sl)) want = job->fp.chunk;
    if ((st = grow_chunk(sl.Y, sl.Y_bytes, want, "slot chunk buffer"))) return -st;
  }
  if (!pp) {
    auto np = std::make_unique<PoolPlan>();
    pkde_plan* pl = nullptr;
    if ((st = make_lane_plan(sv->ctx, r, &pl, sv->slot_chunk))) return -st;
    np->pl = pl;
    np->key = job->key;
    pp = np.get();
    sv->pool.push_back(std::move(np));
  }
  pkde_plan* pl = pp->pl;
  // the slot's chunk buffer first, so that reserving never allocates one
  if (pl->g.N >= 2) plan_borrow_chunk(pl, sl.Y, sl.Y_bytes);
  if (pp->n_cap < job->n_cap) {
    if ((st = plan_reserve(pl, job->n_cap, false))) return -st;
    pp->n_cap = job->n_cap;
    pp->bytes = std::max(pp->bytes, job->fp.plan);
  }
  const int64_t G = pl->g.G;
  if ((st = grow_host(sl, G, r.n * r.ndims))) return -st;
  pp->busy = true;
  publish_bytes(sv);
  hipStream_t s = sl.stream;
  pl->p.rng_seed = r.params.rng_seed;
  memcpy(sl.x_host, job->samples.data(), (size_t)r.n * r.ndims * 8);
  st = check_hip(hipMemcpyAsync(pl->d.x, sl.x_host, (size_t)r.n * r.ndims * 8, hipMemcpyHostToDevice, s),
                 "upload");
  if (!st) st = plan_enqueue(pl, pl->d.x, r.n, nullptr, pl->d.density, pl->d.chosen, s, false);
  if (!st)
    st = check_hip(hipMemcpyAsync(sl.dens_host, pl->d.density, (size_t)G * 8, hipMemcpyDeviceToHost, s), "dl");
  if (!st)
    st = check_hip(hipMemcpyAsync(sl.chosen_host, pl->d.chosen, (size_t)G * 2, hipMemcpyDeviceToHost, s), "dl");
  if (!st) st = check_hip(hipMemcpyAsync(sl.err_host, pl->d.err, 4, hipMemcpyDeviceToHost, s), "dl");
  if (!st) st = check_hip(hipEventRecord(sl.done, s), "event");
  if (st) {
    // the stream may still read the plan's buffers: wait before reusing it
    (void)hipStreamSynchronize(s);
    pp->busy = false;
    return -st;
  }
  sl.pp = pp;
  sl.job = job;
  return 1;
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
    // admissions (continuous batching), shortest estimated work first
    // within the look-ahead window, with bounded bypass
    for (auto& sl : sv->slots) {
      if (sl.job) continue;
      std::shared_ptr<Job> job;
      size_t pick = 0;
      {
        bool filling = false;  // a GPU-filling request is in flight
        for (const auto& o : sv->slots) filling = filling || (o.job && fills_gpu(*o.job));
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->queue.empty()) break;
        const size_t win = std::min<size_t>(sv->queue.size(), kLookahead);
        bool found = false;
        for (size_t i = 0; i < win; ++i) {
          const Job& c = *sv->queue[i];
          if (filling && fills_gpu(c)) continue;  // waits for the one in flight
          if (c.bypassed >= kMaxBypass) {
            pick = i;
            found = true;
            break;
          }
          if (!found || c.fp.work < sv->queue[pick]->fp.work) pick = i;
          found = true;
        }
        if (!found) break;  // only GPU-filling requests wait in the window
        job = sv->queue[pick];
      }
      auto dequeue = [&] {
        std::lock_guard<std::mutex> lk(sv->mu);
        for (size_t i = 0; i < pick; ++i) sv->queue[i]->bypassed++;
        sv->queue.erase(sv->queue.begin() + pick);
      };
      if (job->fp.plan + job->fp.chunk > sv->cfg.memory_budget) {
        {
          std::lock_guard<std::mutex> lk(sv->mu);
          sv->queue.erase(sv->queue.begin() + pick);
        }
        complete(sv, job, PKDE_OUT_OF_MEMORY, true);
        progressed = true;
        continue;
      }
      const int r = start_job(sv, sl, job);
      if (r == 0) {
        bool busy = false;
        for (auto& o : sv->slots) busy = busy || (bool)o.job;
        if (busy) break;  // wait until running requests release memory
      }
      dequeue();
      if (r <= 0) complete(sv, job, r < 0 ? -r : PKDE_OUT_OF_MEMORY, r == 0);
      progressed = true;
    }
    publish_bytes(sv);
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
  // A slot's plans take chunk buffers of up to half the budget over the
  // slots (at least the batched form's lane budget): rungs of 3.4 GB (128^3)
  // then run 8 to a chunk rather than 4, so the spectra are read half as
  // often, and the slots' buffers together leave half the budget to plans.
  sv->slot_chunk = ctx->lane_chunk_bytes;
  if (!ctx->chunk_fixed)
    sv->slot_chunk = std::max<uint64_t>(sv->slot_chunk,
                                        sv->cfg.memory_budget / (2 * (uint64_t)sv->cfg.max_inflight));
  sv->slots.resize(sv->cfg.max_inflight);
  for (auto& sl : sv->slots) {
    if (!rc) rc = check_hip(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "stream");
    if (!rc) rc = check_hip(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "event");
    if (!rc) rc = check_hip(hipHostMalloc((void**)&sl.err_host, 64, hipHostMallocDefault), "pinned");
    if (!rc) rc = grow_host(sl, kServeMinPoints, kServeMinSamples * 3);
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
  if (!sv || !req || !ticket) return fail(PKDE_INVALID_ARGUMENT, "invalid request");
  int st = check_request(*req);
  if (st) return st;
  auto job = std::make_shared<Job>();
  job->req = *req;
  job->key = key_of(*req);
  // footprint on the host, for a plan reserved for the next power of two of
  // samples (plans of a shape are reused by later requests)
  job->n_cap = std::max<int64_t>(pow2_ceil(req->n), kServeMinSamples);
  job->fp = footprint_of(sv->ctx, *req, job->n_cap, sv->slot_chunk);
  if (job->fp.status) return job->fp.status;
  job->samples.assign(req->samples, req->samples + (size_t)req->n * req->ndims);
  job->req.samples = job->samples.data();
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
  if (st == PKDE_OUT_OF_MEMORY) set_error("Not enough memory for the request within the server's memory budget.");
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
  if (job->status == PKDE_OUT_OF_MEMORY)
    set_error("Not enough memory for the request within the server's memory budget.");
  else if (job->status == PKDE_OUT_OF_GRID)
    set_error("Data points lie beyond grid. Remove them or increase grid boundaries.");
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
  for (auto& sl : sv->slots)
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  for (auto& p : sv->pool) pkde_gradepro_plan_destroy(p->pl);
  sv->pool.clear();
  for (auto& sl : sv->slots) {
    if (sl.Y) (void)hipFree(sl.Y);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.err_host) (void)hipHostFree(sl.err_host);
    if (sl.dens_host) (void)hipHostFree(sl.dens_host);
    if (sl.chosen_host) (void)hipHostFree(sl.chosen_host);
    if (sl.x_host) (void)hipHostFree(sl.x_host);
  }
  delete sv;
}

}  // extern "C"
