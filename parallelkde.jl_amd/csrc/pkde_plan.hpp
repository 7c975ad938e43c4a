// Host-side plan of one :gradepro estimate on one device (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/parallelkde_hip.h"
#include "pkde_kernels.hpp"

struct pkde_plan;

// Stream lane of the batched form: its own stream and the plans (device
// buffers) it keeps resident between calls, keyed by grid and parameters.
namespace pkde {
// Launch shape chosen by the autotuner: ladder rungs per chunk (0: as many
// as the chunk buffer holds), the ladder axis pass's workgroup target, and
// the cells per tile of the sorted binning.
struct TuneShape {
  int rungs = 0;
  int split = kAxisSplitTarget;
  int tile = kBinTileCells;
  bool operator==(const TuneShape& o) const {
    return rungs == o.rungs && split == o.split && tile == o.tile;
  }
};
}  // namespace pkde

struct pkde_batch_lane {
  hipStream_t stream = nullptr;
  std::vector<std::pair<std::vector<unsigned char>, pkde_plan*>> plans;
  double work = 0.0;
  void* Y = nullptr;  // the lane's rung-chunk buffer, lent to its plans
  uint64_t Y_bytes = 0;
};

struct pkde_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  double* tw64 = nullptr;  // exp(-2 pi i k / L) for L = 2..4096 at offset L-2
  float* tw32 = nullptr;
  std::mutex mu;
  uint64_t chunk_bytes = 16ull << 30;  // per-plan budget of the rung-chunk buffer
  bool chunk_fixed = false;           // budget set by PKDE_CHUNK_BYTES
  // budget of one stream lane's (batched form) or server slot's shared
  // chunk buffer: the plain plan's, so a request runs the same chunks in
  // every form
  uint64_t lane_chunk_bytes = 16ull << 30;
  uint64_t mem_total = 0;             // device memory, bytes
  // batched form
  std::vector<pkde_batch_lane> lanes;
  unsigned char* stage = nullptr;  // pinned host staging
  size_t stage_bytes = 0;
  // live plans: destroying the context orphans them (plan->ctx = nullptr), so
  // a plan destroyed afterwards frees its own memory without touching this.
  // Guarded by its own lock: the batched form creates and destroys its lane
  // plans while holding `mu`.
  std::mutex live_mu;
  std::vector<pkde_plan*> live;
  // autotuned launch shapes, keyed by (kind, N, B, S, precision, axis
  // sizes); tune_gen counts the entries so plans look the table up again
  // only after a change
  std::mutex tune_mu;
  std::vector<std::pair<std::vector<int64_t>, pkde::TuneShape>> tune;
  std::atomic<int> tune_gen{0};
  // this device's copy of the library's static scratch array (the
  // allocation-free entry point), resolved once when the context is created
  unsigned char* scratch = nullptr;
};

namespace pkde {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

struct PlanBuffers {
  double* x = nullptr;
  int64_t x_cap = 0;
  int32_t* idx = nullptr;
  int64_t idx_cap = 0;
  void* acc = nullptr;            // (2B+1) x G int64 / double
  cplx<double>* specD = nullptr;  // (2B+1) x H1 x Q
  void* spec = nullptr;           // FP32 copy (precision 0)
  void* Y = nullptr;              // rung-chunk buffer
  int64_t Y_elems = 0;
  // Y belongs to the stream lane or server slot the plan runs on (plans of
  // one stream run one after another, so they share one chunk buffer); the
  // plan neither allocates nor frees it
  bool Y_borrowed = false;
  double* vmr = nullptr;  // S x G
  double* conv = nullptr;  // S x G x 2 convergence terms (small grids only)
  int16_t* chosen = nullptr;
  int16_t* sel = nullptr;
  double* density = nullptr;
  int* err = nullptr;
  double* bw = nullptr;
  double* mean = nullptr;
  double* partial = nullptr;
  double* kfreq = nullptr;
  double* times = nullptr;
  double* dlogt = nullptr;
  double* vscale = nullptr;
  double* cV = nullptr;
  double* params = nullptr;
  void* gM = nullptr;
  void* gV = nullptr;
  double* invk = nullptr;  // Welford weights 1 / (b + 1)
  double* twg64 = nullptr;  // per-axis twiddles exp(-2 pi i k / n_i) at koff[i]
  float* twg32 = nullptr;
  void* binwork = nullptr;  // workspace of the sorted binning
  uint64_t binwork_bytes = 0;
  int S_alloc = 0;
};

// With graphs enabled (pkde_gradepro_graph_enable), repeated estimates with
// the same device buffers replay three captured graphs (binning and forward
// transforms; the ladder; convergence and means) instead of launching each
// kernel again.  A call captures when it repeats
// the previous call's key; `sig` changes when any plan buffer is
// reallocated, so replayed kernels never see a stale pointer.
struct GraphKey {
  const void* x = nullptr;
  int64_t n = 0;
  const void* idx = nullptr;
  const void* dens = nullptr;
  const void* chosen = nullptr;
  uint64_t sig = 0;
  bool timed = false;
  bool operator==(const GraphKey& o) const {
    return x == o.x && n == o.n && idx == o.idx && dens == o.dens && chosen == o.chosen &&
           sig == o.sig && timed == o.timed;
  }
};

struct PlanGraphs {
  bool enabled = false;  // opt-in: direct launches were measured faster
  bool failed = false;  // a capture failed: this plan stays on direct launches
  GraphKey last, key;
  hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
};

}  // namespace pkde

struct pkde_plan {
  pkde_ctx* ctx = nullptr;
  int kind = 0;      // 0: :gradepro, 1: :rot (rule-of-thumb)
  int rot_rule = 0;  // :rot: 0 Scott, 1 Silverman
  pkde::GridDev g{};
  pkde_gradepro_params p{};
  int B = 100;
  int S = 251;            // rungs (n_steps + 1)
  int steps_mode = 1;     // 1: n_steps, 0: bw_step
  // bw_step with the Silverman bw_final: S depends on the samples and is set
  // by each estimate (the plan keeps S = 251 for footprints until then)
  bool S_from_data = false;
  double steps_or_step = 250.0;
  bool bw_given = false;
  double bw_host[PKDE_MAX_DIMS] = {0, 0, 0};
  double alpha_s = 0, alpha_os = 0.1, eps = 2.0;
  int prec = 0;
  int fixed = 1;
  bool debug = false;
  pkde::PlanBuffers d;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int64_t last_n = 0;
  const double* last_x = nullptr;     // samples and draws of the last estimate
  const int32_t* last_idx = nullptr;
  int last_rc = PKDE_OK;
  pkde::PlanGraphs graphs;
  // launch shape (pkde_gradepro_autotune); the context table generation and
  // S it was looked up for
  pkde::TuneShape tune;
  int tune_gen = -1;
  int tune_S = -1;
  // chunk budget of this plan (0: the context's); lane and slot plans take
  // the lane budget
  uint64_t chunk_budget = 0;
};

namespace pkde {
// Enqueues one estimate of n device-resident samples; no host synchronisation.
int plan_enqueue(pkde_plan* pl, const double* d_x, int64_t n, const int32_t* d_idx,
                 double* d_density, int16_t* d_chosen, hipStream_t s, bool timed);
int plan_reserve(pkde_plan* pl, int64_t n, bool need_idx);
// Device bytes of a plan for n samples; with_chunk = false leaves out the
// rung-chunk buffer (a lane or slot buffer lent to the plan)
uint64_t plan_bytes(const pkde_plan* pl, int64_t n, bool with_chunk = true);
// Bytes of the rung-chunk buffer the plan uses (0 on 1-D grids)
uint64_t plan_chunk_bytes(const pkde_plan* pl);
// Host-side part of plan creation: validates the grid and keywords and fills
// every host field; allocates nothing on the device.  The plan's device
// tables are then made by plan_device_init.
int plan_setup_host(pkde_ctx* ctx, int ndims, const int64_t* n_per_axis, const double* lo,
                    const double* hi, const double* h, const pkde_gradepro_params* params,
                    pkde_plan* pl);
// Lends a stream's chunk buffer to the plan (bytes >= plan_chunk_bytes()).
void plan_borrow_chunk(pkde_plan* pl, void* Y, uint64_t bytes);
}  // namespace pkde
