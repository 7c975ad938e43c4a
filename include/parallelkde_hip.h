/*
 * libparallelkde_hip -- C ABI of the MI355X (gfx950) :gradepro density
 * estimator.  Plain C types only: no torch, HIP or C++ types cross it.
 *
 * It is the drop-in for ParallelKDE.jl's `estimate_density!(est, :gradepro)`
 * on a grid.  The Julia glue behind a new `:rocm` device calls it through
 * ccall; Python reaches it through ctypes.  Citations are relative to the
 * ParallelKDE.jl v1.0.1 source tree.
 *
 * Layouts follow Julia's column-major convention, so the glue never
 * transposes anything:
 *   samples       ndims x n, coordinate fastest (KDEs.jl:76-93 keeps the same
 *                 order as Vector{SVector{N}}).
 *   bootstrap_idx n x B, draw fastest, 0-based (KDEs.jl:182-191 returns the
 *                 1-based n x B matrix; the glue subtracts 1).
 *   density,
 *   chosen_rung   grid axis 1 fastest (Julia Array{T,N} order).
 *
 * Status codes map onto the reference's error behaviour:
 *   PKDE_INVALID_ARGUMENT, PKDE_OUT_OF_GRID   -> ArgumentError
 *       (DirectSpace.jl:489-492, GradeProEstimator.jl:712-714, 891-893, 955-960)
 *   PKDE_NO_DEVICE                           -> @warn + fall back to :cpu
 *       (ParallelKDE.jl:123-127)
 *   PKDE_OUT_OF_MEMORY                       -> ArgumentError("Not enough memory...")
 *       (GradeProEstimator.jl:1105-1108)
 * pkde_last_error() returns a thread-local message for the last failure.
 */
#ifndef PARALLELKDE_HIP_H
#define PARALLELKDE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PKDE_ABI_VERSION 2
#define PKDE_MAX_DIMS 6

typedef enum {
  PKDE_OK = 0,
  PKDE_INVALID_ARGUMENT = 1,
  PKDE_OUT_OF_GRID = 2,
  PKDE_NO_DEVICE = 3,
  PKDE_OUT_OF_MEMORY = 4,
  PKDE_DEVICE_ERROR = 5,
  PKDE_UNSUPPORTED = 6,
  PKDE_PENDING = 7
} pkde_status;

typedef struct pkde_ctx pkde_ctx;
typedef struct pkde_plan pkde_plan;
typedef struct pkde_server pkde_server;

/* Keyword arguments of estimate_density!(est, :gradepro; ...)
 * (docs/src/estimators/gradepro.md:19-28, GradeProEstimator.jl:700-796).
 * NaN / <= 0 select the reference default. */
typedef struct {
  int32_t n_bootstraps;                /* 100 (:705) */
  int32_t n_steps;                     /* 250 (:902); <= 0 -> bw_step if finite */
  double bw_step;                      /* NaN: unset (:887-899); with the
                                          Silverman bw_final the ladder length
                                          is set per estimate from the samples
                                          (that estimate waits for them) */
  double bw_final[PKDE_MAX_DIMS];      /* NaN: Silverman rule (:771-775) */
  double alpha_s;                      /* NaN: 0.0 (1-D) / 0.06 (>=2-D) (:754-763) */
  double alpha_os;                     /* NaN: 0.1 (:764-766) */
  double eps;                          /* NaN: 2.0 (1-D) / -0.5 (>=2-D) (:541-550) */
  uint64_t rng_seed;                   /* Philox key for in-kernel bootstrap draws */
  int32_t deterministic_binning;       /* 1: fixed-point integer deposits (default) */
  int32_t precision;                   /* 1: FP64 throughout (default)
                                          0: FP32 spectra + ladder FFTs, FP64
                                             binning/forward FFT/statistics */
} pkde_gradepro_params;

/* Timing / accounting of one estimate (optional output). */
typedef struct {
  int32_t n_rungs;
  int32_t steps_low, steps_over;
  double eps;
  double bw_final[PKDE_MAX_DIMS];
  double ms_total;          /* device time of the whole estimate */
  double ms_ladder;         /* device time of the ladder kernels */
  double algorithmic_bytes; /* U*(8B+4) + n*N*8 + (2B+1)*G*4 + G*6 */
} pkde_stats;

int pkde_version(void);
const char *pkde_last_error(void);
void pkde_gradepro_params_default(pkde_gradepro_params *p);

/* One context per device: owns the stream and the autotune cache.
 * Replaces the reference's device registry entry (Devices.jl:54-82) for the
 * new :rocm device.  Returns PKDE_NO_DEVICE without a usable GPU. */
int pkde_context_create(int device, pkde_ctx **out);
void pkde_context_destroy(pkde_ctx *ctx);
/* Free device memory in bytes (Devices.jl:103-108, get_available_memory). */
int pkde_device_free_memory(pkde_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes);

/* Grid bounds from device-resident samples: the device side of find_grid
 * (Grids.jl:252-300).  Per axis, the extrema of the samples (Grids.jl:271-273)
 * widened on both sides by padding[i], or by 0.1 * |max - min| when padding
 * is NULL (Grids.jl:274-281).  d_samples is ndims x n in device memory,
 * coordinate fastest.  The caller picks the point counts (the reference's
 * default is 150 per axis, Grids.jl:284) and the spacing (hi - lo) / (n - 1)
 * of range(lo, stop=hi, length=n).  PKDE_INVALID_ARGUMENT for non-finite
 * samples. */
int pkde_find_grid(pkde_ctx *ctx, int ndims, const double *d_samples, int64_t n,
                   const double *padding, double *lo_out, double *hi_out);

/* A plan binds a grid (Grid: Grids.jl:67-87 -- per-axis lo, hi, spacing and
 * point count) and the estimator keywords.  Axis sizes must lie in [2, 4096]
 * (any factorization: prime factors above 13 take direct-DFT passes; other
 * sizes return PKDE_UNSUPPORTED); ndims in 1..PKDE_MAX_DIMS (the reference's CPU path
 * takes any dimension; its GPU path stops at 3-D, GradeProEstimator.jl:811-814).  Replaces initialize_estimator
 * (GradeProEstimator.jl:700-735) for the :rocm device. */
int pkde_gradepro_plan_create(pkde_ctx *ctx, int ndims, const int64_t *n_per_axis,
                              const double *lo, const double *hi, const double *h,
                              const pkde_gradepro_params *params, pkde_plan **out);
void pkde_gradepro_plan_destroy(pkde_plan *plan);
/* Device bytes an estimate of n samples needs (extends get_necessary_memory,
 * GradeProEstimator.jl:1091-1096). */
uint64_t pkde_gradepro_plan_footprint(const pkde_plan *plan, int64_t n_samples);

/* Rule-of-thumb plan: the :rot estimator (RoTEstimator.jl:116-265 --
 * initialize_estimator, estimate!): full-sample binning, forward FFT, one
 * Gaussian propagation by t_i = factor * sigma_i - h_i / 2, normalised
 * inverse FFT, |.|.  rule: 0 Scott (the reference's default,
 * RoTEstimator.jl:120), 1 Silverman; precision as in pkde_gradepro_params.
 * Estimate with pkde_gradepro_estimate / _estimate_device (bootstrap_idx is
 * ignored; the host form writes chosen rung 0; the device form does not
 * write chosen); pkde_gradepro_last_stats reports the propagation time per
 * axis in bw_final. */
int pkde_rot_plan_create(pkde_ctx *ctx, int ndims, const int64_t *n_per_axis, const double *lo,
                         const double *hi, const double *h, int rule, int precision,
                         pkde_plan **out);

/* The plain form: estimate!(::GradeProEstimator, kde) (GradeProEstimator.jl:
 * 968-1051) + replace_nans! (:1053-1072).  Synchronous; copies the host
 * samples in, writes caller-owned host outputs.  bootstrap_idx may be NULL
 * (in-kernel Philox draws keyed by rng_seed).  chosen_rung_out (nullable):
 * rung 0..S-1 of the last density assignment, or S where the final-rung
 * fallback was used.  stats (nullable). */
int pkde_gradepro_estimate(pkde_plan *plan, const double *samples, int64_t n,
                           const int32_t *bootstrap_idx, double *density_out,
                           int16_t *chosen_rung_out, pkde_stats *stats);

/* Same on device-resident buffers, asynchronous on `stream` (a hipStream_t,
 * NULL: the context stream).  Nothing is copied to or from the host.
 * Out-of-grid samples are reported through pkde_gradepro_check(). */
int pkde_gradepro_estimate_device(pkde_plan *plan, const double *d_samples, int64_t n,
                                  const int32_t *d_bootstrap_idx, double *d_density,
                                  int16_t *d_chosen_rung, void *stream);
/* Allocation-free estimate.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same estimate
 * as pkde_gradepro_estimate_device on a plan with this grid and keywords, on
 * device-resident samples, asynchronous on `stream` (NULL: the context
 * stream); every working buffer comes from a file-scope static device array
 * of 32 MiB per GPU, held until the process ends.  It calls no hipMalloc,
 * hipFree or allocating library routine.  Limits: 1-D power-of-two grids of
 * 16 to 1024 points (grid lo, hi, spacing h), 2 to 131072 samples, an
 * explicit n_steps, FP64 with fixed-point binning, and a workspace that fits
 * the array (B = 100 and S = 251 take about 26 MiB at 1024 points); others
 * return PKDE_UNSUPPORTED / PKDE_INVALID_ARGUMENT / PKDE_OUT_OF_MEMORY.
 * Calls share their GPU's array: calls that overlap on one GPU (other
 * streams, other threads) race and must be serialised by the caller; each
 * call clears what it reads before writing it.  d_chosen may be NULL. */
int pkde_gradepro_estimate_1d_noalloc(pkde_ctx *ctx, int64_t n_points, double lo, double hi,
                                      double h, const pkde_gradepro_params *params,
                                      const double *d_samples, int64_t n, double *d_density,
                                      int16_t *d_chosen, void *stream);
/* Waits for the plan's last device estimate; PKDE_OUT_OF_GRID if it saw a
 * sample outside the grid (check_data, DirectSpace.jl:479-498). */
int pkde_gradepro_check(pkde_plan *plan);
/* Timing and ladder scalars of the plan's last estimate (waits for it). */
int pkde_gradepro_last_stats(pkde_plan *plan, pkde_stats *stats);
/* Graph replay of repeated estimates (off by default): when an estimate
 * repeats the previous call's samples, bootstrap indices and output buffers
 * (the same device pointers and n), the library captures the estimate as
 * three HIP graphs and later such calls replay them instead of launching
 * each kernel; results are identical.  0 turns it off and drops the graphs.
 * Off by default because replay was measured slower than the direct
 * launches on this path (1-D 512: 0.288 against 0.268 ms per estimate).  No
 * counterpart in the reference (its CUDA path launches every kernel of every
 * rung, GradeProEstimator.jl:968-1051). */
int pkde_gradepro_graph_enable(pkde_plan *plan, int enable);

/* Autotuner of the ladder's launch shape for the plan's grid, B, S and
 * precision.  Times the ladder of the plan's last estimate (its spectra stay
 * resident; run one estimate first), `reps` times after one untimed run, for
 * each candidate: rungs per chunk (the even splits of the ladder into the
 * fewest chunks the chunk buffer allows and up to three more) times the axis
 * pass's workgroup target (no split, 4000, 8000, 16000); then, on 2-D and
 * 3-D grids whose last estimate took its samples from the plan's own buffer
 * (pkde_gradepro_estimate), the binning's deposit tile (8192, 4096 or 2048
 * cells), timed with the binning and forward transforms.  The fastest is
 * kept for this plan and recorded in the context, where every plan of the
 * same shape (the batched and serving forms included) picks it up.  Later
 * estimates are bit-identical to untuned ones: the per-rung arithmetic does
 * not depend on the launch shape.  best_ms_out / default_ms_out (nullable):
 * ladder time of the choice and of the default shape.  No counterpart in the
 * reference: its CUDA path sizes each launch with the occupancy heuristic
 * (launch_configuration, src/StatisticsPropagation/FourierSpace.jl:285-287). */
int pkde_gradepro_autotune(pkde_plan *plan, int32_t reps, double *best_ms_out,
                           double *default_ms_out);

/* Per-rung log-VMR table of the last estimate (S x G, rung-major), for
 * primitive-level parity (calculate_scaled_vmr!, DirectSpace.jl:526-584).
 * Copies the first min(capacity / G, S) whole rungs (capacity in doubles;
 * less than one rung is PKDE_INVALID_ARGUMENT).  debug_enable is kept for
 * ABI compatibility: the table of the last estimate is always resident. */
int pkde_gradepro_debug_enable(pkde_plan *plan, int enable);
int pkde_gradepro_debug_vmr(pkde_plan *plan, double *vmr_out, int64_t capacity);

/* Convergence state machine alone (identify_convergence! + find_stability,
 * DirectSpace.jl:829-871, 975-1031) over a given S x G log-VMR table with the
 * given per-rung dlogt; writes chosen rung per point (S = fallback). */
int pkde_gradepro_convergence(pkde_ctx *ctx, int64_t G, int32_t S, const double *vmr,
                              const double *dlogt, double eps, int32_t steps_low,
                              int32_t steps_over, int16_t *chosen_out);

/* Linear binning alone (initialize_dirac_sequence + generate_dirac_cpu!,
 * DirectSpace.jl:97-145, 196-242) for primitive-level parity: writes the
 * B x G sums and squares fields (real parts) in FP64. */
int pkde_dirac_sequences(pkde_ctx *ctx, int ndims, const int64_t *n_per_axis,
                         const double *lo, const double *hi, const double *h,
                         const double *samples, int64_t n, const int32_t *bootstrap_idx,
                         int32_t B, int deterministic, double *sums_out, double *squares_out);

/* Forward transform alone (plan_fft! + fourier_statistics!,
 * GradeProEstimator.jl:90-95, FourierSpace.jl:25-80), for primitive-level
 * parity: the unnormalised forward DFT (exp(-2 pi i j k / n) along every
 * axis) of `count` real fields on the plan's grid, with the kernels the
 * estimate runs (two fields per complex axis-1 transform, then the passes
 * along the other axes).  fields: count x G doubles, grid axis 1 fastest;
 * spectra_out: count x (n1/2 + 1) x (G / n1) complex values as interleaved
 * (re, im) doubles, [field][k1][q] with q = k2 + n2 k3 + ... (the Hermitian
 * half: X(n1 - k1, -q) = conj X(k1, q)).  1 <= count <= 2 B + 1 of the plan.
 * Uses the plan's binning and spectrum buffers (the next estimate overwrites
 * them). */
int pkde_fft_forward(pkde_plan *plan, int32_t count, const double *fields, double *spectra_out);

/* ---- batched form: many independent estimates in one submission ----
 * Requests are spread over up to PKDE_BATCH_STREAMS (default 8) device
 * streams and run concurrently.  Each stream keeps one plan per grid shape
 * and keywords (the seed is passed per request) and one rung-chunk buffer
 * that its plans share, resident between calls; when the plans a call needs
 * do not fit the free memory, the streams first drop the plans it does not
 * use.  A request that fails (invalid, out of grid, out of memory) gets its
 * own status in status_out; the others complete.  Results are written when
 * the call returns.  No counterpart in the reference, which estimates one
 * dataset per call. */
typedef struct {
  int32_t ndims;
  int64_t n_per_axis[PKDE_MAX_DIMS];
  double lo[PKDE_MAX_DIMS], hi[PKDE_MAX_DIMS], h[PKDE_MAX_DIMS];
  const double *samples; /* host, ndims x n */
  int64_t n;
  pkde_gradepro_params params;
  double *density_out;   /* host, G */
  int16_t *chosen_out;   /* host, G, nullable */
} pkde_request;

int pkde_gradepro_estimate_batch(pkde_ctx *ctx, int32_t count, const pkde_request *reqs,
                                 int32_t *status_out);
/* The same on device-resident data: each request's samples, density_out
 * and chosen_out (nullable) are device pointers; nothing but the status
 * flags crosses PCIe.  Returns when every request has finished. */
int pkde_gradepro_estimate_batch_device(pkde_ctx *ctx, int32_t count, const pkde_request *reqs,
                                        int32_t *status_out);
/* Frees what the batched form keeps resident in the context (its streams'
 * plans and chunk buffers); the next batch allocates them again. */
int pkde_context_release_batch(pkde_ctx *ctx);
/* Device bytes a request needs (its plan for req->n samples and the chunk
 * buffer its stream lends the plan), computed on the host without any
 * device call: the admission figure of the batched and serving forms
 * (extends get_necessary_memory, GradeProEstimator.jl:1091-1096). */
int pkde_gradepro_request_footprint(pkde_ctx *ctx, const pkde_request *req, uint64_t *bytes_out);

/* ---- serving form: continuous batching with HBM-footprint admission ----
 * A server thread admits requests as soon as a stream slot is free --
 * among the first 16 queued, the one with the least estimated device work
 * first, but none passed over more than 16 times; a request whose ladder
 * fills the GPU by itself (2-D 256^2, 3-D 64^3 and larger) waits while
 * another such request runs, smaller ones fill the other slots -- and the
 * device memory the request still needs (a plan
 * for its shape, a larger chunk buffer for the slot) fits the budget next to
 * everything resident (bytes_in_use: the pooled plans and the slots' chunk
 * buffers); idle plans, least recently used first, then idle slots' chunk
 * buffers are released to make room.
 * A request whose footprint alone exceeds the budget completes with
 * PKDE_OUT_OF_MEMORY; the others go on.  Plans are pooled per grid shape and
 * keywords and reused by later requests; none is freed while a stream uses
 * it.  Requests that arrive while others run overlap on the device.  Results
 * do not depend on what else is in flight. */
typedef struct {
  int32_t max_inflight;      /* concurrent requests (device streams) */
  uint64_t memory_budget;    /* bytes of device memory for admitted requests */
} pkde_server_cfg;

typedef struct {
  uint64_t completed, rejected, queued, inflight;
  uint64_t bytes_in_use;
  double p50_ms, p99_ms;     /* submit -> completion latency */
} pkde_server_stats;

int pkde_server_create(pkde_ctx *ctx, const pkde_server_cfg *cfg, pkde_server **out);
/* Copies the request's samples; the caller keeps ownership of its outputs,
 * which must stay valid until the ticket completes. */
int pkde_server_submit(pkde_server *srv, const pkde_request *req, uint64_t *ticket);
/* done = 1 when finished; returns the request's status once done. */
int pkde_server_poll(pkde_server *srv, uint64_t ticket, int *done);
int pkde_server_wait(pkde_server *srv, uint64_t ticket);
int pkde_server_stats_get(pkde_server *srv, pkde_server_stats *out);
void pkde_server_destroy(pkde_server *srv);

#ifdef __cplusplus
}
#endif
#endif
