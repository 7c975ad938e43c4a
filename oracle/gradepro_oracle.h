/*
 * FP64 CPU oracle for the :gradepro density estimator of ParallelKDE.jl v1.0.1.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it restates the
 * reference's `:serial` CPU semantics (and its `:threaded` variant, which is
 * bitwise identical per element) in plain C, so that the HIP path can be
 * compared against it on identical inputs.  Product code never links, loads
 * or calls it; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg do.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the ParallelKDE.jl source tree).  Array layouts follow Julia's column-major
 * convention: grid axis 1 is the fastest index, the bootstrap index is the
 * outermost; samples are ndims x n (coordinate fastest).
 */
#ifndef GRADEPRO_ORACLE_H
#define GRADEPRO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OG_MAX_DIMS 6

enum {
  OG_OK = 0,
  OG_INVALID_ARGUMENT = 1,
  OG_OUT_OF_GRID = 2,
  OG_OUT_OF_MEMORY = 4
};

/* Regular Cartesian grid: per axis low bound, high bound, spacing, points.
 * src/Grids/Grids.jl:67-87 (Grid), :189-210 (spacings/bounds). */
typedef struct {
  int ndims;
  int64_t n[OG_MAX_DIMS];
  double lo[OG_MAX_DIMS];
  double hi[OG_MAX_DIMS];
  double h[OG_MAX_DIMS];
} og_grid;

/* Estimator parameters, src/DensityEstimators/GradeProEstimator.jl:700-796
 * and docs/src/estimators/gradepro.md:19-28.  NaN / <=0 mean "default". */
typedef struct {
  int32_t n_bootstraps;        /* 100 */
  int32_t n_steps;             /* 250; <=0 -> bw_step if finite else 250 */
  double bw_step;              /* NaN: unset */
  double bw_final[OG_MAX_DIMS];/* NaN: Silverman rule */
  double alpha_s;              /* NaN: 0.0 (1-D) / 0.06 (>=2-D) */
  double alpha_os;             /* NaN: 0.1 */
  double eps;                  /* NaN: 2.0 (1-D) / -0.5 (>=2-D) */
  int32_t threads;             /* 1: :serial, >1: :threaded with OpenMP */
} og_params;

void og_params_default(og_params *p);
int64_t og_grid_points(const og_grid *g);

/* ---- project bootstrap RNG (Philox4x32-10, Salmon et al. SC'11) ----
 * idx[b*n + s] = bootstrap draw s of replicate b, 0-based, in [0, n). */
void og_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void og_bootstrap_indices(uint64_t seed, int64_t n, int32_t B, int32_t *idx);

/* ---- grid / ladder helpers ---- */
/* 2pi*fftfreq(n, 1/h) per axis (Grids.jl:225-231). out has n entries. */
void og_fftfreq(int64_t n, double h, double *out);
/* Silverman rule per axis (StatisticsPropagation/DirectSpace.jl:1102-1111). */
void og_silverman(int ndims, int64_t n, const double *samples, double *bw_out);
/* Ladder (GradeProEstimator.jl:860-905).  Returns the number of rungs S
 * (n_steps+1) or -1 on invalid time step.  times_out has S*ndims entries
 * (rung-major) and may be NULL to only query S. */
int og_ladder(int ndims, const double *bw_final, int32_t n_steps, double bw_step,
              double *dt_out, double *times_out);
/* calculate_duration_steps (GradeProEstimator.jl:954-966). */
int og_duration_steps(int ndims, const double *t_end, const double *dt,
                      double alpha_s, double alpha_os, int *steps_low, int *steps_over);
/* calculate_dlogt (GradeProEstimator.jl:1074-1089); t_prev NULL for rung 1. */
double og_dlogt(int ndims, const double *t0, const double *t, const double *t_prev);

/* ---- primitives (each restates one reference function) ---- */
/* check_data (DirectSpace.jl:479-498) plus the upper-edge cell check. */
int og_check_data(const og_grid *g, const double *samples, int64_t n);
/* initialize_dirac_sequence :serial + generate_dirac_cpu!
 * (DirectSpace.jl:97-145, 196-242).  Writes real parts only: sums[B*G],
 * squares[B*G] (may be NULL).  idx is n x B, 0-based. */
int og_dirac_sequences(const og_grid *g, const double *samples, int64_t n,
                       const int32_t *idx, int32_t B, double *sums, double *squares,
                       int threads);
/* In-place unnormalised C2C DFT over all axes of a column-major array of
 * shape n[0..ndims-1]; sign -1 forward, +1 inverse.  data is interleaved
 * (re, im).  Stands in for FFTW's plan_fft!/plan_ifft! (FourierSpace.jl:25-52,
 * 445-472); the inverse here is NOT scaled by 1/G. */
int og_fft_nd(int ndims, const int64_t *n, double *data, int sign);
/* Selects the radix order of og_fft_nd for the calling process: 0 (default)
 * applies radix 4, then 2, then odd primes ascending; 1 applies the same
 * factors in reverse order and sums each direct prime-radix DFT from its last
 * term to its first; 2 splits the power of two into radix-2 passes
 * (then odd primes ascending).  All are exact DFTs in exact arithmetic; the
 * others sum in a different order, so running a whole estimate under each
 * measures how far FP64 rounding alone moves the reference's result (its
 * rounding envelope; mode 2 also perturbs lengths whose factors are all 4).
 * Not thread-safe: set it between estimates. */
void og_set_fft_order(int mode);
/* 1: the ladder keeps only the real part of each normalised inverse
 * transform (a complex-to-real inverse: the exact result is real); 0 (the
 * reference): |.| of the complex result, whose imaginary part is rounding
 * noise.  A rounding-envelope probe.  Not thread-safe. */
void og_set_real_output(int on);
/* Deposit mode of the Dirac sequences.  0: FP64 sums in draw order, as the
 * reference does.  1: the same sums from the last draw to the first (equal in
 * exact arithmetic; a rounding-envelope probe for binning that sums in
 * another order, FP64 atomics).  2: the fixed-point deterministic mode: each
 * deposit rounded to llrint(term * scale), scale = 2^fb * n * prod h (squares:
 * 2^fb * (n * prod h)^2), fb = og_fixed_frac_bits(n), summed exactly in 64-bit
 * integers and scaled back by (double)sum * (1 / scale).  Not thread-safe. */
void og_set_deposit_order(int mode);
/* Fraction bits of the fixed-point deposits for n samples: 61 - (bits of n),
 * so a cell (at most n whole draws) stays below 2^61. */
int og_fixed_frac_bits(int64_t n);
/* propagate_time_cpu! (FourierSpace.jl:215-262) on one complex field. vars may
 * be NULL (means-only form). */
void og_propagate(const og_grid *g, const double *kgrid_axes, const double *t,
                  const double *t0, const double *m0, const double *v0,
                  double *mt, double *vt);
/* calculate_scaled_vmr! :serial (DirectSpace.jl:526-584).  sk, s2k are
 * complex [B][G]; writes G log-scaled VMR values to vmr_out. */
void og_scaled_vmr(int ndims, int64_t G, int32_t B, const double *sk, const double *s2k,
                   const double *t, const double *t0, int64_t n_samples, double *vmr_out);
/* calculate_full_means! :serial (DirectSpace.jl:747-764): |z| per element. */
void og_full_means(int64_t G, const double *sk, double *means_out);
/* find_stability + identify_convergence! :serial (DirectSpace.jl:829-871,
 * 975-1031, 1092-1100) for one rung; followed by the history rotation of
 * update_density_state_vmr! (GradeProEstimator.jl:617-625).  `chosen` (may be
 * NULL) receives `rung` wherever density is assigned. */
void og_identify_convergence(int64_t G, double *density, const double *means,
                             const double *vmr, double *f_prev1, double *f_prev2,
                             double dlogt, double eps, int steps_low, int steps_over,
                             double *minima, uint16_t *counters_low,
                             uint16_t *counters_over, uint8_t *low_flags,
                             int16_t *chosen, int16_t rung, int rotate);

/* ---- the whole estimator ----
 * estimate_density!(est, :gradepro) on a grid (ParallelKDE.jl:211-226 ->
 * GradeProEstimator.jl:700-796, 968-1072).  idx: n x B 0-based bootstrap
 * draws (NULL: draw with og_bootstrap_indices(seed)).  Outputs: density[G],
 * chosen[G] (rung 0..S-1 of the last assignment, S for the NaN fallback),
 * optional vmr_table[S*G] and means_table[S*G] (rung-major).  max_rungs > 0
 * stops the ladder after that many rungs (CPU timing samples only). */
typedef struct {
  int32_t n_rungs;
  int32_t steps_low, steps_over;
  double eps;
  double bw_final[OG_MAX_DIMS];
  double seconds_binning_fft;
  double seconds_ladder;
} og_info;

int og_gradepro_estimate(const og_grid *g, const double *samples, int64_t n,
                         const int32_t *idx, uint64_t seed, const og_params *p,
                         double *density, int16_t *chosen, double *vmr_table,
                         double *means_table, int32_t max_rungs, og_info *info);

/* ---- the rule-of-thumb estimator ----
 * estimate_density!(est, :rot) on a grid, :serial (RoTEstimator.jl:116-265):
 * full-sample Dirac sequence and forward FFT (initialize_density, :191-202),
 * propagation of the means by t_i = factor * sigma_i - h_i / 2
 * (propagation_time :22-28 Silverman / :42-47 Scott, RoTEstimator :67-70,
 * propagate_kernels! :219-235), normalised inverse FFT and |.|
 * (ifft_density!, copy_density! :237-265).  rule: 0 Scott, 1 Silverman.
 * time_out (may be NULL) receives t. */
int og_rot_estimate(const og_grid *g, const double *samples, int64_t n, int rule,
                    double *density, double *time_out);

#ifdef __cplusplus
}
#endif
#endif
