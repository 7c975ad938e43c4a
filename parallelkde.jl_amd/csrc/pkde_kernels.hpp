// Kernel-level interface of the :gradepro HIP path (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pkde_fft.hpp"

namespace pkde {

constexpr int kMaxDims = 6;
constexpr int kThreads = 256;

// Grid geometry as the kernels see it (Grids.jl:67-87: lo, hi, spacing, n).
struct GridDev {
  int N;
  int n[kMaxDims];
  int log2n[kMaxDims];
  int64_t G;   // prod n
  int H1;      // n1/2 + 1 Hermitian half along axis 1
  // Row-kernel layout of a rung chunk ("tiled rows"): element (field, row
  // q', k1) at field * fsY + (k1 / TKy) * (Q * TKy) + q' * TKy + k1 % TKy, so
  // the axis pass that writes it stores each block's TKy columns as one
  // contiguous tile.  Rows are q' = q in 2-D and q' = j3 + n3 * j2 in 3-D.
  int TKy, log2TKy, ktY;
  int64_t fsY;
  // Columns k1 a rung chunk carries: H1, or n1/2 when `pack` -- the
  // power-of-two row path carries the k1 = 0 and k1 = n1/2 columns of a real
  // field as one complex pencil X0 + i X(n1/2) (both stay real through the
  // inverse passes along the other axes), which removes the Nyquist tile.
  int Hc, pack;
  int64_t Q;   // n2 * n3 (pencils per axis-1 row set)
  double lo[kMaxDims], hi[kMaxDims], h[kMaxDims];
  int koff[kMaxDims];  // offset of axis i in per-rung Gaussian tables
  int ksum;            // sum of n_i
  // mixed-radix FFT plans of the axes; pow2 != 0 when every axis is a power
  // of two (the wave-local kernels and the tiled row layout need that)
  FftFactors fac[kMaxDims];
  FftFactors facw[kMaxDims];  // radices 8, 4, 2, 3, 5 (np = 0: other primes)
  int pow2;
};

// ---- launchers (implemented in pkde_kernels.hip) ----
// Binning of B bootstrap resamples plus the full sample into acc[2B+1][G]
// (field ids: [0,B) sums, [B,2B) squares, 2B full sample).  fixed != 0 uses
// int64 fixed-point deposits, otherwise FP64 atomics.  err receives 1 if a
// sample is outside the grid.
void launch_binning(const GridDev& g, const double* x, int64_t n, const int32_t* idx, int B,
                    uint64_t seed, int fixed, int frac_bits, void* acc, int* err,
                    hipStream_t s);

// Deterministic fixed-point binning without global atomics (counting sort of
// the samples by cell, per-bootstrap multiplicities, per-tile LDS deposits);
// writes every accumulator cell, so acc needs no zeroing.  Same integers as
// launch_binning with fixed != 0.  `work` must hold binsort_bytes(); returns
// nonzero (nothing launched) if it does not or the sizes are out of range.
uint64_t binsort_bytes(const GridDev& g, int64_t n, int B);
// Grid cells per tile of the sorted binning's deposits (2 int64 fields in
// LDS: 128 KB); tiles span axis 1, so n1 must not exceed it.
constexpr int kBinTileCells = 8192;

int launch_binning_sorted(const GridDev& g, const double* x, int64_t n, const int32_t* idx, int B,
                          uint64_t seed, int frac_bits, void* acc, int* err, void* work,
                          uint64_t work_bytes, hipStream_t s,
                          int tile_cells = kBinTileCells);

// acc -> forward transform along axis 1 in FP64, Hermitian halves written to
// specD[fid][k1][q].
// twg: the plan's per-axis twiddle tables exp(-2 pi i k / n_i) at offsets
// koff[i] (used for axes that are not powers of two).
void launch_forward_axis1(const GridDev& g, const void* acc, int fixed, double scaleM,
                          double scaleV, int B, cplx<double>* specD, const double* tw64,
                          const double* twg, hipStream_t s);

// Generic batched pencil FFT along axis `axis` (>= 1, i.e. axis 2 or 3) of
// [rung][field][k1][q] arrays, in place or out of place, same layout.
// Optional per-rung Gaussian factors.
template <typename TI, typename TO, int DIR>
void launch_axis_fft(const GridDev& g, int axis, const cplx<TI>* src, int64_t src_rung_stride,
                     int src_fid0, cplx<TO>* dst, int64_t dst_rung_stride, int dst_fid0, int fid0,
                     int nf, int R, int r0, const TI* gM, const TI* gV, int fsplit, const TI* tw,
                     const TI* twg, hipStream_t s);

// Ladder form of the first inverse pass along grid axis 2: spectra read once
// per chunk of R rungs (see k_axis_ladder).  mode: 0 [k1][q], 1 tiled rows
// (row-kernel layout, 2-D), 2 [k1][j2][j3] (input of launch_axis3_rows).  False if
// the axis length is outside the wave-local range.
// 2-D ladder for axis lengths that are not powers of two (FP64): a per-rung
// axis-2 pass into the tiled row layout (8-column tiles) and the row kernel on
// the wave-local mixed-radix FFT.  ladder_gen_rows_ok: the grid qualifies
// (N == 2 or 3, some axis not a power of two, every axis <= 512 with radices
// the wave-local FFT covers, GridDev::TKy == 8).
bool ladder_gen_rows_ok(const GridDev& g);
bool launch_axis_gen_rows(const GridDev& g, const cplx<double>* spec, cplx<double>* dst,
                          int64_t dst_rung_stride, int nf, int R, int r0, const double* gM,
                          const double* gV, int fsplit, const double* twg, hipStream_t s);
// 3-D: both inverse axis passes of a rung chunk (axis 2 into y2 as
// [k1][j2][j3], axis 3 from there into the tiled rows of y).
bool launch_gen3_passes(const GridDev& g, const cplx<double>* spec, cplx<double>* y2,
                        cplx<double>* y, int64_t rung_stride, int nf, int R, int r0,
                        const double* gM, const double* gV, int fsplit, const double* twg,
                        hipStream_t s);
bool launch_ladder_rows_gen(const GridDev& g, const cplx<double>* src, int64_t src_rung_stride,
                            int B, int64_t n_samples, int R, int r0, const double* gM,
                            const double* gV, const double* cV, const double* vscale,
                            const double* twg, const double* invk, double* vmr, hipStream_t s);

// Workgroups the ladder axis pass aims at when a chunk has few tiles (its
// rungs are then split over several workgroups per tile); 0: no split.
constexpr int kAxisSplitTarget = 8000;
template <typename T>
bool launch_axis_ladder(const GridDev& g, const cplx<T>* spec, cplx<T>* dst,
                        int64_t dst_rung_stride, int nf, int R, int r0, const T* gM,
                        const T* gV, int fsplit, const T* tw, int mode, hipStream_t s,
                        int split_target = kAxisSplitTarget, bool fold = false);

// 3-D: inverse pass along grid axis 3 from the [k1][j2][j3] layout to the
// tiled row-kernel layout.  False if n3 is outside the wave-local range.
template <typename T>
bool launch_axis3_rows(const GridDev& g, const cplx<T>* src, int64_t src_rung_stride,
                       cplx<T>* dst, int64_t dst_rung_stride, int nf, int R, int r0,
                       const T* gM, const T* gV, int fsplit, const T* tw, hipStream_t s);

template <typename TI, typename TO>
void launch_convert(const cplx<TI>* src, cplx<TO>* dst, int64_t count, hipStream_t s);

// Ladder: per rung r in [r0, r0+R) the bootstrap log-VMR per grid point
// (propagate_bootstraps!, ifft_bootstraps!, calculate_vmr! fused).
template <typename T>
void launch_ladder_vmr(const GridDev& g, const cplx<T>* src, int64_t src_rung_stride, int B,
                       int64_t n_samples, int R, int r0, const T* gM, const T* gV,
                       const double* cV, const double* vscale, const T* tw, const T* twg,
                       double* vmr, hipStream_t s);

// Row-per-wave form of the same ladder step for Q >= 64 rows and
// 16 <= n1 <= 512, reading the tiled row layout (see GridDev) with the
// Nyquist column packed into column 0; false when the plan does not pack.
template <typename T>
bool launch_ladder_rows(const GridDev& g, const cplx<T>* src, int64_t src_rung_stride, int B,
                        int64_t n_samples, int R, int r0, const T* gM, const T* gV,
                        const double* cV, const double* vscale, const T* tw, const double* invk,
                        double* vmr, hipStream_t s);

// 3-D ladder step fused after the axis-2 pass (mode 2 output, [k1][j2][k3]
// with the Nyquist column packed) for grids whose (k1, k3) slab of a field
// pair fits LDS: axis-3 and axis-1 inverse passes, VMR and Welford in one
// kernel (FP64).  slab3_ok tells whether the grid takes it.
bool slab3_ok(const GridDev& g);
// Grids the split form takes (an axis-3 length of 128 whose whole slab does
// not fit LDS): the axis-2 pass stores the first radix-4 step of axis 3
// (k_axis_ladder mode 3, class-major), launch_slab3 runs quarters.
bool slab3_split_ok(const GridDev& g);
bool launch_slab3(const GridDev& g, const cplx<double>* src, int64_t src_rung_stride, int B,
                  int64_t n_samples, int R, int r0, const double* gM, const double* gV,
                  const double* cV, const double* vscale, const double* tw, const double* invk,
                  double* vmr, hipStream_t s);

// :rot propagation time per axis from the sample standard deviations
// (rule 0 Scott, 1 Silverman); and the per-axis Gaussian factor tables alone.
void launch_rot_times(const GridDev& g, int64_t n, int rule, const double* sigma, double* times,
                      hipStream_t s);
template <typename T>
void launch_gauss_tables(const GridDev& g, int S, const double* times, const double* kfreq, T* gM,
                         T* gV, hipStream_t s);

// Convergence state machine over all rungs -> chosen rung, selection rung.
// terms (2 S G doubles, or NULL): scratch for the small-grid form, which
// computes the logarithmic terms of every (rung, point) in parallel first.
void launch_convergence(int64_t G, int S, const double* vmr, const double* dlogt,
                        const double* params /* eps, steps_low, steps_over */,
                        int16_t* chosen, int16_t* sel, double* terms, hipStream_t s);
// Grids up to this many points use the small-grid convergence form.
constexpr int64_t kConvTermsMaxG = 16384;

// Full-sample means at each point's selected rung -> density.
template <typename T>
void launch_means_select(const GridDev& g, const cplx<T>* src, int64_t src_rung_stride, int R,
                         int r0, int S, const T* gM, const T* tw, const T* twg,
                         const int16_t* sel, double* density, hipStream_t s);

// Silverman rule on device (DirectSpace.jl:1102-1111): bw[i] = prefactor * std_i.
void launch_silverman(int N, const double* x, int64_t n, double prefactor, double* partial,
                      double* mean_out, double* bw_out, hipStream_t s);
constexpr int kMomentBlocks = 256;
// Per-axis [min, max] of ndims x n samples into out[2 N] (NaN if any
// coordinate of the axis is NaN); partial holds kMomentBlocks * 2 N doubles.
void launch_extrema(int N, const double* x, int64_t n, double* partial, double* out, hipStream_t s);

// Per-rung scalars (times, dlogt, VMR scaling, variance propagation factor,
// steps_low / steps_over / eps) and the per-axis Gaussian tables.
template <typename T>
void launch_tables(const GridDev& g, int S, int n_steps_mode, double n_steps_or_bwstep,
                   const double* bw, const double* kfreq, int64_t n_samples, double alpha_s,
                   double alpha_os, double eps, double* times, double* dlogt, double* vscale,
                   double* cV, double* params, T* gM, T* gV, hipStream_t s);

void launch_fill_i32(int* p, int v, int64_t count, hipStream_t s);
void launch_zero_bytes(void* p, int64_t bytes, hipStream_t s);

}  // namespace pkde
