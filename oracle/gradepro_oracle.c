/*
 * FP64 CPU oracle for ParallelKDE.jl's :gradepro estimator (v1.0.1).
 *
 * TEST INFRASTRUCTURE ONLY -- this is the checker, never the product.  It is a
 * plain-C restatement of the reference's `:serial` CPU path; `threads > 1`
 * gives the `:threaded` form, which performs the same per-element arithmetic
 * and so returns bitwise-identical results.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   binning      directspace_tests.jl:1-41,191-208 KAT + mass / first-moment
 *                invariants + brute force on asymmetric samples
 *   FFT          numpy.fft golden vectors, brute-force DFT, round trip
 *   propagation  fourierspace_tests.jl:27-96 (Gaussian semigroup identity)
 *   VMR          directspace_tests.jl:47-61,210-230 closed form
 *   means        directspace_tests.jl:63-67,232-245
 *   convergence  directspace_tests.jl:69-189,247-368 snapshots
 *   ladder       gradepro_tests.jl:166-177; defaults from GradeProEstimator.jl
 *   grid         grid_tests.jl:1-36
 *   end to end   gradepro_tests.jl:485-511 (MISE bound, statistical)
 */
#include "gradepro_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double og_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void og_params_default(og_params *p) {
  p->n_bootstraps = 100;
  p->n_steps = 250;
  p->bw_step = NAN;
  for (int i = 0; i < OG_MAX_DIMS; ++i) p->bw_final[i] = NAN;
  p->alpha_s = NAN;
  p->alpha_os = NAN;
  p->eps = NAN;
  p->threads = 1;
}

int64_t og_grid_points(const og_grid *g) {
  int64_t G = 1;
  for (int i = 0; i < g->ndims; ++i) G *= g->n[i];
  return G;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11).  Project RNG for bootstrap draws: the reference
 * draws with Julia's global Xoshiro (KDEs.jl:189), which cannot be replayed
 * here, so both the oracle and the HIP path consume this counter-based stream. */
static void og_mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

void og_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t hi0, lo0, hi1, lo1;
    og_mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
    og_mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* Draw s of replicate b uses word (s & 3) of the block with counter
 * (s >> 2, b); the index is the high word of word * n (multiply-shift). */
void og_bootstrap_indices(uint64_t seed, int64_t n, int32_t B, int32_t *idx) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int32_t b = 0; b < B; ++b) {
    for (int64_t s = 0; s < n; s += 4) {
      uint64_t blk = (uint64_t)s >> 2;
      uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)b, 0u};
      uint32_t w[4];
      og_philox4x32_10(ctr, key, w);
      for (int q = 0; q < 4 && s + q < n; ++q)
        idx[(int64_t)b * n + s + q] = (int32_t)(((uint64_t)w[q] * (uint64_t)n) >> 32);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Grids.jl:225-231: fftgrid = 2pi * fftfreq(n, 1/h); AbstractFFTs keeps the
 * Frequencies type lazy under scalar multiplication, so element j is
 * j_signed * (((1/h)/n) * 2pi). */
void og_fftfreq(int64_t n, double h, double *out) {
  const double two_pi = 6.283185307179586;
  double mult = ((1.0 / h) / (double)n) * two_pi;
  int64_t n_nonneg = (n + 1) >> 1;
  for (int64_t j = 0; j < n; ++j) {
    int64_t js = (j < n_nonneg) ? j : j - n;
    out[j] = (double)js * mult;
  }
}

/* DirectSpace.jl:1102-1111: (4/((N+2)n))^(1/(N+4)) * std_i, std from the
 * n-1-normalised sample covariance (Statistics.cov). */
void og_silverman(int ndims, int64_t n, const double *samples, double *bw_out) {
  double prefactor = pow(4.0 / ((double)(ndims + 2) * (double)n), 1.0 / (double)(ndims + 4));
  for (int i = 0; i < ndims; ++i) {
    double s = 0.0;
    for (int64_t k = 0; k < n; ++k) s += samples[k * ndims + i];
    double mean = s / (double)n;
    double ss = 0.0;
    for (int64_t k = 0; k < n; ++k) {
      double d = samples[k * ndims + i] - mean;
      ss += d * d;
    }
    bw_out[i] = prefactor * sqrt(ss / (double)(n - 1));
  }
}

/* GradeProEstimator.jl:860-905 (get_time, CPU). */
int og_ladder(int ndims, const double *bw_final, int32_t n_steps, double bw_step,
              double *dt_out, double *times_out) {
  double dt[OG_MAX_DIMS];
  double steps;
  if (n_steps > 0) {
    for (int i = 0; i < ndims; ++i) dt[i] = bw_final[i] / (double)n_steps;
    steps = (double)n_steps;
  } else if (isfinite(bw_step)) {
    double m = INFINITY;
    for (int i = 0; i < ndims; ++i) {
      if (bw_step > bw_final[i] || bw_step <= 0.0) return -1; /* :891-893 */
      double q = floor(bw_final[i] / bw_step);
      if (q < m) m = q;
    }
    for (int i = 0; i < ndims; ++i) dt[i] = bw_step;
    steps = m;
  } else {
    return og_ladder(ndims, bw_final, 250, NAN, dt_out, times_out);
  }
  int S = (int)steps + 1;
  if (dt_out)
    for (int i = 0; i < ndims; ++i) dt_out[i] = dt[i];
  if (times_out)
    for (int k = 0; k < S; ++k)
      for (int i = 0; i < ndims; ++i) times_out[k * ndims + i] = dt[i] * (double)k;
  return S;
}

/* GradeProEstimator.jl:954-966; Julia's round(Int, x) is ties-to-even. */
int og_duration_steps(int ndims, const double *t_end, const double *dt,
                      double alpha_s, double alpha_os, int *steps_low, int *steps_over) {
  if (alpha_s < 0 || alpha_s > 1 || alpha_os < 0 || alpha_os > 1) return OG_INVALID_ARGUMENT;
  double sb = 0.0, so = 0.0;
  for (int i = 0; i < ndims; ++i) {
    double ns = t_end[i] / dt[i];
    double b = alpha_s * ns, o = alpha_os * ns;
    sb = (i == 0) ? b : sb + b;
    so = (i == 0) ? o : so + o;
  }
  *steps_low = (int)nearbyint(sb / (double)ndims);
  *steps_over = (int)nearbyint(so / (double)ndims);
  return OG_OK;
}

/* GradeProEstimator.jl:1074-1089. */
double og_dlogt(int ndims, const double *t0, const double *t, const double *t_prev) {
  double det_prev = 1.0, det_curr = 1.0;
  for (int i = 0; i < ndims; ++i) {
    double t0s = t0[i] * t0[i];
    double p = t_prev ? t0s + t_prev[i] * t_prev[i] : t0s;
    double c = t0s + t[i] * t[i];
    det_prev = (i == 0) ? p : det_prev * p;
    det_curr = (i == 0) ? c : det_curr * c;
  }
  return log(det_curr / det_prev);
}

/* ------------------------------------------------------------------------ */
/* check_data (DirectSpace.jl:479-498): every coordinate strictly inside
 * (lo, hi).  A sample whose low cell is the last node would deposit past the
 * grid end (the reference writes out of bounds there); it is rejected too. */
int og_check_data(const og_grid *g, const double *samples, int64_t n) {
  for (int64_t s = 0; s < n; ++s)
    for (int i = 0; i < g->ndims; ++i) {
      double x = samples[s * g->ndims + i];
      if (!(x < g->hi[i]) || !(x > g->lo[i])) return OG_OUT_OF_GRID;
      double l = floor((x - g->lo[i]) / g->h[i]);
      if (l > (double)(g->n[i] - 2)) return OG_OUT_OF_GRID;
    }
  return OG_OK;
}

/* generate_dirac_cpu! (DirectSpace.jl:196-242) for one bootstrap column.
 * Deposit modes (og_set_deposit_order): 0 the reference's FP64 sums in draw
 * order; 1 the same sums last draw first (a rounding-envelope probe); 2 the
 * optional fixed-point deterministic mode: each deposit is rounded to
 * llrint(term * 2^fb * n * prod h) (its square to llrint(term^2 * 2^fb *
 * (n * prod h)^2)) and summed exactly in 64-bit integers, fb = 61 - (bits of
 * n), then scaled back as (double)sum * (1 / scale) -- the same rounding of
 * the same sum, in a form whose result does not depend on the order. */
static int og_deposit_mode = 0;
void og_set_deposit_order(int mode) { og_deposit_mode = (mode == 1 || mode == 2) ? mode : 0; }

int og_fixed_frac_bits(int64_t n) {
  int lg = 0;
  while ((1ll << lg) <= n) ++lg;
  return 61 - lg;
}

static void og_dirac_one(const og_grid *g, const double *samples, int64_t n,
                         const int32_t *col, double *sums, double *squares) {
  const int N = g->ndims;
  double hp = g->h[0];
  for (int i = 1; i < N; ++i) hp *= g->h[i];
  const double spacing_squared = hp * hp;
  const double denom = (double)n * spacing_squared;
  int64_t stride[OG_MAX_DIMS];
  stride[0] = 1;
  for (int i = 1; i < N; ++i) stride[i] = stride[i - 1] * g->n[i - 1];

  const int fixed = og_deposit_mode == 2;
  const double two_fb = ldexp(1.0, og_fixed_frac_bits(n));
  const double scaleM = (double)n * hp * two_fb;
  const double scaleV = ((double)n * hp) * ((double)n * hp) * two_fb;
  int64_t *isums = (int64_t *)sums, *isq = (int64_t *)squares;
  for (int64_t s0 = 0; s0 < n; ++s0) {
    const int64_t s = og_deposit_mode == 1 ? n - 1 - s0 : s0;
    const double *x = samples + (int64_t)col[s] * N;
    int64_t il[OG_MAX_DIMS];
    double rl[OG_MAX_DIMS], rh[OG_MAX_DIMS];
    for (int i = 0; i < N; ++i) {
      double delta = x[i] - g->lo[i];
      il[i] = (int64_t)floor(delta / g->h[i]);
      rh[i] = fmod(delta, g->h[i]);
      rl[i] = g->h[i] - rh[i];
    }
    for (int c = 0; c < (1 << N); ++c) {
      /* mask[j] = bit (N-j) of c for 1-based axis j (:229): axis 1 is the MSB */
      int64_t off = 0;
      double pl = 1.0, ph = 1.0;
      int have_l = 0, have_h = 0;
      for (int j = 0; j < N; ++j) {
        int mask = (c >> (N - 1 - j)) & 1;
        off += (mask ? il[j] : il[j] + 1) * stride[j];
        if (mask) {
          pl = have_l ? pl * rl[j] : rl[j];
          have_l = 1;
        } else {
          ph = have_h ? ph * rh[j] : rh[j];
          have_h = 1;
        }
      }
      double term = (pl * ph) / denom;
      if (fixed) {
        isums[off] += (int64_t)llrint(term * scaleM);
        if (squares) isq[off] += (int64_t)llrint((term * term) * scaleV);
        continue;
      }
      sums[off] += term;
      if (squares) squares[off] += term * term;
    }
  }
  if (fixed) {
    const int64_t G = og_grid_points(g);
    const double iM = 1.0 / scaleM, iV = 1.0 / scaleV;
    for (int64_t j = 0; j < G; ++j) sums[j] = (double)isums[j] * iM;
    if (squares)
      for (int64_t j = 0; j < G; ++j) squares[j] = (double)isq[j] * iV;
  }
}

int og_dirac_sequences(const og_grid *g, const double *samples, int64_t n,
                       const int32_t *idx, int32_t B, double *sums, double *squares,
                       int threads) {
  int rc = og_check_data(g, samples, n);
  if (rc != OG_OK) return rc;
  const int64_t G = og_grid_points(g);
  memset(sums, 0, sizeof(double) * (size_t)(G * B));
  if (squares) memset(squares, 0, sizeof(double) * (size_t)(G * B));
  (void)threads;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
  for (int32_t b = 0; b < B; ++b)
    og_dirac_one(g, samples, n, idx + (int64_t)b * n, sums + (int64_t)b * G,
                 squares ? squares + (int64_t)b * G : NULL);
  return OG_OK;
}

/* ------------------------------------------------------------------------ */
/* FFT stand-in for FFTW: mixed-radix Stockham autosort, any length. */
typedef struct {
  int64_t n;
  int nf;
  int64_t factors[64];
  double *w; /* w[2k], w[2k+1] = cos, sin of 2 pi k / n */
} og_fft1;

static int og_fft_mode = 0;
void og_set_fft_order(int mode) { og_fft_mode = (mode == 1 || mode == 2) ? mode : 0; }

/* 1: the ladder keeps only the real part of each inverse transform (the
 * exact result is real: a complex-to-real inverse, as a Hermitian-packed
 * implementation computes it); 0: the reference's |.| of the complex
 * result, whose imaginary part is rounding noise.  A rounding-envelope
 * probe; not thread-safe. */
static int og_real_output = 0;
void og_set_real_output(int on) { og_real_output = on ? 1 : 0; }

static int og_fft1_init(og_fft1 *p, int64_t n) {
  p->n = n;
  p->nf = 0;
  int64_t m = n;
  if (og_fft_mode != 2)
    while (m % 4 == 0) { p->factors[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->factors[p->nf++] = 2; m /= 2; }
  for (int64_t f = 3; f * f <= m; f += 2)
    while (m % f == 0) { p->factors[p->nf++] = f; m /= f; }
  if (m > 1) p->factors[p->nf++] = m;
  if (og_fft_mode == 1)
    for (int i = 0, j = p->nf - 1; i < j; ++i, --j) {
      const int64_t t = p->factors[i];
      p->factors[i] = p->factors[j];
      p->factors[j] = t;
    }
  p->w = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
  if (!p->w) return OG_OUT_OF_MEMORY;
  for (int64_t k = 0; k < n; ++k) {
    double a = 6.283185307179586476925286766559 * (double)k / (double)n;
    p->w[2 * k] = cos(a);
    p->w[2 * k + 1] = sin(a);
  }
  return OG_OK;
}

static void og_fft1_free(og_fft1 *p) { free(p->w); }

/* exp(sign * 2 pi i k / n) */
static inline void og_tw(const og_fft1 *p, int64_t k, int sign, double *c, double *s) {
  k %= p->n;
  *c = p->w[2 * k];
  *s = sign * p->w[2 * k + 1];
}

/* x (length n, interleaved) -> in-place transform using work (same size). */
static void og_fft1_exec(const og_fft1 *p, double *x, double *work, int sign) {
  const int64_t n = p->n;
  if (n <= 1) return;
  double *in = x, *out = work;
  int64_t Ns = 1;
  double vr[64], vi[64], tr[64], ti[64];
  double *bigr = NULL, *bigi = NULL, *bigtr = NULL, *bigti = NULL;
  for (int f = 0; f < p->nf; ++f) {
    const int64_t R = p->factors[f];
    double *ar = vr, *ai = vi, *br = tr, *bi = ti;
    if (R > 64) {
      bigr = (double *)realloc(bigr, sizeof(double) * R);
      bigi = (double *)realloc(bigi, sizeof(double) * R);
      bigtr = (double *)realloc(bigtr, sizeof(double) * R);
      bigti = (double *)realloc(bigti, sizeof(double) * R);
      ar = bigr; ai = bigi; br = bigtr; bi = bigti;
    }
    const int64_t nR = n / R;
    const int64_t step = n / (Ns * R);
    for (int64_t j = 0; j < nR; ++j) {
      const int64_t jm = j % Ns;
      for (int64_t r = 0; r < R; ++r) {
        double a = in[2 * (j + r * nR)], b = in[2 * (j + r * nR) + 1];
        if (r > 0 && jm > 0) {
          double c, s;
          og_tw(p, jm * r * step, sign, &c, &s);
          double t = a * c - b * s;
          b = a * s + b * c;
          a = t;
        }
        ar[r] = a;
        ai[r] = b;
      }
      if (R == 2) {
        br[0] = ar[0] + ar[1]; bi[0] = ai[0] + ai[1];
        br[1] = ar[0] - ar[1]; bi[1] = ai[0] - ai[1];
      } else if (R == 4) {
        double s0r = ar[0] + ar[2], s0i = ai[0] + ai[2];
        double d0r = ar[0] - ar[2], d0i = ai[0] - ai[2];
        double s1r = ar[1] + ar[3], s1i = ai[1] + ai[3];
        double d1r = ar[1] - ar[3], d1i = ai[1] - ai[3];
        /* multiply d1 by (sign * i) */
        double jr = -sign * d1i, ji = sign * d1r;
        br[0] = s0r + s1r; bi[0] = s0i + s1i;
        br[2] = s0r - s1r; bi[2] = s0i - s1i;
        br[1] = d0r + jr; bi[1] = d0i + ji;
        br[3] = d0r - jr; bi[3] = d0i - ji;
      } else {
        const int64_t wstep = n / R;
        for (int64_t q = 0; q < R; ++q) {
          double sr = 0.0, si = 0.0;
          for (int64_t rr = 0; rr < R; ++rr) {
            /* mode 1 also sums the direct DFT's terms last first */
            const int64_t r = og_fft_mode == 1 ? R - 1 - rr : rr;
            double c, s;
            og_tw(p, ((r * q) % R) * wstep, sign, &c, &s);
            sr += ar[r] * c - ai[r] * s;
            si += ar[r] * s + ai[r] * c;
          }
          br[q] = sr;
          bi[q] = si;
        }
      }
      const int64_t base = (j / Ns) * Ns * R + jm;
      for (int64_t r = 0; r < R; ++r) {
        out[2 * (base + r * Ns)] = br[r];
        out[2 * (base + r * Ns) + 1] = bi[r];
      }
    }
    double *t = in; in = out; out = t;
    Ns *= R;
  }
  if (in != x) memcpy(x, in, sizeof(double) * 2 * (size_t)n);
  free(bigr); free(bigi); free(bigtr); free(bigti);
}

int og_fft_nd(int ndims, const int64_t *n, double *data, int sign) {
  int64_t G = 1;
  for (int i = 0; i < ndims; ++i) G *= n[i];
  int64_t stride = 1;
  for (int ax = 0; ax < ndims; ++ax) {
    const int64_t L = n[ax];
    og_fft1 plan;
    if (og_fft1_init(&plan, L) != OG_OK) return OG_OUT_OF_MEMORY;
    double *pen = (double *)malloc(sizeof(double) * 4 * (size_t)L);
    if (!pen) { og_fft1_free(&plan); return OG_OUT_OF_MEMORY; }
    double *work = pen + 2 * L;
    const int64_t outer = G / (L * stride);
    for (int64_t o = 0; o < outer; ++o)
      for (int64_t in = 0; in < stride; ++in) {
        double *base = data + 2 * (o * L * stride + in);
        for (int64_t k = 0; k < L; ++k) {
          pen[2 * k] = base[2 * k * stride];
          pen[2 * k + 1] = base[2 * k * stride + 1];
        }
        og_fft1_exec(&plan, pen, work, sign);
        for (int64_t k = 0; k < L; ++k) {
          base[2 * k * stride] = pen[2 * k];
          base[2 * k * stride + 1] = pen[2 * k + 1];
        }
      }
    free(pen);
    og_fft1_free(&plan);
    stride *= L;
  }
  return OG_OK;
}

/* ------------------------------------------------------------------------ */
/* propagate_time_cpu! (FourierSpace.jl:215-262).  kgrid_axes holds the
 * concatenated per-axis Fourier coordinates (og_fftfreq of each axis). */
void og_propagate(const og_grid *g, const double *kgrid_axes, const double *t,
                  const double *t0, const double *m0, const double *v0,
                  double *mt, double *vt) {
  const int N = g->ndims;
  const int64_t G = og_grid_points(g);
  double det_t0 = t0[0], det_t2 = t0[0] * t0[0] + t[0] * t[0];
  for (int i = 1; i < N; ++i) {
    det_t0 *= t0[i];
    det_t2 *= t0[i] * t0[i] + t[i] * t[i];
  }
  const double det_t = sqrt(det_t2);
  for (int64_t p = 0; p < G; ++p) {
    int64_t rem = p, koff = 0;
    double prop = 0.0;
    for (int i = 0; i < N; ++i) {
      int64_t ji = rem % g->n[i];
      rem /= g->n[i];
      double a = kgrid_axes[koff + ji] * t[i];
      prop += a * a;
      koff += g->n[i];
    }
    prop = exp(-0.5 * prop);
    mt[2 * p] = m0[2 * p] * prop;
    mt[2 * p + 1] = m0[2 * p + 1] * prop;
    if (v0) {
      double sp = sqrt(prop);
      vt[2 * p] = ((v0[2 * p] * sp) * det_t0) / det_t;
      vt[2 * p + 1] = ((v0[2 * p + 1] * sp) * det_t0) / det_t;
    }
  }
}

/* calculate_scaled_vmr! :serial (DirectSpace.jl:526-584) on points
 * [j0, j1): magnitudes, VMR per element, Welford over bootstraps in order,
 * log of the scaled variance (NaN when not finite). */
static void og_vmr_range(int ndims, int64_t G, int32_t B, const double *sk, const double *s2k,
                         const double *t, const double *t0, int64_t n_samples, int64_t j0,
                         int64_t j1, double *vmr_out) {
  double prodt = 0.0;
  for (int i = 0; i < ndims; ++i) {
    double q = t[i] * t[i] + t0[i] * t0[i];
    prodt = (i == 0) ? q : prodt * q;
  }
  /* n_samples^3 is Int64 arithmetic in Julia, then promoted (:564). */
  const double scaling = pow(prodt, 1.5) * (double)(n_samples * n_samples * n_samples);
  const double dn = (double)n_samples;
  for (int64_t j = j0; j < j1; ++j) {
    double mean = 0.0, m2 = 0.0;
    for (int32_t b = 0; b < B; ++b) {
      const double *a = sk + 2 * ((int64_t)b * G + j);
      const double *v = s2k + 2 * ((int64_t)b * G + j);
      double ski = sqrt(a[0] * a[0] + a[1] * a[1]) / dn;
      double s2ki = sqrt(v[0] * v[0] + v[1] * v[1]) / dn - ski * ski;
      double x = s2ki / ski;
      double delta = x - mean;
      mean += delta / (double)(b + 1);
      m2 += delta * (x - mean);
    }
    double vmr_v = scaling * m2 / (double)(B - 1);
    vmr_out[j] = isfinite(vmr_v) ? log(vmr_v) : NAN;
  }
}

void og_scaled_vmr(int ndims, int64_t G, int32_t B, const double *sk, const double *s2k,
                   const double *t, const double *t0, int64_t n_samples, double *vmr_out) {
  og_vmr_range(ndims, G, B, sk, s2k, t, t0, n_samples, 0, G, vmr_out);
}

/* calculate_full_means! :serial (DirectSpace.jl:747-764). */
void og_full_means(int64_t G, const double *sk, double *means_out) {
  for (int64_t j = 0; j < G; ++j)
    means_out[j] = sqrt(sk[2 * j] * sk[2 * j] + sk[2 * j + 1] * sk[2 * j + 1]);
}

/* find_stability (DirectSpace.jl:975-1031) + identify_convergence! :serial
 * (:829-871) + the history rotation of update_density_state_vmr!
 * (GradeProEstimator.jl:617-625). */
void og_identify_convergence(int64_t G, double *density, const double *means,
                             const double *vmr, double *f_prev1, double *f_prev2,
                             double dlogt, double eps, int steps_low, int steps_over,
                             double *minima, uint16_t *counters_low,
                             uint16_t *counters_over, uint8_t *low_flags,
                             int16_t *chosen, int16_t rung, int rotate) {
  for (int64_t i = 0; i < G; ++i) {
    const double f = vmr[i], f1 = f_prev1[i], f2 = f_prev2[i];
    const double d1 = log(fabs((f - f1) / (2.0 * dlogt)));
    const double d2 = log(fabs((f - 2.0 * f1 + f2) / (dlogt * dlogt)));
    double cur_min = minima[i];
    uint16_t cl = counters_low[i], co = counters_over[i];
    int low = low_flags[i];
    int more_stable = 0;
    double new_min = NAN;
    if (!low) {
      if (d2 < eps) {
        if (((d1 < cur_min) || isnan(cur_min)) && ((int)co <= steps_over)) {
          more_stable = 1;
          new_min = d1;
          co = 0;
        } else {
          co = (uint16_t)(co + 1);
        }
      } else {
        cl = (uint16_t)(cl + 1);
        if ((int)cl > steps_low) {
          low = 1;
          cur_min = NAN;
        }
      }
    } else {
      if ((d2 < eps) && isnan(cur_min)) {
        more_stable = 1;
        new_min = d1;
      }
    }
    if (isnan(new_min)) new_min = cur_min;
    counters_low[i] = cl;
    counters_over[i] = co;
    low_flags[i] = (uint8_t)low;
    minima[i] = new_min;
    if (more_stable) {
      density[i] = means[i];
      if (chosen) chosen[i] = rung;
    }
  }
  if (rotate)
    for (int64_t i = 0; i < G; ++i) {
      f_prev2[i] = f_prev1[i];
      f_prev1[i] = vmr[i];
    }
}

/* ------------------------------------------------------------------------ */
int og_gradepro_estimate(const og_grid *g, const double *samples, int64_t n,
                         const int32_t *idx_in, uint64_t seed, const og_params *p,
                         double *density, int16_t *chosen, double *vmr_table,
                         double *means_table, int32_t max_rungs, og_info *info) {
  const int N = g->ndims;
  if (N < 1 || N > OG_MAX_DIMS || n < 2 || p->n_bootstraps < 2) return OG_INVALID_ARGUMENT;
  for (int i = 0; i < N; ++i)
    if (g->n[i] < 2 || !(g->h[i] > 0)) return OG_INVALID_ARGUMENT;
  const int64_t G = og_grid_points(g);
  const int32_t B = p->n_bootstraps;
  const int threads = p->threads > 0 ? p->threads : 1;
  int rc = og_check_data(g, samples, n);
  if (rc != OG_OK) return rc;

  double t_start = og_now();
  /* ---- initialize_estimator (GradeProEstimator.jl:700-735) ---- */
  int32_t *idx = (int32_t *)idx_in;
  int32_t *idx_own = NULL;
  if (!idx) {
    idx_own = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n * B));
    if (!idx_own) return OG_OUT_OF_MEMORY;
    og_bootstrap_indices(seed, n, B, idx_own);
    idx = idx_own;
  }
  int32_t *ident = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  double *sums = (double *)malloc(sizeof(double) * (size_t)(G * B));
  double *sqs = (double *)malloc(sizeof(double) * (size_t)(G * B));
  double *full = (double *)malloc(sizeof(double) * (size_t)G);
  double *M0 = (double *)malloc(sizeof(double) * 2 * (size_t)(G * B));
  double *V0 = (double *)malloc(sizeof(double) * 2 * (size_t)(G * B));
  double *F0 = (double *)malloc(sizeof(double) * 2 * (size_t)G);
  double *Mt = (double *)malloc(sizeof(double) * 2 * (size_t)(G * B));
  double *Vt = (double *)malloc(sizeof(double) * 2 * (size_t)(G * B));
  double *Ft = (double *)malloc(sizeof(double) * 2 * (size_t)G);
  double *vmr = (double *)malloc(sizeof(double) * (size_t)G);
  double *means = (double *)malloc(sizeof(double) * (size_t)G);
  double *fp1 = (double *)malloc(sizeof(double) * (size_t)G);
  double *fp2 = (double *)malloc(sizeof(double) * (size_t)G);
  double *minima = (double *)malloc(sizeof(double) * (size_t)G);
  uint16_t *cl = (uint16_t *)calloc((size_t)G, sizeof(uint16_t));
  uint16_t *co = (uint16_t *)calloc((size_t)G, sizeof(uint16_t));
  uint8_t *lf = (uint8_t *)calloc((size_t)G, 1);
  int16_t *ch = chosen ? chosen : (int16_t *)malloc(sizeof(int16_t) * (size_t)G);
  rc = OG_OUT_OF_MEMORY;
  if (!ident || !sums || !sqs || !full || !M0 || !V0 || !F0 || !Mt || !Vt || !Ft || !vmr ||
      !means || !fp1 || !fp2 || !minima || !cl || !co || !lf || !ch)
    goto done;
  for (int64_t s = 0; s < n; ++s) ident[s] = (int32_t)s;

  /* initialize_kernels (GradeProEstimator.jl:67-122): Dirac sequences then a
   * forward FFT of every bootstrap slice. */
  og_dirac_sequences(g, samples, n, idx, B, sums, sqs, threads);
  og_dirac_sequences(g, samples, n, ident, 1, full, NULL, 1);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (int32_t b = 0; b < B; ++b) {
    double *m = M0 + 2 * (int64_t)b * G, *v = V0 + 2 * (int64_t)b * G;
    for (int64_t j = 0; j < G; ++j) {
      m[2 * j] = sums[(int64_t)b * G + j];
      m[2 * j + 1] = 0.0;
      v[2 * j] = sqs[(int64_t)b * G + j];
      v[2 * j + 1] = 0.0;
    }
    og_fft_nd(N, g->n, m, -1);
    og_fft_nd(N, g->n, v, -1);
  }
  for (int64_t j = 0; j < G; ++j) {
    F0[2 * j] = full[j];
    F0[2 * j + 1] = 0.0;
  }
  og_fft_nd(N, g->n, F0, -1);

  /* ---- initialize_estimator_propagation (GradeProEstimator.jl:740-796) ---- */
  double bw[OG_MAX_DIMS];
  int have_bw = 1;
  for (int i = 0; i < N; ++i) have_bw &= isfinite(p->bw_final[i]);
  if (have_bw)
    for (int i = 0; i < N; ++i) bw[i] = p->bw_final[i];
  else
    og_silverman(N, n, samples, bw);
  double alpha_s = isnan(p->alpha_s) ? (N == 1 ? 0.0 : 0.06) : p->alpha_s;
  double alpha_os = isnan(p->alpha_os) ? 0.1 : p->alpha_os;
  double eps = isnan(p->eps) ? (N == 1 ? 2.0 : -0.5) : p->eps; /* :541-550 */
  double dt[OG_MAX_DIMS];
  int S = og_ladder(N, bw, p->n_steps, p->bw_step, dt, NULL);
  if (S < 1) { rc = OG_INVALID_ARGUMENT; goto done; }
  double *times = (double *)malloc(sizeof(double) * (size_t)(S * N));
  if (!times) goto done;
  og_ladder(N, bw, p->n_steps, p->bw_step, dt, times);
  int steps_low, steps_over;
  if (og_duration_steps(N, times + (int64_t)(S - 1) * N, dt, alpha_s, alpha_os, &steps_low,
                        &steps_over) != OG_OK) {
    free(times);
    rc = OG_INVALID_ARGUMENT;
    goto done;
  }
  int64_t ksum = 0;
  for (int i = 0; i < N; ++i) ksum += g->n[i];
  double *kgrid = (double *)malloc(sizeof(double) * (size_t)ksum);
  if (!kgrid) { free(times); goto done; }
  {
    int64_t off = 0;
    for (int i = 0; i < N; ++i) {
      og_fftfreq(g->n[i], g->h[i], kgrid + off);
      off += g->n[i];
    }
  }
  double t0[OG_MAX_DIMS];
  for (int i = 0; i < N; ++i) t0[i] = g->h[i] / 2.0; /* Grids.jl:217 */
  for (int64_t j = 0; j < G; ++j) {
    density[j] = NAN; /* ParallelKDE.jl:216 */
    fp1[j] = NAN;
    fp2[j] = NAN;
    minima[j] = NAN;
    ch[j] = (int16_t)S;
  }
  double t_ladder = og_now();

  /* ---- estimate! ladder (GradeProEstimator.jl:968-1051) ---- */
  int S_run = (max_rungs > 0 && max_rungs < S) ? max_rungs : S;
  const double ginv = 1.0 / (double)G;
  for (int k = 0; k < S_run; ++k) {
    const double *tk = times + (int64_t)k * N;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int32_t b = 0; b < B; ++b) {
      double *mt = Mt + 2 * (int64_t)b * G, *vt = Vt + 2 * (int64_t)b * G;
      og_propagate(g, kgrid, tk, t0, M0 + 2 * (int64_t)b * G, V0 + 2 * (int64_t)b * G, mt, vt);
      /* plan_ifft! is the normalised inverse (FourierSpace.jl:454-455) */
      og_fft_nd(N, g->n, mt, +1);
      og_fft_nd(N, g->n, vt, +1);
      for (int64_t j = 0; j < 2 * G; ++j) {
        mt[j] *= ginv;
        vt[j] *= ginv;
      }
      if (og_real_output)
        for (int64_t j = 0; j < G; ++j) {
          mt[2 * j + 1] = 0.0;
          vt[2 * j + 1] = 0.0;
        }
    }
    /* VMR over points; split the point range across threads */
    {
      const int64_t chunk = (G + threads - 1) / threads;
#pragma omp parallel for schedule(static, 1) num_threads(threads)
      for (int c = 0; c < threads; ++c) {
        int64_t j0 = (int64_t)c * chunk, j1 = j0 + chunk < G ? j0 + chunk : G;
        if (j0 >= j1) continue;
        og_vmr_range(N, G, B, Mt, Vt, tk, t0, n, j0, j1, vmr);
      }
    }
    /* full-sample means (propagate_means!, ifft_means!, calculate_means!) */
    og_propagate(g, kgrid, tk, t0, F0, NULL, Ft, NULL);
    og_fft_nd(N, g->n, Ft, +1);
    for (int64_t j = 0; j < 2 * G; ++j) Ft[j] *= ginv;
    if (og_real_output)
      for (int64_t j = 0; j < G; ++j) Ft[2 * j + 1] = 0.0;
    og_full_means(G, Ft, means);
    double dlogt = og_dlogt(N, t0, tk, k == 0 ? NULL : times + (int64_t)(k - 1) * N);
    og_identify_convergence(G, density, means, vmr, fp1, fp2, dlogt, eps, steps_low, steps_over,
                            minima, cl, co, lf, ch, (int16_t)k, 1);
    if (vmr_table) memcpy(vmr_table + (int64_t)k * G, vmr, sizeof(double) * (size_t)G);
    if (means_table) memcpy(means_table + (int64_t)k * G, means, sizeof(double) * (size_t)G);
  }
  /* replace_nans! (GradeProEstimator.jl:1053-1061): last rung's means */
  for (int64_t j = 0; j < G; ++j)
    if (isnan(density[j])) {
      density[j] = means[j];
      ch[j] = (int16_t)S;
    }
  if (info) {
    info->n_rungs = S;
    info->steps_low = steps_low;
    info->steps_over = steps_over;
    info->eps = eps;
    for (int i = 0; i < N; ++i) info->bw_final[i] = bw[i];
    info->seconds_binning_fft = t_ladder - t_start;
    info->seconds_ladder = og_now() - t_ladder;
  }
  free(times);
  free(kgrid);
  rc = OG_OK;
done:
  free(idx_own); free(ident); free(sums); free(sqs); free(full);
  free(M0); free(V0); free(F0); free(Mt); free(Vt); free(Ft);
  free(vmr); free(means); free(fp1); free(fp2); free(minima);
  free(cl); free(co); free(lf);
  if (!chosen) free(ch);
  return rc;
}

/* RoTEstimator.jl:116-265, :serial. */
int og_rot_estimate(const og_grid *g, const double *samples, int64_t n, int rule,
                    double *density, double *time_out) {
  const int N = g->ndims;
  if (N < 1 || N > OG_MAX_DIMS || n < 2 || (rule != 0 && rule != 1)) return OG_INVALID_ARGUMENT;
  int rc = og_check_data(g, samples, n); /* initialize_dirac_sequence -> check_data */
  if (rc != OG_OK) return rc;
  const int64_t G = og_grid_points(g);
  /* propagation_time: factor .* sqrt.(diag(cov(data, dims=2))) (n - 1 normalised) */
  const double ex = -1.0 / (double)(N + 4);
  const double factor = rule == 1 ? pow((double)(n * (int64_t)(N + 2)) / 4.0, ex)
                                  : pow((double)n, ex);
  double t[OG_MAX_DIMS], t0[OG_MAX_DIMS];
  for (int i = 0; i < N; ++i) {
    double s = 0.0;
    for (int64_t k = 0; k < n; ++k) s += samples[k * N + i];
    const double mean = s / (double)n;
    double ss = 0.0;
    for (int64_t k = 0; k < n; ++k) {
      const double d = samples[k * N + i] - mean;
      ss += d * d;
    }
    t0[i] = g->h[i] / 2.0; /* initial_bandwidth, Grids.jl:217 */
    t[i] = factor * sqrt(ss / (double)(n - 1)) - t0[i];
    if (time_out) time_out[i] = t[i];
  }
  int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  double *re = (double *)malloc(sizeof(double) * (size_t)G);
  double *f0 = (double *)malloc(sizeof(double) * 2 * (size_t)G);
  double *ft = (double *)malloc(sizeof(double) * 2 * (size_t)G);
  int64_t ksum = 0;
  for (int i = 0; i < N; ++i) ksum += g->n[i];
  double *kgrid = (double *)malloc(sizeof(double) * (size_t)ksum);
  rc = OG_OUT_OF_MEMORY;
  if (!idx || !re || !f0 || !ft || !kgrid) goto done;
  for (int64_t k = 0; k < n; ++k) idx[k] = (int32_t)k; /* no bootstraps: identity */
  rc = og_dirac_sequences(g, samples, n, idx, 1, re, NULL, 1);
  if (rc != OG_OK) goto done;
  for (int64_t j = 0; j < G; ++j) {
    f0[2 * j] = re[j];
    f0[2 * j + 1] = 0.0;
  }
  og_fft_nd(N, g->n, f0, -1);
  {
    int64_t off = 0;
    for (int i = 0; i < N; ++i) {
      og_fftfreq(g->n[i], g->h[i], kgrid + off);
      off += g->n[i];
    }
  }
  og_propagate(g, kgrid, t, t0, f0, NULL, ft, NULL);
  og_fft_nd(N, g->n, ft, +1);
  {
    const double ginv = 1.0 / (double)G; /* plan_ifft! normalisation */
    for (int64_t j = 0; j < 2 * G; ++j) ft[j] *= ginv;
  }
  og_full_means(G, ft, density); /* copy_density!: abs */
  rc = OG_OK;
done:
  free(idx);
  free(re);
  free(f0);
  free(ft);
  free(kgrid);
  return rc;
}
