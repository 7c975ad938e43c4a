// C ABI of libparallelkde_hip: contexts, plans and the plain estimate.
// The batched and serving forms live in pkde_serve.hip.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "pkde_plan.hpp"

using namespace pkde;

// Device memory of the allocation-free estimate (pkde_gradepro_estimate_1d_
// noalloc): every buffer of its plan is carved from this array.  Each GPU
// holds its own copy; host code reaches it only through the address
// hipGetSymbolAddress returns for the device (pkde_ctx::scratch).
static constexpr uint64_t kScratchBytes = 32ull << 20;
static constexpr int64_t kNoallocMaxSamples = 131072;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace pkde {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return PKDE_OK;
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return fail(PKDE_OUT_OF_MEMORY, m);
  return fail(PKDE_DEVICE_ERROR, m);
}

#define PKDE_TRY(expr, what)                     \
  do {                                           \
    int _rc = check_hip((expr), what);           \
    if (_rc != PKDE_OK) return _rc;              \
  } while (0)

static int dalloc(void** p, size_t bytes, const char* what) {
  if (*p) return PKDE_OK;
  bytes = (bytes + 255) & ~(size_t)255;
  if (bytes == 0) bytes = 256;
  const hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    // a failed allocation is reported here; it must not also surface as the
    // launch error of a later, unrelated estimate
    (void)hipGetLastError();
  }
  return check_hip(e, what);
}

static void dfree(void* p) {
  if (p) (void)hipFree(p);
}

static bool is_pow2(int64_t v) { return v >= 2 && (v & (v - 1)) == 0; }

// Radix list of a length: 16s, one 8/4/2, then 3, 5, 7, 11, 13, then any
// larger prime factors (direct-DFT passes).  False only past 16 passes.
static bool factorize(int L, FftFactors& f) {
  f.L = L;
  f.np = 0;
  int n = L;
  while (n % 16 == 0 && f.np < 16) f.rad[f.np++] = 16, n /= 16;
  if (n % 8 == 0) f.rad[f.np++] = 8, n /= 8;
  else if (n % 4 == 0) f.rad[f.np++] = 4, n /= 4;
  else if (n % 2 == 0) f.rad[f.np++] = 2, n /= 2;
  for (int p : {3, 5, 7, 11, 13})
    while (n % p == 0 && f.np < 16) f.rad[f.np++] = p, n /= p;
  // larger primes: one direct pass each (lds_fft_gen)
  for (int p = 17; p * p <= n; p += 2)
    while (n % p == 0 && f.np < 16) f.rad[f.np++] = p, n /= p;
  if (n > 1 && f.np < 16) f.rad[f.np++] = n, n = 1;
  return n == 1;
}

// Radices 8, 4, 2, 3, 5 for the wave-local mixed-radix FFT; np = 0 when L
// has another prime factor.
static void factorize_wave(int L, FftFactors& f) {
  f.L = L;
  f.np = 0;
  int n = L;
  while (n % 8 == 0 && f.np < 16) f.rad[f.np++] = 8, n /= 8;
  if (n % 4 == 0) f.rad[f.np++] = 4, n /= 4;
  else if (n % 2 == 0) f.rad[f.np++] = 2, n /= 2;
  for (int p : {3, 5})
    while (n % p == 0 && f.np < 16) f.rad[f.np++] = p, n /= p;
  if (n != 1) f.np = 0;
}

// The power-of-two row path of the 2-D/3-D ladder: every axis a power of
// two, Q >= 64 rows, 16 <= n1 <= 512 and the other axes in the wave-local FFT
// range [E, 64 E] (E = 8 in FP64, 16 in FP32), so their passes can write the
// tiled row layout.
static bool rows_path(const GridDev& g, int prec) {
  const int Ew = prec ? 8 : 16;
  auto wave_ok = [&](int L) { return L >= Ew && L <= 64 * Ew; };
  return (g.N == 2 || g.N == 3) && g.pow2 && g.Q >= 64 && g.n[0] >= 16 && g.n[0] <= 512 &&
         wave_ok(g.n[1]) && (g.N == 2 || wave_ok(g.n[2]));
}

static int ilog2(int64_t v) {
  int l = 0;
  while ((1ll << l) < v) ++l;
  return l;
}

static size_t real_size(int prec) { return prec ? sizeof(double) : sizeof(float); }

// Fixed-point fraction bits.  A draw deposits round((w / prod h) 2^fb) per
// corner, with w / prod h in [0, 1] and the corners of one draw summing to 1,
// so a cell of one field holds at most n 2^fb: fb = 61 - ceil(log2(n + 1))
// keeps every cell below 2^61.  Each deposit is exact to 2^-(fb+1) of a whole
// draw, so the most bits the bound allows is the closest match to FP64 sums.
static int frac_bits_for(int64_t n) {
  int lg = 0;
  while ((1ll << lg) <= n) ++lg;
  return 61 - lg;
}

static int64_t fsize(const pkde_plan* pl) { return (int64_t)pl->g.H1 * pl->g.Q; }

// The rung-chunk buffer has y_units() equal parts (two in 3-D: the axis-2
// result and the axis-3 result in the tiled row layout; one where the fused
// slab kernel, whole or in quarters of axis 3, reads the axis-2 result
// directly); each part holds at least one rung of the 2B bootstrap fields.
static bool use_slab3(const pkde_plan* pl) {
  return pl->prec && (slab3_ok(pl->g) || slab3_split_ok(pl->g));
}
static int y_units(const pkde_plan* pl) { return pl->g.N == 3 && !use_slab3(pl) ? 2 : 1; }

// Rungs per chunk: as many as the chunk budget holds (16 GiB, or 1/16 of
// the device memory if that is less), and -- unless the budget was set
// explicitly -- at least 4 when they fit in 1/8 of the device memory; then
// spread evenly over the chunks the ladder needs, so the last chunk is not a
// small remainder.  Every chunk pays the tails of its kernels (the waves of
// its last, partly filled round) and reads the spectra once: at 256^2 the
// ladder takes 11.97 ms with 4 chunks against 12.38 ms with 7; at 64^3
// 45.9 ms with 7 chunks (16 GiB) against 48.4 ms with 14 (8 GiB), and no
// less with 32 GiB.  A plan of its own (not one lent a stream's chunk buffer)
// whose rungs are large takes at least 16 rungs per chunk when they fit in a
// quarter of the device memory: at 128^3 (3.4 GB per rung) 5-rung chunks
// read the spectra 51 times per ladder, and the ladder took 514 ms against
// 485 ms with 19-rung chunks.
static uint64_t y_part_bytes(const pkde_plan* pl) {
  const size_t rs = real_size(pl->prec);
  const uint64_t per_rung = (uint64_t)2 * pl->B * pl->g.fsY * 2 * rs;
  const uint64_t units = (uint64_t)y_units(pl);
  if (per_rung == 0) return (uint64_t)fsize(pl) * 2 * rs * 2;  // :rot: the means pass only
  uint64_t budget = pl->chunk_budget ? pl->chunk_budget : pl->ctx->chunk_bytes;
  // the context's default budget is capped at 1/16 of device memory; a
  // budget given to the plan (a stream lane's or a server slot's, sized by
  // the caller for its streams) at 1/4
  if (!pl->ctx->chunk_fixed && pl->ctx->mem_total)
    budget = std::min<uint64_t>(budget, pl->ctx->mem_total / (pl->chunk_budget ? 4 : 16));
  uint64_t rungs = std::max<uint64_t>(1, budget / units / per_rung);
  // at least 4 rungs when they fit in 1/8 of the device memory
  if (!pl->ctx->chunk_fixed)
    rungs = std::max<uint64_t>(rungs, std::min<uint64_t>(4, pl->ctx->mem_total / 8 / units / per_rung));
  if (!pl->ctx->chunk_fixed && !pl->chunk_budget)
    rungs = std::max<uint64_t>(rungs, std::min<uint64_t>(16, pl->ctx->mem_total / 4 / units / per_rung));
  const uint64_t S = (uint64_t)pl->S;
  rungs = std::min<uint64_t>(rungs, S);
  const uint64_t chunks = (S + rungs - 1) / rungs;
  rungs = (S + chunks - 1) / chunks;
  return std::max<uint64_t>(rungs * per_rung, (uint64_t)fsize(pl) * 2 * rs * 2);
}

uint64_t plan_chunk_bytes(const pkde_plan* pl) {
  return pl->g.N >= 2 ? (uint64_t)y_units(pl) * y_part_bytes(pl) : 0;
}

void plan_borrow_chunk(pkde_plan* pl, void* Y, uint64_t bytes) {
  if (pl->d.Y && !pl->d.Y_borrowed) dfree(pl->d.Y);
  pl->d.Y = Y;
  pl->d.Y_elems = (int64_t)(bytes / (2 * real_size(pl->prec)));
  pl->d.Y_borrowed = true;
}

uint64_t plan_bytes(const pkde_plan* pl, int64_t n, bool with_chunk) {
  const int64_t G = pl->g.G, F = 2 * (int64_t)pl->B + 1, fs = fsize(pl);
  const size_t rs = real_size(pl->prec);
  uint64_t b = 0;
  b += (uint64_t)n * pl->g.N * 8;                 // samples
  b += (uint64_t)F * G * 8;                       // binning accumulators
  b += (uint64_t)F * fs * 16;                     // FP64 spectra
  if (!pl->prec) b += (uint64_t)F * fs * 8;       // FP32 spectra
  b += (uint64_t)pl->S * G * 8;                   // log-VMR table
  if (G <= kConvTermsMaxG) b += (uint64_t)pl->S * G * 16;  // convergence terms
  b += (uint64_t)G * (2 + 2 + 8);                 // chosen, sel, density
  b += (uint64_t)pl->S * pl->g.ksum * 2 * rs;     // Gaussian tables
  if (with_chunk) b += plan_chunk_bytes(pl);
  if (pl->fixed && pl->g.N <= 3) b += binsort_bytes(pl->g, n, pl->B);  // sorted binning
  b += (uint64_t)pl->g.ksum * 2 * 12 + 4096;     // twiddles, Fourier coordinates, scalars
  return b;
}

int plan_reserve(pkde_plan* pl, int64_t n, bool need_idx) {
  PlanBuffers& d = pl->d;
  const int64_t G = pl->g.G, F = 2 * (int64_t)pl->B + 1, fs = fsize(pl);
  const size_t rs = real_size(pl->prec);
  if (!d.acc) {
    // check_memory (GradeProEstimator.jl:1098-1112): the plan's whole
    // footprint (plan_bytes, the footprint model extending get_necessary_memory) against the
    // device's free memory, before anything is allocated
    size_t fr = 0, tot = 0;
    const uint64_t need = plan_bytes(pl, n, !d.Y_borrowed);
    if (hipMemGetInfo(&fr, &tot) == hipSuccess && need > (uint64_t)fr)
      return fail(PKDE_OUT_OF_MEMORY, "Not enough memory available. Required: " +
                                          std::to_string(need >> 20) + " MiB, Available: " +
                                          std::to_string((uint64_t)fr >> 20) + " MiB");
  }
  if (n > d.x_cap) {
    dfree(d.x);
    d.x = nullptr;
    int rc = dalloc((void**)&d.x, (size_t)n * pl->g.N * sizeof(double), "samples");
    if (rc) return rc;
    d.x_cap = n;
  }
  if (need_idx && n * (int64_t)pl->B > d.idx_cap) {
    dfree(d.idx);
    d.idx = nullptr;
    int rc = dalloc((void**)&d.idx, (size_t)n * pl->B * sizeof(int32_t), "bootstrap indices");
    if (rc) return rc;
    d.idx_cap = n * (int64_t)pl->B;
  }
  int rc = PKDE_OK;
  if ((rc = dalloc(&d.acc, (size_t)F * G * 8, "binning accumulators"))) return rc;
  if ((rc = dalloc((void**)&d.specD, (size_t)F * fs * sizeof(cplx<double>), "spectra"))) return rc;
  if (!pl->prec && (rc = dalloc(&d.spec, (size_t)F * fs * sizeof(cplx<float>), "spectra fp32"))) return rc;
  if (pl->g.N >= 2 && !d.Y && !d.Y_borrowed) {
    const uint64_t Y = (uint64_t)y_units(pl) * y_part_bytes(pl);
    if ((rc = dalloc(&d.Y, Y, "rung chunk buffer"))) return rc;
    d.Y_elems = (int64_t)(Y / (2 * rs));
  }
  if (d.S_alloc < pl->S) {
    dfree(d.vmr); dfree(d.times); dfree(d.dlogt); dfree(d.vscale); dfree(d.cV); dfree(d.gM); dfree(d.gV);
    d.vmr = nullptr; d.times = nullptr; d.dlogt = nullptr; d.vscale = nullptr; d.cV = nullptr;
    d.gM = nullptr; d.gV = nullptr;
    const int S = pl->S;
    if ((rc = dalloc((void**)&d.vmr, (size_t)S * G * 8, "vmr table"))) return rc;
    dfree(d.conv);
    d.conv = nullptr;
    if (G <= kConvTermsMaxG &&
        (rc = dalloc((void**)&d.conv, (size_t)S * G * 16, "convergence terms")))
      return rc;
    if ((rc = dalloc((void**)&d.times, (size_t)S * PKDE_MAX_DIMS * 8, "times"))) return rc;
    if ((rc = dalloc((void**)&d.dlogt, (size_t)S * 8, "dlogt"))) return rc;
    if ((rc = dalloc((void**)&d.vscale, (size_t)S * 8, "vscale"))) return rc;
    if ((rc = dalloc((void**)&d.cV, (size_t)S * 8, "cV"))) return rc;
    if ((rc = dalloc(&d.gM, (size_t)S * pl->g.ksum * rs, "gauss M"))) return rc;
    if ((rc = dalloc(&d.gV, (size_t)S * pl->g.ksum * rs, "gauss V"))) return rc;
    d.S_alloc = S;
  }
  if ((rc = dalloc((void**)&d.chosen, (size_t)G * 2, "chosen"))) return rc;
  if ((rc = dalloc((void**)&d.sel, (size_t)G * 2, "sel"))) return rc;
  if ((rc = dalloc((void**)&d.density, (size_t)G * 8, "density"))) return rc;
  if ((rc = dalloc((void**)&d.err, 256, "err"))) return rc;
  if ((rc = dalloc((void**)&d.params, 256, "params"))) return rc;
  if ((rc = dalloc((void**)&d.mean, 256, "mean"))) return rc;
  if ((rc = dalloc((void**)&d.partial, (size_t)kMomentBlocks * PKDE_MAX_DIMS * 8, "partial"))) return rc;
  if (pl->fixed && pl->g.N <= 3) {  // sorted binning: 1-3-D tiles
    const uint64_t need = binsort_bytes(pl->g, n, pl->B);
    if (need > d.binwork_bytes) {
      dfree(d.binwork);
      d.binwork = nullptr;
      d.binwork_bytes = 0;
      if ((rc = dalloc(&d.binwork, need, "binning workspace"))) return rc;
      d.binwork_bytes = need;
    }
  }
  if (!d.invk) {
    // Welford weights 1 / (b + 1), b < B (IEEE division, as on the device)
    std::vector<double> w(pl->B);
    for (int b = 0; b < pl->B; ++b) w[b] = 1.0 / (double)(b + 1);
    if ((rc = dalloc((void**)&d.invk, w.size() * sizeof(double), "welford weights"))) return rc;
    if ((rc = check_hip(hipMemcpy(d.invk, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice),
                        "welford weights")))
      return rc;
  }
  return PKDE_OK;
}

static void graphs_clear(PlanGraphs& G);

static void plan_free(pkde_plan* pl) {
  graphs_clear(pl->graphs);
  PlanBuffers& d = pl->d;
  void* ps[] = {d.x, d.idx, d.acc, d.specD, d.spec, d.Y_borrowed ? nullptr : d.Y, d.vmr, d.conv,
                d.chosen, d.sel, d.density, d.err, d.bw, d.mean, d.partial, d.kfreq, d.times,
                d.dlogt, d.vscale, d.cV, d.params, d.gM, d.gV, d.invk, d.binwork, d.twg64, d.twg32};
  for (void* p : ps) dfree(p);
  d = PlanBuffers();
  for (auto& e : pl->ev)
    if (e) (void)hipEventDestroy(e);
}

// Full-sample means at each point's selected rung (d.sel) -> density: the
// means field's inverse passes for the rungs in use, then |.| at the selected
// rung (calculate_full_means!, DirectSpace.jl:747-764; replace_nans!).
template <typename T>
static int enqueue_means(pkde_plan* pl, const cplx<T>* spec, const T* gM, const T* tw,
                         double* d_density, hipStream_t s) {
  PlanBuffers& d = pl->d;
  const GridDev& g = pl->g;
  const int B = pl->B, S = pl->S, N = g.N;
  const int64_t fs = fsize(pl);
  const T* twg = (const T*)(sizeof(T) == 8 ? (const void*)d.twg64 : (const void*)d.twg32);
  if (N == 1) {
    launch_means_select<T>(g, spec + 2 * (int64_t)B * fs, 0, S, 0, S, gM, tw, twg, d.sel, d_density,
                           s);
  } else {
    if (d.Y_elems < 2 * fs) return fail(PKDE_DEVICE_ERROR, "rung chunk buffer smaller than two rungs");
    int Rm = (int)std::min<int64_t>(S, d.Y_elems / fs);
    if (Rm > 1 && (Rm & 1)) --Rm;
    cplx<T>* Y = (cplx<T>*)d.Y;
    const int fM = 2 * B;
    for (int r0 = 0; r0 < S; r0 += Rm) {
      const int R = std::min(Rm, S - r0);
      launch_axis_fft<T, T, 1>(g, 1, spec, 0, 0, Y, fs, fM, fM, 1, R, r0, gM, gM, fM + 1, tw, twg, s);
      for (int ax = 2; ax < N; ++ax)  // the other axes in place
        launch_axis_fft<T, T, 1>(g, ax, Y, fs, fM, Y, fs, fM, fM, 1, R, r0, gM, gM, fM + 1, tw, twg,
                                 s);
      launch_means_select<T>(g, Y, fs, R, r0, S, gM, tw, twg, d.sel, d_density, s);
    }
  }
  return PKDE_OK;
}

template <typename T>
static int enqueue_typed(pkde_plan* pl, const double* d_x, int64_t n, const int32_t* d_idx,
                         double* d_density, int16_t* d_chosen, hipStream_t s, bool timed,
                         int seg = -1) {
  PlanBuffers& d = pl->d;
  const GridDev& g = pl->g;
  const int B = pl->B, S = pl->S, N = g.N;
  const int64_t G = g.G, F = 2 * (int64_t)B + 1, fs = fsize(pl);
  const T* tw = (const T*)(sizeof(T) == 8 ? (const void*)pl->ctx->tw64 : (const void*)pl->ctx->tw32);
  const T* twg = (const T*)(sizeof(T) == 8 ? (const void*)d.twg64 : (const void*)d.twg32);
  // seg < 0: the whole estimate; 0, 1, 2: only binning and forward
  // transforms, only the ladder, only convergence and means (graph segments)
  const bool all = seg < 0;

  if (all && timed) PKDE_TRY(hipEventRecord(pl->ev[0], s), "event");
  if (all || seg == 0) {
    launch_fill_i32(d.err, 0, 1, s);
    // the sorted fixed-point binning writes every accumulator cell itself
    const bool sorted = pl->fixed && d.binwork;
    if (!sorted) launch_zero_bytes(d.acc, F * G * 8, s);
    if (!pl->bw_given) {
      const double pref = pow(4.0 / ((double)(N + 2) * (double)n), 1.0 / (double)(N + 4));
      launch_silverman(N, d_x, n, pref, d.partial, d.mean, d.bw, s);
    }
    launch_tables<T>(g, S, pl->steps_mode, pl->steps_or_step, d.bw, d.kfreq, n, pl->alpha_s,
                     pl->alpha_os, pl->eps, d.times, d.dlogt, d.vscale, d.cV, d.params, (T*)d.gM,
                     (T*)d.gV, s);
    if (sorted) {
      if (launch_binning_sorted(g, d_x, n, d_idx, B, pl->p.rng_seed, frac_bits_for(n), d.acc, d.err,
                                d.binwork, d.binwork_bytes, s, pl->tune.tile))
        return fail(PKDE_DEVICE_ERROR, "binning workspace does not fit the estimate");
    } else {
      launch_binning(g, d_x, n, d_idx, B, pl->p.rng_seed, pl->fixed, frac_bits_for(n), d.acc, d.err,
                     s);
    }
    {
      double hp = g.h[0];
      for (int i = 1; i < N; ++i) hp *= g.h[i];
      const double two_fb = ldexp(1.0, frac_bits_for(n));
      const double scaleM = (double)n * hp * two_fb;
      const double scaleV = ((double)n * hp) * ((double)n * hp) * two_fb;
      launch_forward_axis1(g, d.acc, pl->fixed, scaleM, scaleV, B, d.specD, pl->ctx->tw64, d.twg64,
                           s);
    }
    for (int ax = 1; ax < N; ++ax)
      launch_axis_fft<double, double, -1>(g, ax, d.specD, 0, 0, d.specD, 0, 0, 0, (int)F, 1, 0,
                                          nullptr, nullptr, 0, pl->ctx->tw64, d.twg64, s);
    if (sizeof(T) == 4) launch_convert<double, float>(d.specD, (cplx<float>*)d.spec, F * fs, s);
  }  // segment 0
  const cplx<T>* spec = sizeof(T) == 4 ? (const cplx<T>*)d.spec : (const cplx<T>*)d.specD;
  if (all && timed) PKDE_TRY(hipEventRecord(pl->ev[1], s), "event");

  // ---- ladder: log-VMR for every rung ----
  const T* gM = (const T*)d.gM;
  const T* gV = (const T*)d.gV;
  if (!all && seg != 1) {
    // not this segment
  } else if (N == 1) {
    launch_ladder_vmr<T>(g, spec, 0, B, n, S, 0, gM, gV, d.cV, d.vscale, tw, twg, d.vmr, s);
  } else {
    // axis 2 (and 3) inverse passes write the tiled row layout the row kernel
    // reads; in 3-D the axis-2 result goes through a second buffer
    const int64_t per_rung = 2 * (int64_t)B * g.fsY;  // tiled rows (>= H1 * Q per field)
    const bool rows = g.pack != 0;  // the power-of-two row path (rows_path)
    const bool rows3 = N == 3 && rows;
    const bool rows_gen = sizeof(T) == 8 && ladder_gen_rows_ok(g);
    const bool slab = sizeof(T) == 8 && use_slab3(pl);  // 3-D: axes 3 and 1 fused
    const int64_t cap = d.Y_elems / y_units(pl);
    if (cap < per_rung) return fail(PKDE_DEVICE_ERROR, "rung chunk buffer smaller than one rung");
    int Rc = (int)std::min<int64_t>(S, cap / per_rung);
    if (pl->tune.rungs > 0) Rc = std::min(Rc, pl->tune.rungs);
    cplx<T>* Y = (cplx<T>*)d.Y;
    cplx<T>* Y2 = Y + cap;
    for (int r0 = 0; r0 < S; r0 += Rc) {
      const int R = std::min(Rc, S - r0);
      // first pass (grid axis 2): spectra read once per chunk.  3-D with the
      // row kernel: the first pass leaves axis 3 contiguous ([k1][j2][j3]),
      // so the axis-3 pass reads whole pencils and writes [q][k1] rows.
      if constexpr (sizeof(T) == 8) {
        if (rows_gen) {
          const bool passes =
              N == 2 ? launch_axis_gen_rows(g, spec, Y, per_rung, 2 * B, R, r0, gM, gV, B, twg, s)
                     : launch_gen3_passes(g, spec, Y2, Y, per_rung, 2 * B, R, r0, gM, gV, B, twg,
                                          s);
          if (!passes || !launch_ladder_rows_gen(g, Y, per_rung, B, n, R, r0, gM, gV, d.cV,
                                                 d.vscale, twg, d.invk, d.vmr, s))
            return fail(PKDE_DEVICE_ERROR, "mixed-radix ladder layout mismatch");
          continue;
        }
      }
      cplx<T>* Y1 = N == 3 && !slab ? Y2 : Y;
      // 3-D: [k1][j2][k3] for the axis-3 pass and the whole slab kernel; the
      // first radix-4 step of axis 3 on store for the slab kernel in quarters
      const int mode1 = N == 2 ? (rows ? 1 : 0)
                               : (rows3 ? (slab && !slab3_ok(g) ? 3 : 2) : 0);
      // the slab kernel takes its axis-3 and axis-1 factors from this pass
      if (!launch_axis_ladder<T>(g, spec, Y1, per_rung, 2 * B, R, r0, gM, gV, B, tw, mode1, s,
                                 pl->tune.split, slab)) {
        if (mode1) return fail(PKDE_DEVICE_ERROR, "ladder pass layout mismatch");
        launch_axis_fft<T, T, 1>(g, 1, spec, 0, 0, Y1, per_rung, 0, 0, 2 * B, R, r0, gM, gV, B, tw,
                                 twg, s);
      }
      if constexpr (sizeof(T) == 8) {
        if (slab) {
          if (!launch_slab3(g, Y1, per_rung, B, n, R, r0, gM, gV, d.cV, d.vscale, tw, d.invk, d.vmr,
                            s))
            return fail(PKDE_DEVICE_ERROR, "slab kernel layout mismatch");
          continue;
        }
      }
      if (N == 3) {
        if (!rows3)
          launch_axis_fft<T, T, 1>(g, 2, Y2, per_rung, 0, Y, per_rung, 0, 0, 2 * B, R, r0, gM, gV,
                                   B, tw, twg, s);
        else if (!launch_axis3_rows<T>(g, Y2, per_rung, Y, per_rung, 2 * B, R, r0, gM, gV, B, tw,
                                       s))
          return fail(PKDE_DEVICE_ERROR, "ladder pass layout mismatch");
      } else {
        // 4-D and up: the remaining axes in place
        for (int ax = 2; ax < N; ++ax)
          launch_axis_fft<T, T, 1>(g, ax, Y, per_rung, 0, Y, per_rung, 0, 0, 2 * B, R, r0, gM, gV,
                                   B, tw, twg, s);
      }
      if (rows) {
        if (!launch_ladder_rows<T>(g, Y, per_rung, B, n, R, r0, gM, gV, d.cV, d.vscale, tw, d.invk,
                                   d.vmr, s))
          return fail(PKDE_DEVICE_ERROR, "row kernel layout mismatch");
      } else
        launch_ladder_vmr<T>(g, Y, per_rung, B, n, R, r0, gM, gV, d.cV, d.vscale, tw, twg, d.vmr,
                             s);
    }
  }
  if (all && timed) PKDE_TRY(hipEventRecord(pl->ev[2], s), "event");

  // ---- convergence + means at the chosen rung ----
  if (all || seg == 2) {
    int16_t* chosen = d_chosen ? d_chosen : d.chosen;
    launch_convergence(G, S, d.vmr, d.dlogt, d.params, chosen, d.sel, d.conv, s);
    const int rc = enqueue_means<T>(pl, spec, gM, tw, d_density, s);
    if (rc) return rc;
  }
  if (all && timed) PKDE_TRY(hipEventRecord(pl->ev[3], s), "event");
  return check_hip(hipGetLastError(), "kernel launch");
}

// :rot (RoTEstimator.jl:147-265): full-sample binning and forward FFT
// (initialize_density, :191-202), one Gaussian propagation by the rule of
// thumb minus h/2 (propagate_kernels!, :219-235), normalised inverse FFT and
// |.| (ifft_density! + copy_density!, :237-265).  The means pass of the
// :gradepro path does the last two steps with every point at rung 0.
template <typename T>
static int enqueue_rot(pkde_plan* pl, const double* d_x, int64_t n, double* d_density,
                       hipStream_t s, bool timed) {
  PlanBuffers& d = pl->d;
  const GridDev& g = pl->g;
  const int N = g.N;
  const int64_t G = g.G, fs = fsize(pl);
  const T* tw = (const T*)(sizeof(T) == 8 ? (const void*)pl->ctx->tw64 : (const void*)pl->ctx->tw32);
  if (timed) PKDE_TRY(hipEventRecord(pl->ev[0], s), "event");
  launch_fill_i32(d.err, 0, 1, s);
  const bool sorted = pl->fixed && d.binwork;
  if (!sorted) launch_zero_bytes(d.acc, G * 8, s);
  // sample standard deviations (prefactor 1) -> propagation time -> factors
  launch_silverman(N, d_x, n, 1.0, d.partial, d.mean, d.bw, s);
  launch_rot_times(g, n, pl->rot_rule, d.bw, d.times, s);
  launch_gauss_tables<T>(g, 1, d.times, d.kfreq, (T*)d.gM, (T*)d.gV, s);
  if (sorted) {
    if (launch_binning_sorted(g, d_x, n, nullptr, 0, 0, frac_bits_for(n), d.acc, d.err, d.binwork,
                              d.binwork_bytes, s))
      return fail(PKDE_DEVICE_ERROR, "binning workspace does not fit the estimate");
  } else {
    launch_binning(g, d_x, n, nullptr, 0, 0, pl->fixed, frac_bits_for(n), d.acc, d.err, s);
  }
  {
    double hp = g.h[0];
    for (int i = 1; i < N; ++i) hp *= g.h[i];
    const double two_fb = ldexp(1.0, frac_bits_for(n));
    const double scaleM = (double)n * hp * two_fb;
    const double scaleV = ((double)n * hp) * ((double)n * hp) * two_fb;
    launch_forward_axis1(g, d.acc, pl->fixed, scaleM, scaleV, 0, d.specD, pl->ctx->tw64, d.twg64,
                         s);
  }
  for (int ax = 1; ax < N; ++ax)
    launch_axis_fft<double, double, -1>(g, ax, d.specD, 0, 0, d.specD, 0, 0, 0, 1, 1, 0, nullptr,
                                        nullptr, 0, pl->ctx->tw64, d.twg64, s);
  const cplx<T>* spec;
  if (sizeof(T) == 4) {
    launch_convert<double, float>(d.specD, (cplx<float>*)d.spec, fs, s);
    spec = (const cplx<T>*)d.spec;
  } else {
    spec = (const cplx<T>*)d.specD;
  }
  if (timed) PKDE_TRY(hipEventRecord(pl->ev[1], s), "event");
  if (timed) PKDE_TRY(hipEventRecord(pl->ev[2], s), "event");
  launch_zero_bytes(d.sel, G * 2, s);     // every point at the one rung
  launch_zero_bytes(d.chosen, G * 2, s);
  const int rc = enqueue_means<T>(pl, spec, (const T*)d.gM, tw, d_density, s);
  if (rc) return rc;
  if (timed) PKDE_TRY(hipEventRecord(pl->ev[3], s), "event");
  return check_hip(hipGetLastError(), "kernel launch");
}

// bw_step with the Silverman bw_final (GradeProEstimator.jl:887-899): the
// ladder length n_steps = min_i floor(bw_final_i / bw_step) needs the
// samples' bandwidths, computed on the device and read back (this estimate
// waits for them), then the plan's S-sized buffers are grown if needed.
static int ladder_from_data(pkde_plan* pl, const double* d_x, int64_t n, hipStream_t s) {
  PlanBuffers& d = pl->d;
  const int N = pl->g.N;
  const double pref = pow(4.0 / ((double)(N + 2) * (double)n), 1.0 / (double)(N + 4));
  launch_silverman(N, d_x, n, pref, d.partial, d.mean, d.bw, s);
  double bw[PKDE_MAX_DIMS];
  PKDE_TRY(hipMemcpyAsync(bw, d.bw, N * sizeof(double), hipMemcpyDeviceToHost, s), "bandwidth");
  PKDE_TRY(hipStreamSynchronize(s), "bandwidth");
  double m = INFINITY;
  for (int i = 0; i < N; ++i) {
    if (!(pl->steps_or_step <= bw[i]))  // also catches a NaN bandwidth
      return fail(PKDE_INVALID_ARGUMENT, "Invalid time step");
    m = std::min(m, floor(bw[i] / pl->steps_or_step));
  }
  if (m + 1 > 32000) return fail(PKDE_INVALID_ARGUMENT, "too many bandwidth steps");
  pl->S = (int)m + 1;
  return plan_reserve(pl, n, false);
}

// Signature of the plan's device buffers: captured graphs hold their
// addresses, so a reallocation must invalidate them.
static uint64_t buffer_sig(const pkde_plan* pl) {
  const PlanBuffers& d = pl->d;
  const void* ps[] = {d.x,      d.idx,    d.acc,   d.specD,  d.spec,   d.Y,      d.vmr,
                      d.conv,   d.chosen, d.sel,   d.density, d.err,   d.bw,     d.mean,
                      d.partial, d.kfreq, d.times, d.dlogt,  d.vscale, d.cV,     d.params,
                      d.gM,     d.gV,     d.invk,  d.binwork, d.twg64, d.twg32};
  uint64_t h = 1469598103934665603ull;  // FNV-1a over the addresses and sizes
  auto mix = [&](uint64_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  for (const void* p : ps) mix((uint64_t)(uintptr_t)p);
  mix(d.binwork_bytes);
  mix((uint64_t)d.Y_elems);
  mix((uint64_t)pl->S);
  return h;
}

static void graphs_clear(PlanGraphs& G) {
  for (auto& e : G.exec)
    if (e) {
      (void)hipGraphExecDestroy(e);
      e = nullptr;
    }
  G.key = GraphKey();
}

// Autotuned ladder launch shapes: the context table holds one entry per
// (kind, N, B, S, precision, axis sizes); a plan looks it up again when the
// table or its S changed, and drops its captured graphs if its shape moved.
static std::vector<int64_t> tune_key(const pkde_plan* pl) {
  std::vector<int64_t> k = {pl->kind, pl->g.N, pl->B, pl->S, pl->prec};
  for (int i = 0; i < pl->g.N; ++i) k.push_back(pl->g.n[i]);
  return k;
}

static void tune_apply(pkde_plan* pl) {
  pkde_ctx* c = pl->ctx;
  const int gen = c->tune_gen.load();
  if (gen == pl->tune_gen && pl->S == pl->tune_S) return;
  TuneShape t;
  {
    std::lock_guard<std::mutex> lk(c->tune_mu);
    const std::vector<int64_t> k = tune_key(pl);
    for (const auto& e : c->tune)
      if (e.first == k) {
        t = e.second;
        break;
      }
  }
  if (!(t == pl->tune)) graphs_clear(pl->graphs);
  pl->tune = t;
  pl->tune_gen = gen;
  pl->tune_S = pl->S;
}

// The estimate as three captured graphs, replayed while the caller repeats
// the same buffers.  The first call with a new key runs directly; the second
// captures; capture failures leave the plan on direct launches.  The timing
// events are recorded between the graph launches, as in the direct form.
template <typename T>
static int enqueue_graphs(pkde_plan* pl, const double* d_x, int64_t n, const int32_t* d_idx,
                          double* d_density, int16_t* d_chosen, hipStream_t s, bool timed) {
  PlanGraphs& G = pl->graphs;
  GraphKey k;
  k.x = d_x;
  k.n = n;
  k.idx = d_idx;
  k.dens = d_density;
  k.chosen = d_chosen;
  k.sig = buffer_sig(pl);
  k.timed = timed;
  if (!(G.exec[0] && k == G.key)) {
    const bool repeat = k == G.last;
    G.last = k;
    if (!repeat) return enqueue_typed<T>(pl, d_x, n, d_idx, d_density, d_chosen, s, timed);
    graphs_clear(G);
    for (int seg = 0; seg < 3; ++seg) {
      hipGraph_t graph = nullptr;
      int rc = PKDE_OK;
      hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        rc = enqueue_typed<T>(pl, d_x, n, d_idx, d_density, d_chosen, s, false, seg);
        e = hipStreamEndCapture(s, &graph);
      }
      if (e == hipSuccess && rc == PKDE_OK && graph)
        e = hipGraphInstantiate(&G.exec[seg], graph, nullptr, nullptr, 0);
      if (graph) (void)hipGraphDestroy(graph);
      if (rc != PKDE_OK || e != hipSuccess || !G.exec[seg]) {
        (void)hipGetLastError();
        graphs_clear(G);
        G.failed = true;
        if (rc != PKDE_OK) return rc;
        return enqueue_typed<T>(pl, d_x, n, d_idx, d_density, d_chosen, s, timed);
      }
    }
    G.key = k;
  }
  for (int seg = 0; seg < 3; ++seg) {
    if (timed) PKDE_TRY(hipEventRecord(pl->ev[seg], s), "event");
    PKDE_TRY(hipGraphLaunch(G.exec[seg], s), "graph launch");
  }
  if (timed) PKDE_TRY(hipEventRecord(pl->ev[3], s), "event");
  return PKDE_OK;
}

int plan_enqueue(pkde_plan* pl, const double* d_x, int64_t n, const int32_t* d_idx,
                 double* d_density, int16_t* d_chosen, hipStream_t s, bool timed) {
  if (n < 2) return fail(PKDE_INVALID_ARGUMENT, "need at least two samples");
  if (n >= (1ll << 31)) return fail(PKDE_INVALID_ARGUMENT, "more than 2^31-1 samples");
  // launch errors are read back after this estimate's launches: drop any
  // error an earlier call left behind
  (void)hipGetLastError();
  int rc = plan_reserve(pl, n, false);
  if (rc) return rc;
  if (pl->S_from_data && pl->kind == 0 && (rc = ladder_from_data(pl, d_x, n, s))) return rc;
  tune_apply(pl);
  pl->last_x = d_x;
  pl->last_idx = d_idx;
  pl->last_n = n;
  if (pl->kind == 1)
    return pl->prec ? enqueue_rot<double>(pl, d_x, n, d_density, s, timed)
                    : enqueue_rot<float>(pl, d_x, n, d_density, s, timed);
  if (pl->graphs.enabled && !pl->graphs.failed && !pl->S_from_data && !pl->debug)
    return pl->prec ? enqueue_graphs<double>(pl, d_x, n, d_idx, d_density, d_chosen, s, timed)
                    : enqueue_graphs<float>(pl, d_x, n, d_idx, d_density, d_chosen, s, timed);
  if (pl->prec) return enqueue_typed<double>(pl, d_x, n, d_idx, d_density, d_chosen, s, timed);
  return enqueue_typed<float>(pl, d_x, n, d_idx, d_density, d_chosen, s, timed);
}

// Launch-shape autotuner.  Ladder candidates: rungs per chunk, the even
// splits of S into the fewest chunks the chunk buffer allows and up to three
// more (more chunks: smaller tails per kernel, more spectra reads), times the
// axis pass's workgroup target (no split, 4000, 8000, 16000); each is timed
// on the plan's resident spectra (the ladder of its last estimate, launched
// alone).  Binning candidates (sorted fixed-point binning, when the last
// estimate's samples and draws are in the plan's own buffers): 8192, 4096 or
// 2048 cells per deposit tile (smaller tiles: more workgroups per CU and
// less imbalance between crowded and empty tiles, more samples read twice at
// tile edges), timed with the binning and forward transforms of that
// estimate.  Each candidate runs once untimed, then `reps` times; the
// fastest goes to the plan and the context table.  Integer deposits and the
// per-rung arithmetic do not depend on the launch shape, so results are
// bit-identical whatever is chosen.
int autotune_plan(pkde_plan* pl, int reps, double* best_ms, double* default_ms) {
  hipStream_t s = pl->ctx->stream;
  PKDE_TRY(hipDeviceSynchronize(), "estimate");
  const GridDev& g = pl->g;
  std::vector<TuneShape> ladder_c(1);
  if (g.N >= 2 && pl->d.Y) {
    const int64_t per_rung = 2 * (int64_t)pl->B * g.fsY;
    const int64_t cap = pl->d.Y_elems / y_units(pl) / per_rung;
    std::vector<int> rungs_c = {0};
    if (cap >= 1) {
      const int S = pl->S;
      const int c0 = (int)((S + cap - 1) / cap);
      for (int c = c0 + 1; c <= c0 + 3 && c <= S; ++c) {
        const int R = (S + c - 1) / c;
        if (std::find(rungs_c.begin(), rungs_c.end(), R) == rungs_c.end()) rungs_c.push_back(R);
      }
    }
    ladder_c.clear();
    for (int R : rungs_c)
      for (int sp : {kAxisSplitTarget, 0, 4000, 16000}) {
        TuneShape t;
        t.rungs = R;
        t.split = sp;
        ladder_c.push_back(t);
      }
  }
  // (1-D tiles are one grid row whatever the capacity)
  const bool bin_tunable = g.N >= 2 && pl->fixed && pl->d.binwork && pl->last_x == pl->d.x &&
                           (pl->last_idx == nullptr || pl->last_idx == pl->d.idx);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PKDE_TRY(hipEventCreate(&e0), "event");
  PKDE_TRY(hipEventCreate(&e1), "event");
  const TuneShape keep = pl->tune;
  // mean time of `reps` runs of segment seg (0: binning + forward
  // transforms, 1: ladder) with shape t, after one untimed run
  auto time_seg = [&](const TuneShape& t, int seg, double& out) -> int {
    pl->tune = t;
    for (int it = 0; it <= reps; ++it) {
      if (it == 1 && hipEventRecord(e0, s) != hipSuccess) return PKDE_DEVICE_ERROR;
      const int rc = pl->prec ? enqueue_typed<double>(pl, pl->last_x, pl->last_n, pl->last_idx,
                                                      nullptr, nullptr, s, false, seg)
                              : enqueue_typed<float>(pl, pl->last_x, pl->last_n, pl->last_idx,
                                                     nullptr, nullptr, s, false, seg);
      if (rc != PKDE_OK) return rc;
    }
    float ms = 0.0f;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      return PKDE_DEVICE_ERROR;
    out = (double)ms / reps;
    return PKDE_OK;
  };
  TuneShape best = keep;
  best.rungs = 0;
  best.split = kAxisSplitTarget;
  double best_t = 1e300, dflt = 0.0;
  int rc = PKDE_OK;
  for (const TuneShape& c : ladder_c) {
    TuneShape t = c;
    t.tile = keep.tile;
    double ms = 0.0;
    if ((rc = time_seg(t, 1, ms)) != PKDE_OK) break;
    if (t.rungs == 0 && t.split == kAxisSplitTarget) dflt = ms;
    if (ms < best_t) {
      best_t = ms;
      best.rungs = t.rungs;
      best.split = t.split;
    }
  }
  // the binning runs recompute the resident spectra and tables bit for bit
  if (rc == PKDE_OK && bin_tunable) {
    double tbest = 1e300;
    for (int tile : {kBinTileCells, kBinTileCells / 2, kBinTileCells / 4}) {
      if (g.n[0] > tile) continue;
      TuneShape t = best;
      t.tile = tile;
      double ms = 0.0;
      if ((rc = time_seg(t, 0, ms)) != PKDE_OK) break;
      if (ms < tbest) {
        tbest = ms;
        best.tile = tile;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != PKDE_OK) {
    pl->tune = keep;
    return rc == PKDE_DEVICE_ERROR ? fail(rc, "autotune timing failed") : rc;
  }
  {
    std::lock_guard<std::mutex> lk(pl->ctx->tune_mu);
    const std::vector<int64_t> k = tune_key(pl);
    bool found = false;
    for (auto& e : pl->ctx->tune)
      if (e.first == k) {
        e.second = best;
        found = true;
      }
    if (!found) pl->ctx->tune.push_back({k, best});
    pl->ctx->tune_gen.fetch_add(1);
  }
  graphs_clear(pl->graphs);
  pl->tune = best;
  pl->tune_gen = pl->ctx->tune_gen.load();
  pl->tune_S = pl->S;
  if (best_ms) *best_ms = best_t;
  if (default_ms) *default_ms = dflt;
  return PKDE_OK;
}

}  // namespace pkde

namespace pkde {

// Host-side part of plan creation (initialize_estimator,
// GradeProEstimator.jl:700-735, and the grid, Grids.jl:67-87): validates the
// grid and keywords and fills every host field of the plan.  No device call.
int plan_setup_host(pkde_ctx* ctx, int ndims, const int64_t* n_per_axis, const double* lo,
                    const double* hi, const double* h, const pkde_gradepro_params* params,
                    pkde_plan* pl) {
  if (!ctx || !pl || !n_per_axis || !lo || !hi || !h)
    return fail(PKDE_INVALID_ARGUMENT, "null argument");
  if (ndims < 1 || ndims > PKDE_MAX_DIMS)
    return fail(PKDE_UNSUPPORTED, "the :rocm estimator supports 1 to 6 dimensions");
  pkde_gradepro_params p;
  if (params)
    p = *params;
  else
    pkde_gradepro_params_default(&p);
  if (p.n_bootstraps < 2) return fail(PKDE_INVALID_ARGUMENT, "n_bootstraps must be at least 2");
  if (p.n_bootstraps > 16383) return fail(PKDE_INVALID_ARGUMENT, "n_bootstraps too large");
  pl->ctx = ctx;
  pl->p = p;
  GridDev& g = pl->g;
  g.N = ndims;
  g.G = 1;
  g.ksum = 0;
  for (int i = 0; i < PKDE_MAX_DIMS; ++i) {
    g.n[i] = 1;
    g.log2n[i] = 0;
    g.lo[i] = g.hi[i] = g.h[i] = 0.0;
    g.koff[i] = 0;
  }
  g.pow2 = 1;
  for (int i = 0; i < ndims; ++i) {
    if (n_per_axis[i] < 2 || n_per_axis[i] > 4096 || !factorize((int)n_per_axis[i], g.fac[i])) {
        return fail(PKDE_UNSUPPORTED,
                  "grid axis sizes must be in [2, 4096]");
    }
    if (!is_pow2(n_per_axis[i])) g.pow2 = 0;
    factorize_wave((int)n_per_axis[i], g.facw[i]);
    if (!(h[i] > 0) || !(hi[i] > lo[i]) || !isfinite(lo[i]) || !isfinite(hi[i])) {
        return fail(PKDE_INVALID_ARGUMENT, "invalid grid bounds or spacing");
    }
    g.n[i] = (int)n_per_axis[i];
    g.log2n[i] = is_pow2(n_per_axis[i]) ? ilog2(n_per_axis[i]) : 0;
    g.lo[i] = lo[i];
    g.hi[i] = hi[i];
    g.h[i] = h[i];
    g.koff[i] = g.ksum;
    g.ksum += g.n[i];
    g.G *= g.n[i];
  }
  g.H1 = g.n[0] / 2 + 1;
  g.Q = g.G / g.n[0];
  pl->B = p.n_bootstraps;
  pl->prec = p.precision ? 1 : 0;
  {
    // tiled row layout: tile width = the k1 columns one block of the axis pass
    // writing it holds (grid axis 2 in 2-D, axis 3 in 3-D; 256 lanes x E values)
    const int E = pl->prec ? 8 : 16;
    const int L = ndims == 3 ? g.n[2] : g.n[1];
    g.pack = rows_path(g, pl->prec) ? 1 : 0;
    g.Hc = g.pack ? (int)(n_per_axis[0] / 2) : g.H1;
    int hp = 1;
    while (hp < g.Hc) hp <<= 1;
    g.TKy = ndims >= 2 ? std::min(hp, std::max(1, 256 * E / L)) : 1;
    // axes that are not powers of two: 8-column tiles (128-byte FP64 rows) of
    // the mixed-radix 2-D ladder
    if (!g.pow2 && ndims >= 2) g.TKy = 8;
    g.log2TKy = ilog2(g.TKy);
    g.ktY = (g.Hc + g.TKy - 1) / g.TKy;
    g.fsY = (int64_t)g.ktY * g.Q * g.TKy;
  }
  pl->fixed = p.deterministic_binning ? 1 : 0;
  // defaults (GradeProEstimator.jl:541-550, 754-766)
  pl->alpha_s = isnan(p.alpha_s) ? (ndims == 1 ? 0.0 : 0.06) : p.alpha_s;
  pl->alpha_os = isnan(p.alpha_os) ? 0.1 : p.alpha_os;
  pl->eps = isnan(p.eps) ? (ndims == 1 ? 2.0 : -0.5) : p.eps;
  if (pl->alpha_s < 0 || pl->alpha_s > 1 || pl->alpha_os < 0 || pl->alpha_os > 1) {
    return fail(PKDE_INVALID_ARGUMENT, "Fraction must be in the range [0, 1]");
  }
  pl->bw_given = true;
  for (int i = 0; i < ndims; ++i) {
    pl->bw_host[i] = p.bw_final[i];
    pl->bw_given = pl->bw_given && isfinite(p.bw_final[i]);
  }
  if (pl->bw_given)
    for (int i = 0; i < ndims; ++i)
      if (!(pl->bw_host[i] > 0)) {
            return fail(PKDE_INVALID_ARGUMENT, "bw_final must be positive");
      }
  // ladder length (GradeProEstimator.jl:860-905)
  if (p.n_steps > 0) {
    pl->steps_mode = 1;
    pl->steps_or_step = (double)p.n_steps;
    pl->S = p.n_steps + 1;
  } else if (isfinite(p.bw_step) && !pl->bw_given) {
    // the ladder length floor(bw_final / bw_step) waits for the samples
    if (!(p.bw_step > 0.0)) {
        return fail(PKDE_INVALID_ARGUMENT, "Invalid time step");
    }
    pl->steps_mode = 0;
    pl->steps_or_step = p.bw_step;
    pl->S = 251;
    pl->S_from_data = true;
  } else if (isfinite(p.bw_step)) {
    double m = INFINITY;
    for (int i = 0; i < ndims; ++i) {
      if (p.bw_step > pl->bw_host[i] || p.bw_step <= 0.0) {
            return fail(PKDE_INVALID_ARGUMENT, "Invalid time step");
      }
      m = std::min(m, floor(pl->bw_host[i] / p.bw_step));
    }
    pl->steps_mode = 0;
    pl->steps_or_step = p.bw_step;
    pl->S = (int)m + 1;
  } else {
    pl->steps_mode = 1;
    pl->steps_or_step = 250.0;
    pl->S = 251;
  }
  if (pl->S > 32000) {
    return fail(PKDE_INVALID_ARGUMENT, "too many bandwidth steps");
  }
  return PKDE_OK;
}

}  // namespace pkde

// ====================================================================== C ABI
extern "C" {

int pkde_version(void) { return PKDE_ABI_VERSION; }

const char* pkde_last_error(void) { return pkde::g_last_error.c_str(); }

void pkde_gradepro_params_default(pkde_gradepro_params* p) {
  p->n_bootstraps = 100;
  p->n_steps = 250;
  p->bw_step = NAN;
  for (int i = 0; i < PKDE_MAX_DIMS; ++i) p->bw_final[i] = NAN;
  p->alpha_s = NAN;
  p->alpha_os = NAN;
  p->eps = NAN;
  p->rng_seed = 0x5eed;
  p->deterministic_binning = 1;
  p->precision = 1;
}

int pkde_context_create(int device, pkde_ctx** out) {
  if (!out) return fail(PKDE_INVALID_ARGUMENT, "null output pointer");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return fail(PKDE_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(PKDE_NO_DEVICE, "device index out of range");
  PKDE_TRY(hipSetDevice(device), "hipSetDevice");
  pkde_ctx* c = new pkde_ctx();
  c->device = device;
  int rc = check_hip(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream");
  if (rc) {
    delete c;
    return rc;
  }
  // twiddles exp(-2 pi i k / L), L = 2..4096, computed in FP64 on the host
  const int total = 8192 - 2;
  std::vector<double> t64(2 * (size_t)total);
  std::vector<float> t32(2 * (size_t)total);
  for (int lg = 1; lg <= 12; ++lg) {
    const int L = 1 << lg;
    for (int k = 0; k < L; ++k) {
      const double a = 6.283185307179586476925286766559 * (double)k / (double)L;
      const double cr = cos(a), ci = -sin(a);
      t64[2 * (size_t)(L - 2 + k)] = cr;
      t64[2 * (size_t)(L - 2 + k) + 1] = ci;
      t32[2 * (size_t)(L - 2 + k)] = (float)cr;
      t32[2 * (size_t)(L - 2 + k) + 1] = (float)ci;
    }
  }
  rc = check_hip(hipMalloc((void**)&c->tw64, t64.size() * sizeof(double)), "twiddles");
  if (!rc) rc = check_hip(hipMalloc((void**)&c->tw32, t32.size() * sizeof(float)), "twiddles");
  if (!rc) rc = check_hip(hipMemcpy(c->tw64, t64.data(), t64.size() * 8, hipMemcpyHostToDevice), "twiddles");
  if (!rc) rc = check_hip(hipMemcpy(c->tw32, t32.data(), t32.size() * 4, hipMemcpyHostToDevice), "twiddles");
  if (rc) {
    pkde_context_destroy(c);
    return rc;
  }
  if (const char* e = getenv("PKDE_CHUNK_BYTES")) {
    c->chunk_bytes = strtoull(e, nullptr, 10);
    c->chunk_fixed = true;
  }
  {
    // this device's copy of the static scratch array, with the device current
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess) {
      c->scratch = (unsigned char*)ptr;
    } else {
      c->scratch = nullptr;  // the allocation-free entry point then refuses to run
      (void)hipGetLastError();
    }
  }
  {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) c->mem_total = tot;
  }
  *out = c;
  return PKDE_OK;
}

void pkde_context_destroy(pkde_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  for (auto& ln : ctx->lanes) {
    if (ln.stream) (void)hipStreamSynchronize(ln.stream);
    for (auto& kv : ln.plans) pkde_gradepro_plan_destroy(kv.second);
    if (ln.Y) (void)hipFree(ln.Y);
    if (ln.stream) (void)hipStreamDestroy(ln.stream);
  }
  ctx->lanes.clear();
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  {
    std::lock_guard<std::mutex> lk(ctx->live_mu);
    for (pkde_plan* pl : ctx->live) pl->ctx = nullptr;  // the caller still owns them
    ctx->live.clear();
  }
  if (ctx->stage) (void)hipHostFree(ctx->stage);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->tw64) (void)hipFree(ctx->tw64);
  if (ctx->tw32) (void)hipFree(ctx->tw32);
  delete ctx;
}

int pkde_context_release_batch(pkde_ctx* ctx) {
  if (!ctx) return fail(PKDE_INVALID_ARGUMENT, "null context");
  std::lock_guard<std::mutex> lk(ctx->mu);
  PKDE_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  for (auto& ln : ctx->lanes) {
    if (ln.stream) PKDE_TRY(hipStreamSynchronize(ln.stream), "batch lane");
    for (auto& kv : ln.plans) pkde_gradepro_plan_destroy(kv.second);
    ln.plans.clear();
    if (ln.Y) (void)hipFree(ln.Y);
    ln.Y = nullptr;
    ln.Y_bytes = 0;
  }
  return PKDE_OK;
}

int pkde_device_free_memory(pkde_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
  if (!ctx) return fail(PKDE_INVALID_ARGUMENT, "null context");
  PKDE_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  size_t f = 0, t = 0;
  PKDE_TRY(hipMemGetInfo(&f, &t), "hipMemGetInfo");
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return PKDE_OK;
}

int pkde_find_grid(pkde_ctx* ctx, int ndims, const double* d_samples, int64_t n,
                   const double* padding, double* lo_out, double* hi_out) {
  if (!ctx || !d_samples || !lo_out || !hi_out) return fail(PKDE_INVALID_ARGUMENT, "null argument");
  if (ndims < 1 || ndims > PKDE_MAX_DIMS || n < 1)
    return fail(PKDE_INVALID_ARGUMENT, "ndims must be in 1..3 and n >= 1");
  std::lock_guard<std::mutex> lk(ctx->mu);
  PKDE_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  double* dp = nullptr;
  const size_t part = (size_t)kMomentBlocks * 2 * ndims;
  int rc = dalloc((void**)&dp, (part + 2 * (size_t)ndims) * sizeof(double), "extrema");
  double ext[2 * PKDE_MAX_DIMS];
  if (!rc) {
    launch_extrema(ndims, d_samples, n, dp, dp + part, ctx->stream);
    rc = check_hip(hipGetLastError(), "extrema launch");
  }
  if (!rc) rc = check_hip(hipMemcpyAsync(ext, dp + part, 2 * ndims * sizeof(double),
                                         hipMemcpyDeviceToHost, ctx->stream), "extrema");
  if (!rc) rc = check_hip(hipStreamSynchronize(ctx->stream), "extrema");
  (void)hipFree(dp);
  if (rc) return rc;
  for (int i = 0; i < ndims; ++i) {
    const double lo = ext[2 * i], hi = ext[2 * i + 1];
    if (!isfinite(lo) || !isfinite(hi))
      return fail(PKDE_INVALID_ARGUMENT, "samples must be finite to find a grid");
    // Grids.jl:275-281: padding 0.1 * |hi - lo| unless given; bounds +/- padding
    const double pad = padding ? padding[i] : 0.1 * fabs(hi - lo);
    lo_out[i] = lo - pad;
    hi_out[i] = hi + pad;
  }
  return PKDE_OK;
}

int pkde_gradepro_plan_create(pkde_ctx* ctx, int ndims, const int64_t* n_per_axis, const double* lo,
                              const double* hi, const double* h, const pkde_gradepro_params* params,
                              pkde_plan** out) {
  if (!ctx || !out || !n_per_axis || !lo || !hi || !h)
    return fail(PKDE_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  pkde_plan* pl = new pkde_plan();
  {
    const int hrc = plan_setup_host(ctx, ndims, n_per_axis, lo, hi, h, params, pl);
    if (hrc) {
      delete pl;
      return hrc;
    }
  }
  const GridDev& g = pl->g;
  int rc = check_hip(hipSetDevice(ctx->device), "hipSetDevice");
  // Fourier coordinates 2pi*fftfreq(n, 1/h) (Grids.jl:225-231)
  std::vector<double> kf(g.ksum);
  for (int i = 0; i < ndims; ++i) {
    const int64_t ni = g.n[i];
    const double mult = ((1.0 / g.h[i]) / (double)ni) * 6.283185307179586;
    const int64_t nn = (ni + 1) >> 1;
    for (int64_t j = 0; j < ni; ++j) kf[g.koff[i] + j] = (double)(j < nn ? j : j - ni) * mult;
  }
  // per-axis twiddles exp(-2 pi i k / n_i) for the mixed-radix kernels
  {
    std::vector<double> t64(2 * (size_t)g.ksum);
    std::vector<float> t32(2 * (size_t)g.ksum);
    for (int i = 0; i < ndims; ++i)
      for (int k = 0; k < g.n[i]; ++k) {
        const double a = 6.283185307179586476925286766559 * (double)k / (double)g.n[i];
        const double cr = cos(a), ci = -sin(a);
        t64[2 * (size_t)(g.koff[i] + k)] = cr;
        t64[2 * (size_t)(g.koff[i] + k) + 1] = ci;
        t32[2 * (size_t)(g.koff[i] + k)] = (float)cr;
        t32[2 * (size_t)(g.koff[i] + k) + 1] = (float)ci;
      }
    if (!rc) rc = dalloc((void**)&pl->d.twg64, t64.size() * 8, "twiddles");
    if (!rc) rc = dalloc((void**)&pl->d.twg32, t32.size() * 4, "twiddles");
    if (!rc) rc = check_hip(hipMemcpy(pl->d.twg64, t64.data(), t64.size() * 8, hipMemcpyHostToDevice), "twiddles");
    if (!rc) rc = check_hip(hipMemcpy(pl->d.twg32, t32.data(), t32.size() * 4, hipMemcpyHostToDevice), "twiddles");
  }
  if (!rc) rc = dalloc((void**)&pl->d.kfreq, kf.size() * 8, "kfreq");
  if (!rc) rc = check_hip(hipMemcpy(pl->d.kfreq, kf.data(), kf.size() * 8, hipMemcpyHostToDevice), "kfreq");
  if (!rc) rc = dalloc((void**)&pl->d.bw, 256, "bw");
  if (!rc && pl->bw_given)
    rc = check_hip(hipMemcpy(pl->d.bw, pl->bw_host, ndims * 8, hipMemcpyHostToDevice), "bw");
  for (int k = 0; k < 4 && !rc; ++k) rc = check_hip(hipEventCreate(&pl->ev[k]), "event");
  if (rc) {
    plan_free(pl);
    delete pl;
    return rc;
  }
  {
    std::lock_guard<std::mutex> lk(ctx->live_mu);
    ctx->live.push_back(pl);
  }
  *out = pl;
  return PKDE_OK;
}

int pkde_rot_plan_create(pkde_ctx* ctx, int ndims, const int64_t* n_per_axis, const double* lo,
                         const double* hi, const double* h, int rule, int precision,
                         pkde_plan** out) {
  if (rule != 0 && rule != 1) return fail(PKDE_INVALID_ARGUMENT, "Rule of thumb not implemented.");
  pkde_gradepro_params p;
  pkde_gradepro_params_default(&p);
  p.n_steps = 1;
  p.precision = precision ? 1 : 0;
  const int rc = pkde_gradepro_plan_create(ctx, ndims, n_per_axis, lo, hi, h, &p, out);
  if (rc) return rc;
  pkde_plan* pl = *out;
  pl->kind = 1;
  pl->rot_rule = rule;
  pl->B = 0;
  pl->S = 1;
  pl->bw_given = false;
  return PKDE_OK;
}

void pkde_gradepro_plan_destroy(pkde_plan* plan) {
  if (!plan) return;
  if (pkde_ctx* ctx = plan->ctx) {
    {
      std::lock_guard<std::mutex> lk(ctx->live_mu);
      auto it = std::find(ctx->live.begin(), ctx->live.end(), plan);
      if (it != ctx->live.end()) ctx->live.erase(it);
    }
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
  }
  plan_free(plan);
  delete plan;
}

uint64_t pkde_gradepro_plan_footprint(const pkde_plan* plan, int64_t n_samples) {
  return (plan && plan->ctx) ? plan_bytes(plan, n_samples) : 0;
}

static int finish_stats(pkde_plan* pl, int64_t n, pkde_stats* st) {
  if (!st) return PKDE_OK;
  memset(st, 0, sizeof(*st));
  st->n_rungs = pl->S;
  double par[3] = {0, 0, 0};
  PKDE_TRY(hipMemcpy(par, pl->d.params, sizeof(par), hipMemcpyDeviceToHost), "params");
  // :gradepro: bw_final of the ladder; :rot: the propagation time per axis
  PKDE_TRY(hipMemcpy(st->bw_final, pl->kind == 1 ? pl->d.times : pl->d.bw, pl->g.N * 8,
                     hipMemcpyDeviceToHost),
           "bw");
  st->eps = par[0];
  st->steps_low = (int32_t)par[1];
  st->steps_over = (int32_t)par[2];
  float ms = 0;
  PKDE_TRY(hipEventElapsedTime(&ms, pl->ev[0], pl->ev[3]), "event");
  st->ms_total = ms;
  PKDE_TRY(hipEventElapsedTime(&ms, pl->ev[1], pl->ev[2]), "event");
  st->ms_ladder = ms;
  const double G = (double)pl->g.G, U = G * pl->S, B = pl->B;
  st->algorithmic_bytes = U * (8.0 * B + 4.0) + (double)n * pl->g.N * 8.0 + (2 * B + 1) * G * 4.0 + G * 6.0;
  return PKDE_OK;
}

int pkde_gradepro_estimate(pkde_plan* pl, const double* samples, int64_t n, const int32_t* bootstrap_idx,
                           double* density_out, int16_t* chosen_rung_out, pkde_stats* stats) {
  if (!pl || !samples || !density_out) return fail(PKDE_INVALID_ARGUMENT, "null argument");
  if (!pl->ctx) return fail(PKDE_INVALID_ARGUMENT, "the plan's context was destroyed");
  if (n < 2) return fail(PKDE_INVALID_ARGUMENT, "need at least two samples");
  std::lock_guard<std::mutex> lk(pl->ctx->mu);
  PKDE_TRY(hipSetDevice(pl->ctx->device), "hipSetDevice");
  hipStream_t s = pl->ctx->stream;
  int rc = plan_reserve(pl, n, bootstrap_idx != nullptr);
  if (rc) return rc;
  PKDE_TRY(hipMemcpyAsync(pl->d.x, samples, (size_t)n * pl->g.N * 8, hipMemcpyHostToDevice, s), "upload samples");
  if (bootstrap_idx)
    PKDE_TRY(hipMemcpyAsync(pl->d.idx, bootstrap_idx, (size_t)n * pl->B * 4, hipMemcpyHostToDevice, s),
             "upload indices");
  rc = plan_enqueue(pl, pl->d.x, n, bootstrap_idx ? pl->d.idx : nullptr, pl->d.density, pl->d.chosen, s, true);
  if (rc) return rc;
  const int64_t G = pl->g.G;
  PKDE_TRY(hipMemcpyAsync(density_out, pl->d.density, (size_t)G * 8, hipMemcpyDeviceToHost, s), "download density");
  if (chosen_rung_out)
    PKDE_TRY(hipMemcpyAsync(chosen_rung_out, pl->d.chosen, (size_t)G * 2, hipMemcpyDeviceToHost, s), "download chosen");
  int err = 0;
  PKDE_TRY(hipMemcpyAsync(&err, pl->d.err, 4, hipMemcpyDeviceToHost, s), "download flag");
  PKDE_TRY(hipStreamSynchronize(s), "estimate");
  if (err & 1) return fail(PKDE_OUT_OF_GRID, "Data points lie beyond grid. Remove them or increase grid boundaries.");
  if (err & 2) return fail(PKDE_INVALID_ARGUMENT, "bootstrap index outside 0..n-1");
  return finish_stats(pl, n, stats);
}

int pkde_gradepro_estimate_device(pkde_plan* pl, const double* d_samples, int64_t n,
                                  const int32_t* d_bootstrap_idx, double* d_density,
                                  int16_t* d_chosen_rung, void* stream) {
  if (!pl || !d_samples || !d_density) return fail(PKDE_INVALID_ARGUMENT, "null argument");
  if (!pl->ctx) return fail(PKDE_INVALID_ARGUMENT, "the plan's context was destroyed");
  PKDE_TRY(hipSetDevice(pl->ctx->device), "hipSetDevice");
  hipStream_t s = stream ? (hipStream_t)stream : pl->ctx->stream;
  return plan_enqueue(pl, d_samples, n, d_bootstrap_idx, d_density, d_chosen_rung, s, true);
}

int pkde_gradepro_check(pkde_plan* pl) {
  if (!pl) return fail(PKDE_INVALID_ARGUMENT, "null plan");
  if (!pl->ctx) return fail(PKDE_INVALID_ARGUMENT, "the plan's context was destroyed");
  if (!pl->d.err || pl->last_n == 0) return fail(PKDE_INVALID_ARGUMENT, "no estimate has run");
  PKDE_TRY(hipSetDevice(pl->ctx->device), "hipSetDevice");
  PKDE_TRY(hipEventSynchronize(pl->ev[3]), "estimate");
  int err = 0;
  PKDE_TRY(hipMemcpy(&err, pl->d.err, 4, hipMemcpyDeviceToHost), "flag");
  if (err & 1) return fail(PKDE_OUT_OF_GRID, "Data points lie beyond grid. Remove them or increase grid boundaries.");
  if (err & 2) return fail(PKDE_INVALID_ARGUMENT, "bootstrap index outside 0..n-1");
  return PKDE_OK;
}

int pkde_gradepro_last_stats(pkde_plan* pl, pkde_stats* stats) {
  if (!pl || !stats) return fail(PKDE_INVALID_ARGUMENT, "null argument");
  if (!pl->ctx) return fail(PKDE_INVALID_ARGUMENT, "the plan's context was destroyed");
  PKDE_TRY(hipSetDevice(pl->ctx->device), "hipSetDevice");
  PKDE_TRY(hipEventSynchronize(pl->ev[3]), "estimate");
  return finish_stats(pl, pl->last_n, stats);
}

int pkde_gradepro_graph_enable(pkde_plan* pl, int enable) {
  if (!pl) return fail(PKDE_INVALID_ARGUMENT, "null plan");
  if (pl->ctx) PKDE_TRY(hipSetDevice(pl->ctx->device), "hipSetDevice");
  pkde::graphs_clear(pl->graphs);
  pl->graphs.enabled = enable != 0;
  pl->graphs.failed = false;
  pl->graphs.last = pkde::GraphKey();
  return PKDE_OK;
}

int pkde_gradepro_autotune(pkde_plan* pl, int32_t reps, double* best_ms_out,
                           double* default_ms_out) {
  if (!pl) return fail(PKDE_INVALID_ARGUMENT, "null plan");
  if (!pl->ctx) return fail(PKDE_INVALID_ARGUMENT, "the plan's context was destroyed");
  if (pl->kind != 0) return fail(PKDE_INVALID_ARGUMENT, "autotune applies to :gradepro plans");
  if (pl->last_n < 2 || !pl->d.vmr)
    return fail(PKDE_INVALID_ARGUMENT, "run one estimate on the plan before tuning it");
  if (reps < 1 || reps > 100) return fail(PKDE_INVALID_ARGUMENT, "reps must be 1..100");
  std::lock_guard<std::mutex> lk(pl->ctx->mu);
  PKDE_TRY(hipSetDevice(pl->ctx->device), "hipSetDevice");
  return pkde::autotune_plan(pl, reps, best_ms_out, default_ms_out);
}

int pkde_gradepro_debug_enable(pkde_plan* pl, int enable) {
  if (!pl) return fail(PKDE_INVALID_ARGUMENT, "null plan");
  pl->debug = enable != 0;
  return PKDE_OK;
}

int pkde_gradepro_debug_vmr(pkde_plan* pl, double* vmr_out, int64_t capacity) {
  if (!pl || !vmr_out) return fail(PKDE_INVALID_ARGUMENT, "null argument");
  // whole rungs only: the first min(capacity / G, S) rungs of the table
  const int64_t G = pl->g.G;
  if (capacity < G) return fail(PKDE_INVALID_ARGUMENT, "vmr buffer holds less than one rung");
  const int64_t need = std::min<int64_t>(capacity / G, (int64_t)pl->S) * G;
  if (!pl->d.vmr) return fail(PKDE_INVALID_ARGUMENT, "no estimate has run");
  if (!pl->ctx) return fail(PKDE_INVALID_ARGUMENT, "the plan's context was destroyed");
  PKDE_TRY(hipSetDevice(pl->ctx->device), "hipSetDevice");
  PKDE_TRY(hipStreamSynchronize(pl->ctx->stream), "sync");
  PKDE_TRY(hipMemcpy(vmr_out, pl->d.vmr, (size_t)need * 8, hipMemcpyDeviceToHost), "vmr");
  return PKDE_OK;
}

int pkde_gradepro_convergence(pkde_ctx* ctx, int64_t G, int32_t S, const double* vmr, const double* dlogt,
                              double eps, int32_t steps_low, int32_t steps_over, int16_t* chosen_out) {
  if (!ctx || !vmr || !dlogt || !chosen_out || G < 1 || S < 1)
    return fail(PKDE_INVALID_ARGUMENT, "invalid argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  PKDE_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  double *dv = nullptr, *dd = nullptr, *dp = nullptr, *dt = nullptr;
  int16_t *dc = nullptr, *ds = nullptr;
  int rc = dalloc((void**)&dv, (size_t)S * G * 8, "vmr");
  // the form an estimate of this grid size uses
  if (!rc && G <= kConvTermsMaxG) rc = dalloc((void**)&dt, (size_t)S * G * 16, "terms");
  if (!rc) rc = dalloc((void**)&dd, (size_t)S * 8, "dlogt");
  if (!rc) rc = dalloc((void**)&dp, 256, "params");
  if (!rc) rc = dalloc((void**)&dc, (size_t)G * 2, "chosen");
  if (!rc) rc = dalloc((void**)&ds, (size_t)G * 2, "sel");
  const double par[3] = {eps, (double)steps_low, (double)steps_over};
  if (!rc) rc = check_hip(hipMemcpy(dv, vmr, (size_t)S * G * 8, hipMemcpyHostToDevice), "upload");
  if (!rc) rc = check_hip(hipMemcpy(dd, dlogt, (size_t)S * 8, hipMemcpyHostToDevice), "upload");
  if (!rc) rc = check_hip(hipMemcpy(dp, par, sizeof(par), hipMemcpyHostToDevice), "upload");
  if (!rc) {
    launch_convergence(G, S, dv, dd, dp, dc, ds, dt, ctx->stream);
    rc = check_hip(hipStreamSynchronize(ctx->stream), "convergence");
  }
  if (!rc) rc = check_hip(hipMemcpy(chosen_out, dc, (size_t)G * 2, hipMemcpyDeviceToHost), "download");
  dfree(dv); dfree(dd); dfree(dp); dfree(dc); dfree(ds); dfree(dt);
  return rc;
}

int pkde_dirac_sequences(pkde_ctx* ctx, int ndims, const int64_t* n_per_axis, const double* lo,
                         const double* hi, const double* h, const double* samples, int64_t n,
                         const int32_t* bootstrap_idx, int32_t B, int deterministic, double* sums_out,
                         double* squares_out) {
  if (!ctx || !n_per_axis || !samples || !bootstrap_idx || !sums_out || B < 1 || n < 1)
    return fail(PKDE_INVALID_ARGUMENT, "invalid argument");
  if (ndims < 1 || ndims > PKDE_MAX_DIMS) return fail(PKDE_UNSUPPORTED, "1 to 6 dimensions");
  std::lock_guard<std::mutex> lk(ctx->mu);
  PKDE_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  GridDev g{};
  g.N = ndims;
  g.G = 1;
  for (int i = 0; i < PKDE_MAX_DIMS; ++i) g.n[i] = 1;
  for (int i = 0; i < ndims; ++i) {
    if (n_per_axis[i] < 2 || !(h[i] > 0)) return fail(PKDE_INVALID_ARGUMENT, "invalid grid");
    g.n[i] = (int)n_per_axis[i];
    g.lo[i] = lo[i];
    g.hi[i] = hi[i];
    g.h[i] = h[i];
    g.G *= g.n[i];
  }
  const int64_t G = g.G, F = 2 * (int64_t)B + 1;
  double* dx = nullptr;
  int32_t* di = nullptr;
  void* acc = nullptr;
  int* derr = nullptr;
  void* work = nullptr;
  const uint64_t wbytes = deterministic ? binsort_bytes(g, n, B) : 0;
  int rc = dalloc((void**)&dx, (size_t)n * ndims * 8, "samples");
  if (!rc && wbytes) rc = dalloc(&work, wbytes, "binning workspace");
  if (!rc) rc = dalloc((void**)&di, (size_t)n * B * 4, "indices");
  if (!rc) rc = dalloc(&acc, (size_t)F * G * 8, "acc");
  if (!rc) rc = dalloc((void**)&derr, 256, "err");
  if (!rc) rc = check_hip(hipMemcpy(dx, samples, (size_t)n * ndims * 8, hipMemcpyHostToDevice), "upload");
  if (!rc) rc = check_hip(hipMemcpy(di, bootstrap_idx, (size_t)n * B * 4, hipMemcpyHostToDevice), "upload");
  int err = 0;
  if (!rc) {
    hipStream_t s = ctx->stream;
    launch_fill_i32(derr, 0, 1, s);
    const int fb = frac_bits_for(n);
    if (!deterministic ||
        launch_binning_sorted(g, dx, n, di, B, 0, fb, acc, derr, work, wbytes, s)) {
      launch_zero_bytes(acc, F * G * 8, s);
      launch_binning(g, dx, n, di, B, 0, deterministic, fb, acc, derr, s);
    }
    rc = check_hip(hipStreamSynchronize(s), "binning");
    if (!rc) rc = check_hip(hipMemcpy(&err, derr, 4, hipMemcpyDeviceToHost), "flag");
    if (!rc) {
      std::vector<double> host((size_t)F * G);
      rc = check_hip(hipMemcpy(host.data(), acc, (size_t)F * G * 8, hipMemcpyDeviceToHost), "download");
      if (!rc) {
        double hp = g.h[0];
        for (int i = 1; i < ndims; ++i) hp *= g.h[i];
        const double two_fb = ldexp(1.0, fb);
        const double scM = (double)n * hp * two_fb;
        const double scV = ((double)n * hp) * ((double)n * hp) * two_fb;
        // the conversion the forward transform applies (k_forward_axis1)
        const double iM = 1.0 / scM, iV = 1.0 / scV;
        const long long* q = (const long long*)host.data();
        for (int64_t i = 0; i < (int64_t)B * G; ++i) {
          sums_out[i] = deterministic ? (double)q[i] * iM : host[i];
          if (squares_out)
            squares_out[i] = deterministic ? (double)q[(int64_t)B * G + i] * iV : host[(int64_t)B * G + i];
        }
      }
    }
  }
  dfree(dx); dfree(di); dfree(acc); dfree(derr); dfree(work);
  if (rc) return rc;
  if (err & 1) return fail(PKDE_OUT_OF_GRID, "Data points lie beyond grid. Remove them or increase grid boundaries.");
  if (err & 2) return fail(PKDE_INVALID_ARGUMENT, "bootstrap index outside 0..n-1");
  return PKDE_OK;
}

// plan_fft! + fourier_statistics! (GradeProEstimator.jl:90-95,
// FourierSpace.jl:25-80) alone: the estimate's forward kernels on given real
// fields (FP64 accumulator mode, scale 1), for primitive-level parity.
int pkde_fft_forward(pkde_plan* pl, int32_t count, const double* fields, double* spectra_out) {
  if (!pl || !fields || !spectra_out) return fail(PKDE_INVALID_ARGUMENT, "null argument");
  if (!pl->ctx) return fail(PKDE_INVALID_ARGUMENT, "the plan's context was destroyed");
  const int64_t F = 2 * (int64_t)pl->B + 1;
  if (count < 1 || count > F) return fail(PKDE_INVALID_ARGUMENT, "count must be in 1 .. 2 B + 1");
  std::lock_guard<std::mutex> lk(pl->ctx->mu);
  PKDE_TRY(hipSetDevice(pl->ctx->device), "hipSetDevice");
  hipStream_t s = pl->ctx->stream;
  (void)hipGetLastError();
  int rc = plan_reserve(pl, 2, false);
  if (rc) return rc;
  PlanBuffers& d = pl->d;
  const GridDev& g = pl->g;
  const int64_t G = g.G, fs = fsize(pl);
  PKDE_TRY(hipMemcpyAsync(d.acc, fields, (size_t)count * G * 8, hipMemcpyHostToDevice, s),
           "upload fields");
  if (count < F) launch_zero_bytes((unsigned char*)d.acc + (size_t)count * G * 8, (F - count) * G * 8, s);
  launch_forward_axis1(g, d.acc, 0, 1.0, 1.0, pl->B, d.specD, pl->ctx->tw64, d.twg64, s);
  for (int ax = 1; ax < g.N; ++ax)
    launch_axis_fft<double, double, -1>(g, ax, d.specD, 0, 0, d.specD, 0, 0, 0, (int)F, 1, 0,
                                        nullptr, nullptr, 0, pl->ctx->tw64, d.twg64, s);
  PKDE_TRY(hipGetLastError(), "kernel launch");
  PKDE_TRY(hipMemcpyAsync(spectra_out, d.specD, (size_t)count * fs * sizeof(cplx<double>),
                          hipMemcpyDeviceToHost, s),
           "download spectra");
  PKDE_TRY(hipStreamSynchronize(s), "forward transform");
  return PKDE_OK;
}

}  // extern "C"

// ------------------------------------------------- allocation-free estimate
namespace {

// The host-computed tables of plan creation, computed on the device with the
// same IEEE operations (so the results are bit-identical): Fourier
// coordinates 2 pi fftfreq(n, 1/h) (Grids.jl:225-231), Welford weights
// 1 / (b + 1), the given bandwidth, zeroed per-axis twiddles (power-of-two
// lengths take the context's table).
__global__ void k_noalloc_tables(int n1, double h1, double* kfreq, int B, double* invk, int bw_given,
                                 double bw0, double* bw, double* twg) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid < n1) {
    const double mult = ((1.0 / h1) / (double)n1) * 6.283185307179586;
    const int nn = (n1 + 1) >> 1;
    kfreq[tid] = (double)(tid < nn ? tid : tid - n1) * mult;
    twg[2 * tid] = 0.0;
    twg[2 * tid + 1] = 0.0;
  }
  if (tid < B) invk[tid] = 1.0 / (double)(tid + 1);
  if (tid == 0 && bw_given) bw[0] = bw0;
}

struct Carve {
  unsigned char* base;
  uint64_t used = 0;
  template <typename T>
  T* take(uint64_t bytes) {
    T* p = (T*)(base + used);
    used += (bytes + 255) & ~(uint64_t)255;
    return p;
  }
};

}  // namespace

extern "C" int pkde_gradepro_estimate_1d_noalloc(pkde_ctx* ctx, int64_t n_points, double lo,
                                                 double hi, double h,
                                                 const pkde_gradepro_params* params,
                                                 const double* d_samples, int64_t n,
                                                 double* d_density, int16_t* d_chosen,
                                                 void* stream) {
  if (!ctx || !d_samples || !d_density) return fail(PKDE_INVALID_ARGUMENT, "null argument");
  if (!ctx->scratch)
    return fail(PKDE_DEVICE_ERROR, "the static scratch array has no address on this device");
  if (n_points < 16 || n_points > 1024 || (n_points & (n_points - 1)))
    return fail(PKDE_UNSUPPORTED, "the allocation-free estimate takes 1-D power-of-two grids of 16 to "
                                  "1024 points");
  if (n < 2 || n > kNoallocMaxSamples)
    return fail(PKDE_INVALID_ARGUMENT, "the allocation-free estimate takes 2 to 131072 samples");
  pkde_gradepro_params p;
  if (params)
    p = *params;
  else
    pkde_gradepro_params_default(&p);
  if (p.n_steps <= 0)
    return fail(PKDE_UNSUPPORTED, "the allocation-free estimate needs n_steps (a bw_step ladder is "
                                  "sized from the samples)");
  if (p.precision != 1 || p.deterministic_binning != 1)
    return fail(PKDE_UNSUPPORTED, "the allocation-free estimate runs the FP64 fixed-point path");
  pkde_plan pl;
  int rc = plan_setup_host(ctx, 1, &n_points, &lo, &hi, &h, &p, &pl);
  if (rc) return rc;
  const GridDev& g = pl.g;
  const int B = pl.B, S = pl.S;
  const int64_t G = g.G, F = 2 * (int64_t)B + 1, fs = fsize(&pl);
  Carve c{ctx->scratch};
  PlanBuffers& d = pl.d;
  d.acc = c.take<void>((uint64_t)F * G * 8);
  d.specD = c.take<cplx<double>>((uint64_t)F * fs * 16);
  d.vmr = c.take<double>((uint64_t)S * G * 8);
  if (G <= kConvTermsMaxG) d.conv = c.take<double>((uint64_t)S * G * 16);
  d.chosen = c.take<int16_t>((uint64_t)G * 2);
  d.sel = c.take<int16_t>((uint64_t)G * 2);
  d.err = c.take<int>(256);
  d.params = c.take<double>(256);
  d.mean = c.take<double>(256);
  d.bw = c.take<double>(256);
  d.partial = c.take<double>((uint64_t)kMomentBlocks * 2 * PKDE_MAX_DIMS * 8);
  d.times = c.take<double>((uint64_t)S * PKDE_MAX_DIMS * 8);
  d.dlogt = c.take<double>((uint64_t)S * 8);
  d.vscale = c.take<double>((uint64_t)S * 8);
  d.cV = c.take<double>((uint64_t)S * 8);
  d.gM = c.take<void>((uint64_t)S * g.ksum * 8);
  d.gV = c.take<void>((uint64_t)S * g.ksum * 8);
  d.invk = c.take<double>((uint64_t)B * 8);
  d.kfreq = c.take<double>((uint64_t)g.ksum * 8);
  d.twg64 = c.take<double>((uint64_t)g.ksum * 16);
  d.binwork_bytes = binsort_bytes(g, n, B);
  d.binwork = c.take<void>(d.binwork_bytes);
  d.S_alloc = S;
  if (c.used > kScratchBytes)
    return fail(PKDE_OUT_OF_MEMORY, "the estimate needs " + std::to_string(c.used >> 20) +
                                        " MiB of scratch; the static array holds " +
                                        std::to_string(kScratchBytes >> 20) + " MiB");
  rc = check_hip(hipSetDevice(ctx->device), "hipSetDevice");
  if (rc) return rc;
  (void)hipGetLastError();
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  const int thr = 256;
  const int blocks = (int)((std::max<int64_t>(g.ksum, B) + thr - 1) / thr);
  k_noalloc_tables<<<blocks, thr, 0, s>>>(g.n[0], g.h[0], d.kfreq, B, d.invk, pl.bw_given ? 1 : 0,
                                          pl.bw_host[0], d.bw, d.twg64);
  return enqueue_typed<double>(&pl, d_samples, n, nullptr, d_density,
                               d_chosen ? d_chosen : d.chosen, s, false);
}
