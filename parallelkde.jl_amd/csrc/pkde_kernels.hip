// HIP kernels of the :gradepro path for gfx950.
//
// Data flow of one estimate (ParallelKDE.jl GradeProEstimator.jl:700-1072):
//   samples (FP64)
//     -> k_binning         linear binning of B resamples + full sample
//                          (DirectSpace.jl:196-242), exact FP64 cell index,
//                          int64 fixed-point or FP64 atomic deposits
//     -> k_forward_axis1   FP64 forward DFT along axis 1, two real fields per
//                          complex transform, Hermitian halves kept
//     -> k_axis_fft        FP64 forward DFT along axes 2..N
//     -> spectra [field][k1][q] in FP32 (or FP64), resident for all rungs
//   per rung chunk:
//     -> k_axis_fft        Gaussian multiply + inverse DFT along axes 2..N
//     -> k_ladder_vmr      Gaussian multiply + inverse DFT along axis 1 of two
//                          bootstraps per complex transform, then the VMR and
//                          an FP64 Welford over bootstraps (calculate_scaled_vmr!)
//   -> k_convergence       per-point state machine over all rungs (find_stability)
//   -> k_means_select      full-sample means at each point's chosen rung
//
// Rungs are independent until the state machine, so every rung of a chunk is
// processed in parallel; the state machine only needs the per-rung log-VMR.
#include "pkde_kernels.hpp"

#include <math.h>

#include <algorithm>

namespace pkde {

// ------------------------------------------------------------------ utilities
__global__ void k_fill_i32(int* p, int v, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_zero_u64(unsigned long long* p, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

static inline int grid_for(int64_t work, int threads = kThreads, int cap = 65536) {
  int64_t b = (work + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

void launch_fill_i32(int* p, int v, int64_t count, hipStream_t s) {
  k_fill_i32<<<grid_for(count), kThreads, 0, s>>>(p, v, count);
}

void launch_zero_bytes(void* p, int64_t bytes, hipStream_t s) {
  // all buffers are allocated with 256-byte granularity
  int64_t words = (bytes + 7) / 8;
  k_zero_u64<<<grid_for(words, kThreads, 1 << 16), kThreads, 0, s>>>((unsigned long long*)p, words);
}

// -------------------------------------------------------------------- binning
// One thread handles four consecutive draws of one replicate (one Philox
// block).  Replicate index B is the full sample (identity draws); only those
// threads raise the out-of-grid flag, so every sample is checked exactly once
// (check_data, DirectSpace.jl:479-498).  Cell index and corner weights use the
// same FP64 operations in the same order as generate_dirac_cpu!
// (DirectSpace.jl:215-237): floor(delta/h), rem(delta, h), h - rem, and the
// corner weight prod(low weights) * prod(high weights) with axis 1 as the most
// significant mask bit.
template <bool FIXED>
__global__ __launch_bounds__(kThreads) void k_binning(GridDev g, const double* __restrict__ x,
                                                       int64_t n, const int32_t* __restrict__ idx,
                                                       int B, uint32_t key0, uint32_t key1,
                                                       double denom, double scaleM, double scaleV,
                                                       void* __restrict__ acc, int* err) {
  const int N = g.N;
  const int64_t nq = (n + 3) >> 2;
  const int64_t total = (int64_t)(B + 1) * nq;
  int64_t stride[kMaxDims];
  stride[0] = 1;
  for (int i = 1; i < N; ++i) stride[i] = stride[i - 1] * g.n[i - 1];

  for (int64_t it = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; it < total;
       it += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(it / nq);
    const int64_t blk = it - (int64_t)b * nq;
    const int64_t s0 = blk * 4;
    int64_t src[4];
    bool src_ok[4] = {true, true, true, true};
    if (b == B) {
#pragma unroll
      for (int q = 0; q < 4; ++q) src[q] = s0 + q;
    } else if (idx) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int64_t v = (s0 + q < n) ? (int64_t)idx[(int64_t)b * n + s0 + q] : 0;
        src_ok[q] = (v >= 0 && v < n);
        src[q] = src_ok[q] ? v : 0;
      }
    } else {
      uint32_t c[4] = {(uint32_t)blk, (uint32_t)((uint64_t)blk >> 32), (uint32_t)b, 0u};
      philox4x32_10(c, key0, key1);
#pragma unroll
      for (int q = 0; q < 4; ++q) src[q] = (int64_t)(((uint64_t)c[q] * (uint64_t)n) >> 32);
    }
    for (int q = 0; q < 4; ++q) {
      if (s0 + q >= n) break;
      if (!src_ok[q]) {
        atomicOr(err, 2);
        continue;
      }
      const double* xs = x + src[q] * N;
      int64_t il[kMaxDims];
      double rl[kMaxDims], rh[kMaxDims];
      bool ok = true;
      for (int i = 0; i < N; ++i) {
        const double xi = xs[i];
        const double delta = xi - g.lo[i];
        const double l = floor(delta / g.h[i]);
        rh[i] = fmod(delta, g.h[i]);
        rl[i] = g.h[i] - rh[i];
        ok = ok && (xi > g.lo[i]) && (xi < g.hi[i]) && (l <= (double)(g.n[i] - 2));
        il[i] = ok ? (int64_t)l : 0;
      }
      if (!ok) {
        if (b == B) atomicOr(err, 1);
        continue;
      }
      for (int c = 0; c < (1 << N); ++c) {
        int64_t off = 0;
        double pl = 1.0, ph = 1.0;
        bool hl = false, hh = false;
        for (int j = 0; j < N; ++j) {
          const int mask = (c >> (N - 1 - j)) & 1;
          off += (mask ? il[j] : il[j] + 1) * stride[j];
          if (mask) {
            pl = hl ? pl * rl[j] : rl[j];
            hl = true;
          } else {
            ph = hh ? ph * rh[j] : rh[j];
            hh = true;
          }
        }
        const double term = (pl * ph) / denom;
        const int64_t fM = (b == B) ? (int64_t)2 * B : (int64_t)b;
        if (FIXED) {
          unsigned long long* a = (unsigned long long*)acc;
          atomicAdd(a + fM * g.G + off, (unsigned long long)llrint(term * scaleM));
          if (b < B)
            atomicAdd(a + ((int64_t)B + b) * g.G + off, (unsigned long long)llrint((term * term) * scaleV));
        } else {
          double* a = (double*)acc;
          atomicAdd(a + fM * g.G + off, term);
          if (b < B) atomicAdd(a + ((int64_t)B + b) * g.G + off, term * term);
        }
      }
    }
  }
}

void launch_binning(const GridDev& g, const double* x, int64_t n, const int32_t* idx, int B,
                    uint64_t seed, int fixed, int frac_bits, void* acc, int* err,
                    hipStream_t s) {
  double hp = g.h[0];
  for (int i = 1; i < g.N; ++i) hp *= g.h[i];
  const double spacing_squared = hp * hp;
  const double denom = (double)n * spacing_squared;
  const double two_fb = ldexp(1.0, frac_bits);
  const double scaleM = (double)n * hp * two_fb;
  const double scaleV = ((double)n * hp) * ((double)n * hp) * two_fb;
  const int64_t work = (int64_t)(B + 1) * ((n + 3) >> 2);
  const int blocks = grid_for(work, kThreads, 1 << 20);
  if (fixed)
    k_binning<true><<<blocks, kThreads, 0, s>>>(g, x, n, idx, B, (uint32_t)seed,
                                                (uint32_t)(seed >> 32), denom, scaleM, scaleV,
                                                acc, err);
  else
    k_binning<false><<<blocks, kThreads, 0, s>>>(g, x, n, idx, B, (uint32_t)seed,
                                                 (uint32_t)(seed >> 32), denom, scaleM, scaleV,
                                                 acc, err);
}

// ------------------------------------------------------ forward DFT, axis 1
// Real fields are paired within their group (sums with sums, squares with
// squares) so both members have the same magnitude; z = a + i b is
// transformed once and split with the Hermitian symmetry of real inputs.
__device__ __forceinline__ void pair_fields(int pi, int B, int& fa, int& fb) {
  const int nMp = (B + 2) / 2;  // sums + full sample: B + 1 fields
  if (pi < nMp) {
    const int a = 2 * pi, c = 2 * pi + 1;
    fa = (a < B) ? a : 2 * B;
    fb = (c < B) ? c : (c == B ? 2 * B : -1);
  } else {
    const int pv = pi - nMp;
    fa = B + 2 * pv;
    fb = (2 * pv + 1 < B) ? fa + 1 : -1;
  }
}

template <bool FIXED>
__device__ __forceinline__ double acc_value(const void* acc, int f, int B, int64_t G, int64_t off,
                                            double invM, double invV) {
  if (FIXED) {
    const double v = (double)((const long long*)acc)[(int64_t)f * G + off];
    return v * ((f >= B && f < 2 * B) ? invV : invM);
  }
  return ((const double*)acc)[(int64_t)f * G + off];
}

// GEN: mixed-radix FFT with the plan's twiddle table (axis length not a power
// of two); otherwise the power-of-two FFT with the context table.
template <bool FIXED, bool GEN>
__global__ __launch_bounds__(kThreads) void k_forward_axis1(GridDev g, const void* __restrict__ acc,
                                                             double invM, double invV, int B, int TQ,
                                                             cplx<double>* __restrict__ specD,
                                                             const cplx<double>* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<double>* z = (cplx<double>*)smem;
  const int n1 = g.n[0], H1 = g.H1;
  const int64_t Q = g.Q;
  const int ntile = (int)((Q + TQ - 1) / TQ);
  const int pi = blockIdx.x / ntile;
  const int64_t q0 = (int64_t)(blockIdx.x % ntile) * TQ;
  int fa, fb;
  pair_fields(pi, B, fa, fb);
  const int tid = threadIdx.x, nthr = blockDim.x;
  for (int id = tid; id < TQ * n1; id += nthr) {
    const int t = id / n1, j = id - t * n1;
    const int64_t off = (q0 + t) * n1 + j;
    double va = 0.0, vb = 0.0;
    if (q0 + t < Q) {
      va = acc_value<FIXED>(acc, fa, B, g.G, off, invM, invV);
      vb = fb >= 0 ? acc_value<FIXED>(acc, fb, B, g.G, off, invM, invV) : 0.0;
    }
    z[t * n1 + j] = {va, vb};
  }
  __syncthreads();
  if constexpr (GEN)
    lds_fft_gen<-1, double>(z, g.fac[0], TQ, n1, tw, tid, nthr);
  else
    lds_fft<-1, double>(z, g.log2n[0], TQ, n1, tw + (n1 - 2), tid, nthr);
  for (int id = tid; id < TQ * H1; id += nthr) {
    const int k = id / TQ, t = id - k * TQ;
    if (q0 + t >= Q) continue;
    const cplx<double> Zk = z[t * n1 + k];
    const cplx<double> Zm = z[t * n1 + (k == 0 ? 0 : n1 - k)];
    const cplx<double> A = {0.5 * (Zk.re + Zm.re), 0.5 * (Zk.im - Zm.im)};
    const cplx<double> Bv = {0.5 * (Zk.im + Zm.im), -0.5 * (Zk.re - Zm.re)};
    const int64_t o = (int64_t)k * Q + q0 + t;
    specD[(int64_t)fa * H1 * Q + o] = A;
    if (fb >= 0) specD[(int64_t)fb * H1 * Q + o] = Bv;
  }
}

void launch_forward_axis1(const GridDev& g, const void* acc, int fixed, double scaleM,
                          double scaleV, int B, cplx<double>* specD, const double* tw64,
                          const double* twg, hipStream_t s) {
  const int n1 = g.n[0];
  int TQ = 2048 / n1;
  if (TQ < 1) TQ = 1;
  if (TQ > g.Q) TQ = (int)g.Q;
  const int npairs = (B + 2) / 2 + (B + 1) / 2;
  const int blocks = npairs * (int)((g.Q + TQ - 1) / TQ);
  const size_t lds = (size_t)TQ * n1 * sizeof(cplx<double>);
  const bool gen = (n1 & (n1 - 1)) != 0;
  const cplx<double>* tw = (const cplx<double>*)(gen ? twg + 2 * g.koff[0] : tw64);
  const double iM = fixed ? 1.0 / scaleM : 1.0, iV = fixed ? 1.0 / scaleV : 1.0;
  if (fixed && gen)
    k_forward_axis1<true, true><<<blocks, kThreads, lds, s>>>(g, acc, iM, iV, B, TQ, specD, tw);
  else if (fixed)
    k_forward_axis1<true, false><<<blocks, kThreads, lds, s>>>(g, acc, iM, iV, B, TQ, specD, tw);
  else if (gen)
    k_forward_axis1<false, true><<<blocks, kThreads, lds, s>>>(g, acc, iM, iV, B, TQ, specD, tw);
  else
    k_forward_axis1<false, false><<<blocks, kThreads, lds, s>>>(g, acc, iM, iV, B, TQ, specD, tw);
}

// ------------------------------------------------------------- axis passes
// Pencils along grid axis 2 or 3 of fields stored [k1][q], q = j2 + n2*j3.
// One field is viewed as [k1][o][e][i]: e runs along the pencil (stride I),
// o and i are the remaining outer / inner indices.  A workgroup owns TK
// consecutive k1 x TI consecutive i for one (rung, field, o); each pencil
// lives in one wavefront (L/16 lanes) so the FFT passes need no block barrier.
// The result is written either in the same layout or in the tiled row layout
// of the row kernel: k1 is cut into tiles of TKy columns, and tile kt of a
// field holds rows q = 0..Q-1 of TKy contiguous values, at kt*Q*TKy + q*TKy.
// Columns of a tile row to store when TKe of its TK columns hold data: up to
// the end of the 128-byte line, so every store covers whole lines.
template <typename T>
__device__ __forceinline__ int tile_cols_written(int TKe, int TK) {
  constexpr int CL = 128 / (int)sizeof(cplx<T>);
  const int w = (TKe + CL - 1) / CL * CL;
  return w < TK ? w : TK;
}

// Streaming (non-temporal) accesses to the rung chunk: it is written once by
// the axis pass and read once by the row or slab kernel, and is far larger
// than the caches.  256^2 FP64: stores 11.41 -> 11.07 ms per ladder, loads
// too 10.92 ms; 128^3 763 -> 746 ms.
constexpr bool kNtStore = true;
constexpr bool kNtLoad = true;

template <typename T>
__device__ __forceinline__ void chunk_store(cplx<T>* p, cplx<T> v) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  if constexpr (kNtStore) {
    v2 x = {v.re, v.im};
    __builtin_nontemporal_store(x, (v2*)p);
  } else {
    *p = v;
  }
}

template <typename T>
__device__ __forceinline__ cplx<T> chunk_load(const cplx<T>* p) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  if constexpr (kNtLoad) {
    const v2 x = __builtin_nontemporal_load((const v2*)p);
    return {x.x, x.y};
  } else {
    return *p;
  }
}

template <typename T>
struct AxisArgs {
  const cplx<T>* src;
  cplx<T>* dst;
  int64_t src_rs, dst_rs;  // rung strides (elements)
  int64_t fsize;           // H1 * Q
  int64_t dst_fsize;       // H1 * Q, or fsY for the tiled row layout
  int fid0, nf, src_fid0, dst_fid0;
  int R, r0;
  const T* gM;
  const T* gV;
  int fsplit, ksum, goff;
  int L, log2L;
  int64_t I, O, Q;
  int H1, TK, TI, ktiles, itiles;
  int log2TK, log2TI, log2itiles, log2O;
  int transposed;  // store the tiled row layout (TK == GridDev::TKy)
  // row index of the tiled store: q' = o * oO + e * oE
  // (o: outer pencil index, e: position along the axis)
  int64_t oO, oE;
  // ladder pass: TJ outer pencils per block, store mode 0 [k1][q], 1 tiled
  // rows, 2 [k1][j2][j3] (axis 3 contiguous, for the 3-D axis-3 pass), 3 the
  // first radix-4 step of the axis-3 inverse applied on store (see
  // k_axis_ladder); pack: column 0 carries X(k1 = 0) + i X(k1 = H1 columns),
  // H1 = n1 / 2
  int TJ, log2TJ, log2otiles, mode, pack;
  int goff3;  // offset of grid axis 3 in the per-rung Gaussian tables
  // 3-D input of k_slab3 (modes 2 and 3): the pencil's axis-3 factor g3(k3)
  // and, except on the packed column 0, its axis-1 factor g1(k1) (both fixed
  // along the pencil) go into the first pass with the axis-2 factors
  int fold;
  int nsplit;  // ladder pass: the chunk's rungs are split over nsplit blocks per tile
  const cplx<T>* tw;
  // mixed-radix plan and twiddle table of the axis (GEN kernels)
  FftFactors fac;
  const cplx<T>* twg;
};

__host__ __device__ constexpr int gen_pencil_stride(int L) { return L + L / 8 + 1; }

// Whether every pass of the plan f fits the wave-local mixed-radix FFT with
// 64 lanes per pencil.
// Whether every pass of the plan f fits the wave-local mixed-radix FFT with
// `lp` lanes per pencil.
static bool wave_gen_fits_lp(const FftFactors& f, int lp) {
  if (f.L < 2 || f.L > 512) return false;
  int Ns = 1;
  for (int p = 0; p < f.np; ++p) {
    const int R = f.rad[p];
    const int per_lane = (f.L / R + lp - 1) / lp;
    if (wave_gen_nb(R) == 0 || per_lane > wave_gen_nb(R)) return false;
    Ns *= R;
  }
  return Ns == f.L;
}

bool wave_gen_fits(const FftFactors& f) { return wave_gen_fits_lp(f, 64); }

// A wavefront's two pencils p0 and p1 (a0, a1: whether each is in use): with
// LP = 32 the two halves of the wave transform them together (lengths whose
// passes fit 32 lanes, e.g. 150 = 5 5 3 2), with LP = 64 one after the
// other.  Each butterfly does the same arithmetic either way.
template <int LP, typename T = double>
__device__ __forceinline__ void wave_fft_gen_pair(cplx<T>* p0, bool a0, cplx<T>* p1, bool a1,
                                                  int lane, const FftFactors& fac,
                                                  const cplx<T>* twl) {
  if constexpr (LP == 32) {
    const bool hi = lane >= 32;
    wave_fft_gen<1, T, 32>(hi ? p1 : p0, lane & 31, hi ? a1 : a0, fac, twl);
  } else {
    if (a0) wave_fft_gen<1, T, 64>(p0, lane, true, fac, twl);
    if (a1) wave_fft_gen<1, T, 64>(p1, lane, true, fac, twl);
  }
}

// Per-rung axis pass for lengths that are not powers of two and fit the
// wave-local mixed-radix FFT (a.fac: radices 2, 3, 4, 5, 8): the block's
// TK x TI pencils sit in LDS in the padded layout, each wavefront transforms
// one pencil at a time, twiddles come from LDS.  Same indexing as k_axis_fft.
template <typename T, int DIR>
__global__ __launch_bounds__(kThreads) void k_axis_fft_wg(AxisArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* buf = (cplx<T>*)smem;
  const int L = a.L;
  const int PL = gen_pencil_stride(L);
  cplx<T>* twl = buf + (int64_t)a.TK * a.TI * PL;
  for (int k = threadIdx.x; k < L; k += kThreads) twl[k] = a.twg[k];
  int64_t bid = blockIdx.x;
  const int it = (int)(bid % a.itiles);
  bid /= a.itiles;
  const int kt = (int)(bid % a.ktiles);
  bid /= a.ktiles;
  const int64_t o = bid % a.O;
  bid /= a.O;
  const int fl = (int)(bid % a.nf);
  const int rr = (int)(bid / a.nf);
  const int f = a.fid0 + fl;
  const int k0 = kt * a.TK;
  const int TKe = (a.H1 - k0) < a.TK ? (a.H1 - k0) : a.TK;
  const int64_t i0 = (int64_t)it * a.TI;
  const int TI = a.TI;
  const int npen = TKe * TI;
  const T* gt = nullptr;
  if (a.gM) gt = (f < a.fsplit ? a.gM : a.gV) + (int64_t)(a.r0 + rr) * a.ksum + a.goff;
  const cplx<T>* src = a.src + (int64_t)rr * a.src_rs + (int64_t)(f - a.src_fid0) * a.fsize;
  cplx<T>* dst = a.dst + (int64_t)rr * a.dst_rs + (int64_t)(f - a.dst_fid0) * a.dst_fsize;
  const int64_t obase = o * L * a.I;
  const int tid = threadIdx.x;
  const int total = npen * L;
  for (int id = tid; id < total; id += kThreads) {
    int kk, ii, e;
    if (TI == 1) {
      kk = id / L;
      e = id - kk * L;
      ii = 0;
    } else {
      ii = id % TI;
      e = (id / TI) % L;
      kk = id / (TI * L);
    }
    cplx<T> v = {(T)0, (T)0};
    if (i0 + ii < a.I) {
      const int64_t q = obase + (int64_t)e * a.I + i0 + ii;
      v = src[(int64_t)(k0 + kk) * a.Q + q];
      if (gt) v = cscale(v, gt[e]);
    }
    buf[(kk * TI + ii) * PL + pad8(e)] = v;
  }
  __syncthreads();
  const int lane = tid & 63, wv = tid >> 6;
  for (int p = wv; p < npen; p += kThreads / 64)
    wave_fft_gen<DIR, T, 64>(buf + p * PL, lane, true, a.fac, twl);
  __syncthreads();
  for (int id = tid; id < total; id += kThreads) {
    int kk, ii, e;
    if (TI == 1) {
      kk = id / L;
      e = id - kk * L;
      ii = 0;
    } else {
      ii = id % TI;
      e = (id / TI) % L;
      kk = id / (TI * L);
    }
    if (i0 + ii >= a.I) continue;
    const int64_t q = obase + (int64_t)e * a.I + i0 + ii;
    dst[(int64_t)(k0 + kk) * a.Q + q] = buf[(kk * TI + ii) * PL + pad8(e)];
  }
}

// Generic form for pencil lengths outside the wave-local range or not powers
// of two (GEN: mixed-radix FFT with the plan's twiddle table).  Inner tiles
// may run past I (non-power-of-two axes); those pencils are not stored.
template <typename T, int DIR, bool GEN>
__global__ __launch_bounds__(kThreads) void k_axis_fft(AxisArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* buf = (cplx<T>*)smem;
  int64_t bid = blockIdx.x;
  const int it = (int)(bid % a.itiles);
  bid /= a.itiles;
  const int kt = (int)(bid % a.ktiles);
  bid /= a.ktiles;
  const int64_t o = bid % a.O;
  bid /= a.O;
  const int fl = (int)(bid % a.nf);
  const int rr = (int)(bid / a.nf);
  const int f = a.fid0 + fl;
  const int k0 = kt * a.TK;
  const int TKe = (a.H1 - k0) < a.TK ? (a.H1 - k0) : a.TK;
  const int64_t i0 = (int64_t)it * a.TI;
  const int TI = a.TI, L = a.L;
  const int npen = TKe * TI;
  const T* gt = nullptr;
  if (a.gM) gt = (f < a.fsplit ? a.gM : a.gV) + (int64_t)(a.r0 + rr) * a.ksum + a.goff;
  const cplx<T>* src = a.src + (int64_t)rr * a.src_rs + (int64_t)(f - a.src_fid0) * a.fsize;
  cplx<T>* dst = a.dst + (int64_t)rr * a.dst_rs + (int64_t)(f - a.dst_fid0) * a.dst_fsize;
  const int64_t obase = o * L * a.I;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int total = npen * L;
  // load: lanes run along the contiguous index (e if I == 1, else i)
  for (int id = tid; id < total; id += nthr) {
    int kk, ii, e;
    if (TI == 1) {
      kk = id / L;
      e = id - kk * L;
      ii = 0;
    } else {
      ii = id % TI;
      e = (id / TI) % L;
      kk = id / (TI * L);
    }
    cplx<T> v = {(T)0, (T)0};
    if (i0 + ii < a.I) {
      const int64_t q = obase + (int64_t)e * a.I + i0 + ii;
      v = src[(int64_t)(k0 + kk) * a.Q + q];
      if (gt) v = cscale(v, gt[e]);
    }
    buf[(kk * TI + ii) * L + e] = v;
  }
  __syncthreads();
  if constexpr (GEN)
    lds_fft_gen<DIR, T>(buf, a.fac, npen, L, a.twg, tid, nthr);
  else
    lds_fft<DIR, T>(buf, a.log2L, npen, L, a.tw + (L - 2), tid, nthr);
  for (int id = tid; id < total; id += nthr) {
    int kk, ii, e;
    if (TI == 1) {
      kk = id / L;
      e = id - kk * L;
      ii = 0;
    } else {
      ii = id % TI;
      e = (id / TI) % L;
      kk = id / (TI * L);
    }
    if (i0 + ii >= a.I) continue;
    const int64_t q = obase + (int64_t)e * a.I + i0 + ii;
    dst[(int64_t)(k0 + kk) * a.Q + q] = buf[(kk * TI + ii) * L + e];
  }
}

// Wave-local form: pencils of compile-time length L = 2^LOG2L in the padded
// LDS layout, one wave-wide FFT call per wave.  A block holds TK x TI
// pencils (both powers of two; pencils with k1 beyond the Hermitian half are
// skipped).  Index arithmetic is shifts and masks only.
template <typename T, int DIR, int LOG2L>
__global__ __launch_bounds__(kThreads) void k_axis_wave(AxisArgs<T> a) {
  constexpr int E = sizeof(T) == 8 ? 8 : 16;  // elements per lane in the FFT
  constexpr int L = 1 << LOG2L;
  constexpr int PL = pencil_stride(L);
  constexpr int lp = L / E;                    // lanes per pencil
  constexpr int ppw = 64 / lp;                 // pencils per wave
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* buf = (cplx<T>*)smem;
  uint32_t bid = blockIdx.x;
  const uint32_t it = bid & (uint32_t)(a.itiles - 1);
  bid >>= a.log2itiles;
  const uint32_t kt = bid % (uint32_t)a.ktiles;
  bid /= (uint32_t)a.ktiles;
  const uint32_t o = bid & (uint32_t)(a.O - 1);
  bid >>= a.log2O;
  const int fl = (int)(bid % (uint32_t)a.nf);
  const int rr = (int)(bid / (uint32_t)a.nf);
  const int f = a.fid0 + fl;
  const int log2TI = a.log2TI, log2TK = a.log2TK;
  const int TI = 1 << log2TI, TK = 1 << log2TK;
  const int k0 = (int)kt * TK;
  const int TKe = (a.H1 - k0) < TK ? (a.H1 - k0) : TK;
  const int TKw = tile_cols_written<T>(TKe, TK);
  const int64_t i0 = (int64_t)it * TI;
  const T* gt = nullptr;
  if (a.gM) gt = (f < a.fsplit ? a.gM : a.gV) + (int64_t)(a.r0 + rr) * a.ksum + a.goff;
  const cplx<T>* src = a.src + (int64_t)rr * a.src_rs + (int64_t)(f - a.src_fid0) * a.fsize;
  cplx<T>* dst = a.dst + (int64_t)rr * a.dst_rs + (int64_t)(f - a.dst_fid0) * a.dst_fsize;
  const int64_t obase = (int64_t)o * L * a.I;
  const int tid = threadIdx.x;
  const int total = (TK * TI) << LOG2L;
  // load: lanes run along the contiguous index (e if TI == 1, else i)
  for (int id = tid; id < total; id += kThreads) {
    const int ii = id & (TI - 1);
    const int e = (id >> log2TI) & (L - 1);
    const int kk = id >> (log2TI + LOG2L);
    if (kk < TKe) {
      const int64_t q = obase + (int64_t)e * a.I + i0 + ii;
      cplx<T> v = src[(int64_t)(k0 + kk) * a.Q + q];
      if (gt) v = cscale(v, gt[e]);
      buf[(kk * TI + ii) * PL + pad8(e)] = v;
    } else if (a.transposed && kk < TKw) {
      // columns past H1 up to the end of the 128-byte line: zeros
      buf[(kk * TI + ii) * PL + pad8(e)] = {(T)0, (T)0};
    }
  }
  __syncthreads();
  {
    const int lane = tid & 63, wv = tid >> 6;
    const int npen = TKe * TI;
    for (int pb = wv * ppw; pb < TK * TI; pb += (kThreads / 64) * ppw) {
      const int p = pb + lane / lp;
      const bool act = p < npen;
      wave_fft<LOG2L, DIR, T, E>(buf + (act ? p : 0) * PL, lane % lp, act, a.tw + (L - 2));
    }
  }
  __syncthreads();
  if (a.transposed) {
    // tiled rows (TI == 1): this block's tile is kt, its rows o * oO + e * oE
    cplx<T>* tile = dst + (int64_t)kt * a.Q * TK;
    // whole 128-byte lines: the last tile's row is padded with zero columns
    for (int id = tid; id < total; id += kThreads) {
      const int kk = id & (TK - 1);
      const int e = id >> log2TK;
      if (kk < TKw) {
        const int64_t q = (int64_t)o * a.oO + (int64_t)e * a.oE;
        chunk_store(tile + q * TK + kk, buf[kk * PL + pad8(e)]);
      }
    }
  } else {
    for (int id = tid; id < total; id += kThreads) {
      const int ii = id & (TI - 1);
      const int e = (id >> log2TI) & (L - 1);
      const int kk = id >> (log2TI + LOG2L);
      if (kk < TKe) {
        const int64_t q = obase + (int64_t)e * a.I + i0 + ii;
        dst[(int64_t)(k0 + kk) * a.Q + q] = buf[(kk * TI + ii) * PL + pad8(e)];
      }
    }
  }
}

// Pencil stride of the ladder axis pass: the padded layout, or in mode 1
// (FP64, L >= 256) unpadded pencils with the XOR swizzle (lds_slot<1>), L + 2
// values apart.  There a pencil's 32 or more lanes read 16 values 32 apart
// per group in the later passes, which the padded layout serves with a 2-way
// bank conflict in every read; swizzled they fill distinct banks, and with
// the pencils 2 (mod 16) values apart the tile store's reads (8 columns at
// two neighbouring rows per 16-lane group) do too.  256^2: ladder 10.79 ->
// 10.72 ms.  At L = 128 the slot arithmetic costs the pass a wave per SIMD
// (126 -> 154 VGPRs) for no measured gain, so shorter axes stay padded.
template <typename T>
constexpr int axis_ladder_swizzled(int mode, int L) {
  return mode == 1 && sizeof(T) == 8 && L >= 256 ? 1 : 0;
}
template <typename T>
constexpr int axis_ladder_stride(int mode, int L) {
  return axis_ladder_swizzled<T>(mode, L) ? L + 2 : pencil_stride(L);
}

// Ladder form of the first inverse axis pass (grid axis 2, contiguous): the
// block keeps TK x TJ spectrum pencils (TK columns k1, TJ outer indices j3)
// of one field in registers -- each lane holds exactly the inputs of its
// first Stockham pass -- and loops over its share of the chunk's R rungs
// (R / nsplit), so the spectra are read once per share rather than once per
// rung.  Per rung:
// Gaussian factor, first pass from registers, remaining passes through LDS,
// store of the block's pencils in layout MODE: 0 [k1][q], 1 the tiled rows
// of the row kernel (2-D; the block's tile is one contiguous run), 2
// [k1][j2][j3] (axis 3 contiguous, for the 3-D axis-3 pass or k_slab3; for
// k_slab3 (a.fold) the pencil's axis-3 and axis-1 factors go in with the
// axis-2 factors, except the axis-1 factors of the packed column 0), 3 the first
// radix-4 decimation-in-frequency step of the axis-3 inverse transform applied
// on store (NT = 512 threads): the block's 32 pencils are k3 = k' + s n3/4
// for 8 consecutive k' and s = 0..3 (each scaled by its axis-3 and axis-1
// factors with the axis-2 factors), and it stores, for each j2 and residue class
// p = j3 mod 4, c_p(k') = w^(p k') sum_s i^(p s) Y(k' + s n3/4), w =
// exp(2 pi i / n3), class-major at [k1][j2][p][k']: each class's values of a
// (k1, j2) are one contiguous run, written by the blocks in whole 128-byte
// lines (8 values of k') and read by the class's workgroup in whole lines;
// the n3/4-point inverse transform of c_p over k' gives the outputs
// j3 = 4 m + p (k_slab3 with SPLIT = 4).  The arithmetic is that of
// k_axis_wave with DIR = +1.
template <typename T, int LOG2L, int MODE, int NT = kThreads>
__global__ __launch_bounds__(NT, NT == 512 ? 4 : 1) void k_axis_ladder(AxisArgs<T> a) {
  constexpr int E = sizeof(T) == 8 ? 8 : 16;
  constexpr int L = 1 << LOG2L;
  constexpr int SWZ = axis_ladder_swizzled<T>(MODE, L);
  constexpr int PL = axis_ladder_stride<T>(MODE, L);
  constexpr int lp = L / E;                  // lanes per pencil
  constexpr int ppw = 64 / lp;               // pencils per wave
  constexpr int LE = ilog2c(E);
  constexpr int R1 = LOG2L >= LE ? E : L;    // radix of the first pass
  constexpr int LOG2R1 = ilog2c(R1);
  constexpr int nb1 = L / R1;
  constexpr int tpp = L / E;
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* buf = (cplx<T>*)smem;
  uint32_t bid = blockIdx.x;
  // rung range of this block (consecutive blocks share a tile's spectra)
  const int sp = (int)(bid % (uint32_t)a.nsplit);
  bid /= (uint32_t)a.nsplit;
  const int rbeg = sp * a.R / a.nsplit, rend = (sp + 1) * a.R / a.nsplit;
  uint32_t kt, ot;
  if constexpr (MODE == 3) {
    // outer tiles fastest: blocks that run together fill one k1 column
    ot = bid & ((1u << a.log2otiles) - 1u);
    bid >>= a.log2otiles;
    kt = bid % (uint32_t)a.ktiles;
    bid /= (uint32_t)a.ktiles;
  } else {
    kt = bid % (uint32_t)a.ktiles;
    bid /= (uint32_t)a.ktiles;
    ot = bid & ((1u << a.log2otiles) - 1u);
    bid >>= a.log2otiles;
  }
  const int f = a.fid0 + (int)bid;
  const int log2TK = a.log2TK, log2TJ = a.log2TJ;
  const int TK = 1 << log2TK, TJ = 1 << log2TJ;
  const int k0 = (int)kt * TK;
  const int TKe = (a.H1 - k0) < TK ? (a.H1 - k0) : TK;
  const int64_t o0 = (int64_t)ot * TJ;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int p = wv * ppw + lane / lp;   // this lane's pencil: column k0 + kk, outer o0 + jj
  const int kk = p >> log2TJ, jj = p & (TJ - 1);
  const int t = lane % lp;
  const bool act = kk < TKe && p < TK * TJ;
  const cplx<T>* src = a.src + (int64_t)(f - a.src_fid0) * a.fsize;
  const T* gbase = (f < a.fsplit ? a.gM : a.gV) + a.goff;
  // the pencil's outer index: o0 + jj, or in mode 3 k3 = k' + s n3/4 with
  // k' = Cq ot + (jj mod Cq), s = jj / Cq, Cq = TJ / 4 values of k'
  const int log2Cq = log2TJ - 2, Cq = 1 << log2Cq;
  const int64_t oq = MODE == 3 ? (int64_t)ot * Cq + (jj & (Cq - 1)) +
                                       (int64_t)(jj >> log2Cq) * (a.O >> 2)
                                 : o0 + jj;
  cplx<T> x[E];
  if (act) {
    const cplx<T>* col = src + (int64_t)(k0 + kk) * a.Q + oq * L;
#pragma unroll
    for (int u = 0; u < E / R1; ++u)
#pragma unroll
      for (int r = 0; r < R1; ++r) x[u * R1 + r] = col[t + u * tpp + r * nb1];
    if (a.pack && k0 + kk == 0) {
      // X(0) + i X(n1/2): the Nyquist column shares column 0's pencil
      const cplx<T>* colN = src + (int64_t)a.H1 * a.Q + oq * L;
#pragma unroll
      for (int u = 0; u < E / R1; ++u)
#pragma unroll
        for (int r = 0; r < R1; ++r) {
          const cplx<T> h = colN[t + u * tpp + r * nb1];
          cplx<T>& v = x[u * R1 + r];
          v = {v.re - h.im, v.im + h.re};
        }
    }
  }
  cplx<T>* pb = buf + (act ? p : 0) * PL;
  const int log2np = log2TK + log2TJ;
  const int total = 1 << (log2np + LOG2L);
  const int TKw = tile_cols_written<T>(TKe, TK);
  if (MODE == 1 && TKe < TKw) {
    // columns past H1 up to the end of the 128-byte line: zero pencils,
    // stored with the tile (no pass writes them: their lanes are inactive)
    for (int id = tid; id < ((TKw - TKe) << LOG2L); id += NT)
      buf[(TKe + (id >> LOG2L)) * PL + lds_slot<SWZ, LE>(id & (L - 1))] = {(T)0, (T)0};
  }

  // the 64-point [k1][j2][j3] pass (64^3 ladders) holds the twiddles of its
  // second pass in registers, loaded once, instead of reading them every rung
  constexpr bool RT = MODE == 2 && sizeof(T) == 8 && LOG2L == 6;
  constexpr int NTW = RT ? wave_tw_count<LOG2L, LOG2R1, E>() : 1;
  cplx<T> twr[NTW];
  if constexpr (RT) wave_tw_load<LOG2L, LOG2R1, E, 1, T, 0, NTW>(t, a.tw + (L - 2), twr);
  for (int rr = rbeg; rr < rend; ++rr) {
    const T* gt = gbase + (int64_t)(a.r0 + rr) * a.ksum;
    // fold: the pencil's axis-3 factor (k3 = oq) times its axis-1 factor
    // (k1 = k0 + kk; the axis-1 tables are at offset 0), one scalar per rung
    T g13 = (T)1;
    const bool fold = MODE == 3 || (MODE == 2 && a.fold);
    if (fold) {
      const T* gr = gt - a.goff;
      const int kc = act ? k0 + kk : 0;
      const T g3 = gr[a.goff3 + (act ? oq : 0)];
      g13 = a.pack && kc == 0 ? g3 : g3 * gr[kc];
    }
    if (act) {
#pragma unroll
      for (int u = 0; u < E / R1; ++u) {
        const int j = t + u * tpp;
        cplx<T> w[R1];
#pragma unroll
        for (int r = 0; r < R1; ++r)
          w[r] = cscale(x[u * R1 + r], fold ? gt[j + r * nb1] * g13 : gt[j + r * nb1]);
        dft_reg<R1, 1, T>(w);
#pragma unroll
        for (int r = 0; r < R1; ++r) pb[lds_slot<SWZ, LE>((j << LOG2R1) + r)] = w[r];
      }
    }
    wave_sync();
    if constexpr (RT)
      wave_fft_from_rt<LOG2L, LOG2R1, E, 1, T, 0, NTW>(pb, t, act, twr);
    else
      wave_fft_from<LOG2L, LOG2R1, E, 1, T, SWZ>(pb, t, act, a.tw + (L - 2));
    __syncthreads();
    cplx<T>* dst = a.dst + (int64_t)rr * a.dst_rs + (int64_t)(f - a.dst_fid0) * a.dst_fsize;
    if constexpr (MODE == 1) {
      // tiled rows (2-D, TJ == 1, TK == TKy): the block's tile is contiguous;
      // rows are written in whole 128-byte lines (zero columns pad the last
      // tile), since stores that cover only part of a line cost far more
      // HBM time
      cplx<T>* tile = dst + (int64_t)kt * a.Q * TK;
      for (int id = tid; id < total; id += NT) {
        const int k = id & (TK - 1);
        const int e = id >> log2TK;
        if (k < TKw) chunk_store(tile + (o0 * L + e) * TK + k, buf[k * PL + lds_slot<SWZ, LE>(e)]);
      }
    } else if constexpr (MODE == 2) {
      // [k1][j2][j3]: consecutive j3 (outer index) for one (k1, j2)
      for (int id = tid; id < total; id += NT) {
        const int j = id & (TJ - 1);
        const int e = (id >> log2TJ) & (L - 1);
        const int k = id >> (log2TJ + LOG2L);
        if (k < TKe)
          chunk_store(dst + (int64_t)(k0 + k) * a.Q + (int64_t)e * a.O + o0 + j,
                      buf[((k << log2TJ) + j) * PL + pad8(e)]);
      }
    } else if constexpr (MODE == 3) {
      // class-major radix-4 step: a thread takes one (k1, k' = Cq ot + c, j2),
      // reads the four values Y(k' + s n3/4) once and stores all four classes
      // at [p][k']; 8 lanes write the 8 values of k' of one class, so every
      // store instruction covers whole 128-byte lines
      const int c = tid & (Cq - 1);
      const int64_t q0 = (int64_t)ot * Cq + c, cs = a.O >> 2;
      for (int id = tid; id < (total >> 2); id += NT) {
        const int e = (id >> log2Cq) & (L - 1);
        const int k = id >> (log2Cq + LOG2L);
        if (k < TKe) {
          cplx<T> y[4];
#pragma unroll
          for (int sg = 0; sg < 4; ++sg) y[sg] = buf[((k << log2TJ) + c + Cq * sg) * PL + pad8(e)];
          cplx<T>* o = dst + (int64_t)(k0 + k) * a.Q + (int64_t)e * a.O + q0;
#pragma unroll
          for (int pc = 0; pc < 4; ++pc) {
            cplx<T> acc = {(T)0, (T)0};
#pragma unroll
            for (int sg = 0; sg < 4; ++sg) {
              switch ((pc * sg) & 3) {  // times i^(p s)
                case 0: acc = {acc.re + y[sg].re, acc.im + y[sg].im}; break;
                case 1: acc = {acc.re - y[sg].im, acc.im + y[sg].re}; break;
                case 2: acc = {acc.re - y[sg].re, acc.im - y[sg].im}; break;
                default: acc = {acc.re + y[sg].im, acc.im - y[sg].re}; break;
              }
            }
            // exp(+2 pi i p k' / n3), from the table (held in registers it spills)
            const cplx<T> t4 = a.tw[(a.O - 2) + ((pc * (int)q0) & (int)(a.O - 1))];
            const cplx<T> wp = {t4.re, -t4.im};
            const cplx<T> out = {acc.re * wp.re - acc.im * wp.im, acc.re * wp.im + acc.im * wp.re};
            chunk_store(o + pc * cs, out);
          }
        }
      }
    } else {
      for (int id = tid; id < total; id += NT) {
        const int e = id & (L - 1);
        const int j = (id >> LOG2L) & (TJ - 1);
        const int k = id >> (LOG2L + log2TJ);
        if (k < TKe)
          dst[(int64_t)(k0 + k) * a.Q + (o0 + j) * L + e] = buf[((k << log2TJ) + j) * PL + pad8(e)];
      }
    }
    __syncthreads();
  }
}

template <int LOG2L>
static bool axis_ladder512(unsigned blocks, size_t lds, hipStream_t s,
                           const AxisArgs<double>& a) {
  static const bool attr = [] {
    (void)hipFuncSetAttribute((const void*)k_axis_ladder<double, LOG2L, 3, 512>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
    return true;
  }();
  (void)attr;
  k_axis_ladder<double, LOG2L, 3, 512><<<blocks, 512, lds, s>>>(a);
  return true;
}

// The store layout is a template parameter, so each kernel carries only its
// own store (with the layout chosen at run time the 2-D pass ran 1.6 % slower).
template <typename T, int MODE>
static bool axis_ladder_launch_mode(int log2L, unsigned blocks, size_t lds, hipStream_t s,
                                    const AxisArgs<T>& a) {
  if constexpr (sizeof(T) == 8) {
    switch (log2L) {
      case 3: k_axis_ladder<T, 3, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      case 4: k_axis_ladder<T, 4, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      case 5: k_axis_ladder<T, 5, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      case 6: k_axis_ladder<T, 6, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      case 7: k_axis_ladder<T, 7, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      case 8: k_axis_ladder<T, 8, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      case 9: k_axis_ladder<T, 9, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      default: return false;
    }
  } else {
    switch (log2L) {
      case 4: k_axis_ladder<T, 4, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      case 5: k_axis_ladder<T, 5, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      case 6: k_axis_ladder<T, 6, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      case 7: k_axis_ladder<T, 7, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      case 8: k_axis_ladder<T, 8, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      case 9: k_axis_ladder<T, 9, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      case 10: k_axis_ladder<T, 10, MODE><<<blocks, kThreads, lds, s>>>(a); return true;
      default: return false;
    }
  }
}

template <typename T>
static bool axis_ladder_launch(int log2L, unsigned blocks, size_t lds, hipStream_t s,
                               const AxisArgs<T>& a) {
  switch (a.mode) {
    case 0: return axis_ladder_launch_mode<T, 0>(log2L, blocks, lds, s, a);
    case 1: return axis_ladder_launch_mode<T, 1>(log2L, blocks, lds, s, a);
    case 2: return axis_ladder_launch_mode<T, 2>(log2L, blocks, lds, s, a);
    default: break;
  }
  if constexpr (sizeof(T) == 8) {
    if (a.mode == 3) {
      // 32 pencils per block, 512 threads, 72 KB of LDS (above the 64 KB
      // default: the attribute is set once per instantiation)
      if (lds > 96 * 1024) return false;
      switch (log2L) {
        case 4: return axis_ladder512<4>(blocks, lds, s, a);
        case 5: return axis_ladder512<5>(blocks, lds, s, a);
        case 6: return axis_ladder512<6>(blocks, lds, s, a);
        case 7: return axis_ladder512<7>(blocks, lds, s, a);
        default: return false;
      }
    }
  }
  return false;
}

// First inverse pass of a rung chunk along grid axis 2 for the 2B bootstrap
// fields (fields [0, B) use the means factors, [B, 2B) the variance factors):
// spec [field][k1][q] -> dst rungs [r0, r0 + R) in layout `mode` (see
// k_axis_ladder).  Returns false when the axis length is outside the
// wave-local range (the caller then uses the per-rung pass).
template <typename T>
bool launch_axis_ladder(const GridDev& g, const cplx<T>* spec, cplx<T>* dst,
                        int64_t dst_rung_stride, int nf, int R, int r0, const T* gM,
                        const T* gV, int fsplit, const T* tw, int mode, hipStream_t s,
                        int split_target, bool fold) {
  AxisArgs<T> a{};
  const int Ef = sizeof(T) == 8 ? 8 : 16;
  a.L = g.n[1];
  a.log2L = g.log2n[1];
  if (g.N > 3 || a.L < Ef || a.L > 64 * Ef || (a.L & (a.L - 1))) return false;
  a.src = spec;
  a.dst = dst;
  a.src_rs = 0;
  a.dst_rs = dst_rung_stride;
  a.fsize = (int64_t)g.H1 * g.Q;  // the spectra: H1 columns per field
  a.fid0 = 0;
  a.nf = nf;
  a.src_fid0 = 0;
  a.dst_fid0 = 0;
  a.R = R;
  a.r0 = r0;
  a.gM = gM;
  a.gV = gV;
  a.fsplit = fsplit;
  a.ksum = g.ksum;
  a.goff = g.koff[1];
  a.Q = g.Q;
  // columns carried: H1, or n1/2 with the Nyquist column packed into column 0;
  // the output holds only the carried columns (mode 1: tiled rows, fsY)
  a.pack = mode != 0 && g.pack;
  a.H1 = a.pack ? g.Hc : g.H1;
  a.dst_fsize = mode == 1 ? g.fsY : (int64_t)a.H1 * g.Q;
  a.I = 1;
  a.O = g.n[2];
  // the block index is split with shifts and masks: O must be a power of two
  if (a.O & (a.O - 1)) return false;
  a.TI = 1;
  a.mode = mode;
  a.goff3 = g.N == 3 ? g.koff[2] : 0;
  // the class-major split store (mode 3) feeds only k_slab3
  a.fold = (fold || mode == 3) && g.N == 3 && (mode == 2 || mode == 3) ? 1 : 0;
  if (mode == 3 && !a.fold) return false;
  const int nt = mode == 3 ? 512 : kThreads;  // threads per block
  if (mode == 3 && (sizeof(T) != 8 || a.L < 16 || a.L > 128)) return false;
  const int per_block = (nt * Ef) / a.L;
  int hp = 1;
  while (hp < a.H1) hp <<= 1;
  // mode 2: TJ pencils along j3 give TJ-value contiguous runs of the output;
  // 16 (256-byte runs) measured 0.7 % faster than 8 at 64^3 and 128^3, 32 slower
  a.TJ = mode == 2   ? (int)std::min<int64_t>(a.O, std::min(16, per_block))
         : mode == 3 ? (int)std::min<int64_t>(a.O, std::min(32, per_block))
                     : 1;
  // mode 3: 32 pencils (8 values of k' times the 4 quarters of axis 3)
  if (mode == 3 && (a.TJ != 32 || g.N != 3 || a.O < 32 || (a.O & 31))) return false;
  a.TK = std::min(hp, per_block / a.TJ);
  if (mode == 1 && a.TK != g.TKy) return false;  // tile width fixed by the layout
  a.ktiles = (a.H1 + a.TK - 1) / a.TK;
  a.itiles = 1;
  a.log2TI = 0;
  a.log2TK = ilog2c(a.TK);
  a.log2TJ = ilog2c(a.TJ);
  a.log2otiles = ilog2c((int)(a.O / a.TJ));
  a.log2itiles = 0;
  a.log2O = ilog2c((int)a.O);
  a.transposed = mode == 1;
  a.tw = (const cplx<T>*)tw;
  // With few tiles the last, partly filled round of blocks is a large share
  // of the pass: split the chunk's rungs so that about split_target blocks
  // run, each keeping at least 8 rungs (it reads its spectra once per split).
  // 256^2, 63-rung chunks, 8000: 3200 blocks -> 3 splits, ladder 12.13 -> 11.8 ms.
  const int64_t tiles = (int64_t)nf * (a.O / a.TJ) * a.ktiles;
  a.nsplit = split_target <= 0 ? 1
                               : (int)std::max<int64_t>(1, std::min<int64_t>(
                                                               (split_target + tiles - 1) / tiles,
                                                               R / 8));
  const int64_t blocks = tiles * a.nsplit;
  const size_t lds = (size_t)a.TK * a.TJ * axis_ladder_stride<T>(mode, a.L) * sizeof(cplx<T>);
  return axis_ladder_launch<T>(a.log2L, (unsigned)blocks, lds, s, a);
}

template <typename T, int DIR>
static void axis_wave_launch(int log2L, unsigned blocks, size_t lds, hipStream_t s,
                             const AxisArgs<T>& a);

// Second inverse pass of a 3-D rung chunk (grid axis 3) on the [k1][j2][j3]
// layout written by launch_axis_ladder (mode 2): pencils along j3 are
// contiguous; the output is the tiled row layout of the row kernel with rows
// ordered j3-fastest (q = j3 + n3 * j2).  False when n3 is outside the wave-local range.
template <typename T>
bool launch_axis3_rows(const GridDev& g, const cplx<T>* src, int64_t src_rung_stride,
                       cplx<T>* dst, int64_t dst_rung_stride, int nf, int R, int r0,
                       const T* gM, const T* gV, int fsplit, const T* tw, hipStream_t s) {
  AxisArgs<T> a{};
  const int Ef = sizeof(T) == 8 ? 8 : 16;
  a.L = g.n[2];
  a.log2L = g.log2n[2];
  if (g.N != 3 || a.L < Ef || a.L > 64 * Ef || (a.L & (a.L - 1))) return false;
  a.src = src;
  a.dst = dst;
  a.src_rs = src_rung_stride;
  a.dst_rs = dst_rung_stride;
  a.fsize = (int64_t)g.Hc * g.Q;  // fields of the first pass's output
  a.dst_fsize = g.fsY;
  a.fid0 = 0;
  a.nf = nf;
  a.src_fid0 = 0;
  a.dst_fid0 = 0;
  a.R = R;
  a.r0 = r0;
  a.gM = gM;
  a.gV = gV;
  a.fsplit = fsplit;
  a.ksum = g.ksum;
  a.goff = g.koff[2];
  a.Q = g.Q;
  a.H1 = g.Hc;  // the first pass packed the Nyquist column when g.pack
  // input: pencil (k1, j2) at k1 * Q + j2 * n3, contiguous along j3
  a.I = 1;
  a.O = g.n[1];
  a.TI = 1;
  a.transposed = 1;
  a.oO = g.n[2];     // row q' = j3 + n3 * j2: the block's rows are contiguous
  a.oE = 1;
  const int per_block = (kThreads * Ef) / a.L;
  int hp = 1;
  while (hp < g.Hc) hp <<= 1;
  a.TK = std::min(hp, per_block);
  if (a.TK != g.TKy) return false;  // tile width fixed by the layout
  a.ktiles = (g.Hc + a.TK - 1) / a.TK;
  a.itiles = 1;
  a.log2TI = 0;
  a.log2TK = ilog2c(a.TK);
  a.log2itiles = 0;
  a.log2O = ilog2c((int)a.O);
  a.tw = (const cplx<T>*)tw;
  const int64_t blocks = (int64_t)R * nf * a.O * a.ktiles;
  const size_t lds = (size_t)a.TK * pencil_stride(a.L) * sizeof(cplx<T>);
  axis_wave_launch<T, 1>(a.log2L, (unsigned)blocks, lds, s, a);
  return true;
}
template bool launch_axis3_rows<double>(const GridDev&, const cplx<double>*, int64_t,
                                        cplx<double>*, int64_t, int, int, int, const double*,
                                        const double*, int, const double*, hipStream_t);
template bool launch_axis3_rows<float>(const GridDev&, const cplx<float>*, int64_t, cplx<float>*,
                                       int64_t, int, int, int, const float*, const float*, int,
                                       const float*, hipStream_t);
template bool launch_axis_ladder<double>(const GridDev&, const cplx<double>*, cplx<double>*,
                                         int64_t, int, int, int, const double*, const double*,
                                         int, const double*, int, hipStream_t, int, bool);
template bool launch_axis_ladder<float>(const GridDev&, const cplx<float>*, cplx<float>*,
                                        int64_t, int, int, int, const float*, const float*, int,
                                        const float*, int, hipStream_t, int, bool);

template <typename T, int DIR>
static void axis_wave_launch(int log2L, unsigned blocks, size_t lds, hipStream_t s,
                             const AxisArgs<T>& a) {
  if constexpr (sizeof(T) == 8) {
    switch (log2L) {
      case 3: k_axis_wave<T, DIR, 3><<<blocks, kThreads, lds, s>>>(a); break;
      case 4: k_axis_wave<T, DIR, 4><<<blocks, kThreads, lds, s>>>(a); break;
      case 5: k_axis_wave<T, DIR, 5><<<blocks, kThreads, lds, s>>>(a); break;
      case 6: k_axis_wave<T, DIR, 6><<<blocks, kThreads, lds, s>>>(a); break;
      case 7: k_axis_wave<T, DIR, 7><<<blocks, kThreads, lds, s>>>(a); break;
      case 8: k_axis_wave<T, DIR, 8><<<blocks, kThreads, lds, s>>>(a); break;
      case 9: k_axis_wave<T, DIR, 9><<<blocks, kThreads, lds, s>>>(a); break;
      default: break;
    }
  } else {
    switch (log2L) {
      case 4: k_axis_wave<T, DIR, 4><<<blocks, kThreads, lds, s>>>(a); break;
      case 5: k_axis_wave<T, DIR, 5><<<blocks, kThreads, lds, s>>>(a); break;
      case 6: k_axis_wave<T, DIR, 6><<<blocks, kThreads, lds, s>>>(a); break;
      case 7: k_axis_wave<T, DIR, 7><<<blocks, kThreads, lds, s>>>(a); break;
      case 8: k_axis_wave<T, DIR, 8><<<blocks, kThreads, lds, s>>>(a); break;
      case 9: k_axis_wave<T, DIR, 9><<<blocks, kThreads, lds, s>>>(a); break;
      case 10: k_axis_wave<T, DIR, 10><<<blocks, kThreads, lds, s>>>(a); break;
      default: break;
    }
  }
}

template <typename TI, typename TO, int DIR>
void launch_axis_fft(const GridDev& g, int axis, const cplx<TI>* src, int64_t src_rung_stride,
                     int src_fid0, cplx<TO>* dst, int64_t dst_rung_stride, int dst_fid0, int fid0,
                     int nf, int R, int r0, const TI* gM, const TI* gV, int fsplit, const TI* tw,
                     const TI* twg, hipStream_t s) {
  static_assert(sizeof(TI) == sizeof(TO), "axis passes keep the element type");
  AxisArgs<TI> a;
  a.src = src;
  a.dst = (cplx<TI>*)dst;
  a.src_rs = src_rung_stride;
  a.dst_rs = dst_rung_stride;
  a.fsize = (int64_t)g.H1 * g.Q;
  a.fid0 = fid0;
  a.nf = nf;
  a.src_fid0 = src_fid0;
  a.dst_fid0 = dst_fid0;
  a.R = R;
  a.r0 = r0;
  a.gM = gM;
  a.gV = gV;
  a.fsplit = fsplit;
  a.ksum = g.ksum;
  a.goff = g.koff[axis];
  a.L = g.n[axis];
  a.log2L = g.log2n[axis];
  a.Q = g.Q;
  a.H1 = g.H1;
  a.dst_fsize = a.fsize;
  // q = j2 + n2 (j3 + n3 ...): inner index = the grid axes 2..axis, outer =
  // the axes after it
  a.I = 1;
  a.O = 1;
  for (int i = 1; i < axis; ++i) a.I *= g.n[i];
  for (int i = axis + 1; i < g.N; ++i) a.O *= g.n[i];
  // pencils per workgroup: one per L/E lanes (E = 8 in FP64, 16 in FP32)
  const int Ef = sizeof(TI) == 8 ? 8 : 16;
  const int per_block = (kThreads * Ef) / a.L > 0 ? (kThreads * Ef) / a.L : 1;
  a.transposed = 0;
  a.tw = (const cplx<TI>*)tw;
  const bool gen = (a.L & (a.L - 1)) != 0;
  a.fac = g.fac[axis];
  a.twg = twg ? (const cplx<TI>*)twg + g.koff[axis] : nullptr;
  // the wave-local form splits the block index with shifts and masks, so the
  // other axes' extents I and O must be powers of two as well
  const bool pow2_io = (a.I & (a.I - 1)) == 0 && (a.O & (a.O - 1)) == 0;
  if (!gen && pow2_io && a.L >= Ef && a.L <= 64 * Ef) {
    // wave-local form: TI and TK powers of two, TK * TI <= per_block
    a.TI = a.I == 1 ? 1 : (int)std::min<int64_t>(a.I, std::min(8, per_block));
    int hp = 1;
    while (hp < g.H1) hp <<= 1;
    a.TK = std::min(hp, per_block / a.TI);
    a.ktiles = (g.H1 + a.TK - 1) / a.TK;
    a.itiles = (int)(a.I / a.TI);
    a.log2TI = ilog2c(a.TI);
    a.log2TK = ilog2c(a.TK);
    a.log2itiles = ilog2c(a.itiles);
    a.log2O = ilog2c((int)a.O);
    const int64_t blocks = (int64_t)R * nf * a.O * a.ktiles * a.itiles;
    const size_t lds = (size_t)a.TK * a.TI * pencil_stride(a.L) * sizeof(cplx<TI>);
    axis_wave_launch<TI, DIR>(a.log2L, (unsigned)blocks, lds, s, a);
    return;
  }
  if (gen && twg && wave_gen_fits(g.facw[axis])) {
    // wave-local mixed-radix form: 8 pencils per block (8 columns, or 8
    // consecutive inner indices along the contiguous direction)
    a.fac = g.facw[axis];
    a.TI = a.I == 1 ? 1 : (int)std::min<int64_t>(a.I, 8);
    a.TK = std::max(1, std::min(g.H1, 8 / a.TI));
    a.ktiles = (g.H1 + a.TK - 1) / a.TK;
    a.itiles = (int)((a.I + a.TI - 1) / a.TI);
    const int64_t blocks = (int64_t)R * nf * a.O * a.ktiles * a.itiles;
    const size_t lds = ((size_t)a.TK * a.TI * gen_pencil_stride(a.L) + a.L) * sizeof(cplx<TI>);
    k_axis_fft_wg<TI, DIR><<<(unsigned)blocks, kThreads, lds, s>>>(a);
    return;
  }
  if (a.I == 1) {
    a.TI = 1;
  } else {
    a.TI = (int)std::min<int64_t>(a.I, std::max(1, std::min(8, per_block)));
  }
  a.TK = std::max(1, std::min(g.H1, per_block / a.TI));
  a.ktiles = (g.H1 + a.TK - 1) / a.TK;
  a.itiles = (int)((a.I + a.TI - 1) / a.TI);
  const int64_t blocks = (int64_t)R * nf * a.O * a.ktiles * a.itiles;
  const size_t lds = (size_t)a.TK * a.TI * a.L * sizeof(cplx<TI>);
  if (gen)
    k_axis_fft<TI, DIR, true><<<(unsigned)blocks, kThreads, lds, s>>>(a);
  else
    k_axis_fft<TI, DIR, false><<<(unsigned)blocks, kThreads, lds, s>>>(a);
}

#define PKDE_AXIS_INST(TI, TO, DIR)                                                            \
  template void launch_axis_fft<TI, TO, DIR>(const GridDev&, int, const cplx<TI>*, int64_t, int,  \
                                             cplx<TO>*, int64_t, int, int, int, int, int,        \
                                             const TI*, const TI*, int, const TI*, const TI*,    \
                                             hipStream_t);
PKDE_AXIS_INST(double, double, -1)
PKDE_AXIS_INST(double, double, 1)
PKDE_AXIS_INST(float, float, 1)
#undef PKDE_AXIS_INST

template <typename TI, typename TO>
__global__ void k_convert(const cplx<TI>* __restrict__ src, cplx<TO>* __restrict__ dst, int64_t count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const cplx<TI> v = src[i];
    dst[i] = {(TO)v.re, (TO)v.im};
  }
}

template <typename TI, typename TO>
void launch_convert(const cplx<TI>* src, cplx<TO>* dst, int64_t count, hipStream_t s) {
  k_convert<TI, TO><<<grid_for(count, kThreads, 1 << 16), kThreads, 0, s>>>(src, dst, count);
}
template void launch_convert<double, float>(const cplx<double>*, cplx<float>*, int64_t, hipStream_t);

// ----------------------------------------------- Hermitian pair construction
// Z(k) = A(k) + i B(k) over the full axis-1 frequency range from the halves
// k in [0, n1/2]; the transform of Z then holds a in its real part and b in
// its imaginary part for real fields a, b.
template <typename T>
__device__ __forceinline__ void put_pair(cplx<T>* row, int n1, int k1, cplx<T> A, cplx<T> Bv) {
  row[k1] = {A.re - Bv.im, A.im + Bv.re};
  if (k1 > 0 && 2 * k1 < n1) row[n1 - k1] = {A.re + Bv.im, Bv.re - A.im};
}

// r = |s|/|m| - |m|/n, the per-bootstrap variance-to-mean ratio
// (DirectSpace.jl:550-552 rearranged: (|s|/n - (|m|/n)^2) / (|m|/n)).  The
// quotient uses the hardware reciprocal refined by two Newton steps (within
// 1 ulp) instead of the IEEE division sequence; |m| = 0 gives NaN as the
// division does once it reaches the Welford update.
__device__ __forceinline__ double vmr_ratio(double av, double am, double invn) {
  double inv = __builtin_amdgcn_rcp(am);
  inv = fma(fma(-am, inv, 1.0), inv, inv);
  inv = fma(fma(-am, inv, 1.0), inv, inv);
  return fma(av, inv, -am * invn);
}

// --------------------------------------------------------------- ladder VMR
// One workgroup owns TQ consecutive axis-1 rows (TQ*n1 grid points) of one
// rung and walks all bootstraps in order, two at a time per complex
// transform, NBP bootstrap pairs per LDS batch.  Per point it forms
//   mu = |m|/n, v = |s|/n - mu^2, r = v/mu  (= |s|/|m| - |m|/n)
// with m, s the propagated sums / squares fields (FourierSpace.jl:215-241,
// normalised inverse FFT) and accumulates an FP64 Welford mean/M2 over the
// bootstrap index (DirectSpace.jl:567-581).
template <typename T>
struct VmrArgs {
  const cplx<T>* src;
  int64_t src_rs, fsize;
  int B;
  double invn, invG;
  int R, r0;
  const T* gM;
  const T* gV;
  int ksum;
  const double* cV;
  const double* vscale;
  const cplx<T>* tw;
  double* vmr;
  int64_t G, Q;
  int n1, log2n1, H1, TQ, NBP, ntiles;
  // tiled row layout (rows kernel): tile width 2^log2TKy, tile stride tstride;
  // rows3: rows are q' = j3 + n3 * j2 (3-D), grid row q = j2 + n2 * j3
  int log2TKy, rows3, log2n2, log2n3;
  int n2r, n3r;  // 3-D row remap for lengths that are not powers of two
  int goff3;     // offset of grid axis 3 in the per-rung Gaussian tables
  int64_t tstride;
  const double* invk;  // 1 / (b + 1), b < B (Welford weights)
  FftFactors fac;      // axis-1 plan and twiddle table (GEN kernels)
  const cplx<T>* twg;
};

template <typename T, int PPT, bool GEN>
__global__ __launch_bounds__(kThreads) void k_ladder_vmr(VmrArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* Z = (cplx<T>*)smem;
  const int rr = blockIdx.x / a.ntiles;
  const int tile = blockIdx.x - rr * a.ntiles;
  const int r = a.r0 + rr;
  const int64_t q0 = (int64_t)tile * a.TQ;
  const int n1 = a.n1, H1 = a.H1, TQ = a.TQ, NBP = a.NBP, B = a.B;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const T* gm1 = a.gM + (int64_t)r * a.ksum;
  const T* gv1 = a.gV + (int64_t)r * a.ksum;
  const cplx<T>* src = a.src + (int64_t)rr * a.src_rs;
  const double vfac = a.cV[r] * a.invG;
  const int npts = TQ * n1;
  const int npen = 2 * NBP * TQ;

  double mean[PPT], m2[PPT];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    mean[u] = 0.0;
    m2[u] = 0.0;
  }

  for (int bp = 0; bp < B; bp += 2 * NBP) {
    const int nload = NBP * 2 * H1 * TQ;
    for (int id = tid; id < nload; id += nthr) {
      const int t = id % TQ;
      int rest = id / TQ;
      const int k1 = rest % H1;
      rest /= H1;
      const int ty = rest & 1;
      const int slot = rest >> 1;
      const int b = bp + 2 * slot;
      cplx<T>* row = Z + (int64_t)((slot * 2 + ty) * TQ + t) * n1;
      if (b >= B || q0 + t >= a.Q) continue;
      const int fa = ty * B + b;
      const T gk = ty ? gv1[k1] : gm1[k1];
      const int64_t o = (int64_t)k1 * a.Q + q0 + t;
      cplx<T> A = cscale(src[(int64_t)fa * a.fsize + o], gk);
      cplx<T> Bv = {(T)0, (T)0};
      if (b + 1 < B) Bv = cscale(src[(int64_t)(fa + 1) * a.fsize + o], gk);
      put_pair(row, n1, k1, A, Bv);
    }
    __syncthreads();
    if constexpr (GEN)
      lds_fft_gen<1, T>(Z, a.fac, npen, n1, a.twg, tid, nthr);
    else
      lds_fft<1, T>(Z, a.log2n1, npen, n1, a.tw + (n1 - 2), tid, nthr);
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
      const int w = tid + u * nthr;
      if (w < npts) {
        const int t = w / n1, j1 = w - t * n1;
        for (int slot = 0; slot < NBP; ++slot) {
          const int b = bp + 2 * slot;
          if (b >= B) break;
          const cplx<T> zm = Z[(int64_t)((slot * 2) * TQ + t) * n1 + j1];
          const cplx<T> zv = Z[(int64_t)((slot * 2 + 1) * TQ + t) * n1 + j1];
          {
            const double am = fabs((double)zm.re) * a.invG;
            const double av = fabs((double)zv.re) * vfac;
            const double xv = vmr_ratio(av, am, a.invn);
            const double d = xv - mean[u];
            mean[u] += d / (double)(b + 1);
            m2[u] += d * (xv - mean[u]);
          }
          if (b + 1 < B) {
            const double am = fabs((double)zm.im) * a.invG;
            const double av = fabs((double)zv.im) * vfac;
            const double xv = vmr_ratio(av, am, a.invn);
            const double d = xv - mean[u];
            mean[u] += d / (double)(b + 2);
            m2[u] += d * (xv - mean[u]);
          }
        }
      }
    }
    __syncthreads();
  }
  const double sc = a.vscale[r];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int w = tid + u * nthr;
    if (w < npts) {
      const int t = w / n1, j1 = w - t * n1;
      if (q0 + t >= a.Q) continue;
      const double vv = sc * m2[u] / (double)(B - 1);
      a.vmr[(int64_t)r * a.G + (q0 + t) * n1 + j1] = isfinite(vv) ? log(vv) : (double)NAN;
    }
  }
}

// 1-D ladder for power-of-two grids in the wave-local FFT range: one block
// of kWaves1d wavefronts per rung.  Each round, wavefront w takes bootstrap
// pair (b, b + 1), b = 2 (kWaves1d round + w): it loads the sums and squares
// fields of both, builds and transforms the two complex pencils and leaves
// the two VMR ratios of every point in its LDS slice; after a barrier each
// thread applies the Welford updates of its points over the round's
// bootstraps in order.  Same arithmetic as k_ladder_vmr.
constexpr int kWaves1d = 8;

template <typename T, int LOG2N1>
__global__ __launch_bounds__(64 * kWaves1d) void k_ladder_1d(VmrArgs<T> a) {
  constexpr int n1 = 1 << LOG2N1;
  constexpr int H = n1 / 2;
  constexpr int E = sizeof(T) == 8 ? 8 : 16;
  constexpr int PL = pencil_stride(n1);
  constexpr int lp = n1 / E;      // lanes per pencil
  constexpr int ppc = 64 / lp;    // pencils per wave-wide FFT call
  constexpr int NT = 64 * kWaves1d;
  constexpr int PPT = (n1 + NT - 1) / NT;  // points per thread in the Welford phase
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  cplx<T>* twl = (cplx<T>*)smem + kWaves1d * 2 * PL;
  for (int k = tid; k < n1; k += NT) twl[k] = a.tw[n1 - 2 + k];
  __syncthreads();
  cplx<T>* Z = (cplx<T>*)smem + wv * 2 * PL;  // sums pencil; squares pencil at Z + PL
  const int r = a.r0 + (int)blockIdx.x;
  const T* gm1 = a.gM + (int64_t)r * a.ksum;
  const T* gv1 = a.gV + (int64_t)r * a.ksum;
  const cplx<T>* src = a.src + (int64_t)blockIdx.x * a.src_rs;
  const double vfac = a.cV[r] * a.invG;
  const int B = a.B;
  double mean[PPT], m2[PPT];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    mean[u] = 0.0;
    m2[u] = 0.0;
  }
  const cplx<T> zero = {(T)0, (T)0};
  // The wave's loads of a round (sums and squares fields of its pair, the
  // half spectrum) go to registers all at once, one round ahead: they are in
  // flight during the previous round's Welford updates.  A wave without a
  // pair in the round reloads bootstrap 0 (valid addresses, values unused).
  constexpr int NIT = (2 * (H + 1) + 63) / 64;
  cplx<T> la[NIT], lb[NIT];
  auto issue = [&](int b) {
    const int bb = b < B ? b : 0;
    const int o2 = bb + 1 < B ? 1 : 0;
#pragma unroll
    for (int i = 0; i < NIT; ++i) {
      const int id = lane + 64 * i;
      const int idc = id < 2 * (H + 1) ? id : 0;
      const int ty = idc > H;
      const int k1 = ty ? idc - H - 1 : idc;
      const int fa = ty * B + bb;
      la[i] = src[(int64_t)fa * a.fsize + k1];
      lb[i] = src[(int64_t)(fa + o2) * a.fsize + k1];
    }
  };
  issue(2 * wv);
  for (int b0 = 0; b0 < B; b0 += 2 * kWaves1d) {
    const int b = b0 + 2 * wv;
    if (b < B) {
      const bool has2 = b + 1 < B;
#pragma unroll
      for (int i = 0; i < NIT; ++i) {
        const int id = lane + 64 * i;
        if (id < 2 * (H + 1)) {
          const int ty = id > H;
          const int k1 = ty ? id - H - 1 : id;
          const T gk = ty ? gv1[k1] : gm1[k1];
          const cplx<T> A = cscale(la[i], gk);
          const cplx<T> Bv = has2 ? cscale(lb[i], gk) : zero;
          cplx<T>* row = Z + ty * PL;
          row[pad8(k1)] = {A.re - Bv.im, A.im + Bv.re};
          if (k1 > 0 && k1 < H) row[pad8(n1 - k1)] = {A.re + Bv.im, Bv.re - A.im};
        }
      }
      wave_sync();
#pragma unroll
      for (int c = 0; c < 2; c += ppc) {
        const int p = c + lane / lp;
        const bool act = p < 2;
        wave_fft<LOG2N1, 1, T, E>(Z + (act ? p : 0) * PL, lane % lp, act, twl);
      }
      // the ratios of bootstraps b and b + 1 replace the pencils (as doubles
      // in the first 8 bytes of each slot; every lane rewrites what it read)
      for (int j = lane; j < n1; j += 64) {
        const cplx<T> zm = Z[pad8(j)];
        const cplx<T> zv = Z[PL + pad8(j)];
        const double x0 = vmr_ratio(fabs((double)zv.re) * vfac, fabs((double)zm.re) * a.invG, a.invn);
        const double x1 =
            has2 ? vmr_ratio(fabs((double)zv.im) * vfac, fabs((double)zm.im) * a.invG, a.invn) : 0.0;
        *(double*)(Z + pad8(j)) = x0;
        *(double*)(Z + PL + pad8(j)) = x1;
      }
    }
    issue(b + 2 * kWaves1d);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
      const int j = tid + u * NT;
      if (j < n1) {
        for (int w = 0; w < kWaves1d; ++w) {
          const int bw = b0 + 2 * w;
          if (bw >= B) break;
          const cplx<T>* Zw = (const cplx<T>*)smem + w * 2 * PL;
          {
            const double xv = *(const double*)(Zw + pad8(j));
            const double d = xv - mean[u];
            mean[u] += d / (double)(bw + 1);
            m2[u] += d * (xv - mean[u]);
          }
          if (bw + 1 < B) {
            const double xv = *(const double*)(Zw + PL + pad8(j));
            const double d = xv - mean[u];
            mean[u] += d / (double)(bw + 2);
            m2[u] += d * (xv - mean[u]);
          }
        }
      }
    }
    __syncthreads();
  }
  const double sc = a.vscale[r];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int j = tid + u * NT;
    if (j < n1) {
      const double vv = sc * m2[u] / (double)(B - 1);
      a.vmr[(int64_t)r * a.G + j] = isfinite(vv) ? log(vv) : (double)NAN;
    }
  }
}

// 1-D ladder for 512- and 1024-point grids in FP64, one block of NW
// wavefronts per rung.  Each wavefront takes ONE bootstrap b per round and
// transforms z = m + i s, the sums field m and the squares field s of that
// bootstrap (both real), as one complex pencil: the real part of the result
// is the sums field and the imaginary part the squares field, the two terms
// of b's VMR ratio.  At 1024 points a pencil pair per wave does not fit LDS
// (2 pencils of 18 KB per wave): 7 wavefronts with 16 values per lane
// (radix-16 passes) leave room for the twiddle table.  At 512 points the
// single pencils let 16 wavefronts share the CU where k_ladder_1d's pairs
// fit 8 (radix-8 passes).  After a barrier every thread runs the Welford
// updates of its points over the round's bootstraps in order.  Same
// arithmetic as k_ladder_1d otherwise.
template <int LOG2N1>
constexpr int ladder_1d_ms_waves() { return LOG2N1 == 10 ? 8 : 16; }

template <int LOG2N1, int NW>
__global__ __launch_bounds__(64 * NW) void k_ladder_1d_ms(VmrArgs<double> a) {
  constexpr int n1 = 1 << LOG2N1;
  constexpr int H = n1 / 2;
  constexpr int E = n1 / 64;  // one pencil per wavefront
  constexpr int PL = n1;  // unpadded pencils, XOR-swizzled slots (lds_slot<1>)
  constexpr int LE = ilog2c(E);
  constexpr int NT = 64 * NW;
  constexpr int PPT = (n1 + NT - 1) / NT;  // points per thread in the Welford phase
  static_assert(E == 8 || E == 16, "512- or 1024-point pencils, 8 or 16 values per lane");
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  cplx<double>* twl = (cplx<double>*)smem + NW * PL;
  for (int k = tid; k < n1; k += NT) twl[k] = a.tw[n1 - 2 + k];
  __syncthreads();
  cplx<double>* Z = (cplx<double>*)smem + wv * PL;
  const int r = a.r0 + (int)blockIdx.x;
  const double* gm1 = a.gM + (int64_t)r * a.ksum;
  const double* gv1 = a.gV + (int64_t)r * a.ksum;
  const cplx<double>* src = a.src + (int64_t)blockIdx.x * a.src_rs;
  const double vfac = a.cV[r] * a.invG;
  const int B = a.B;
  double mean[PPT], m2[PPT];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    mean[u] = 0.0;
    m2[u] = 0.0;
  }
  // The wave's loads of a round (the half spectra of its bootstrap's sums
  // and squares fields) go to registers all at once, one round ahead.  A
  // wave without a bootstrap in the round reloads bootstrap 0 (valid
  // addresses, values unused).
  constexpr int NIT = (2 * (H + 1) + 63) / 64;
  cplx<double> la[NIT];
  auto issue = [&](int b) {
    const int bb = b < B ? b : 0;
#pragma unroll
    for (int i = 0; i < NIT; ++i) {
      const int id = lane + 64 * i;
      const int idc = id < 2 * (H + 1) ? id : 0;
      const int ty = idc > H;
      const int k1 = ty ? idc - H - 1 : idc;
      la[i] = src[(int64_t)(ty * B + bb) * a.fsize + k1];
    }
  };
  issue(wv);
  for (int b0 = 0; b0 < B; b0 += NW) {
    const int b = b0 + wv;
    if (b < B) {
      // Z[k] = M[k] + i S[k], Z[n1 - k] = conj(M[k]) + i conj(S[k]), with M
      // the sums half spectrum and S the squares one, each with its factor:
      // the sums first, then the squares added in
#pragma unroll
      for (int i = 0; i < NIT; ++i) {
        const int id = lane + 64 * i;
        if (id <= H) {
          const cplx<double> M = cscale(la[i], gm1[id]);
          Z[lds_slot<1, LE>(id)] = M;
          if (id > 0 && id < H) Z[lds_slot<1, LE>(n1 - id)] = cconj(M);
        }
      }
      wave_sync();
#pragma unroll
      for (int i = 0; i < NIT; ++i) {
        const int id = lane + 64 * i;
        if (id > H && id < 2 * (H + 1)) {
          const int k1 = id - H - 1;
          const cplx<double> S = cscale(la[i], gv1[k1]);
          const cplx<double> m = Z[lds_slot<1, LE>(k1)];
          Z[lds_slot<1, LE>(k1)] = {m.re - S.im, m.im + S.re};
          if (k1 > 0 && k1 < H) {
            const cplx<double> mc = Z[lds_slot<1, LE>(n1 - k1)];
            Z[lds_slot<1, LE>(n1 - k1)] = {mc.re + S.im, mc.im + S.re};
          }
        }
      }
      wave_sync();
      wave_fft<LOG2N1, 1, double, E, 1>(Z, lane, true, twl);
      // the ratio of bootstrap b replaces the pencil (as doubles in the
      // first 8 bytes of each slot; every lane rewrites what it read)
      for (int j = lane; j < n1; j += 64) {
        const cplx<double> z = Z[lds_slot<1, LE>(j)];
        *(double*)(Z + lds_slot<1, LE>(j)) = vmr_ratio(fabs(z.im) * vfac, fabs(z.re) * a.invG, a.invn);
      }
    }
    issue(b + NW);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
      const int j = tid + u * NT;
      if (j < n1) {
        for (int w = 0; w < NW; ++w) {
          const int bw = b0 + w;
          if (bw >= B) break;
          const cplx<double>* Zw = (const cplx<double>*)smem + w * PL;
          const double xv = *(const double*)(Zw + lds_slot<1, LE>(j));
          const double d = xv - mean[u];
          mean[u] += d / (double)(bw + 1);
          m2[u] += d * (xv - mean[u]);
        }
      }
    }
    __syncthreads();
  }
  const double sc = a.vscale[r];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int j = tid + u * NT;
    if (j < n1) {
      const double vv = sc * m2[u] / (double)(B - 1);
      a.vmr[(int64_t)r * a.G + j] = isfinite(vv) ? log(vv) : (double)NAN;
    }
  }
}

// The same for lengths that are not powers of two but fit the wave-local
// mixed-radix FFT (a.fac: radices 2, 3, 4, 5, 8; n1 <= 512): PPT = points per
// thread of the Welford phase.
template <typename T, int LP>
__global__ __launch_bounds__(64 * kWaves1d) void k_ladder_1d_gen(VmrArgs<T> a) {
  constexpr int NT = 64 * kWaves1d;
  extern __shared__ __align__(16) unsigned char smem[];
  const int n1 = a.n1, H1 = a.H1;
  const int PL = gen_pencil_stride(n1);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  cplx<T>* twl = (cplx<T>*)smem + kWaves1d * 2 * PL;
  for (int k = tid; k < n1; k += NT) twl[k] = a.twg[k];
  __syncthreads();
  cplx<T>* Z = (cplx<T>*)smem + wv * 2 * PL;
  const int r = a.r0 + (int)blockIdx.x;
  const T* gm1 = a.gM + (int64_t)r * a.ksum;
  const T* gv1 = a.gV + (int64_t)r * a.ksum;
  const cplx<T>* src = a.src + (int64_t)blockIdx.x * a.src_rs;
  const double vfac = a.cV[r] * a.invG;
  const int B = a.B;
  const int j = tid;  // n1 <= 512 = NT: one point per thread
  double mean = 0.0, m2 = 0.0;
  const cplx<T> zero = {(T)0, (T)0};
  // loads of a round in registers, one round ahead (as k_ladder_1d); the
  // half spectrum has 2 H1 <= 2 * 257 items over the wave's 64 lanes
  constexpr int NIT = (2 * 257 + 63) / 64;
  cplx<T> la[NIT], lb[NIT];
  auto issue = [&](int b) {
    const int bb = b < B ? b : 0;
    const int o2 = bb + 1 < B ? 1 : 0;
#pragma unroll
    for (int i = 0; i < NIT; ++i) {
      const int id = lane + 64 * i;
      const int idc = id < 2 * H1 ? id : 0;
      const int ty = idc >= H1;
      const int k1 = ty ? idc - H1 : idc;
      const int fa = ty * B + bb;
      la[i] = src[(int64_t)fa * a.fsize + k1];
      lb[i] = src[(int64_t)(fa + o2) * a.fsize + k1];
    }
  };
  issue(2 * wv);
  for (int b0 = 0; b0 < B; b0 += 2 * kWaves1d) {
    const int b = b0 + 2 * wv;
    if (b < B) {
      const bool has2 = b + 1 < B;
#pragma unroll
      for (int i = 0; i < NIT; ++i) {
        const int id = lane + 64 * i;
        if (id < 2 * H1) {
          const int ty = id >= H1;
          const int k1 = ty ? id - H1 : id;
          const T gk = ty ? gv1[k1] : gm1[k1];
          const cplx<T> A = cscale(la[i], gk);
          const cplx<T> Bv = has2 ? cscale(lb[i], gk) : zero;
          cplx<T>* row = Z + ty * PL;
          row[pad8(k1)] = {A.re - Bv.im, A.im + Bv.re};
          if (k1 > 0 && 2 * k1 < n1) row[pad8(n1 - k1)] = {A.re + Bv.im, Bv.re - A.im};
        }
      }
      wave_sync();
      wave_fft_gen_pair<LP, T>(Z, true, Z + PL, true, lane, a.fac, twl);
      for (int jj = lane; jj < n1; jj += 64) {
        const cplx<T> zm = Z[pad8(jj)];
        const cplx<T> zv = Z[PL + pad8(jj)];
        const double x0 = vmr_ratio(fabs((double)zv.re) * vfac, fabs((double)zm.re) * a.invG, a.invn);
        const double x1 =
            has2 ? vmr_ratio(fabs((double)zv.im) * vfac, fabs((double)zm.im) * a.invG, a.invn) : 0.0;
        *(double*)(Z + pad8(jj)) = x0;
        *(double*)(Z + PL + pad8(jj)) = x1;
      }
    }
    issue(b + 2 * kWaves1d);
    __syncthreads();
    if (j < n1) {
      for (int w = 0; w < kWaves1d; ++w) {
        const int bw = b0 + 2 * w;
        if (bw >= B) break;
        const cplx<T>* Zw = (const cplx<T>*)smem + w * 2 * PL;
        {
          const double xv = *(const double*)(Zw + pad8(j));
          const double d = xv - mean;
          mean += d / (double)(bw + 1);
          m2 += d * (xv - mean);
        }
        if (bw + 1 < B) {
          const double xv = *(const double*)(Zw + PL + pad8(j));
          const double d = xv - mean;
          mean += d / (double)(bw + 2);
          m2 += d * (xv - mean);
        }
      }
    }
    __syncthreads();
  }
  if (j < n1) {
    const double vv = a.vscale[r] * m2 / (double)(B - 1);
    a.vmr[(int64_t)r * a.G + j] = isfinite(vv) ? log(vv) : (double)NAN;
  }
}

template <typename T>
static bool ladder_1d_gen_launch(const GridDev& g, unsigned blocks, hipStream_t s,
                                 const VmrArgs<T>& a) {
  const int n1 = g.n[0];
  if (n1 > 64 * kWaves1d || !wave_gen_fits(g.facw[0]) || !a.twg) return false;
  const size_t lds = ((size_t)kWaves1d * 2 * gen_pencil_stride(n1) + n1) * sizeof(cplx<T>);
  static const bool attr = [] {
    // beyond 64 KB of LDS per block (gfx950 has 160 KB per CU); set once
    const int most = (int)(((size_t)kWaves1d * 2 * gen_pencil_stride(512) + 512) * sizeof(cplx<T>));
    (void)hipFuncSetAttribute((const void*)k_ladder_1d_gen<T, 32>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, most);
    (void)hipFuncSetAttribute((const void*)k_ladder_1d_gen<T, 64>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, most);
    return true;
  }();
  (void)attr;
  VmrArgs<T> b = a;
  b.fac = g.facw[0];
  // the sums and squares pencils together in the two halves of the wave
  // when the length fits 32 lanes
  if (wave_gen_fits_lp(b.fac, 32))
    k_ladder_1d_gen<T, 32><<<blocks, 64 * kWaves1d, lds, s>>>(b);
  else
    k_ladder_1d_gen<T, 64><<<blocks, 64 * kWaves1d, lds, s>>>(b);
  return true;
}

template <typename T, int LOG2N1>
static void ladder_1d_launch(unsigned blocks, hipStream_t s, const VmrArgs<T>& a) {
  constexpr int n1 = 1 << LOG2N1;
  const size_t lds = ((size_t)kWaves1d * 2 * pencil_stride(n1) + n1) * sizeof(cplx<T>);
  static const bool attr = [lds] {
    // beyond 64 KB of LDS per block (gfx950 has 160 KB per CU); set once
    (void)hipFuncSetAttribute((const void*)k_ladder_1d<T, LOG2N1>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return true;
  }();
  (void)attr;
  k_ladder_1d<T, LOG2N1><<<blocks, 64 * kWaves1d, lds, s>>>(a);
}

// True if the 1-D ladder took the wave-parallel form (n1 a power of two in
// the wave-local FFT range, at least 16).
template <int LOG2N1>
static void ladder_1d_ms_launch(unsigned blocks, hipStream_t s, const VmrArgs<double>& a) {
  constexpr int n1 = 1 << LOG2N1, NW = ladder_1d_ms_waves<LOG2N1>();
  constexpr size_t lds = ((size_t)NW * n1 + n1) * sizeof(cplx<double>);  // unpadded pencils
  static_assert(lds <= 160 * 1024, "the wavefronts' pencils fit the LDS of a CU");
  static const bool attr = [] {
    // beyond 64 KB of LDS per block (gfx950 has 160 KB per CU); set once
    (void)hipFuncSetAttribute((const void*)k_ladder_1d_ms<LOG2N1, NW>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return true;
  }();
  (void)attr;
  k_ladder_1d_ms<LOG2N1, NW><<<blocks, 64 * NW, lds, s>>>(a);
}

template <typename T>
static bool ladder_1d_dispatch(int log2n1, unsigned blocks, hipStream_t s, const VmrArgs<T>& a) {
  switch (log2n1) {
    case 4: ladder_1d_launch<T, 4>(blocks, s, a); return true;
    case 5: ladder_1d_launch<T, 5>(blocks, s, a); return true;
    case 6: ladder_1d_launch<T, 6>(blocks, s, a); return true;
    case 7: ladder_1d_launch<T, 7>(blocks, s, a); return true;
    case 8: ladder_1d_launch<T, 8>(blocks, s, a); return true;
    case 9:
      if constexpr (sizeof(T) == 4) ladder_1d_launch<T, 9>(blocks, s, a);
      else ladder_1d_ms_launch<9>(blocks, s, a);
      return true;
    case 10:
      if constexpr (sizeof(T) == 4) ladder_1d_launch<T, 10>(blocks, s, a);
      else ladder_1d_ms_launch<10>(blocks, s, a);
      return true;
    default: return false;
  }
}

template <typename T, bool GEN>
static void ladder_vmr_dispatch(int ppt, unsigned blocks, size_t lds, hipStream_t s,
                                const VmrArgs<T>& a) {
  if (ppt <= 1)
    k_ladder_vmr<T, 1, GEN><<<blocks, kThreads, lds, s>>>(a);
  else if (ppt <= 2)
    k_ladder_vmr<T, 2, GEN><<<blocks, kThreads, lds, s>>>(a);
  else if (ppt <= 4)
    k_ladder_vmr<T, 4, GEN><<<blocks, kThreads, lds, s>>>(a);
  else if (ppt <= 8)
    k_ladder_vmr<T, 8, GEN><<<blocks, kThreads, lds, s>>>(a);
  else
    k_ladder_vmr<T, 16, GEN><<<blocks, kThreads, lds, s>>>(a);
}

template <typename T>
void launch_ladder_vmr(const GridDev& g, const cplx<T>* src, int64_t src_rung_stride, int B,
                       int64_t n_samples, int R, int r0, const T* gM, const T* gV,
                       const double* cV, const double* vscale, const T* tw, const T* twg,
                       double* vmr, hipStream_t s) {
  VmrArgs<T> a;
  const int n1 = g.n[0];
  // 2048 points per workgroup where the grid allows it; PPT = points/thread
  int TQ = 2048 / n1;
  if (TQ < 1) TQ = 1;
  if (TQ > g.Q) TQ = (int)g.Q;
  const int npts = TQ * n1;
  int ppt = (npts + kThreads - 1) / kThreads;
  int NBP = 2048 / npts;  // bootstrap pairs per LDS batch
  if (NBP < 1) NBP = 1;
  if (NBP > (B + 1) / 2) NBP = (B + 1) / 2;
  a.src = src;
  a.src_rs = src_rung_stride;
  a.fsize = (int64_t)g.H1 * g.Q;
  a.B = B;
  a.invn = 1.0 / (double)n_samples;
  a.invG = 1.0 / (double)g.G;
  a.R = R;
  a.r0 = r0;
  a.gM = gM;
  a.gV = gV;
  a.ksum = g.ksum;
  a.cV = cV;
  a.vscale = vscale;
  a.tw = (const cplx<T>*)tw;
  a.vmr = vmr;
  a.G = g.G;
  a.Q = g.Q;
  a.n1 = n1;
  a.log2n1 = g.log2n[0];
  a.H1 = g.H1;
  a.TQ = TQ;
  a.NBP = NBP;
  a.ntiles = (int)((g.Q + TQ - 1) / TQ);
  a.fac = g.fac[0];
  a.twg = twg ? (const cplx<T>*)twg + g.koff[0] : nullptr;
  const bool gen = (n1 & (n1 - 1)) != 0;
  if (g.N == 1 && !gen && ladder_1d_dispatch<T>(g.log2n[0], (unsigned)R, s, a)) return;
  if (g.N == 1 && gen && ladder_1d_gen_launch<T>(g, (unsigned)R, s, a)) return;
  const unsigned blocks = (unsigned)((int64_t)R * a.ntiles);
  const size_t lds = (size_t)2 * NBP * TQ * n1 * sizeof(cplx<T>);
  if (gen)
    ladder_vmr_dispatch<T, true>(ppt, blocks, lds, s, a);
  else
    ladder_vmr_dispatch<T, false>(ppt, blocks, lds, s, a);
}

template void launch_ladder_vmr<float>(const GridDev&, const cplx<float>*, int64_t, int, int64_t, int,
                                       int, const float*, const float*, const double*,
                                       const double*, const float*, const float*, double*,
                                       hipStream_t);
template void launch_ladder_vmr<double>(const GridDev&, const cplx<double>*, int64_t, int, int64_t,
                                        int, int, const double*, const double*, const double*,
                                        const double*, const double*, const double*, double*,
                                        hipStream_t);

// ------------------------------------------------- ladder VMR, row per wave
// Same computation as k_ladder_vmr for grids with Q >= 64 rows and
// 16 <= n1 <= 512, reading the axis-2/3-transformed fields in the tiled row
// layout [rung][field][k1 tile][q][k1 within tile], so each row segment is
// contiguous.  Each wavefront owns W = 64 PPT / n1 rows of one rung, builds
// the M and V pencils of two bootstraps per row in its own LDS slice,
// transforms them with wave-local passes (twiddles read from LDS) and
// accumulates the FP64 Welford statistics of its PPT points per lane; after
// the block stages the twiddles, its wavefronts never synchronise with each
// other.  Global loads run one bootstrap pair ahead of the arithmetic.
template <typename T, int E, int PPT, int LOG2N1>
__global__ __launch_bounds__(kThreads) void k_ladder_rows(VmrArgs<T> a) {
  constexpr int n1 = 1 << LOG2N1;
  constexpr int H = n1 / 2;                // Hermitian half without the Nyquist term
  constexpr int W = 64 * PPT / n1;         // rows per wave (PPT points per lane)
  constexpr int LOG2W = ilog2c(W);
  constexpr int PL = pencil_stride(n1);
  constexpr int lp = n1 / E;               // lanes per pencil
  constexpr int ppc = 64 / lp;             // pencils per wave-wide FFT call
  // FP64 pencils use the XOR-swizzled slots (lds_slot<1>): the transform's
  // 16-lane reads of values 32 or 64 apart and the Welford reads of 16
  // consecutive values fill distinct banks, where the padded layout has a
  // 2-way conflict in each of them
  constexpr int SWR = sizeof(T) == 8 ? 1 : 0;
  constexpr int LE = ilog2c(E);
  static_assert(W >= 1 && (W & (W - 1)) == 0, "rows per wave must be a power of two");
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  // twiddles exp(-2 pi i k / n1) staged once per block after the waves' rows
  cplx<T>* twl = (cplx<T>*)smem + (kThreads / 64) * (2 * W * PL);
  for (int k = threadIdx.x; k < n1; k += kThreads) twl[k] = a.tw[n1 - 2 + k];
  __syncthreads();
  cplx<T>* Z = (cplx<T>*)smem + wv * (2 * W * PL);
  const int64_t rows_per_rung = a.Q >> LOG2W;
  const int64_t gw = (int64_t)blockIdx.x * (blockDim.x >> 6) + wv;
  if (gw >= (int64_t)a.R * rows_per_rung) return;
  const int rr = (int)(gw / rows_per_rung);
  const int64_t q0 = (gw - (int64_t)rr * rows_per_rung) * W;
  const int r = a.r0 + rr;
  const T* gm1 = a.gM + (int64_t)r * a.ksum;
  const T* gv1 = a.gV + (int64_t)r * a.ksum;
  const cplx<T>* src = a.src + (int64_t)rr * a.src_rs;
  const double vfac = a.cV[r] * a.invG;
  // tiled rows: (row, k1) at (k1 >> log2TKy) * tstride + row * TKy + (k1 & (TKy - 1))
  const int ltk = a.log2TKy;
  const int tkm = (1 << ltk) - 1;
  auto toff = [&](int row, int k1) -> int64_t {
    return (int64_t)(k1 >> ltk) * a.tstride + ((q0 + row) << ltk) + (k1 & tkm);
  };

  // Per lane: PPT (row, k1) items of the half spectrum k1 < n1/2.  Row index
  // ty * W + w: ty = 0 sums field, ty = 1 squares field.  The Nyquist column
  // k1 = n1/2 is packed into column 0 (imaginary part); its factors are the
  // same for every lane of a field type.
  T ld_g[PPT];
  int64_t ld_off[PPT];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int id = lane + 64 * u;
    const int k1 = id & (H - 1);
    const int rw = id >> (LOG2N1 - 1);
    ld_g[u] = (rw >> LOG2W) ? gv1[k1] : gm1[k1];
    // squares rows sit B fields after the sums rows
    ld_off[u] = toff(rw & (W - 1), k1) + ((rw >> LOG2W) ? (int64_t)a.B * a.fsize : 0);
  }
  const T gNm = gm1[H], gNv = gv1[H];

  double mean[PPT], m2[PPT];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    mean[u] = 0.0;
    m2[u] = 0.0;
  }
  const int B = a.B;
  // Loads of one bootstrap pair (fields b and b + 1 of the sums and squares
  // blocks), issued one pair ahead so they are in flight during the current
  // pair's transform and Welford updates.  Branch-free: a pair without a
  // second field reloads the first and discards it.
  cplx<T> pa[PPT], pb2[PPT];
  double pi1, pi2;
  auto issue = [&](int b) {
    const cplx<T>* f = src + (int64_t)b * a.fsize;
    const int64_t o2 = b + 1 < B ? a.fsize : 0;
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
      pa[u] = chunk_load(f + ld_off[u]);
      pb2[u] = chunk_load(f + ld_off[u] + o2);
    }
    pi1 = a.invk[b];
    pi2 = a.invk[b + 1 < B ? b + 1 : b];
  };
  issue(0);
  for (int b = 0; b < B; b += 2) {
    const bool has2 = b + 1 < B;
    const cplx<T> zero = {(T)0, (T)0};
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
      const int id = lane + 64 * u;
      const int k1 = id & (H - 1);
      const int rw = id >> (LOG2N1 - 1);
      cplx<T>* row = Z + rw * PL;
      if (k1 == 0) {
        // packed column: real part k1 = 0, imaginary part k1 = n1/2 (both
        // real-valued per field); A + i B for the pair at each
        const T gN = (rw >> LOG2W) ? gNv : gNm;
        const T br = has2 ? pb2[u].re : (T)0, bi = has2 ? pb2[u].im : (T)0;
        row[lds_slot<SWR, LE>(0)] = {pa[u].re * ld_g[u], br * ld_g[u]};
        row[lds_slot<SWR, LE>(H)] = {pa[u].im * gN, bi * gN};
      } else {
        const cplx<T> A = cscale(pa[u], ld_g[u]);
        const cplx<T> Bv = has2 ? cscale(pb2[u], ld_g[u]) : zero;
        row[lds_slot<SWR, LE>(k1)] = {A.re - Bv.im, A.im + Bv.re};
        if (k1 > 0) row[lds_slot<SWR, LE>(n1 - k1)] = {A.re + Bv.im, Bv.re - A.im};
      }
    }
    const double inv1 = pi1;
    const double inv2 = has2 ? pi2 : 0.0;
    wave_sync();
#pragma unroll
    for (int c = 0; c < 2 * W; c += ppc) {
      const int p = c + lane / lp;
      const bool act = p < 2 * W;
      wave_fft<LOG2N1, 1, T, E, SWR>(Z + (act ? p : 0) * PL, lane % lp, act, twl);
    }
    // next pair's loads (the last iteration reloads its own pair: valid
    // addresses, values unused)
    issue(b + 2 < B ? b + 2 : b);
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
      const int pt = lane + 64 * u;
      const int w = pt >> LOG2N1, j1 = pt & (n1 - 1);
      const cplx<T> zm = Z[w * PL + lds_slot<SWR, LE>(j1)];
      const cplx<T> zv = Z[(W + w) * PL + lds_slot<SWR, LE>(j1)];
      {
        const double am = fabs((double)zm.re) * a.invG;
        const double av = fabs((double)zv.re) * vfac;
        const double xv = vmr_ratio(av, am, a.invn);
        const double d = xv - mean[u];
        mean[u] += d * inv1;
        m2[u] += d * (xv - mean[u]);
      }
      if (has2) {
        const double am = fabs((double)zm.im) * a.invG;
        const double av = fabs((double)zv.im) * vfac;
        const double xv = vmr_ratio(av, am, a.invn);
        const double d = xv - mean[u];
        mean[u] += d * inv2;
        m2[u] += d * (xv - mean[u]);
      }
    }
    wave_sync();
  }
  const double sc = a.vscale[r];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int pt = lane + 64 * u;
    const int w = pt >> LOG2N1, j1 = pt & (n1 - 1);
    const int64_t qr = q0 + w;  // row of the layout -> grid row
    const int64_t qg = a.rows3 ? (qr >> a.log2n3) + ((qr & ((1 << a.log2n3) - 1)) << a.log2n2) : qr;
    const double vv = sc * m2[u] / (double)(B - 1);
    a.vmr[(int64_t)r * a.G + qg * n1 + j1] = isfinite(vv) ? log(vv) : (double)NAN;
  }
}

// Points per lane: 4 in FP64 for rows up to 256 points (8 KB of LDS per wave
// and about half the registers of 8 points), 8 otherwise; a 512-point row
// needs 8.
template <typename T, int E, int PPT, int LOG2N1>
static void rows_launch(unsigned blocks, size_t lds, hipStream_t s, const VmrArgs<T>& a) {
  k_ladder_rows<T, E, PPT, LOG2N1><<<blocks, kThreads, lds, s>>>(a);
}

// Row lengths 16..512 with W = 64 * PPT / n1 >= 1 rows per wave.
template <typename T, int E, int PPT>
static void rows_dispatch(int log2n1, unsigned blocks, size_t lds, hipStream_t s,
                          const VmrArgs<T>& a) {
  switch (log2n1) {
    case 4: rows_launch<T, E, PPT, 4>(blocks, lds, s, a); break;
    case 5: rows_launch<T, E, PPT, 5>(blocks, lds, s, a); break;
    case 6: rows_launch<T, E, PPT, 6>(blocks, lds, s, a); break;
    case 7: rows_launch<T, E, PPT, 7>(blocks, lds, s, a); break;
    case 8: rows_launch<T, E, PPT, 8>(blocks, lds, s, a); break;
    case 9:
      if constexpr (PPT == 8) rows_launch<T, E, PPT, 9>(blocks, lds, s, a);
      break;
    default: break;
  }
}

template <typename T>
bool launch_ladder_rows(const GridDev& g, const cplx<T>* src, int64_t src_rung_stride, int B,
                        int64_t n_samples, int R, int r0, const T* gM, const T* gV,
                        const double* cV, const double* vscale, const T* tw, const double* invk,
                        double* vmr, hipStream_t s) {
  if (!g.pack) return false;  // the kernel reads the packed Nyquist column
  VmrArgs<T> a;
  const int n1 = g.n[0];
  constexpr int E = sizeof(T) == 8 ? 8 : 16;
  const int ppt = (sizeof(T) == 8 && n1 <= 256) ? 4 : 8;
  const int W = 64 * ppt / n1;
  a.src = src;
  a.src_rs = src_rung_stride;
  a.fsize = g.fsY;
  a.log2TKy = g.log2TKy;
  a.tstride = g.Q << g.log2TKy;
  a.rows3 = g.N == 3;
  a.log2n2 = g.log2n[1];
  a.log2n3 = g.log2n[2];
  a.B = B;
  a.invn = 1.0 / (double)n_samples;
  a.invG = 1.0 / (double)g.G;
  a.R = R;
  a.r0 = r0;
  a.gM = gM;
  a.gV = gV;
  a.ksum = g.ksum;
  a.cV = cV;
  a.vscale = vscale;
  a.tw = (const cplx<T>*)tw;
  a.invk = invk;
  a.vmr = vmr;
  a.G = g.G;
  a.Q = g.Q;
  a.n1 = n1;
  a.log2n1 = g.log2n[0];
  a.H1 = g.H1;
  a.TQ = W;
  a.NBP = 1;
  a.ntiles = 0;
  const int64_t waves = (int64_t)R * (g.Q / W);
  const int wpb = kThreads / 64;
  const unsigned blocks = (unsigned)((waves + wpb - 1) / wpb);
  const size_t lds = ((size_t)wpb * 2 * W * pencil_stride(n1) + n1) * sizeof(cplx<T>);
  if (ppt == 4)
    rows_dispatch<T, E, 4>(g.log2n[0], blocks, lds, s, a);
  else
    rows_dispatch<T, E, 8>(g.log2n[0], blocks, lds, s, a);
  return true;
}

// ------------------------------------------------ 3-D ladder, small slabs
// The 3-D ladder step after the axis-2 pass for grids whose (k1, k3) slab of
// a field pair fits LDS (n1 n3 up to 64 x 64, 128 x 32 or 256 x 16 in FP64):
// one workgroup of 8 waves per (rung, j2) walks the bootstraps in order.
// Per bootstrap b, for column j2 of its sums and squares fields ([k1][j2][k3]
// from the axis-2 pass, k3 contiguous, with the axis-3 and axis-1 Gaussian
// factors already applied by that pass, except the axis-1 factors of the
// packed column 0):
//  - axis 3: the lanes of the pencil (plane, k1) load exactly the inputs of
//    their first radix-8 pass, one bootstrap ahead, and run that pass from
//    registers; the other passes go through the pencil in LDS;
//  - axis 1: the lanes of row j3 read, per input of their first radix-8
//    pass, the pair (sums, squares) at the column that input comes from and
//    form the row with the sums field in the real part and the squares
//    field in the imaginary part (Nyquist column packed into column 0, as
//    k_ladder_rows, with its two axis-1 factors); the first pass runs from
//    registers and the last one leaves its outputs in registers, where the
//    VMR ratio and Welford update of each point are formed.
// Against the axis-3 pass into tiled rows followed by k_ladder_rows, the
// rung's fields are read once and not written again.  Eight waves keep every
// lane of both transforms busy at 64 x 64 (64 pencils of 8 lanes, 8 rows of
// 8 lanes per wave); one workgroup runs per CU.
//
// SPLIT = 4: the axis-2 pass stored the first radix-4 step of the axis-3
// inverse (k_axis_ladder mode 3: its axis-3 factors applied, class p of
// k' at [k1][j2][p][k']); a workgroup per (rung, j2, p) takes
// the n3 = (axis-3 length) / 4 values of class p of every pencil, runs the
// rest of the axis-3 transform on them and writes the rows j3 = 4 m + p.  So
// a grid whose (k1, k3) slab pair is 4x too large for LDS (128^3: 295 KB)
// runs the same fused step on a quarter of it.
constexpr int kSlabThreads = 512;
// Twiddles of the axis-3 / axis-1 passes held in registers when a lane needs
// at most this many (else read from the LDS table in every pass); axis 1
// only for 64-point rows, where they fit the registers without spilling.
constexpr int kSlabRegTw3 = 8;
constexpr int kSlabRegTw1 = 8;
// Bootstraps whose planes are in flight while one is transformed: two for
// the 64 x 64 slab (64^3 ladder 46.3 -> 43.4 ms, 254 VGPRs, no spills), one
// elsewhere (the other shapes spill at two; 128^3 in quarters 490 -> 505 ms,
// 32^3 slower too).
template <int LOG2N1, int LOG2N3, int SPLIT>
constexpr int slab_depth() {
  return SPLIT == 1 && LOG2N1 == 6 && LOG2N3 == 6 ? 2 : 1;
}

template <int LOG2N1, int LOG2N3, int SPLIT>
__global__ __launch_bounds__(kSlabThreads) void k_slab3(VmrArgs<double> a) {
  constexpr int n1 = 1 << LOG2N1, H = n1 / 2, n3 = 1 << LOG2N3;
  constexpr int E = 8;
  // LDS banks on this chip, measured: a wave's 16-byte reads are served 16
  // lanes (256 bytes) at a time and its 16-byte writes 8 lanes at a time, and
  // lanes of one group that hit the same 4-byte bank with different
  // addresses conflict.  At 64 x 64 the planes sit 72 values apart (8 mod 16),
  // so the two 8-lane pencils of a read group fill 16 different 16-byte
  // slots, and the rows' lanes are numbered column-major (t1 = lane / 8), so
  // that the gather of the row inputs from the planes reads 8 rows of two
  // columns per group; with 73 and row-major lanes both had a 2-way conflict.
  // Writes (a pencil's 8 lanes at stride 9, or 8 rows 73 apart) stay
  // conflict-free.  64^3 ladder 45.0 -> 43.4 ms.
  constexpr bool kBanks64 = LOG2N1 == 6 && LOG2N3 == 6 && SPLIT == 1;
  // The 128 x 32 quarter: pencils of 4 lanes 36 values apart (4 mod 16), rows
  // 148 apart (4 mod 16), and each 16-lane group of the row phase takes 4 rows
  // of 4 consecutive columns, so that the axis-3 passes (4 pencils per read
  // group, 2 per write group), the gather (4 rows x 4 columns) and the row
  // passes all fill distinct slots; with 37 / 145 and 16-lane rows the axis-3
  // accesses and the row reads had 2-way conflicts.
  constexpr bool kBanksQ = LOG2N1 == 7 && LOG2N3 == 5 && SPLIT == 4;
  constexpr int PL1 = kBanksQ ? n1 + n1 / 8 + 4 : pencil_stride(n1);
  constexpr int PL3 = kBanks64 || kBanksQ ? n3 + n3 / 8 : pencil_stride(n3);
  static_assert(2 * H * PL3 + n3 * PL1 <= n1 * pencil_stride(n3) + n3 * pencil_stride(n1),
                "the planes and rows fit the LDS that slab3_lds allocates");
  constexpr int NW = kSlabThreads / 64;
  constexpr int RW = n3 / NW;                  // rows (values of j3) per wave
  constexpr int lp1 = n1 / E, lp3 = n3 / E;    // lanes per row / per pencil
  constexpr int nb1 = n1 / E, nb3 = n3 / E;    // input stride of the first passes
  constexpr int ppc3 = 64 / lp3;               // pencils per wave
  // the last axis-1 pass: radix and sub-transform length
  constexpr int LR1 = (LOG2N1 - 3) % 3 == 0 ? 8 : (1 << ((LOG2N1 - 3) % 3));
  constexpr int NS1 = n1 / LR1;
  static_assert(RW >= 1 && RW * lp1 <= 64 && NW * ppc3 >= 2 * H, "slab shape");
  static_assert(n1 > E && n3 > E, "slab axes longer than one radix-8 pass");
  static_assert(SPLIT == 1 || (SPLIT == 4 && nb3 == 4), "the split layout has 4-value runs");
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<double>* P = (cplx<double>*)smem;       // [2][H][PL3]: the pair's planes
  cplx<double>* Zall = P + 2 * H * PL3;        // [NW][RW][PL1]: axis-1 rows
  cplx<double>* twl1 = Zall + n3 * PL1;
  cplx<double>* twl3 = twl1 + n1;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int k = tid; k < n1; k += kSlabThreads) twl1[k] = a.tw[n1 - 2 + k];
  for (int k = tid; k < n3; k += kSlabThreads) twl3[k] = a.tw[n3 - 2 + k];
  const int n2 = a.n2r;
  int rr, j2, pcls = 0;
  if constexpr (SPLIT == 1) {
    rr = (int)blockIdx.x / n2;
    j2 = (int)blockIdx.x - rr * n2;
  } else {
    // the 4 classes of one (rung, j2) in consecutive workgroups
    const uint32_t bid = blockIdx.x;
    pcls = (int)(bid & 3u);
    const uint32_t colid = bid >> 2;
    rr = (int)colid / n2;
    j2 = (int)colid - rr * n2;
  }
  const int r = a.r0 + rr;
  const double* gm = a.gM + (int64_t)r * a.ksum;
  const double* gv = a.gV + (int64_t)r * a.ksum;
  const double vfac = a.cV[r] * a.invG;
  const int B = a.B;
  // axis-3 pencil of this lane: plane pl (0 sums field b, 1 squares field
  // B + b), column k1; its thread t3
  const int pen = wv * ppc3 + lane / lp3, t3 = lane % lp3;
  const bool act3 = pen < 2 * H;
  const int pl = act3 ? pen / H : 0, k1p = act3 ? pen % H : 0;
  // input i of this lane: k3 (or k' in class pcls, stored class-major at
  // [p][k']) = t3 + i nb3
  const cplx<double>* src = a.src + (int64_t)rr * a.src_rs + (int64_t)j2 * (n3 * SPLIT) +
                            (int64_t)pl * B * a.fsize + (int64_t)k1p * a.Q + pcls * n3 + t3;
  cplx<double>* pb3 = P + (act3 ? pen : 0) * PL3;
  cplx<double> nx[E], ny[E];
#pragma unroll
  for (int i = 0; i < E; ++i) nx[i] = ny[i] = {0.0, 0.0};
  auto load = [&](cplx<double>(&dst)[E], int bb) {
    if (act3) {
#pragma unroll
      for (int i = 0; i < E; ++i) dst[i] = chunk_load(src + (int64_t)bb * a.fsize + i * nb3);
    }
  };
  // axis-1 row of this lane: j3 = wv RW + w, its thread t1; the column that
  // input i (= t1 + i nb1) comes from
  const int w = kBanks64 ? lane % RW : kBanksQ ? (lane >> 2) & 3 : lane / lp1;
  const int t1 = kBanks64 ? lane / RW : kBanksQ ? ((lane >> 4) << 2) | (lane & 3) : lane % lp1;
  const bool act1 = w < RW;
  const int j3 = wv * RW + (act1 ? w : 0);
  cplx<double>* pb1 = Zall + (wv * RW + (act1 ? w : 0)) * PL1;
  int col[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int k = t1 + i * nb1;
    col[i] = k < H ? k : (k == H ? 0 : n1 - k);
  }
  // the axis-2 pass applied the axis-3 and axis-1 factors, except the
  // axis-1 factors of the packed column (k1 = 0 and n1/2: inputs 0 and E/2
  // of the lanes with t1 = 0)
  const double fm0 = gm[0], fv0 = gv[0], fmH = gm[H], fvH = gv[H];
  double mean[E], m2[E];
#pragma unroll
  for (int u = 0; u < E; ++u) {
    mean[u] = 0.0;
    m2[u] = 0.0;
  }
  // loop-invariant twiddles of the passes after the first, in registers
  constexpr int NT3 = wave_tw_count<LOG2N3, 3, E>(), NT1 = wave_tw_count<LOG2N1, 3, E>();
  // the 128-point rows of the quarter kernel (11 twiddles) fit too, now that
  // the Gaussian factors come from the axis-2 pass
  constexpr bool rt3 = NT3 <= kSlabRegTw3,
                 rt1 = (NT1 <= kSlabRegTw1 && LOG2N1 == 6) || (kBanksQ && NT1 <= 12);
  cplx<double> tw3r[rt3 ? NT3 : 1], tw1r[rt1 ? NT1 : 1];
  if constexpr (rt3)
    wave_tw_load<LOG2N3, 3, E, 1, double, 0, NT3>(t3, (const cplx<double>*)a.tw + n3 - 2, tw3r);
  if constexpr (rt1)
    wave_tw_load<LOG2N1, 3, E, 1, double, 0, NT1>(t1, (const cplx<double>*)a.tw + n1 - 2, tw1r);
  // one bootstrap: its planes were loaded into cur kDepth bootstraps ahead;
  // cur then takes the loads of bootstrap b + kDepth (or b's again at the end:
  // valid addresses, values unused)
  constexpr int kSlabDepth = slab_depth<LOG2N1, LOG2N3, SPLIT>();
  auto step = [&](int b, cplx<double>(&cur)[E]) {
    const double inv = a.invk[b];
    cplx<double> v[E];
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = cur[i];
    load(cur, b + kSlabDepth < B ? b + kSlabDepth : b);
    __syncthreads();  // every wave is done reading the previous planes
    wave_first_pass<LOG2N3, E, 1, double>(pb3, t3, act3, v);
    if constexpr (rt3)
      wave_fft_from_rt<LOG2N3, 3, E, 1, double, 0, NT3>(pb3, t3, act3, tw3r);
    else
      wave_fft_from<LOG2N3, 3, E, 1, double>(pb3, t3, act3, twl3);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const cplx<double> A = P[col[i] * PL3 + pad8(j3)];
      const cplx<double> Bv = P[(H + col[i]) * PL3 + pad8(j3)];
      if (i < E / 2)
        v[i] = {A.re - Bv.im, A.im + Bv.re};  // column k = t1 + i nb1 < n1/2
      else
        v[i] = {A.re + Bv.im, Bv.re - A.im};  // column n1 - k
    }
    if (t1 == 0) {
      // packed column: real part k1 = 0, imaginary part k1 = n1/2
      const cplx<double> pa = P[pad8(j3)], pq = P[H * PL3 + pad8(j3)];
      v[0] = {pa.re * fm0, pq.re * fv0};
      v[E / 2] = {pa.im * fmH, pq.im * fvH};
    }
    wave_first_pass<LOG2N1, E, 1, double>(pb1, t1, act1, v);
    if constexpr (rt1)
      wave_fft_to_regs_rt<LOG2N1, 3, E, 1, double, 0, NT1>(pb1, t1, act1, tw1r, v);
    else
      wave_fft_to_regs<LOG2N1, 3, E, 1, double>(pb1, t1, act1, twl1, v);
    if (act1) {
#pragma unroll
      for (int u = 0; u < E; ++u) {
        const cplx<double> z = v[u];  // sums + i squares
        const double xv = vmr_ratio(fabs(z.im) * vfac, fabs(z.re) * a.invG, a.invn);
        const double d = xv - mean[u];
        mean[u] += d * inv;
        m2[u] += d * (xv - mean[u]);
      }
    }
  };
  if constexpr (kSlabDepth == 1) {
    load(nx, 0);
    for (int b = 0; b < B; ++b) step(b, nx);
  } else {
    load(nx, 0);
    load(ny, B > 1 ? 1 : 0);
    for (int b = 0; b < B; b += 2) {
      step(b, nx);
      if (b + 1 < B) step(b + 1, ny);
    }
  }
  if (!act1) return;
  const double sc = a.vscale[r];
  const int64_t q = (int64_t)j2 + (int64_t)n2 * (SPLIT * j3 + pcls);  // grid row j2 + n2 j3
  double* out = a.vmr + (int64_t)r * a.G + q * n1;
#pragma unroll
  for (int u = 0; u < E; ++u) {
    // output u of the last pass: X[j + s NS1], j = t1 + (u / LR1) nb1, s = u % LR1
    const int j1 = t1 + (u / LR1) * nb1 + (u % LR1) * NS1;
    const double vv = sc * m2[u] / (double)(B - 1);
    out[j1] = isfinite(vv) ? log(vv) : (double)NAN;
  }
}

static size_t slab3_lds(int n1, int n3) {
  // planes [2][n1/2][PL3], rows [n3][PL1], twiddles
  return ((size_t)n1 * pencil_stride(n3) + (size_t)n3 * pencil_stride(n1) + n1 + n3) *
         sizeof(cplx<double>);
}

// Shapes (log2 n1, log2 n3) with an instantiation of k_slab3.
static bool slab3_shape(int l1, int l3) {
  switch (l1 * 16 + l3) {
    case 4 * 16 + 6: case 5 * 16 + 5: case 5 * 16 + 6: case 6 * 16 + 4: case 6 * 16 + 5:
    case 6 * 16 + 6: case 7 * 16 + 4: case 7 * 16 + 5: case 8 * 16 + 4:
      return true;
    default:
      return false;
  }
}

bool slab3_ok(const GridDev& g) {
  return g.N == 3 && g.pack && g.pow2 && slab3_shape(g.log2n[0], g.log2n[2]) &&
         slab3_lds(g.n[0], g.n[2]) <= 160 * 1024;
}

// The split form: a 32-value quarter of axis 3 per workgroup (axis-3 length
// 128), where the whole slab does not fit; the axis-2 pass (mode 3, 16
// pencils per block) takes axis-2 lengths 16..128 in FP64.
bool slab3_split_ok(const GridDev& g) {
  return g.N == 3 && g.pack && g.pow2 && !slab3_ok(g) && g.n[2] == 128 && g.n[1] >= 16 &&
         g.n[1] <= 128 && (g.log2n[0] >= 5 && g.log2n[0] <= 7) &&
         slab3_shape(g.log2n[0], 5) && slab3_lds(g.n[0], 32) <= 160 * 1024;
}

template <int LOG2N1, int LOG2N3, int SPLIT>
static void slab3_launch(unsigned blocks, hipStream_t s, const VmrArgs<double>& a) {
  constexpr int n1 = 1 << LOG2N1, n3 = 1 << LOG2N3;
  static const bool attr = [] {
    // more than 64 KB of LDS per workgroup; set once per instantiation
    (void)hipFuncSetAttribute((const void*)k_slab3<LOG2N1, LOG2N3, SPLIT>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)slab3_lds(n1, n3));
    return true;
  }();
  (void)attr;
  k_slab3<LOG2N1, LOG2N3, SPLIT><<<blocks, kSlabThreads, slab3_lds(n1, n3), s>>>(a);
}

bool launch_slab3(const GridDev& g, const cplx<double>* src, int64_t src_rung_stride, int B,
                  int64_t n_samples, int R, int r0, const double* gM, const double* gV,
                  const double* cV, const double* vscale, const double* tw, const double* invk,
                  double* vmr, hipStream_t s) {
  const bool split = !slab3_ok(g);
  if (split && !slab3_split_ok(g)) return false;
  VmrArgs<double> a{};
  a.src = src;
  a.src_rs = src_rung_stride;
  a.fsize = (int64_t)g.Hc * g.Q;  // the axis-2 pass's output: n1/2 columns per field
  a.B = B;
  a.invn = 1.0 / (double)n_samples;
  a.invG = 1.0 / (double)g.G;
  a.R = R;
  a.r0 = r0;
  a.gM = gM;
  a.gV = gV;
  a.ksum = g.ksum;
  a.goff3 = g.koff[2];
  a.cV = cV;
  a.vscale = vscale;
  a.tw = (const cplx<double>*)tw;
  a.invk = invk;
  a.vmr = vmr;
  a.G = g.G;
  a.Q = g.Q;
  a.n1 = g.n[0];
  a.n2r = g.n[1];
  if (split) {
    const int64_t cols = (int64_t)R * g.n[1];
    const unsigned blocks = (unsigned)(cols * 4);
    switch (g.log2n[0]) {
      case 5: slab3_launch<5, 5, 4>(blocks, s, a); break;
      case 6: slab3_launch<6, 5, 4>(blocks, s, a); break;
      case 7: slab3_launch<7, 5, 4>(blocks, s, a); break;
      default: return false;
    }
    return true;
  }
  const unsigned blocks = (unsigned)((int64_t)R * g.n[1]);
  switch (g.log2n[0] * 16 + g.log2n[2]) {
    case 4 * 16 + 6: slab3_launch<4, 6, 1>(blocks, s, a); break;
    case 5 * 16 + 5: slab3_launch<5, 5, 1>(blocks, s, a); break;
    case 5 * 16 + 6: slab3_launch<5, 6, 1>(blocks, s, a); break;
    case 6 * 16 + 4: slab3_launch<6, 4, 1>(blocks, s, a); break;
    case 6 * 16 + 5: slab3_launch<6, 5, 1>(blocks, s, a); break;
    case 6 * 16 + 6: slab3_launch<6, 6, 1>(blocks, s, a); break;
    case 7 * 16 + 4: slab3_launch<7, 4, 1>(blocks, s, a); break;
    case 7 * 16 + 5: slab3_launch<7, 5, 1>(blocks, s, a); break;
    case 8 * 16 + 4: slab3_launch<8, 4, 1>(blocks, s, a); break;
    default: return false;
  }
  return true;
}

template bool launch_ladder_rows<float>(const GridDev&, const cplx<float>*, int64_t, int, int64_t,
                                        int, int, const float*, const float*, const double*,
                                        const double*, const float*, const double*, double*,
                                        hipStream_t);
template bool launch_ladder_rows<double>(const GridDev&, const cplx<double>*, int64_t, int,
                                         int64_t, int, int, const double*, const double*,
                                         const double*, const double*, const double*,
                                         const double*, double*, hipStream_t);

// ------------------------------------------------- mixed-radix 2-D ladder
// The 2-D ladder for axis lengths that are not powers of two (FP64), in the
// same two steps as the power-of-two path: a per-rung axis-2 pass writing the
// tiled row layout (tiles of 8 columns, whole 128-byte lines), then one row
// per wavefront for the axis-1 transform, VMR and Welford.  Both use the
// wave-local mixed-radix FFT, one pencil at a time over the 64 lanes of a
// wavefront, with the plan's twiddle tables staged in LDS.
constexpr int kGenTK = 8;  // tile width of the tiled row layout (FP64)

// Axis-2 pass of the mixed-radix 2-D ladder: one block per (field, 8-column
// tile, share of the chunk's rungs) reads its spectrum columns once into LDS
// and, for every rung of its share, applies that rung's Gaussian factor, runs
// the wave-local FFT and stores the tile's rows into the tiled row layout
// (whole 128-byte lines: columns past H1 up to the tile width are zero).
template <int LP>
__global__ __launch_bounds__(kThreads) void k_axis_gen_ladder(AxisArgs<double> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int L = a.L;
  const int PL = gen_pencil_stride(L);
  cplx<double>* buf = (cplx<double>*)smem;          // working pencils [kGenTK][PL]
  cplx<double>* raw = buf + kGenTK * PL;             // spectrum columns [kGenTK][L]
  cplx<double>* twl = raw + kGenTK * L;
  for (int k = threadIdx.x; k < L; k += kThreads) twl[k] = a.twg[k];
  uint32_t bid = blockIdx.x;
  const int sp = (int)(bid % (uint32_t)a.nsplit);
  bid /= (uint32_t)a.nsplit;
  const int rbeg = sp * a.R / a.nsplit, rend = (sp + 1) * a.R / a.nsplit;
  const int kt = (int)(bid % (uint32_t)a.ktiles);
  const int f = a.fid0 + (int)(bid / (uint32_t)a.ktiles);
  const int k0 = kt * kGenTK;
  const int TKe = (a.H1 - k0) < kGenTK ? (a.H1 - k0) : kGenTK;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const cplx<double>* src = a.src + (int64_t)(f - a.src_fid0) * a.fsize + (int64_t)k0 * a.Q;
  for (int id = tid; id < kGenTK * L; id += kThreads) {
    const int kk = id / L, e = id - kk * L;
    raw[id] = kk < TKe ? src[(int64_t)kk * a.Q + e] : cplx<double>{0.0, 0.0};
  }
  const double* gbase = (f < a.fsplit ? a.gM : a.gV) + a.goff;
  for (int rr = rbeg; rr < rend; ++rr) {
    const double* gt = gbase + (int64_t)(a.r0 + rr) * a.ksum;
    __syncthreads();  // the columns are loaded; the previous rung's stores have read buf
    for (int id = tid; id < kGenTK * L; id += kThreads) {
      const int kk = id / L, e = id - kk * L;
      buf[kk * PL + pad8(e)] = kk < TKe ? cscale(raw[id], gt[e]) : cplx<double>{0.0, 0.0};
    }
    __syncthreads();
    // wave wv transforms columns 2 wv and 2 wv + 1 of the tile
    wave_fft_gen_pair<LP>(buf + 2 * wv * PL, 2 * wv < TKe, buf + (2 * wv + 1) * PL,
                          2 * wv + 1 < TKe, lane, a.fac, twl);
    __syncthreads();
    cplx<double>* tile = a.dst + (int64_t)rr * a.dst_rs + (int64_t)(f - a.dst_fid0) * a.dst_fsize +
                         (int64_t)kt * a.Q * kGenTK;
    for (int id = tid; id < kGenTK * L; id += kThreads) {
      const int k = id & (kGenTK - 1);
      const int e = id >> 3;
      tile[(int64_t)e * kGenTK + k] = buf[k * PL + pad8(e)];
    }
  }
}

static size_t gen_ladder_lds(int L) {
  return ((size_t)kGenTK * gen_pencil_stride(L) + (size_t)kGenTK * L + L) * sizeof(cplx<double>);
}

bool ladder_gen_rows_ok(const GridDev& g) {
  return (g.N == 2 || g.N == 3) && !g.pow2 && g.TKy == kGenTK && (g.n[0] + 63) / 64 <= 8 &&
         wave_gen_fits(g.facw[0]) && wave_gen_fits(g.facw[1]) &&
         (g.N == 2 || wave_gen_fits(g.facw[2]));
}

// 3-D, first pass (grid axis 2, length n2 = a.I), ladder form: one block per
// (field, k1, 8 consecutive j3, share of the chunk's rungs) reads its 8
// spectrum pencils along j2 (contiguous in [k1][j2 + n2 j3]) once into LDS
// and, for every rung of its share, applies the rung's axis-2 factor, runs
// the wave-local FFT and stores [k1][j2][j3], 8 consecutive j3 (one 128-byte
// run) per (k1, j2).
template <int LP>
__global__ __launch_bounds__(kThreads) void k_gen3_pass1(AxisArgs<double> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int L = a.L;  // n2
  const int PL = gen_pencil_stride(L);
  cplx<double>* buf = (cplx<double>*)smem;  // working pencils [kGenTK][PL]
  cplx<double>* raw = buf + kGenTK * PL;     // spectrum pencils [kGenTK][L]
  cplx<double>* twl = raw + kGenTK * L;
  for (int k = threadIdx.x; k < L; k += kThreads) twl[k] = a.twg[k];
  const int64_t n3 = a.O;
  const int jtiles = (int)((n3 + kGenTK - 1) / kGenTK);
  uint32_t bid = blockIdx.x;
  const int sp = (int)(bid % (uint32_t)a.nsplit);
  bid /= (uint32_t)a.nsplit;
  const int rbeg = sp * a.R / a.nsplit, rend = (sp + 1) * a.R / a.nsplit;
  const int jt = (int)(bid % (uint32_t)jtiles);
  bid /= (uint32_t)jtiles;
  const int k1 = (int)(bid % (uint32_t)a.H1);
  const int f = a.fid0 + (int)(bid / (uint32_t)a.H1);
  const int j30 = jt * kGenTK;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const cplx<double>* src = a.src + (int64_t)(f - a.src_fid0) * a.fsize + (int64_t)k1 * a.Q;
  for (int id = tid; id < kGenTK * L; id += kThreads) {
    const int j = id / L, e = id - j * L;  // pencil j: j3 = j30 + j
    raw[id] = j30 + j < n3 ? src[(int64_t)(j30 + j) * L + e] : cplx<double>{0.0, 0.0};
  }
  const double* gbase = (f < a.fsplit ? a.gM : a.gV) + a.goff;
  for (int rr = rbeg; rr < rend; ++rr) {
    const double* gt = gbase + (int64_t)(a.r0 + rr) * a.ksum;
    __syncthreads();  // the pencils are loaded; the previous rung's stores have read buf
    for (int id = tid; id < kGenTK * L; id += kThreads) {
      const int j = id / L, e = id - j * L;
      buf[j * PL + pad8(e)] = j30 + j < n3 ? cscale(raw[id], gt[e]) : cplx<double>{0.0, 0.0};
    }
    __syncthreads();
    wave_fft_gen_pair<LP>(buf + 2 * wv * PL, j30 + 2 * wv < n3, buf + (2 * wv + 1) * PL,
                          j30 + 2 * wv + 1 < n3, lane, a.fac, twl);
    __syncthreads();
    cplx<double>* dst = a.dst + (int64_t)rr * a.dst_rs + (int64_t)(f - a.dst_fid0) * a.dst_fsize +
                        (int64_t)k1 * a.Q;
    for (int id = tid; id < kGenTK * L; id += kThreads) {
      const int j = id & (kGenTK - 1), e = id >> 3;
      if (j30 + j < n3) dst[(int64_t)e * n3 + j30 + j] = buf[j * PL + pad8(e)];
    }
  }
}

// 3-D, second pass (grid axis 3, length n3): one block per (rung, field, j2,
// 8-column k1 tile); pencils along j3 are contiguous in [k1][j2][j3]; the
// result goes to the tiled row layout with rows q' = j3 + n3 j2 (whole
// 128-byte lines: columns past H1 are zero).
template <int LP>
__global__ __launch_bounds__(kThreads) void k_gen3_pass2(AxisArgs<double> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<double>* buf = (cplx<double>*)smem;
  const int L = a.L;  // n3
  const int PL = gen_pencil_stride(L);
  cplx<double>* twl = buf + kGenTK * PL;
  for (int k = threadIdx.x; k < L; k += kThreads) twl[k] = a.twg[k];
  uint32_t bid = blockIdx.x;
  const int kt = (int)(bid % (uint32_t)a.ktiles);
  bid /= (uint32_t)a.ktiles;
  const int64_t j2 = bid % (uint32_t)a.I;
  bid /= (uint32_t)a.I;
  const int f = a.fid0 + (int)(bid % (uint32_t)a.nf);
  const int rr = (int)(bid / (uint32_t)a.nf);
  const int k0 = kt * kGenTK;
  const int TKe = (a.H1 - k0) < kGenTK ? (a.H1 - k0) : kGenTK;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const double* gt = (f < a.fsplit ? a.gM : a.gV) + (int64_t)(a.r0 + rr) * a.ksum + a.goff;
  const cplx<double>* src = a.src + (int64_t)rr * a.src_rs + (int64_t)(f - a.src_fid0) * a.fsize +
                            j2 * L;
  for (int id = tid; id < kGenTK * L; id += kThreads) {
    const int kk = id / L, e = id - kk * L;
    cplx<double> v = {0.0, 0.0};
    if (kk < TKe) v = cscale(src[(int64_t)(k0 + kk) * a.Q + e], gt[e]);
    buf[kk * PL + pad8(e)] = v;
  }
  __syncthreads();
  wave_fft_gen_pair<LP>(buf + 2 * wv * PL, 2 * wv < TKe, buf + (2 * wv + 1) * PL,
                        2 * wv + 1 < TKe, lane, a.fac, twl);
  __syncthreads();
  cplx<double>* tile = a.dst + (int64_t)rr * a.dst_rs + (int64_t)(f - a.dst_fid0) * a.dst_fsize +
                       (int64_t)kt * a.Q * kGenTK + j2 * L * kGenTK;
  for (int id = tid; id < kGenTK * L; id += kThreads)
    tile[id] = buf[(id & (kGenTK - 1)) * PL + pad8(id >> 3)];
}

bool launch_gen3_passes(const GridDev& g, const cplx<double>* spec, cplx<double>* y2,
                        cplx<double>* y, int64_t rung_stride, int nf, int R, int r0,
                        const double* gM, const double* gV, int fsplit, const double* twg,
                        hipStream_t s) {
  if (g.N != 3 || !ladder_gen_rows_ok(g)) return false;
  AxisArgs<double> a{};
  a.src = spec;
  a.dst = y2;
  a.dst_rs = rung_stride;
  a.fsize = (int64_t)g.H1 * g.Q;
  a.dst_fsize = a.fsize;
  a.nf = nf;
  a.R = R;
  a.r0 = r0;
  a.gM = gM;
  a.gV = gV;
  a.fsplit = fsplit;
  a.ksum = g.ksum;
  a.Q = g.Q;
  a.H1 = g.H1;
  a.I = g.n[1];
  a.O = g.n[2];
  // pass 1: axis 2
  a.L = g.n[1];
  a.goff = g.koff[1];
  a.fac = g.facw[1];
  a.twg = (const cplx<double>*)twg + g.koff[1];
  const int jtiles = (g.n[2] + kGenTK - 1) / kGenTK;
  {
    static const bool attr = [] {
      (void)hipFuncSetAttribute((const void*)k_gen3_pass1<32>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      (void)hipFuncSetAttribute((const void*)k_gen3_pass1<64>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      return true;
    }();
    (void)attr;
    // with few tiles, split the rungs so that about 8000 blocks run, each
    // keeping at least 8 rungs (it reads its pencils once per share)
    const int64_t tiles = (int64_t)nf * g.H1 * jtiles;
    a.nsplit = (int)std::max<int64_t>(1, std::min<int64_t>((kAxisSplitTarget + tiles - 1) / tiles,
                                                           R / 8));
    const unsigned blocks = (unsigned)(tiles * a.nsplit);
    if (wave_gen_fits_lp(a.fac, 32))
      k_gen3_pass1<32><<<blocks, kThreads, gen_ladder_lds(a.L), s>>>(a);
    else
      k_gen3_pass1<64><<<blocks, kThreads, gen_ladder_lds(a.L), s>>>(a);
  }
  size_t lds;
  // pass 2: axis 3, into the tiled rows
  a.src = y2;
  a.src_rs = rung_stride;
  a.dst = y;
  a.dst_fsize = g.fsY;
  a.L = g.n[2];
  a.goff = g.koff[2];
  a.fac = g.facw[2];
  a.twg = (const cplx<double>*)twg + g.koff[2];
  a.ktiles = g.ktY;
  lds = ((size_t)kGenTK * gen_pencil_stride(a.L) + a.L) * sizeof(cplx<double>);
  if (wave_gen_fits_lp(a.fac, 32))
    k_gen3_pass2<32><<<(unsigned)((int64_t)R * nf * g.n[1] * g.ktY), kThreads, lds, s>>>(a);
  else
    k_gen3_pass2<64><<<(unsigned)((int64_t)R * nf * g.n[1] * g.ktY), kThreads, lds, s>>>(a);
  return true;
}

bool launch_axis_gen_rows(const GridDev& g, const cplx<double>* spec, cplx<double>* dst,
                          int64_t dst_rung_stride, int nf, int R, int r0, const double* gM,
                          const double* gV, int fsplit, const double* twg, hipStream_t s) {
  if (!ladder_gen_rows_ok(g)) return false;
  AxisArgs<double> a{};
  a.src = spec;
  a.dst = dst;
  a.dst_rs = dst_rung_stride;
  a.fsize = (int64_t)g.H1 * g.Q;
  a.dst_fsize = g.fsY;
  a.nf = nf;
  a.R = R;
  a.r0 = r0;
  a.gM = gM;
  a.gV = gV;
  a.fsplit = fsplit;
  a.ksum = g.ksum;
  a.goff = g.koff[1];
  a.L = g.n[1];
  a.Q = g.Q;
  a.H1 = g.H1;
  a.ktiles = g.ktY;
  a.fac = g.facw[1];
  a.twg = (const cplx<double>*)twg + g.koff[1];
  {
    // with few tiles, split the rungs so that about 8000 blocks run, each
    // keeping at least 8 rungs (it reads its columns once per share); every
    // length that fits the wave-local FFT (<= 512) fits LDS
    static const bool attr = [] {
      (void)hipFuncSetAttribute((const void*)k_axis_gen_ladder<32>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      (void)hipFuncSetAttribute((const void*)k_axis_gen_ladder<64>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      return true;
    }();
    (void)attr;
    const int64_t tiles = (int64_t)nf * g.ktY;
    a.nsplit = (int)std::max<int64_t>(1, std::min<int64_t>((kAxisSplitTarget + tiles - 1) / tiles,
                                                           R / 8));
    const unsigned blocks = (unsigned)(tiles * a.nsplit);
    if (wave_gen_fits_lp(a.fac, 32))
      k_axis_gen_ladder<32><<<blocks, kThreads, gen_ladder_lds(a.L), s>>>(a);
    else
      k_axis_gen_ladder<64><<<blocks, kThreads, gen_ladder_lds(a.L), s>>>(a);
  }
  return true;
}

// One wavefront per (rung, row q): the row's sums and squares fields of a
// bootstrap pair form two complex pencils, transformed one after the other;
// then VMR and Welford over b in order for its PPT points per lane.  Loads
// run one pair ahead.  Same arithmetic as k_ladder_rows.
template <int PPT, int LP>
__global__ __launch_bounds__(kThreads) void k_ladder_rows_gen(VmrArgs<double> a) {
  constexpr int IT = PPT + 1;  // half-spectrum items per lane (2 H1 <= 64 PPT + 2)
  extern __shared__ __align__(16) unsigned char smem[];
  const int n1 = a.n1, H1 = a.H1;
  const int PL = gen_pencil_stride(n1);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  cplx<double>* twl = (cplx<double>*)smem + (kThreads / 64) * 2 * PL;
  for (int k = threadIdx.x; k < n1; k += kThreads) twl[k] = a.twg[k];
  __syncthreads();
  cplx<double>* Z = (cplx<double>*)smem + wv * 2 * PL;
  const int64_t gw = (int64_t)blockIdx.x * (kThreads / 64) + wv;
  if (gw >= (int64_t)a.R * a.Q) return;
  const int rr = (int)(gw / a.Q);
  const int64_t q = gw - (int64_t)rr * a.Q;
  const int r = a.r0 + rr;
  const double* gm1 = a.gM + (int64_t)r * a.ksum;
  const double* gv1 = a.gV + (int64_t)r * a.ksum;
  const cplx<double>* src = a.src + (int64_t)rr * a.src_rs;
  const double vfac = a.cV[r] * a.invG;
  // item u: k1 = id mod H1 of the sums (ty 0) or squares (ty 1) row
  int it_k1[IT];
  bool it_sq[IT], it_ok[IT];
  double it_g[IT];
  int64_t it_off[IT];
#pragma unroll
  for (int u = 0; u < IT; ++u) {
    const int id = lane + 64 * u;
    it_ok[u] = id < 2 * H1;
    it_sq[u] = id >= H1;
    const int k1 = it_ok[u] ? (it_sq[u] ? id - H1 : id) : 0;
    it_k1[u] = k1;
    it_g[u] = it_sq[u] ? gv1[k1] : gm1[k1];
    it_off[u] = (int64_t)(k1 >> 3) * a.tstride + q * kGenTK + (k1 & 7) +
                (it_sq[u] ? (int64_t)a.B * a.fsize : 0);
  }
  double mean[PPT], m2[PPT];
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    mean[u] = 0.0;
    m2[u] = 0.0;
  }
  const int B = a.B;
  cplx<double> pa[IT], pb2[IT];
  double pi1, pi2;
  auto issue = [&](int b) {
    const cplx<double>* fb = src + (int64_t)b * a.fsize;
    const int64_t o2 = b + 1 < B ? a.fsize : 0;
#pragma unroll
    for (int u = 0; u < IT; ++u) {
      pa[u] = fb[it_off[u]];
      pb2[u] = fb[it_off[u] + o2];
    }
    pi1 = a.invk[b];
    pi2 = a.invk[b + 1 < B ? b + 1 : b];
  };
  issue(0);
  for (int b = 0; b < B; b += 2) {
    const bool has2 = b + 1 < B;
    const cplx<double> zero = {0.0, 0.0};
#pragma unroll
    for (int u = 0; u < IT; ++u) {
      if (it_ok[u]) {
        const int k1 = it_k1[u];
        const cplx<double> A = cscale(pa[u], it_g[u]);
        const cplx<double> Bv = has2 ? cscale(pb2[u], it_g[u]) : zero;
        cplx<double>* row = Z + (it_sq[u] ? PL : 0);
        row[pad8(k1)] = {A.re - Bv.im, A.im + Bv.re};
        if (k1 > 0 && 2 * k1 < n1) row[pad8(n1 - k1)] = {A.re + Bv.im, Bv.re - A.im};
      }
    }
    const double inv1 = pi1;
    const double inv2 = has2 ? pi2 : 0.0;
    wave_sync();
    wave_fft_gen_pair<LP>(Z, true, Z + PL, true, lane, a.fac, twl);
    issue(b + 2 < B ? b + 2 : b);
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
      const int j1 = lane + 64 * u;
      if (j1 < n1) {
        const cplx<double> zm = Z[pad8(j1)];
        const cplx<double> zv = Z[PL + pad8(j1)];
        {
          const double am = fabs(zm.re) * a.invG;
          const double av = fabs(zv.re) * vfac;
          const double xv = vmr_ratio(av, am, a.invn);
          const double d = xv - mean[u];
          mean[u] += d * inv1;
          m2[u] += d * (xv - mean[u]);
        }
        if (has2) {
          const double am = fabs(zm.im) * a.invG;
          const double av = fabs(zv.im) * vfac;
          const double xv = vmr_ratio(av, am, a.invn);
          const double d = xv - mean[u];
          mean[u] += d * inv2;
          m2[u] += d * (xv - mean[u]);
        }
      }
    }
    wave_sync();
  }
  const double sc = a.vscale[r];
  // 3-D: layout rows q' = j3 + n3 j2 -> grid row j2 + n2 j3
  const int64_t qg = a.rows3 ? q / a.n3r + (q % a.n3r) * a.n2r : q;
#pragma unroll
  for (int u = 0; u < PPT; ++u) {
    const int j1 = lane + 64 * u;
    if (j1 < n1) {
      const double vv = sc * m2[u] / (double)(B - 1);
      a.vmr[(int64_t)r * a.G + qg * n1 + j1] = isfinite(vv) ? log(vv) : (double)NAN;
    }
  }
}

bool launch_ladder_rows_gen(const GridDev& g, const cplx<double>* src, int64_t src_rung_stride,
                            int B, int64_t n_samples, int R, int r0, const double* gM,
                            const double* gV, const double* cV, const double* vscale,
                            const double* twg, const double* invk, double* vmr, hipStream_t s) {
  const int n1 = g.n[0];
  const int ppt = (n1 + 63) / 64;
  if (!ladder_gen_rows_ok(g)) return false;
  VmrArgs<double> a{};
  a.src = src;
  a.src_rs = src_rung_stride;
  a.fsize = g.fsY;
  a.tstride = (int64_t)g.Q * kGenTK;
  a.B = B;
  a.invn = 1.0 / (double)n_samples;
  a.invG = 1.0 / (double)g.G;
  a.R = R;
  a.r0 = r0;
  a.gM = gM;
  a.gV = gV;
  a.ksum = g.ksum;
  a.cV = cV;
  a.vscale = vscale;
  a.invk = invk;
  a.vmr = vmr;
  a.G = g.G;
  a.Q = g.Q;
  a.n1 = n1;
  a.H1 = g.H1;
  a.fac = g.facw[0];
  a.twg = (const cplx<double>*)twg + g.koff[0];
  a.rows3 = g.N == 3;
  a.n2r = g.n[1];
  a.n3r = g.N == 3 ? g.n[2] : 1;
  const int64_t waves = (int64_t)R * g.Q;
  const int wpb = kThreads / 64;
  const unsigned blocks = (unsigned)((waves + wpb - 1) / wpb);
  const size_t lds = ((size_t)wpb * 2 * gen_pencil_stride(n1) + n1) * sizeof(cplx<double>);
  // the sums and squares pencils of a row together in the two halves of the
  // wave when the row length fits 32 lanes
  if (wave_gen_fits_lp(a.fac, 32)) {
    switch (ppt) {
      case 1: k_ladder_rows_gen<1, 32><<<blocks, kThreads, lds, s>>>(a); break;
      case 2: k_ladder_rows_gen<2, 32><<<blocks, kThreads, lds, s>>>(a); break;
      case 3: k_ladder_rows_gen<3, 32><<<blocks, kThreads, lds, s>>>(a); break;
      default: k_ladder_rows_gen<4, 32><<<blocks, kThreads, lds, s>>>(a); break;
    }
    return true;
  }
  switch (ppt) {
    case 1: k_ladder_rows_gen<1, 64><<<blocks, kThreads, lds, s>>>(a); break;
    case 2: k_ladder_rows_gen<2, 64><<<blocks, kThreads, lds, s>>>(a); break;
    case 3: k_ladder_rows_gen<3, 64><<<blocks, kThreads, lds, s>>>(a); break;
    case 4: k_ladder_rows_gen<4, 64><<<blocks, kThreads, lds, s>>>(a); break;
    case 5: k_ladder_rows_gen<5, 64><<<blocks, kThreads, lds, s>>>(a); break;
    case 6: k_ladder_rows_gen<6, 64><<<blocks, kThreads, lds, s>>>(a); break;
    case 7: k_ladder_rows_gen<7, 64><<<blocks, kThreads, lds, s>>>(a); break;
    default: k_ladder_rows_gen<8, 64><<<blocks, kThreads, lds, s>>>(a); break;
  }
  return true;
}

// -------------------------------------------------------------- convergence
// find_stability (DirectSpace.jl:975-1031) applied rung after rung with the
// history rotation of update_density_state_vmr! (GradeProEstimator.jl:617-625);
// state stays in registers.  chosen = rung of the last density assignment, or
// S if the point never converged (replace_nans!, GradeProEstimator.jl:1053-1061).
// The two terms of find_stability (DirectSpace.jl:1092-1100, first and
// second differences of the log-VMR over dlogt) for the value f at one rung
// and its two predecessors (NaN before the first rungs).
__device__ __forceinline__ void stability_terms(double f, double f1, double f2, double dl,
                                                double& d1, double& d2) {
  d1 = log(fabs(__ddiv_rn(__dsub_rn(f, f1), __dmul_rn(2.0, dl))));
  d2 = log(fabs(__ddiv_rn(__dadd_rn(__dsub_rn(f, __dmul_rn(2.0, f1)), f2), __dmul_rn(dl, dl))));
}

// One rung of the per-point state machine (find_stability and the state
// update, DirectSpace.jl:975-1031, GradeProEstimator.jl:617-636).
struct StabilityState {
  double cur_min = NAN;
  unsigned cl = 0, co = 0;
  bool low = false;
  int ch;
};

__device__ __forceinline__ void stability_step(StabilityState& st, double d1, double d2, double eps,
                                               int steps_low, int steps_over, int r) {
  bool more_stable = false;
  double new_min = NAN;
  if (!st.low) {
    if (d2 < eps) {
      if (((d1 < st.cur_min) || isnan(st.cur_min)) && ((int)st.co <= steps_over)) {
        more_stable = true;
        new_min = d1;
        st.co = 0;
      } else {
        st.co = (st.co + 1) & 0xFFFFu;
      }
    } else {
      st.cl = (st.cl + 1) & 0xFFFFu;
      if ((int)st.cl > steps_low) {
        st.low = true;
        st.cur_min = NAN;
      }
    }
  } else {
    if ((d2 < eps) && isnan(st.cur_min)) {
      more_stable = true;
      new_min = d1;
    }
  }
  if (isnan(new_min)) new_min = st.cur_min;
  st.cur_min = new_min;
  if (more_stable) st.ch = r;
}

__global__ __launch_bounds__(kThreads) void k_convergence(int64_t G, int S, const double* __restrict__ vmr,
                                                          const double* __restrict__ dlogt,
                                                          const double* __restrict__ params,
                                                          int16_t* __restrict__ chosen,
                                                          int16_t* __restrict__ sel) {
  const double eps = params[0];
  const int steps_low = (int)params[1];
  const int steps_over = (int)params[2];
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < G;
       p += (int64_t)gridDim.x * blockDim.x) {
    double f1 = NAN, f2 = NAN;
    StabilityState st;
    st.ch = S;
    auto step = [&](const double f, const int r) {
      double d1, d2;
      stability_terms(f, f1, f2, dlogt[r], d1, d2);
      stability_step(st, d1, d2, eps, steps_low, steps_over, r);
      f2 = f1;
      f1 = f;
    };
    // the table is read RB rungs ahead of the state machine, so a thread keeps
    // RB loads in flight (one thread per point leaves few waves per SIMD)
    constexpr int RB = 16;
    double nxt[RB];
    auto fetch = [&](int r0) {
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        const int r = r0 + i < S ? r0 + i : S - 1;
        nxt[i] = vmr[(int64_t)r * G + p];
      }
    };
    fetch(0);
    for (int r0 = 0; r0 < S; r0 += RB) {
      double cur[RB];
#pragma unroll
      for (int i = 0; i < RB; ++i) cur[i] = nxt[i];
      if (r0 + RB < S) fetch(r0 + RB);
#pragma unroll
      for (int i = 0; i < RB; ++i)
        if (r0 + i < S) step(cur[i], r0 + i);
    }
    chosen[p] = (int16_t)st.ch;
    sel[p] = (int16_t)(st.ch == S ? S - 1 : st.ch);
  }
}

// Small grids (a few waves in all): the logarithms and divisions of every
// (rung, point) are computed in parallel first (k_conv_terms), so the serial
// walk over the rungs (k_convergence_walk) only compares.  Same expressions,
// same results as k_convergence.
__global__ __launch_bounds__(kThreads) void k_conv_terms(int64_t G, int S, const double* __restrict__ vmr,
                                                         const double* __restrict__ dlogt,
                                                         double* __restrict__ terms) {
  const int64_t total = (int64_t)S * G;
  for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(id / G);
    const double f = vmr[id];
    const double f1 = r >= 1 ? vmr[id - G] : (double)NAN;
    const double f2 = r >= 2 ? vmr[id - 2 * G] : (double)NAN;
    double d1, d2;
    stability_terms(f, f1, f2, dlogt[r], d1, d2);
    terms[2 * id] = d1;
    terms[2 * id + 1] = d2;
  }
}

__global__ __launch_bounds__(64) void k_convergence_walk(int64_t G, int S, const double* __restrict__ terms,
                                                         const double* __restrict__ params,
                                                         int16_t* __restrict__ chosen,
                                                         int16_t* __restrict__ sel) {
  const double eps = params[0];
  const int steps_low = (int)params[1];
  const int steps_over = (int)params[2];
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= G) return;
  StabilityState st;
  st.ch = S;
  constexpr int RB = 16;
  double2 nxt[RB];
  const double2* t2 = (const double2*)terms;
  auto fetch = [&](int r0) {
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int r = r0 + i < S ? r0 + i : S - 1;
      nxt[i] = t2[(int64_t)r * G + p];
    }
  };
  fetch(0);
  for (int r0 = 0; r0 < S; r0 += RB) {
    double2 cur[RB];
#pragma unroll
    for (int i = 0; i < RB; ++i) cur[i] = nxt[i];
    if (r0 + RB < S) fetch(r0 + RB);
#pragma unroll
    for (int i = 0; i < RB; ++i)
      if (r0 + i < S) stability_step(st, cur[i].x, cur[i].y, eps, steps_low, steps_over, r0 + i);
  }
  chosen[p] = (int16_t)st.ch;
  sel[p] = (int16_t)(st.ch == S ? S - 1 : st.ch);
}

void launch_convergence(int64_t G, int S, const double* vmr, const double* dlogt,
                        const double* params, int16_t* chosen, int16_t* sel, double* terms,
                        hipStream_t s) {
  if (terms) {
    k_conv_terms<<<grid_for((int64_t)S * G, kThreads, 1 << 16), kThreads, 0, s>>>(G, S, vmr, dlogt,
                                                                                 terms);
    k_convergence_walk<<<(unsigned)((G + 63) / 64), 64, 0, s>>>(G, S, terms, params, chosen, sel);
    return;
  }
  k_convergence<<<grid_for(G, kThreads, 1 << 16), kThreads, 0, s>>>(G, S, vmr, dlogt, params, chosen,
                                                                     sel);
}

// ------------------------------------------------------------ means select
// Full-sample means (propagate_means!, ifft_means!, calculate_full_means!,
// GradeProEstimator.jl:388-511) for two rungs per complex transform; each
// point keeps the value of its selected rung.
template <typename T>
struct MeansArgs {
  const cplx<T>* src;
  int64_t src_rs;
  int R, r0, S;
  const T* gM;
  int ksum;
  const cplx<T>* tw;
  const int16_t* sel;
  double* density;
  double invG;
  int64_t Q;
  int n1, log2n1, H1, TQ, ntiles;
  FftFactors fac;  // axis-1 plan and twiddle table (GEN kernels)
  const cplx<T>* twg;
};

template <typename T, bool GEN>
__global__ __launch_bounds__(kThreads) void k_means_select(MeansArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* Z = (cplx<T>*)smem;
  const int rp = blockIdx.x / a.ntiles;
  const int tile = blockIdx.x - rp * a.ntiles;
  const int ra = a.r0 + 2 * rp;
  const int rb = ra + 1;
  const bool has_b = (2 * rp + 1) < a.R;
  const int64_t q0 = (int64_t)tile * a.TQ;
  const int n1 = a.n1, H1 = a.H1, TQ = a.TQ;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const cplx<T>* sa = a.src + (int64_t)(2 * rp) * a.src_rs;
  const cplx<T>* sb = a.src + (int64_t)(2 * rp + 1) * a.src_rs;
  const T* ga = a.gM + (int64_t)ra * a.ksum;
  const T* gb = a.gM + (int64_t)rb * a.ksum;
  for (int id = tid; id < H1 * TQ; id += nthr) {
    const int t = id % TQ, k1 = id / TQ;
    if (q0 + t >= a.Q) continue;
    const int64_t o = (int64_t)k1 * a.Q + q0 + t;
    const cplx<T> A = cscale(sa[o], ga[k1]);
    cplx<T> Bv = {(T)0, (T)0};
    if (has_b) Bv = cscale(sb[o], gb[k1]);
    put_pair(Z + (int64_t)t * n1, n1, k1, A, Bv);
  }
  __syncthreads();
  if constexpr (GEN)
    lds_fft_gen<1, T>(Z, a.fac, TQ, n1, a.twg, tid, nthr);
  else
    lds_fft<1, T>(Z, a.log2n1, TQ, n1, a.tw + (n1 - 2), tid, nthr);
  for (int w = tid; w < TQ * n1; w += nthr) {
    const int t = w / n1, j1 = w - t * n1;
    if (q0 + t >= a.Q) continue;
    const int64_t p = (q0 + t) * n1 + j1;
    const int sp = a.sel[p];
    const cplx<T> z = Z[(int64_t)t * n1 + j1];
    if (sp == ra)
      a.density[p] = fabs((double)z.re) * a.invG;
    else if (has_b && sp == rb)
      a.density[p] = fabs((double)z.im) * a.invG;
  }
}

template <typename T>
void launch_means_select(const GridDev& g, const cplx<T>* src, int64_t src_rung_stride, int R,
                         int r0, int S, const T* gM, const T* tw, const T* twg,
                         const int16_t* sel, double* density, hipStream_t s) {
  MeansArgs<T> a;
  const int n1 = g.n[0];
  int TQ = 2048 / n1;
  if (TQ < 1) TQ = 1;
  if (TQ > g.Q) TQ = (int)g.Q;
  a.src = src;
  a.src_rs = src_rung_stride;
  a.R = R;
  a.r0 = r0;
  a.S = S;
  a.gM = gM;
  a.ksum = g.ksum;
  a.tw = (const cplx<T>*)tw;
  a.sel = sel;
  a.density = density;
  a.invG = 1.0 / (double)g.G;
  a.Q = g.Q;
  a.n1 = n1;
  a.log2n1 = g.log2n[0];
  a.H1 = g.H1;
  a.TQ = TQ;
  a.ntiles = (int)((g.Q + TQ - 1) / TQ);
  a.fac = g.fac[0];
  a.twg = twg ? (const cplx<T>*)twg + g.koff[0] : nullptr;
  const unsigned blocks = (unsigned)(((R + 1) / 2) * (int64_t)a.ntiles);
  const size_t lds = (size_t)TQ * n1 * sizeof(cplx<T>);
  if ((n1 & (n1 - 1)) != 0)
    k_means_select<T, true><<<blocks, kThreads, lds, s>>>(a);
  else
    k_means_select<T, false><<<blocks, kThreads, lds, s>>>(a);
}

template void launch_means_select<float>(const GridDev&, const cplx<float>*, int64_t, int, int, int,
                                         const float*, const float*, const float*, const int16_t*,
                                         double*, hipStream_t);
template void launch_means_select<double>(const GridDev&, const cplx<double>*, int64_t, int, int, int,
                                          const double*, const double*, const double*,
                                          const int16_t*, double*, hipStream_t);

// ---------------------------------------------------------------- Silverman
// Two-pass sample moments with a fixed reduction tree (deterministic).
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (tid < o) red[tid] += red[tid + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kThreads) void k_moment_partial(int N, const double* __restrict__ x, int64_t n,
                                                             const double* __restrict__ mean,
                                                             double* __restrict__ partial) {
  __shared__ double red[kThreads];
  for (int i = 0; i < N; ++i) {
    double acc = 0.0;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
      const double v = x[k * N + i];
      if (mean) {
        const double d = v - mean[i];
        acc += d * d;
      } else {
        acc += v;
      }
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x * N + i] = s;
  }
}

__global__ __launch_bounds__(kThreads) void k_moment_final(int N, int nblocks, int64_t n, double prefactor,
                                                           const double* __restrict__ partial,
                                                           double* mean_out, double* bw_out) {
  __shared__ double red[kThreads];
  for (int i = 0; i < N; ++i) {
    double acc = 0.0;
    for (int k = threadIdx.x; k < nblocks; k += blockDim.x) acc += partial[k * N + i];
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) {
      if (bw_out)
        bw_out[i] = prefactor * sqrt(s / (double)(n - 1));
      else
        mean_out[i] = s / (double)n;
    }
  }
}

void launch_silverman(int N, const double* x, int64_t n, double prefactor, double* partial,
                      double* mean_out, double* bw_out, hipStream_t s) {
  k_moment_partial<<<kMomentBlocks, kThreads, 0, s>>>(N, x, n, nullptr, partial);
  k_moment_final<<<1, kThreads, 0, s>>>(N, kMomentBlocks, n, prefactor, partial, mean_out, nullptr);
  k_moment_partial<<<kMomentBlocks, kThreads, 0, s>>>(N, x, n, mean_out, partial);
  k_moment_final<<<1, kThreads, 0, s>>>(N, kMomentBlocks, n, prefactor, partial, nullptr, bw_out);
}

// Per-axis extrema of the samples (find_grid, Grids.jl:271-273: extrema of
// each row of the ndims x n data); a NaN coordinate makes both extrema NaN,
// as Julia's extrema does.  partial: [block][axis][min, max].
__device__ __forceinline__ double nan_min(double a, double b) { return (b < a || b != b) ? b : a; }
__device__ __forceinline__ double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }

__global__ __launch_bounds__(kThreads) void k_extrema_partial(int N, const double* __restrict__ x,
                                                              int64_t n, double* __restrict__ partial) {
  __shared__ double lo_s[kThreads], hi_s[kThreads];
  for (int i = 0; i < N; ++i) {
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
      const double v = x[k * N + i];
      lo = nan_min(lo, v);
      hi = nan_max(hi, v);
    }
    lo_s[threadIdx.x] = lo;
    hi_s[threadIdx.x] = hi;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) {
        lo_s[threadIdx.x] = nan_min(lo_s[threadIdx.x], lo_s[threadIdx.x + w]);
        hi_s[threadIdx.x] = nan_max(hi_s[threadIdx.x], hi_s[threadIdx.x + w]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      partial[(blockIdx.x * N + i) * 2] = lo_s[0];
      partial[(blockIdx.x * N + i) * 2 + 1] = hi_s[0];
    }
    __syncthreads();
  }
}

__global__ void k_extrema_final(int N, int nblocks, const double* __restrict__ partial,
                                double* __restrict__ out) {
  const int i = threadIdx.x;
  if (i >= N) return;
  double lo = INFINITY, hi = -INFINITY;
  for (int b = 0; b < nblocks; ++b) {
    lo = nan_min(lo, partial[(b * N + i) * 2]);
    hi = nan_max(hi, partial[(b * N + i) * 2 + 1]);
  }
  out[2 * i] = lo;
  out[2 * i + 1] = hi;
}

void launch_extrema(int N, const double* x, int64_t n, double* partial, double* out, hipStream_t s) {
  k_extrema_partial<<<kMomentBlocks, kThreads, 0, s>>>(N, x, n, partial);
  k_extrema_final<<<1, 64, 0, s>>>(N, kMomentBlocks, partial, out);
}

// Rule-of-thumb propagation time of the :rot estimator (RoTEstimator.jl:
// 22-28 Silverman, 42-47 Scott, 67-70): t_i = factor * sigma_i - h_i / 2,
// factor = (n (N + 2) / 4)^(-1 / (N + 4)) (Silverman) or n^(-1 / (N + 4))
// (Scott); sigma_i is the sample standard deviation (n - 1 normalisation).
__global__ void k_rot_times(GridDev g, int64_t n, int rule, const double* __restrict__ sigma,
                            double* __restrict__ times) {
  if (threadIdx.x != 0) return;
  const int N = g.N;
  const double ex = __ddiv_rn(-1.0, (double)(N + 4));
  const double base = rule == 1 ? __ddiv_rn((double)(n * (int64_t)(N + 2)), 4.0) : (double)n;
  const double factor = pow(base, ex);
  for (int i = 0; i < N; ++i)
    times[i] = __dsub_rn(__dmul_rn(factor, sigma[i]), __ddiv_rn(g.h[i], 2.0));
}

void launch_rot_times(const GridDev& g, int64_t n, int rule, const double* sigma, double* times,
                      hipStream_t s) {
  k_rot_times<<<1, 64, 0, s>>>(g, n, rule, sigma, times);
}

template <typename T>
void launch_gauss_tables(const GridDev& g, int S, const double* times, const double* kfreq, T* gM,
                         T* gV, hipStream_t s);

// ------------------------------------------------------------------- tables
// Ladder scalars per rung (GradeProEstimator.jl:860-905, 954-966, 1074-1089;
// DirectSpace.jl:564; FourierSpace.jl:224-225), with explicit rounding
// intrinsics so the FP64 values follow the reference's operation order.
__global__ void k_ladder_scalars(GridDev g, int S, int mode_steps, double steps_or_step,
                                 const double* __restrict__ bw, int64_t n_samples, double alpha_s,
                                 double alpha_os, double eps, double* times, double* dlogt,
                                 double* vscale, double* cV, double* params) {
  const int N = g.N;
  double dt[kMaxDims], t0[kMaxDims];
  for (int i = 0; i < N; ++i) {
    dt[i] = mode_steps ? __ddiv_rn(bw[i], steps_or_step) : steps_or_step;
    t0[i] = __ddiv_rn(g.h[i], 2.0);
  }
  const double n3 = (double)(n_samples * n_samples * n_samples);
  for (int r = threadIdx.x; r < S; r += blockDim.x) {
    double t[kMaxDims], tp[kMaxDims];
    for (int i = 0; i < N; ++i) {
      t[i] = __dmul_rn(dt[i], (double)r);
      tp[i] = __dmul_rn(dt[i], (double)(r - 1));
      times[r * N + i] = t[i];
    }
    double det_prev = 1.0, det_curr = 1.0, prodt = 1.0, det_t0 = 1.0;
    for (int i = 0; i < N; ++i) {
      const double t0s = __dmul_rn(t0[i], t0[i]);
      const double p = (r == 0) ? t0s : __dadd_rn(t0s, __dmul_rn(tp[i], tp[i]));
      const double c = __dadd_rn(t0s, __dmul_rn(t[i], t[i]));
      const double q = __dadd_rn(__dmul_rn(t[i], t[i]), t0s);
      det_prev = (i == 0) ? p : __dmul_rn(det_prev, p);
      det_curr = (i == 0) ? c : __dmul_rn(det_curr, c);
      prodt = (i == 0) ? q : __dmul_rn(prodt, q);
      det_t0 = (i == 0) ? t0[i] : __dmul_rn(det_t0, t0[i]);
    }
    dlogt[r] = log(__ddiv_rn(det_curr, det_prev));
    vscale[r] = __dmul_rn(pow(prodt, 1.5), n3);
    cV[r] = __ddiv_rn(det_t0, sqrt(det_curr));
  }
  if (threadIdx.x == 0) {
    // calculate_duration_steps on the last rung (ties-to-even rounding)
    double sb = 0.0, so = 0.0;
    for (int i = 0; i < N; ++i) {
      const double tend = __dmul_rn(dt[i], (double)(S - 1));
      const double ns = __ddiv_rn(tend, dt[i]);
      const double b = __dmul_rn(alpha_s, ns), o = __dmul_rn(alpha_os, ns);
      sb = (i == 0) ? b : __dadd_rn(sb, b);
      so = (i == 0) ? o : __dadd_rn(so, o);
    }
    params[0] = eps;
    params[1] = rint(__ddiv_rn(sb, (double)N));
    params[2] = rint(__ddiv_rn(so, (double)N));
  }
}

// Per-axis Gaussian factors: gM = exp(-(k t)^2 / 2), gV = sqrt(gM) so that
// prod_i gM_i = exp(-sum_i (k_i t_i)^2 / 2) (propagate_time_cpu!).
template <typename T>
__global__ void k_gauss_tables(GridDev g, int S, const double* __restrict__ times,
                               const double* __restrict__ kfreq, T* gM, T* gV) {
  const int64_t total = (int64_t)S * g.ksum;
  for (int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(id / g.ksum);
    const int kk = (int)(id - (int64_t)r * g.ksum);
    int ax = 0;
    while (ax + 1 < g.N && kk >= g.koff[ax + 1]) ++ax;
    const double a = __dmul_rn(kfreq[kk], times[r * g.N + ax]);
    const double e = exp(__dmul_rn(-0.5, __dmul_rn(a, a)));
    gM[id] = (T)e;
    gV[id] = (T)sqrt(e);
  }
}

template <typename T>
void launch_tables(const GridDev& g, int S, int n_steps_mode, double n_steps_or_bwstep,
                   const double* bw, const double* kfreq, int64_t n_samples, double alpha_s,
                   double alpha_os, double eps, double* times, double* dlogt, double* vscale,
                   double* cV, double* params, T* gM, T* gV, hipStream_t s) {
  k_ladder_scalars<<<1, kThreads, 0, s>>>(g, S, n_steps_mode, n_steps_or_bwstep, bw, n_samples,
                                          alpha_s, alpha_os, eps, times, dlogt, vscale, cV, params);
  const int64_t total = (int64_t)S * g.ksum;
  k_gauss_tables<T><<<grid_for(total), kThreads, 0, s>>>(g, S, times, kfreq, gM, gV);
}

template <typename T>
void launch_gauss_tables(const GridDev& g, int S, const double* times, const double* kfreq, T* gM,
                         T* gV, hipStream_t s) {
  const int64_t total = (int64_t)S * g.ksum;
  k_gauss_tables<T><<<grid_for(total), kThreads, 0, s>>>(g, S, times, kfreq, gM, gV);
}
template void launch_gauss_tables<float>(const GridDev&, int, const double*, const double*, float*,
                                         float*, hipStream_t);
template void launch_gauss_tables<double>(const GridDev&, int, const double*, const double*,
                                          double*, double*, hipStream_t);

template void launch_tables<float>(const GridDev&, int, int, double, const double*, const double*,
                                   int64_t, double, double, double, double*, double*, double*,
                                   double*, double*, float*, float*, hipStream_t);
template void launch_tables<double>(const GridDev&, int, int, double, const double*, const double*,
                                    int64_t, double, double, double, double*, double*, double*,
                                    double*, double*, double*, double*, hipStream_t);

}  // namespace pkde
