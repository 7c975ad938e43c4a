// Device-side building blocks for the :gradepro HIP path on gfx950:
// complex arithmetic, in-register radix-2/4/8/16 DFTs and a batched Stockham
// FFT whose passes exchange data through LDS.  Pencils of power-of-two
// length L <= 4096 live in LDS; every thread owns E = min(L, 16) elements per
// pass, so a pass is one LDS read, the radix-R butterflies in registers, and
// one LDS write.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pkde {

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return {a.re, -a.im}; }

// cos / sin of 2*pi*k/16, k = 0..15 (exact decimal expansions of the FP64 values)
__device__ constexpr double kCos16[16] = {
    1.0, 0.92387953251128674, 0.70710678118654757, 0.38268343236508978,
    0.0, -0.38268343236508978, -0.70710678118654757, -0.92387953251128674,
    -1.0, -0.92387953251128674, -0.70710678118654757, -0.38268343236508978,
    0.0, 0.38268343236508978, 0.70710678118654757, 0.92387953251128674};
__device__ constexpr double kSin16[16] = {
    0.0, 0.38268343236508978, 0.70710678118654757, 0.92387953251128674,
    1.0, 0.92387953251128674, 0.70710678118654757, 0.38268343236508978,
    0.0, -0.38268343236508978, -0.70710678118654757, -0.92387953251128674,
    -1.0, -0.92387953251128674, -0.70710678118654757, -0.38268343236508978};

// x * exp(DIR * 2 pi i k / R) for compile-time R <= 16; k resolves to a
// constant after unrolling, so the 0 and quarter-turn cases cost nothing.
template <int R, int DIR, typename T>
__device__ __forceinline__ cplx<T> twiddle_const(cplx<T> x, int k) {
  const int j = (k * (16 / R)) & 15;
  if (j == 0) return x;
  if (j == 4) return DIR > 0 ? cplx<T>{-x.im, x.re} : cplx<T>{x.im, -x.re};
  if (j == 8) return {-x.re, -x.im};
  if (j == 12) return DIR > 0 ? cplx<T>{x.im, -x.re} : cplx<T>{-x.im, x.re};
  const T c = (T)kCos16[j];
  const T s = (T)(DIR * kSin16[j]);
  return {x.re * c - x.im * s, x.re * s + x.im * c};
}

// In-register DFT of R = 2^m points, v[r] -> V[s] = sum_r v[r] exp(DIR 2pi i r s / R).
template <int R, int DIR, typename T>
__device__ __forceinline__ void dft_reg(cplx<T>* v) {
  if constexpr (R == 1) {
    return;
  } else if constexpr (R == 2) {
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else {
    constexpr int H = R / 2;
    cplx<T> e[H], o[H];
#pragma unroll
    for (int k = 0; k < H; ++k) {
      e[k] = v[2 * k];
      o[k] = v[2 * k + 1];
    }
    dft_reg<H, DIR, T>(e);
    dft_reg<H, DIR, T>(o);
#pragma unroll
    for (int k = 0; k < H; ++k) {
      cplx<T> t = twiddle_const<R, DIR, T>(o[k], k);
      v[k] = cadd(e[k], t);
      v[k + H] = csub(e[k], t);
    }
  }
}

// Radix plan for L = 2^log2L: passes of 16 and one trailing 8/4/2.
struct FftPlanDev {
  int L, log2L, E, tpp;  // E: elements per thread, tpp: threads per pencil
};

__device__ __forceinline__ FftPlanDev make_plan(int log2L) {
  FftPlanDev p;
  p.log2L = log2L;
  p.L = 1 << log2L;
  p.E = p.L < 16 ? p.L : 16;
  p.tpp = p.L / p.E;
  return p;
}

// One Stockham pass of radix R over `npen` pencils of length L stored at
// buf[p * pitch + e].  Each work item is a (pencil, thread-in-pencil) pair that
// performs E/R butterflies; items are processed in rounds of nthr so every
// thread meets the same barriers.
// tw[k] = exp(-2 pi i k / L), k in [0, L) (FP64-accurate, rounded to T).
template <int R, int E, int DIR, typename T>
__device__ __forceinline__ void stockham_pass(cplx<T>* buf, int L, int npen, int pitch, int Ns,
                                              const cplx<T>* __restrict__ tw, int tid, int nthr) {
  const int tpp = L / E;
  const int items = npen * tpp;
  const int nbfly = L / R;  // butterflies per pencil
  const int twstep = L / (Ns * R);
  for (int base = 0; base < items; base += nthr) {
    const int item = base + tid;
    const bool active = item < items;
    cplx<T> v[E];
    int pen = 0, t = 0;
    if (active) {
      pen = item / tpp;
      t = item - pen * tpp;
      const cplx<T>* pb = buf + pen * pitch;
#pragma unroll
      for (int u = 0; u < E / R; ++u) {
        const int j = t + u * tpp;
#pragma unroll
        for (int r = 0; r < R; ++r) v[u * R + r] = pb[j + r * nbfly];
      }
    }
    __syncthreads();
    if (active) {
      cplx<T>* pb = buf + pen * pitch;
#pragma unroll
      for (int u = 0; u < E / R; ++u) {
        const int j = t + u * tpp;
        const int jm = j % Ns;
        cplx<T>* w = v + u * R;
        if (jm != 0) {
#pragma unroll
          for (int r = 1; r < R; ++r) {
            cplx<T> f = tw[(jm * r * twstep) & (L - 1)];
            if (DIR > 0) f.im = -f.im;
            w[r] = cmul(w[r], f);
          }
        }
        dft_reg<R, DIR, T>(w);
        const int out0 = (j / Ns) * Ns * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) pb[out0 + r * Ns] = w[r];
      }
    }
    __syncthreads();
  }
}

// Batched in-LDS FFT: npen pencils of length L = 2^log2L (L <= 4096) at
// buf[p * pitch + e], transformed in place (natural order in and out).
// Unnormalised; DIR = -1 forward, +1 inverse.  Must be called by the whole
// workgroup (contains barriers).
template <int DIR, typename T>
__device__ void lds_fft(cplx<T>* buf, int log2L, int npen, int pitch,
                        const cplx<T>* __restrict__ tw, int tid, int nthr) {
  const int L = 1 << log2L;
  if (L == 1) return;
  int Ns = 1;
  int rem = log2L;
  if (L >= 16) {
    while (rem >= 4) {
      stockham_pass<16, 16, DIR, T>(buf, L, npen, pitch, Ns, tw, tid, nthr);
      Ns *= 16;
      rem -= 4;
    }
    if (rem == 3) stockham_pass<8, 16, DIR, T>(buf, L, npen, pitch, Ns, tw, tid, nthr);
    else if (rem == 2) stockham_pass<4, 16, DIR, T>(buf, L, npen, pitch, Ns, tw, tid, nthr);
    else if (rem == 1) stockham_pass<2, 16, DIR, T>(buf, L, npen, pitch, Ns, tw, tid, nthr);
  } else if (L == 8) {
    stockham_pass<8, 8, DIR, T>(buf, L, npen, pitch, Ns, tw, tid, nthr);
  } else if (L == 4) {
    stockham_pass<4, 4, DIR, T>(buf, L, npen, pitch, Ns, tw, tid, nthr);
  } else if (L == 2) {
    stockham_pass<2, 2, DIR, T>(buf, L, npen, pitch, Ns, tw, tid, nthr);
  }
}

// ------------------------------------------------------------ mixed radix
// Pencil length L = prod(rad[0..np)) with radices 2, 3, 4, 5, 7, 8, 11, 13,
// 16 or any larger prime; tw[k] = exp(-2 pi i k / L), k < L.
struct FftFactors {
  int L, np;
  int rad[16];
};

// v[s] = sum_r v[r] exp(DIR 2 pi i r s / R) for an odd prime R, with the
// roots taken from the length-L table: exp(-2 pi i m / R) = tw[m * (L / R)].
template <int R, int DIR, typename T>
__device__ __forceinline__ void dft_prime(cplx<T>* v, const cplx<T>* __restrict__ tw, int LoverR) {
  cplx<T> out[R];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    cplx<T> acc = v[0];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      cplx<T> f = tw[((r * s) % R) * LoverR];
      if (DIR > 0) f.im = -f.im;
      acc = cadd(acc, cmul(v[r], f));
    }
    out[s] = acc;
  }
#pragma unroll
  for (int s = 0; s < R; ++s) v[s] = out[s];
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void dft_any(cplx<T>* v, const cplx<T>* __restrict__ tw, int L) {
  if constexpr (R == 2 || R == 4 || R == 8 || R == 16)
    dft_reg<R, DIR, T>(v);
  else
    dft_prime<R, DIR, T>(v, tw, L / R);
}

// One mixed-radix Stockham pass over npen pencils of length L at
// buf[p * pitch + e]: npen * L / R butterflies, at most BPT per thread.  All
// inputs are read before the barrier, so the pass is done in place; the
// caller keeps npen * L <= BPT * R * nthr (npen * L <= 4096 with 256 threads).
template <int R, int BPT, int DIR, typename T>
__device__ __forceinline__ void stockham_pass_gen(cplx<T>* buf, int L, int npen, int pitch, int Ns,
                                                  const cplx<T>* __restrict__ tw, int tid,
                                                  int nthr) {
  const int nb = L / R;
  const int items = npen * nb;
  const int twstep = L / (Ns * R);
  cplx<T> v[BPT][R];
#pragma unroll
  for (int k = 0; k < BPT; ++k) {
    const int it = tid + k * nthr;
    if (it < items) {
      const int pen = it / nb, j = it - pen * nb;
      const cplx<T>* pb = buf + (int64_t)pen * pitch;
#pragma unroll
      for (int r = 0; r < R; ++r) v[k][r] = pb[j + r * nb];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < BPT; ++k) {
    const int it = tid + k * nthr;
    if (it < items) {
      const int pen = it / nb, j = it - pen * nb;
      const int jm = j % Ns;
      cplx<T>* w = v[k];
      if (jm != 0) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cplx<T> f = tw[jm * r * twstep];
          if (DIR > 0) f.im = -f.im;
          w[r] = cmul(w[r], f);
        }
      }
      dft_any<R, DIR, T>(w, tw, L);
      cplx<T>* pb = buf + (int64_t)pen * pitch;
      const int out0 = (j / Ns) * Ns * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) pb[out0 + r * Ns] = w[r];
    }
  }
  __syncthreads();
}

// Stockham pass of a prime radix R above 13, known only at run time: each
// output is formed directly, y = sum_r x_r exp(DIR 2 pi i r (jm + s Ns) /
// (Ns R)) over the butterfly's R inputs (the pass twiddle and the DFT root
// folded into one table entry, index r (jm + s Ns) L / (Ns R) mod L).  A
// thread forms at most 16 outputs (npen * L <= 4096 with 256 threads), all
// of them before the barrier, so the pass is in place.  O(L R) work per
// pencil: the axis lengths that need it are small.
template <int DIR, typename T>
__device__ __forceinline__ void stockham_pass_any(cplx<T>* buf, int L, int npen, int pitch, int Ns,
                                                  int R, const cplx<T>* __restrict__ tw, int tid,
                                                  int nthr) {
  constexpr int OPT = 16;
  const int nb = L / R;
  const int span = Ns * R;
  const int twstep = L / span;
  const int total = npen * L;
  cplx<T> y[OPT];
  int dst[OPT];
#pragma unroll
  for (int k = 0; k < OPT; ++k) {
    const int o = tid + k * nthr;
    dst[k] = -1;
    if (o < total) {
      const int pen = o / L, po = o - pen * L;
      const int q = po / span, rem = po - q * span;
      const int s = rem / Ns, jm = rem - s * Ns;
      const int j = q * Ns + jm;
      const cplx<T>* pb = buf + (int64_t)pen * pitch;
      const int64_t dk = (int64_t)(jm + s * Ns) * twstep;  // < L
      cplx<T> acc = pb[j];
      int64_t e = 0;
      for (int r = 1; r < R; ++r) {
        e += dk;
        if (e >= L) e -= L;
        cplx<T> f = tw[e];
        if (DIR > 0) f.im = -f.im;
        acc = cadd(acc, cmul(pb[j + r * nb], f));
      }
      y[k] = acc;
      dst[k] = pen * pitch + po;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < OPT; ++k)
    if (dst[k] >= 0) buf[dst[k]] = y[k];
  __syncthreads();
}

// Batched in-LDS FFT of any length f.L <= 4096: natural order in and out,
// unnormalised; DIR = -1 forward, +1 inverse.  Radices 2-16 have fixed
// butterflies; a prime factor above 13 takes the direct pass above.  Pencils
// are done in groups of at most 4096 points so every pass fits one round of
// 256 threads.  Must be called by the whole workgroup of 256 threads.
template <int DIR, typename T>
__device__ void lds_fft_gen(cplx<T>* buf, const FftFactors& f, int npen, int pitch,
                            const cplx<T>* __restrict__ tw, int tid, int nthr) {
  const int L = f.L;
  const int pg = L <= 4096 ? 4096 / L : 1;
  for (int p0 = 0; p0 < npen; p0 += pg) {
    cplx<T>* b = buf + (int64_t)p0 * pitch;
    const int cnt = npen - p0 < pg ? npen - p0 : pg;
    int Ns = 1;
    for (int p = 0; p < f.np; ++p) {
      switch (f.rad[p]) {
        case 2: stockham_pass_gen<2, 8, DIR, T>(b, L, cnt, pitch, Ns, tw, tid, nthr); break;
        case 3: stockham_pass_gen<3, 6, DIR, T>(b, L, cnt, pitch, Ns, tw, tid, nthr); break;
        case 4: stockham_pass_gen<4, 4, DIR, T>(b, L, cnt, pitch, Ns, tw, tid, nthr); break;
        case 5: stockham_pass_gen<5, 4, DIR, T>(b, L, cnt, pitch, Ns, tw, tid, nthr); break;
        case 7: stockham_pass_gen<7, 3, DIR, T>(b, L, cnt, pitch, Ns, tw, tid, nthr); break;
        case 8: stockham_pass_gen<8, 2, DIR, T>(b, L, cnt, pitch, Ns, tw, tid, nthr); break;
        case 11: stockham_pass_gen<11, 2, DIR, T>(b, L, cnt, pitch, Ns, tw, tid, nthr); break;
        case 13: stockham_pass_gen<13, 2, DIR, T>(b, L, cnt, pitch, Ns, tw, tid, nthr); break;
        case 16: stockham_pass_gen<16, 1, DIR, T>(b, L, cnt, pitch, Ns, tw, tid, nthr); break;
        default: stockham_pass_any<DIR, T>(b, L, cnt, pitch, Ns, f.rad[p], tw, tid, nthr); break;
      }
      Ns *= f.rad[p];
    }
  }
}

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon et al., SC'11): counter-based bootstrap draws.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = (uint32_t)p1;
    c[2] = n2;
    c[3] = (uint32_t)p0;
  }
}

}  // namespace pkde

namespace pkde {

// ------------------------------------------------------ wave-local passes
// When a pencil is owned by lanes of a single wavefront, LDS exchanges
// between passes need only wavefront-scope ordering: LDS operations of one
// wave execute in order, so a fence that stops the compiler from moving
// LDS accesses across the pass boundary is sufficient.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-local pencils in LDS use a padded layout: element p of a pencil sits
// at pad8(p) = p + p/8 (one spare slot after every 8 values), so the first
// Stockham pass, whose lanes write R values apart, spreads its stores over
// all banks.  Pencils of length L sit pencil_stride(L) values apart; the odd
// remainder modulo 8 keeps accesses that walk across pencils conflict-free.
__host__ __device__ constexpr int pad8(int p) { return p + (p >> 3); }
__host__ __device__ constexpr int pencil_stride(int L) { return L + L / 8 + 1; }

__host__ __device__ constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }

// Slot of element p of a wave-local pencil of E values per lane: SW = 0 the
// padded layout (pad8); SW = 1 an unpadded XOR swizzle, p ^ ((p / E) mod 8).
// LDS serves a wave's 16-byte reads 16 lanes at a time and its writes 8 at a
// time, and lanes of one group on the same bank conflict.  The padded layout
// keeps stride-8 writes conflict-free but puts one 2-way conflict in every
// read of 16 consecutive values; the swizzle permutes only the low 3 bits
// within each run of 8, so reads of 16 aligned consecutive values and the
// first pass's stride-E writes both fill distinct banks.
template <int SW, int LE>
__host__ __device__ constexpr int lds_slot(int p) {
  return SW ? (p ^ ((p >> LE) & 7)) : pad8(p);
}

// One Stockham pass of radix R over the lane's own pencil `pb` (length
// L = 2^LOG2L, padded layout), the lane being thread t of the pencil's L/E
// threads; Ns = 2^LOG2NS is the length of the sub-transforms already done.
// All sizes are compile-time, so index arithmetic reduces to shifts and masks.
// All lanes of the wave call it (active = false for lanes without a pencil).
template <int R, int E, int LOG2L, int LOG2NS, int DIR, typename T, int SW = 0>
__device__ __forceinline__ void wave_pass(cplx<T>* pb, int t, bool active,
                                          const cplx<T>* __restrict__ tw) {
  constexpr int LE = ilog2c(E);
  constexpr int L = 1 << LOG2L;
  constexpr int tpp = L / E;
  constexpr int nbfly = L / R;
  constexpr int Ns = 1 << LOG2NS;
  constexpr int LOG2R = ilog2c(R);
  constexpr int twstep = L / (Ns * R);
  cplx<T> v[E];
  if (active) {
#pragma unroll
    for (int u = 0; u < E / R; ++u) {
      const int j = t + u * tpp;
#pragma unroll
      for (int r = 0; r < R; ++r) v[u * R + r] = pb[lds_slot<SW, LE>(j + r * nbfly)];
    }
  }
  wave_sync();
  if (active) {
#pragma unroll
    for (int u = 0; u < E / R; ++u) {
      const int j = t + u * tpp;
      const int jm = j & (Ns - 1);
      cplx<T>* w = v + u * R;
      if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cplx<T> f = tw[jm * r * twstep];
          if (DIR > 0) f.im = -f.im;
          w[r] = cmul(w[r], f);
        }
      }
      dft_reg<R, DIR, T>(w);
      const int out0 = ((j >> LOG2NS) << (LOG2NS + LOG2R)) + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) pb[lds_slot<SW, LE>(out0 + r * Ns)] = w[r];
    }
  }
  wave_sync();
}

template <int LOG2L, int LOG2NS, int E, int DIR, typename T, int SW = 0>
__device__ __forceinline__ void wave_fft_from(cplx<T>* pb, int t, bool active,
                                              const cplx<T>* __restrict__ tw) {
  constexpr int LE = ilog2c(E);
  constexpr int rem = LOG2L - LOG2NS;
  if constexpr (rem >= LE) {
    wave_pass<E, E, LOG2L, LOG2NS, DIR, T, SW>(pb, t, active, tw);
    wave_fft_from<LOG2L, LOG2NS + LE, E, DIR, T, SW>(pb, t, active, tw);
  } else if constexpr (rem > 0) {
    wave_pass<(1 << rem), E, LOG2L, LOG2NS, DIR, T, SW>(pb, t, active, tw);
  }
}

// The first pass (radix E, Ns = 1) of a wave-local transform from registers:
// v[r] = x[t + r L/E] (the inputs wave_pass would read); the results go to
// the pencil in LDS as wave_pass stores them.  The caller continues with
// wave_fft_from<LOG2L, log2 E>.  Needs L > E.
template <int LOG2L, int E, int DIR, typename T>
__device__ __forceinline__ void wave_first_pass(cplx<T>* pb, int t, bool active, cplx<T>* v) {
  constexpr int LE = ilog2c(E);
  static_assert((1 << LOG2L) > E, "first pass from registers needs L > E");
  if (active) {
    dft_reg<E, DIR, T>(v);
#pragma unroll
    for (int r = 0; r < E; ++r) pb[pad8((t << LE) + r)] = v[r];
  }
  wave_sync();
}

// The last pass (Ns R = L) of a wave-local transform, its outputs kept in
// registers: out[u R + r] = X[j + r Ns] with j = t + u L/E.
template <int R, int E, int LOG2L, int LOG2NS, int DIR, typename T>
__device__ __forceinline__ void wave_last_pass_regs(const cplx<T>* pb, int t, bool active,
                                                    const cplx<T>* __restrict__ tw, cplx<T>* out) {
  constexpr int L = 1 << LOG2L;
  constexpr int tpp = L / E;
  constexpr int nbfly = L / R;
  constexpr int Ns = 1 << LOG2NS;
  static_assert(Ns * R == L, "not the last pass");
  if (active) {
#pragma unroll
    for (int u = 0; u < E / R; ++u) {
      const int j = t + u * tpp;
      cplx<T>* w = out + u * R;
#pragma unroll
      for (int r = 0; r < R; ++r) w[r] = pb[pad8(j + r * nbfly)];
      if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cplx<T> f = tw[j * r];
          if (DIR > 0) f.im = -f.im;
          w[r] = cmul(w[r], f);
        }
      }
      dft_reg<R, DIR, T>(w);
    }
  }
}

// Passes LOG2NS.. of a wave-local transform, the last one into registers
// (wave_last_pass_regs).  Needs LOG2L > LOG2NS.
template <int LOG2L, int LOG2NS, int E, int DIR, typename T>
__device__ __forceinline__ void wave_fft_to_regs(cplx<T>* pb, int t, bool active,
                                                 const cplx<T>* __restrict__ tw, cplx<T>* out) {
  constexpr int LE = ilog2c(E);
  constexpr int rem = LOG2L - LOG2NS;
  static_assert(rem > 0, "no pass left");
  if constexpr (rem > LE) {
    wave_pass<E, E, LOG2L, LOG2NS, DIR, T>(pb, t, active, tw);
    wave_fft_to_regs<LOG2L, LOG2NS + LE, E, DIR, T>(pb, t, active, tw, out);
  } else {
    wave_last_pass_regs<(1 << rem), E, LOG2L, LOG2NS, DIR, T>(pb, t, active, tw, out);
  }
}

// Wave-local FFT of the lane's pencil (padded layout): L = 2^LOG2L with
// E <= L <= 64 * E, so the L/E threads of a pencil sit in one wavefront.
// E = 16 gives radix-16 passes; E = 8 (radix-8 passes) halves the registers
// per lane, which the FP64 kernels need for occupancy.  Natural order in and
// out; `tw` is the table exp(-2 pi i k / L), k < L.
template <int LOG2L, int DIR, typename T, int E, int SW = 0>
__device__ __forceinline__ void wave_fft(cplx<T>* pb, int t, bool active,
                                         const cplx<T>* __restrict__ tw) {
  static_assert((1 << LOG2L) >= E && (1 << LOG2L) <= 64 * E, "pencil must fit one wavefront");
  wave_fft_from<LOG2L, 0, E, DIR, T, SW>(pb, t, active, tw);
}

// ------------------------------- wave-local passes, twiddles in registers
// The same passes as wave_fft_from / wave_fft_to_regs for a lane that keeps
// the twiddles of every pass after the first in registers (loaded once with
// wave_tw_load, outside a loop over many transforms of the same length), so
// that the passes read only data from LDS.  twc holds, pass by pass, for
// each butterfly u of the lane and r = 1..R-1, tw[jm r L/(Ns R)] (conjugated
// for the forward direction), exactly the factors the LDS-table passes use.
template <int LOG2L, int LOG2NS, int E>
__host__ __device__ constexpr int wave_tw_count() {
  constexpr int LE = ilog2c(E);
  constexpr int rem = LOG2L - LOG2NS;
  if constexpr (rem <= 0) {
    return 0;
  } else {
    constexpr int LR = rem >= LE ? LE : rem;
    constexpr int R = 1 << LR;
    return (E / R) * (R - 1) + wave_tw_count<LOG2L, LOG2NS + LR, E>();
  }
}

template <int LOG2L, int LOG2NS, int E, int DIR, typename T, int OFF, int NT>
__device__ __forceinline__ void wave_tw_load(int t, const cplx<T>* __restrict__ tw,
                                             cplx<T> (&twc)[NT]) {
  constexpr int LE = ilog2c(E);
  constexpr int rem = LOG2L - LOG2NS;
  if constexpr (rem > 0) {
    constexpr int LR = rem >= LE ? LE : rem;
    constexpr int R = 1 << LR;
    constexpr int L = 1 << LOG2L, tpp = L / E, Ns = 1 << LOG2NS;
    constexpr int twstep = L / (Ns * R);
#pragma unroll
    for (int u = 0; u < E / R; ++u) {
      const int jm = (t + u * tpp) & (Ns - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cplx<T> f = tw[jm * r * twstep];
        if (DIR > 0) f.im = -f.im;
        twc[OFF + u * (R - 1) + r - 1] = f;
      }
    }
    wave_tw_load<LOG2L, LOG2NS + LR, E, DIR, T, OFF + (E / R) * (R - 1), NT>(t, tw, twc);
  }
}

template <int R, int E, int LOG2L, int LOG2NS, int DIR, typename T, int OFF, int NT>
__device__ __forceinline__ void wave_pass_rt(cplx<T>* pb, int t, bool active,
                                             const cplx<T> (&twc)[NT]) {
  constexpr int L = 1 << LOG2L;
  constexpr int tpp = L / E;
  constexpr int nbfly = L / R;
  constexpr int Ns = 1 << LOG2NS;
  constexpr int LOG2R = ilog2c(R);
  static_assert(Ns > 1, "the first pass has no twiddles");
  cplx<T> v[E];
  if (active) {
#pragma unroll
    for (int u = 0; u < E / R; ++u) {
      const int j = t + u * tpp;
#pragma unroll
      for (int r = 0; r < R; ++r) v[u * R + r] = pb[pad8(j + r * nbfly)];
    }
  }
  wave_sync();
  if (active) {
#pragma unroll
    for (int u = 0; u < E / R; ++u) {
      const int j = t + u * tpp;
      const int jm = j & (Ns - 1);
      cplx<T>* w = v + u * R;
#pragma unroll
      for (int r = 1; r < R; ++r) w[r] = cmul(w[r], twc[OFF + u * (R - 1) + r - 1]);
      dft_reg<R, DIR, T>(w);
      const int out0 = ((j >> LOG2NS) << (LOG2NS + LOG2R)) + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) pb[pad8(out0 + r * Ns)] = w[r];
    }
  }
  wave_sync();
}

template <int LOG2L, int LOG2NS, int E, int DIR, typename T, int OFF, int NT>
__device__ __forceinline__ void wave_fft_from_rt(cplx<T>* pb, int t, bool active,
                                                 const cplx<T> (&twc)[NT]) {
  constexpr int LE = ilog2c(E);
  constexpr int rem = LOG2L - LOG2NS;
  if constexpr (rem > 0) {
    constexpr int LR = rem >= LE ? LE : rem;
    constexpr int R = 1 << LR;
    wave_pass_rt<R, E, LOG2L, LOG2NS, DIR, T, OFF, NT>(pb, t, active, twc);
    wave_fft_from_rt<LOG2L, LOG2NS + LR, E, DIR, T, OFF + (E / R) * (R - 1), NT>(pb, t, active,
                                                                                 twc);
  }
}

template <int LOG2L, int LOG2NS, int E, int DIR, typename T, int OFF, int NT>
__device__ __forceinline__ void wave_fft_to_regs_rt(cplx<T>* pb, int t, bool active,
                                                    const cplx<T> (&twc)[NT], cplx<T>* out) {
  constexpr int LE = ilog2c(E);
  constexpr int rem = LOG2L - LOG2NS;
  static_assert(rem > 0 && LOG2NS > 0, "a pass with twiddles left");
  if constexpr (rem > LE) {
    wave_pass_rt<E, E, LOG2L, LOG2NS, DIR, T, OFF, NT>(pb, t, active, twc);
    wave_fft_to_regs_rt<LOG2L, LOG2NS + LE, E, DIR, T, OFF + (E - 1), NT>(pb, t, active, twc,
                                                                          out);
  } else {
    constexpr int R = 1 << rem;
    constexpr int L = 1 << LOG2L;
    constexpr int tpp = L / E;
    constexpr int nbfly = L / R;
    if (active) {
#pragma unroll
      for (int u = 0; u < E / R; ++u) {
        const int j = t + u * tpp;
        cplx<T>* w = out + u * R;
#pragma unroll
        for (int r = 0; r < R; ++r) w[r] = pb[pad8(j + r * nbfly)];
#pragma unroll
        for (int r = 1; r < R; ++r) w[r] = cmul(w[r], twc[OFF + u * (R - 1) + r - 1]);
        dft_reg<R, DIR, T>(w);
      }
    }
  }
}

// ------------------------------------------- wave-local mixed-radix passes
// Radix-3 and radix-5 butterflies with the roots taken from the length-L
// table (w_m = tw[m L / R], conjugated for the inverse), in the symmetric
// form: y_s and y_{R-s} share their real combinations.
template <int DIR, typename T>
__device__ __forceinline__ void dft3(cplx<T>* v, const cplx<T>* __restrict__ tw, int LoverR) {
  cplx<T> w = tw[LoverR];
  if (DIR > 0) w.im = -w.im;
  const cplx<T> t1 = cadd(v[1], v[2]);
  const cplx<T> t2 = csub(v[1], v[2]);
  const cplx<T> m = {v[0].re + w.re * t1.re, v[0].im + w.re * t1.im};
  const cplx<T> r = {-w.im * t2.im, w.im * t2.re};  // i * Im(w) * (v1 - v2)
  v[0] = cadd(v[0], t1);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft5(cplx<T>* v, const cplx<T>* __restrict__ tw, int LoverR) {
  cplx<T> w1 = tw[LoverR], w2 = tw[2 * LoverR];
  if (DIR > 0) {
    w1.im = -w1.im;
    w2.im = -w2.im;
  }
  const cplx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
  const cplx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  const cplx<T> m1 = {v[0].re + w1.re * a1.re + w2.re * a2.re, v[0].im + w1.re * a1.im + w2.re * a2.im};
  const cplx<T> m2 = {v[0].re + w2.re * a1.re + w1.re * a2.re, v[0].im + w2.re * a1.im + w1.re * a2.im};
  // i * (s1 b1 + s2 b2) and i * (s2 b1 - s1 b2)
  const cplx<T> n1 = {-(w1.im * b1.im + w2.im * b2.im), w1.im * b1.re + w2.im * b2.re};
  const cplx<T> n2 = {-(w2.im * b1.im - w1.im * b2.im), w2.im * b1.re - w1.im * b2.re};
  v[0] = cadd(v[0], cadd(a1, a2));
  v[1] = cadd(m1, n1);
  v[4] = csub(m1, n1);
  v[2] = cadd(m2, n2);
  v[3] = csub(m2, n2);
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void dft_gen(cplx<T>* v, const cplx<T>* __restrict__ tw, int L) {
  if constexpr (R == 2 || R == 4 || R == 8 || R == 16)
    dft_reg<R, DIR, T>(v);
  else if constexpr (R == 3)
    dft3<DIR, T>(v, tw, L / 3);
  else if constexpr (R == 5)
    dft5<DIR, T>(v, tw, L / 5);
  else
    dft_prime<R, DIR, T>(v, tw, L / R);
}

// One Stockham pass of radix R over a pencil of length L owned by LP lanes
// of one wavefront (padded layout), lane t doing butterflies t, t + LP, ...
// (at most NB of them).  All inputs are read before the outputs are written,
// so the pass is in place.  tw[k] = exp(-2 pi i k / L), k < L.
template <int R, int NB, int LP, int DIR, typename T>
__device__ __forceinline__ void wave_pass_gen(cplx<T>* pb, int t, bool active, int L, int Ns,
                                              const cplx<T>* __restrict__ tw) {
  const int nb = L / R;
  const int twstep = L / (Ns * R);
  const double invNs = 1.0 / (double)Ns;  // b / Ns for b < 2^20 exactly
  cplx<T> v[NB][R];
  if (active) {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int b = t + i * LP;
      if (b < nb) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[i][r] = pb[pad8(b + r * nb)];
      }
    }
  }
  wave_sync();
  if (active) {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int b = t + i * LP;
      if (b < nb) {
        const int q = (int)(((double)b + 0.5) * invNs);
        const int jm = b - q * Ns;
        cplx<T>* w = v[i];
        if (jm != 0) {
#pragma unroll
          for (int r = 1; r < R; ++r) {
            cplx<T> f = tw[jm * r * twstep];
            if (DIR > 0) f.im = -f.im;
            w[r] = cmul(w[r], f);
          }
        }
        dft_gen<R, DIR, T>(w, tw, L);
        const int out0 = q * Ns * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) pb[pad8(out0 + r * Ns)] = w[r];
      }
    }
  }
  wave_sync();
}

// Butterflies per lane the wave-local mixed-radix passes provide for radix R
// (at most 8 values per lane); a length fits when ceil(L / R / LP) does not
// exceed it for every pass.
__host__ __device__ constexpr int wave_gen_nb(int R) {
  return R == 2 ? 4 : R == 3 ? 2 : R == 4 ? 2 : R == 5 ? 1 : R == 8 ? 1 : 0;
}

// Wave-local FFT of a length f.L = 2^a 3^b 5^c (natural order in and out,
// unnormalised), the pencil owned by LP lanes of one wavefront, with the
// radices of f (each 2, 3, 4, 5 or 8; uniform per wave, so the switch does
// not diverge).  Larger radices are left to the block-level form: the
// registers of the largest case bound the kernel's occupancy.
template <int DIR, typename T, int LP>
__device__ __forceinline__ void wave_fft_gen(cplx<T>* pb, int t, bool active, const FftFactors& f,
                                             const cplx<T>* __restrict__ tw) {
  const int L = f.L;
  int Ns = 1;
  for (int p = 0; p < f.np; ++p) {
    const int R = f.rad[p];
    switch (R) {
      case 2: wave_pass_gen<2, wave_gen_nb(2), LP, DIR, T>(pb, t, active, L, Ns, tw); break;
      case 3: wave_pass_gen<3, wave_gen_nb(3), LP, DIR, T>(pb, t, active, L, Ns, tw); break;
      case 4: wave_pass_gen<4, wave_gen_nb(4), LP, DIR, T>(pb, t, active, L, Ns, tw); break;
      case 5: wave_pass_gen<5, wave_gen_nb(5), LP, DIR, T>(pb, t, active, L, Ns, tw); break;
      default: wave_pass_gen<8, wave_gen_nb(8), LP, DIR, T>(pb, t, active, L, Ns, tw); break;
    }
    Ns *= R;
  }
}

}  // namespace pkde
