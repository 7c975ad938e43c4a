// Deterministic binning without global atomics (fixed-point mode).
//
// The linear binning of generate_dirac_cpu! (DirectSpace.jl:196-242) deposits,
// for every draw of every bootstrap resample, the corner weights of the drawn
// sample onto its 2^N neighbouring grid nodes.  A draw's deposit depends only
// on the sample, so a resample's field is sum_i c_{b,i} * deposit_i, with
// c_{b,i} the number of times sample i was drawn.  In int64 fixed point the
// per-deposit value llrint(term * scale) times c equals the sum of c separate
// deposits, so this form gives exactly the integers of the per-draw atomic
// form.  Steps:
//   1. cell of every sample (FP64, the operations of the reference), counting
//      sort of the samples by cell;
//   2. per bootstrap, the multiplicities c_{b,i} of the samples, counted in
//      LDS by sample index (uint8; an overflow marks the bootstrap, which is
//      then binned by the per-draw atomic kernel);
//   3. per (bootstrap, grid tile), the deposits of the samples whose corners
//      reach the tile, accumulated in LDS and written with plain stores:
//      every accumulator cell is written exactly once, nothing needs zeroing.
#include "pkde_kernels.hpp"

#include <math.h>

#include <algorithm>

namespace pkde {

namespace {

constexpr int kBinThreads = 1024;
constexpr int kTileCells = kBinTileCells;  // at most: 2 x 8192 int64 = 128 KB of LDS
constexpr int kCountChunk = 131072;        // samples per counting block (uint8)
constexpr int kScanBlock = 4096;           // elements per scan block (1024 x 4)

// Cell index and remainders exactly as k_binning / generate_dirac_cpu!:
// l = floor(delta / h), rh = rem(delta, h) with delta = x - lo.
__device__ __forceinline__ bool sample_cell(const GridDev& g, const double* xs, int64_t& cell,
                                            double* rh) {
  bool ok = true;
  int64_t c = 0, stride = 1;
  for (int i = 0; i < g.N; ++i) {
    const double xi = xs[i];
    const double delta = xi - g.lo[i];
    const double l = floor(delta / g.h[i]);
    rh[i] = fmod(delta, g.h[i]);
    ok = ok && (xi > g.lo[i]) && (xi < g.hi[i]) && (l <= (double)(g.n[i] - 2));
    c += (ok ? (int64_t)l : 0) * stride;
    stride *= g.n[i];
  }
  cell = ok ? c : g.G;
  return ok;
}

__global__ __launch_bounds__(kThreads) void k_bin_cells(GridDev g, const double* __restrict__ x,
                                                        int64_t n, int32_t* __restrict__ cell,
                                                        int32_t* __restrict__ hist, int* err) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t c;
    double rh[kMaxDims];
    if (!sample_cell(g, x + i * g.N, c, rh)) atomicOr(err, 1);
    cell[i] = (int32_t)c;
    atomicAdd(hist + c, 1);
  }
}

// Exclusive scan of int32 counts, kScanBlock elements per block; block totals
// go to `sums` (scanned recursively by the host driver).
__global__ __launch_bounds__(1024) void k_scan_block(const int32_t* __restrict__ in,
                                                     int32_t* __restrict__ out, int64_t m,
                                                     int32_t* __restrict__ sums) {
  __shared__ int32_t wsum[16];
  const int64_t base = (int64_t)blockIdx.x * kScanBlock;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int32_t v[4];
  int32_t t = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t idx = base + tid * 4 + k;
    v[k] = idx < m ? in[idx] : 0;
    t += v[k];
  }
  // inclusive scan of t across the wave
  int32_t inc = t;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t y = __shfl_up(inc, off, 64);
    if (lane >= off) inc += y;
  }
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  int32_t wbase = 0;
  for (int k = 0; k < wv; ++k) wbase += wsum[k];
  int32_t run = wbase + inc - t;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t idx = base + tid * 4 + k;
    if (idx < m) out[idx] = run;
    run += v[k];
  }
  if (tid == 1023) sums[blockIdx.x] = wbase + inc;
}

__global__ __launch_bounds__(1024) void k_scan_add(int32_t* __restrict__ out, int64_t m,
                                                   const int32_t* __restrict__ offs) {
  const int64_t base = (int64_t)blockIdx.x * kScanBlock;
  const int32_t o = offs[blockIdx.x];
  for (int k = threadIdx.x; k < kScanBlock; k += 1024) {
    const int64_t idx = base + k;
    if (idx < m) out[idx] += o;
  }
}

__global__ __launch_bounds__(kThreads) void k_bin_scatter(GridDev g, const double* __restrict__ x,
                                                          int64_t n, const int32_t* __restrict__ cell,
                                                          const int32_t* __restrict__ start,
                                                          int32_t* __restrict__ cursor,
                                                          int32_t* __restrict__ order,
                                                          int32_t* __restrict__ scell,
                                                          double* __restrict__ srh) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = cell[i];
    const int32_t pos = start[c] + atomicAdd(cursor + c, 1);
    order[pos] = (int32_t)i;
    scell[pos] = c;
    int64_t cc;
    double rh[kMaxDims];
    sample_cell(g, x + i * g.N, cc, rh);
    for (int k = 0; k < g.N; ++k) srh[(int64_t)pos * g.N + k] = rh[k];
  }
}

// Multiplicities of bootstrap b for samples [c0, c0 + kCountChunk):
// every block regenerates all n draws of its bootstrap (Philox4x32-10 with
// counter {block, block >> 32, b, 0}, as the atomic kernel) or reads them
// from idx, and counts those that fall in its chunk.
__global__ __launch_bounds__(kBinThreads) void k_bin_count(int64_t n, int64_t npad, int nchunks,
                                                           const int32_t* __restrict__ idx,
                                                           uint32_t key0, uint32_t key1,
                                                           uint8_t* __restrict__ cnt,
                                                           int32_t* __restrict__ ovf, int* err) {
  extern __shared__ uint32_t words[];
  const int b = blockIdx.x / nchunks;
  const int64_t c0 = (int64_t)(blockIdx.x - b * nchunks) * kCountChunk;
  const int64_t clen = std::min<int64_t>(kCountChunk, n - c0);
  const int nwords = (int)((clen + 3) >> 2);
  for (int k = threadIdx.x; k < nwords; k += kBinThreads) words[k] = 0u;
  __syncthreads();
  bool over = false;
  const int64_t nq = (n + 3) >> 2;
  for (int64_t blk = threadIdx.x; blk < nq; blk += kBinThreads) {
    const int64_t s0 = blk * 4;
    int64_t src[4];
    bool ok[4] = {true, true, true, true};
    if (idx) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t v = (s0 + q < n) ? (int64_t)idx[(int64_t)b * n + s0 + q] : 0;
        ok[q] = v >= 0 && v < n;
        src[q] = ok[q] ? v : 0;
      }
    } else {
      uint32_t c[4] = {(uint32_t)blk, (uint32_t)((uint64_t)blk >> 32), (uint32_t)b, 0u};
      philox4x32_10(c, key0, key1);
#pragma unroll
      for (int q = 0; q < 4; ++q) src[q] = (int64_t)(((uint64_t)c[q] * (uint64_t)n) >> 32);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (s0 + q >= n) break;
      if (!ok[q]) {
        if (c0 == 0) atomicOr(err, 2);
        continue;
      }
      const int64_t rel = src[q] - c0;
      if (rel >= 0 && rel < clen) {
        const int sh = (int)(rel & 3) * 8;
        const uint32_t old = atomicAdd(words + (rel >> 2), 1u << sh);
        over |= ((old >> sh) & 0xFFu) == 0xFFu;
      }
    }
  }
  if (over) atomicOr(ovf + b, 1);
  __syncthreads();
  uint32_t* out = (uint32_t*)(cnt + (int64_t)b * npad + c0);
  for (int k = threadIdx.x; k < nwords; k += kBinThreads) out[k] = words[k];
}

// LDS of k_bin_deposit: the tile's two accumulator blocks and the run
// bounds of its tz + 1 planes.
size_t deposit_lds(int T, int tz) {
  return (size_t)2 * T * sizeof(unsigned long long) + (((size_t)2 * (tz + 1) * 4 + 7) & ~(size_t)7);
}

struct TileGeom {
  int ty, tz;            // tile rows along axes 2, 3 (axis 1 spans the grid)
  int tiles_y, tiles_z;  // tiles per axis
  int T;                 // cells per tile
};

// Deposits of bootstrap b (b == B: the full sample, multiplicity 1) into one
// grid tile, N = 1, 2 or 3.  Corner weights and fixed-point values follow
// k_binning operation for operation.  Each thread takes kBatch samples of the
// tile's sorted run per trip and issues all their loads (sorted cell,
// remainders, sample index, then the multiplicities through the sample
// indices) before depositing any of them, so that one trip waits for one
// round of memory latency rather than kBatch: with one 128 KB workgroup per
// CU the loop is latency-bound.
//
// Blocks run tile-major (block = tile (B + 1) + b).  Workgroups are handed to
// the XCDs, and within an XCD to its shader engines, in turn, and each
// engine takes its next workgroup in order; with bootstrap-major blocks every
// bootstrap of one tile went to the same engine, so the engine of a dense
// tile (the grid's centre: 15x the samples of an edge tile) ran all of that
// tile's workgroups in series while the rest of its XCD waited behind it in
// the dispatch order.  Tile-major blocks spread each tile over every engine.
constexpr int kBatch = 4;

template <int N>
__global__ __launch_bounds__(kBinThreads) void k_bin_deposit(GridDev g, int B, TileGeom tg,
                                                             const int32_t* __restrict__ start,
                                                             const int32_t* __restrict__ scell,
                                                             const int32_t* __restrict__ order,
                                                             const double* __restrict__ srh,
                                                             const uint8_t* __restrict__ cnt,
                                                             int64_t npad,
                                                             const int32_t* __restrict__ ovf,
                                                             double denom, double scaleM,
                                                             double scaleV,
                                                             unsigned long long* __restrict__ acc) {
  extern __shared__ unsigned long long lacc[];
  const int ntiles = tg.tiles_y * tg.tiles_z;
  const int tile = blockIdx.x / (B + 1);
  const int b = blockIdx.x - tile * (B + 1);
  const int tyi = tile % tg.tiles_y, tzi = tile / tg.tiles_y;
  const int n1 = g.n[0], n2 = N >= 2 ? g.n[1] : 1;
  const int y0 = tyi * tg.ty, z0 = tzi * tg.tz;
  const int T = tg.T;
  const bool sq = b < B;
  unsigned long long* lM = lacc;
  unsigned long long* lV = lacc + T;
  // [plane][2]: the sorted run of each plane's rows [ylo, yhi)
  int32_t* lrun = (int32_t*)(lacc + 2 * T);
  const int n3 = N == 3 ? g.n[2] : 1;
  const int zlo = N == 3 ? std::max(0, z0 - 1) : 0;
  const int zhi = N == 3 ? std::min(z0 + tg.tz, n3) : 1;  // exclusive
  const int ylo = N >= 2 ? std::max(0, y0 - 1) : 0;
  const int yhi = N >= 2 ? std::min(y0 + tg.ty, n2) : 1;
  for (int k = threadIdx.x; k < 2 * T; k += kBinThreads) lacc[k] = 0ull;
  // every plane's run bounds in one round of loads, not one round per plane
  for (int k = threadIdx.x; k < 2 * (zhi - zlo); k += kBinThreads) {
    const int64_t c = ((int64_t)(zlo + (k >> 1)) * n2 + ((k & 1) ? yhi : ylo)) * n1;
    lrun[k] = start[std::min<int64_t>(c, g.G)];
  }
  (void)ntiles;
  __syncthreads();
  const bool skip = sq && ovf[b] != 0;  // binned by the per-draw kernel instead
  if (!skip) {
    const uint8_t* cntb = cnt + (int64_t)b * npad;
    for (int z = zlo; z < zhi; ++z) {
      // base cells of rows [ylo, yhi) of plane z: one contiguous run
      const int32_t pA = lrun[2 * (z - zlo)], pB = lrun[2 * (z - zlo) + 1];
      for (int32_t p0 = pA + threadIdx.x; p0 < pB; p0 += kBatch * kBinThreads) {
        int32_t cl[kBatch], ord[kBatch];
        double rh[kBatch][N];
        unsigned c[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          const int32_t pos = p0 + u * kBinThreads;
          const bool in = pos < pB;
          cl[u] = in ? scell[pos] : 0;
          ord[u] = in && sq ? order[pos] : -1;
#pragma unroll
          for (int i = 0; i < N; ++i) rh[u][i] = in ? srh[(int64_t)pos * N + i] : 0.0;
          c[u] = in ? 1u : 0u;
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
          if (ord[u] >= 0) c[u] = (unsigned)cntb[ord[u]];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          if (c[u] == 0u) continue;
          int il[N];
          il[0] = cl[u] % n1;
          if (N >= 2) il[1] = (cl[u] / n1) % n2;
          if (N == 3) il[2] = cl[u] / (n1 * n2);
          double rl[N];
#pragma unroll
          for (int i = 0; i < N; ++i) rl[i] = g.h[i] - rh[u][i];
#pragma unroll
          for (int cc = 0; cc < (1 << N); ++cc) {
            double pl = 1.0, ph = 1.0;
            bool hl = false, hh = false;
            int li[N];
#pragma unroll
            for (int j = 0; j < N; ++j) {
              const int mask = (cc >> (N - 1 - j)) & 1;
              li[j] = mask ? il[j] : il[j] + 1;
              if (mask) {
                pl = hl ? pl * rl[j] : rl[j];
                hl = true;
              } else {
                ph = hh ? ph * rh[u][j] : rh[u][j];
                hh = true;
              }
            }
            // local cell in the tile (axis 1 always inside)
            int loc = li[0];
            if (N >= 2) {
              const int yy = li[1] - y0;
              if (yy < 0 || yy >= tg.ty) continue;
              loc += yy * n1;
            }
            if (N == 3) {
              const int zz = li[2] - z0;
              if (zz < 0 || zz >= tg.tz) continue;
              loc += zz * n1 * tg.ty;
            }
            const double term = (pl * ph) / denom;
            atomicAdd(lM + loc, (unsigned long long)c[u] * (unsigned long long)llrint(term * scaleM));
            if (sq)
              atomicAdd(lV + loc,
                        (unsigned long long)c[u] * (unsigned long long)llrint((term * term) * scaleV));
          }
        }
      }
    }
  }
  __syncthreads();
  // write the tile: rows y0.., planes z0.. (edge tiles may be partial)
  const int64_t fM = sq ? b : 2 * (int64_t)B;
  for (int k = threadIdx.x; k < T; k += kBinThreads) {
    const int x = k % n1;
    const int rest = k / n1;
    const int yy = rest % tg.ty, zz = rest / tg.ty;
    if (y0 + yy >= n2 || z0 + zz >= n3) continue;
    const int64_t cellid = (((int64_t)(z0 + zz) * n2) + (y0 + yy)) * n1 + x;
    acc[fM * g.G + cellid] = lM[k];
    if (sq) acc[((int64_t)B + b) * g.G + cellid] = lV[k];
  }
}

// Bootstraps whose multiplicities overflowed: per-draw atomic deposits of all
// their draws (one block per bootstrap; returns at once for the others).
__global__ __launch_bounds__(kThreads) void k_bin_fallback(GridDev g, const double* __restrict__ x,
                                                           int64_t n, const int32_t* __restrict__ idx,
                                                           int B, uint32_t key0, uint32_t key1,
                                                           const int32_t* __restrict__ ovf,
                                                           double denom, double scaleM,
                                                           double scaleV,
                                                           unsigned long long* __restrict__ acc) {
  const int b = blockIdx.x;
  if (b >= B || ovf[b] == 0) return;
  const int N = g.N;
  const int64_t nq = (n + 3) >> 2;
  int64_t stride[kMaxDims];
  stride[0] = 1;
  for (int i = 1; i < N; ++i) stride[i] = stride[i - 1] * g.n[i - 1];
  for (int64_t blk = threadIdx.x; blk < nq; blk += blockDim.x) {
    const int64_t s0 = blk * 4;
    int64_t src[4];
    bool ok4[4] = {true, true, true, true};
    if (idx) {
      for (int q = 0; q < 4; ++q) {
        const int64_t v = (s0 + q < n) ? (int64_t)idx[(int64_t)b * n + s0 + q] : 0;
        ok4[q] = v >= 0 && v < n;
        src[q] = ok4[q] ? v : 0;
      }
    } else {
      uint32_t c[4] = {(uint32_t)blk, (uint32_t)((uint64_t)blk >> 32), (uint32_t)b, 0u};
      philox4x32_10(c, key0, key1);
      for (int q = 0; q < 4; ++q) src[q] = (int64_t)(((uint64_t)c[q] * (uint64_t)n) >> 32);
    }
    for (int q = 0; q < 4; ++q) {
      if (s0 + q >= n) break;
      if (!ok4[q]) continue;
      int64_t cell;
      double rh[kMaxDims], rl[kMaxDims];
      if (!sample_cell(g, x + src[q] * N, cell, rh)) continue;
      int64_t il[kMaxDims];
      int64_t rest = cell;
      for (int i = 0; i < N; ++i) {
        il[i] = rest % g.n[i];
        rest /= g.n[i];
        rl[i] = g.h[i] - rh[i];
      }
      for (int cc = 0; cc < (1 << N); ++cc) {
        int64_t off = 0;
        double pl = 1.0, ph = 1.0;
        bool hl = false, hh = false;
        for (int j = 0; j < N; ++j) {
          const int mask = (cc >> (N - 1 - j)) & 1;
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
        atomicAdd(acc + (int64_t)b * g.G + off, (unsigned long long)llrint(term * scaleM));
        atomicAdd(acc + ((int64_t)B + b) * g.G + off,
                  (unsigned long long)llrint((term * term) * scaleV));
      }
    }
  }
}

// Tiles span axis 1 and hold at most `cells` cells (n1 <= cells).
TileGeom tile_geom(const GridDev& g, int cells) {
  TileGeom t{};
  const int n1 = g.n[0];
  const int n2 = g.N >= 2 ? g.n[1] : 1, n3 = g.N == 3 ? g.n[2] : 1;
  const int rows = std::max(1, cells / n1);  // rows of n1 cells per tile
  if (g.N == 1) {
    t.ty = 1;
    t.tz = 1;
  } else if (g.N == 2) {
    t.ty = std::min(n2, rows);
    t.tz = 1;
  } else {
    int ty = 1;
    while (ty * ty < rows) ty <<= 1;
    t.ty = std::min(n2, std::min(ty, rows));
    t.tz = std::min(n3, std::max(1, rows / t.ty));
  }
  t.tiles_y = (n2 + t.ty - 1) / t.ty;
  t.tiles_z = (n3 + t.tz - 1) / t.tz;
  t.T = n1 * t.ty * t.tz;
  return t;
}

// Scratch of scan_exclusive: per level, block sums and their scan.
int64_t scan_levels_elems(int64_t m) {
  int64_t total = 0;
  while (m > 1) {
    m = (m + kScanBlock - 1) / kScanBlock;
    total += 2 * (m + 1);
  }
  return total + 2;
}

void scan_exclusive(const int32_t* in, int32_t* out, int64_t m, int32_t* tmp, hipStream_t s) {
  const int64_t nb = (m + kScanBlock - 1) / kScanBlock;
  int32_t* sums = tmp;
  k_scan_block<<<(unsigned)nb, 1024, 0, s>>>(in, out, m, sums);
  if (nb > 1) {
    int32_t* offs = tmp + nb + 1;
    scan_exclusive(sums, offs, nb, offs + nb + 1, s);
    k_scan_add<<<(unsigned)nb, 1024, 0, s>>>(out, m, offs);
  }
}

}  // namespace

uint64_t binsort_bytes(const GridDev& g, int64_t n, int B) {
  const int64_t npad = (n + 3) & ~(int64_t)3;
  uint64_t b = 0;
  auto add = [&](uint64_t v) { b += (v + 255) & ~(uint64_t)255; };
  add((uint64_t)n * 4);                 // cell
  add((uint64_t)(g.G + 2) * 4);         // hist / start
  add((uint64_t)(g.G + 2) * 4);         // cursor
  add((uint64_t)n * 4);                 // order
  add((uint64_t)n * 4);                 // scell
  add((uint64_t)n * g.N * 8);           // srh
  add((uint64_t)B * npad);              // cnt
  add((uint64_t)B * 4);                 // ovf
  add((uint64_t)scan_levels_elems(g.G + 2) * 4 + 4096);  // scan temporaries
  return b;
}

int launch_binning_sorted(const GridDev& g, const double* x, int64_t n, const int32_t* idx, int B,
                          uint64_t seed, int frac_bits, void* acc, int* err, void* work,
                          uint64_t work_bytes, hipStream_t s, int tile_cells) {
  if (g.N > 3 || binsort_bytes(g, n, B) > work_bytes) return -1;
  if (tile_cells <= 0 || tile_cells > kTileCells) tile_cells = kTileCells;
  if (g.G + 2 >= (1ll << 31) || n >= (1ll << 31) || g.n[0] > tile_cells) return -1;
  const TileGeom tg = tile_geom(g, tile_cells);
  if (deposit_lds(tg.T, tg.tz) > 160 * 1024) return -1;
  const int64_t npad = (n + 3) & ~(int64_t)3;
  unsigned char* p = (unsigned char*)work;
  auto take = [&](uint64_t v) {
    unsigned char* r = p;
    p += (v + 255) & ~(uint64_t)255;
    return r;
  };
  int32_t* cell = (int32_t*)take((uint64_t)n * 4);
  int32_t* start = (int32_t*)take((uint64_t)(g.G + 2) * 4);
  int32_t* cursor = (int32_t*)take((uint64_t)(g.G + 2) * 4);
  int32_t* order = (int32_t*)take((uint64_t)n * 4);
  int32_t* scell = (int32_t*)take((uint64_t)n * 4);
  double* srh = (double*)take((uint64_t)n * g.N * 8);
  uint8_t* cnt = (uint8_t*)take((uint64_t)B * npad);
  int32_t* ovf = (int32_t*)take((uint64_t)B * 4);
  int32_t* tmp = (int32_t*)take((uint64_t)scan_levels_elems(g.G + 2) * 4 + 4096);

  double hp = g.h[0];
  for (int i = 1; i < g.N; ++i) hp *= g.h[i];
  const double denom = (double)n * (hp * hp);
  const double two_fb = ldexp(1.0, frac_bits);
  const double scaleM = (double)n * hp * two_fb;
  const double scaleV = ((double)n * hp) * ((double)n * hp) * two_fb;
  const uint32_t key0 = (uint32_t)seed, key1 = (uint32_t)(seed >> 32);

  // LDS beyond 64 KB per block (gfx950 has 160 KB per CU); set once
  static const bool attrs = [] {
    (void)hipFuncSetAttribute((const void*)k_bin_count, hipFuncAttributeMaxDynamicSharedMemorySize,
                              kCountChunk);
    const void* dep[3] = {(const void*)k_bin_deposit<1>, (const void*)k_bin_deposit<2>,
                          (const void*)k_bin_deposit<3>};
    for (const void* f : dep)
      (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                160 * 1024);
    return true;
  }();
  (void)attrs;
  // 1. cells, histogram (cursor doubles as the histogram), scan, scatter
  launch_fill_i32(cursor, 0, g.G + 2, s);
  launch_fill_i32(ovf, 0, B, s);
  const unsigned nb = (unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, 1 << 16);
  k_bin_cells<<<nb, kThreads, 0, s>>>(g, x, n, cell, cursor, err);
  scan_exclusive(cursor, start, g.G + 2, tmp, s);
  launch_fill_i32(cursor, 0, g.G + 2, s);
  k_bin_scatter<<<nb, kThreads, 0, s>>>(g, x, n, cell, start, cursor, order, scell, srh);
  // 2. multiplicities
  if (B > 0) {
    const int nchunks = (int)((n + kCountChunk - 1) / kCountChunk);
    const int64_t clen = std::min<int64_t>(kCountChunk, n);
    const size_t lds = (size_t)((clen + 3) >> 2) * 4;
    k_bin_count<<<(unsigned)(B * nchunks), kBinThreads, lds, s>>>(n, npad, nchunks, idx, key0,
                                                                   key1, cnt, ovf, err);
  }
  // 3. tile deposits for every bootstrap and the full sample
  const unsigned nblk = (unsigned)((B + 1) * tg.tiles_y * tg.tiles_z);
  const size_t lds = deposit_lds(tg.T, tg.tz);
  auto deposit = g.N == 1 ? k_bin_deposit<1> : g.N == 2 ? k_bin_deposit<2> : k_bin_deposit<3>;
  deposit<<<nblk, kBinThreads, lds, s>>>(g, B, tg, start, scell, order, srh, cnt, npad, ovf, denom,
                                         scaleM, scaleV, (unsigned long long*)acc);
  if (B > 0)
    k_bin_fallback<<<(unsigned)B, kThreads, 0, s>>>(g, x, n, idx, B, key0, key1, ovf, denom,
                                                    scaleM, scaleV, (unsigned long long*)acc);
  return 0;
}

}  // namespace pkde
