// t-SNE device kernels for gfx950.
//
// Per-iteration pipeline (six launches, all state resident in HBM):
//   1. splat            bilinear spread of every point onto the G x G grid
//                       (mass, mass*x, mass*y) with 64-bit fixed-point atomics
//                       (LDS-aggregated per chunk when the chunk is compact),
//                       so the grid sums do not depend on arrival order;
//   2. conv_rows_fwd    row FFTs of z1 = mass + i*mass_x/x_scale and
//                       z2 = mass_y (zero padded to F = 2G), and of the G
//                       kernel rows Kq + i*Kq^2 (both real and even, so the
//                       spectra come out as real + i*real);
//   3. conv_cols        per column pair (c, F-c): column FFTs, Hermitian
//                       split of z1 into FFT(mass) and FFT(mass_x), the
//                       pointwise products with the kernel spectra, inverse
//                       column FFTs, and Z by Parseval in FP64;
//   4. conv_rows_inv    inverse row FFTs; writes the (q2, xq2, yq2, q) field;
//   5. attract          CSR attraction over the symmetric P (no atomics);
//   6. update           per point: bilinear gather of the field, the
//                       gains/momentum/clip update, and the FP64 layout
//                       statistics (mean, bounds) for the next iteration.
// Centring is applied lazily: the layout buffer holds uncentred coordinates
// and every reader subtracts the float mean kept in TsneStats, which gives
// bit-for-bit the values an explicit centring pass would have written.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "fe_fft.hpp"

namespace fe {
// kernels have internal linkage: every translation unit that includes this
// header gets its own copy of the device code
namespace {

constexpr int kThreads = 256;

struct TsneStats {
  double mean_x, mean_y;   // FP64 mean of the raw layout buffer
  double sum_q_total;      // sum_i bilinear(q)(y_i) = Z + n (Parseval)
  double fx_mass, fx_xy;   // fixed-point scales (powers of two)
  float mean_xf, mean_yf;  // float(mean): what every reader subtracts
  float lower_x, lower_y;  // grid corner, centred coordinates
  float spacing, inv_spacing;
  float min_x, max_x, min_y, max_y;  // bounds of the centred layout
  // power of two >= max |x|: the mass*x channel shares a complex FFT with
  // the mass channel, and is carried as mass*x / x_scale so both parts have
  // the magnitude of the mass (the Hermitian split then loses no digits)
  double x_scale;
  unsigned int done_cols;  // last-block counters (reset by the last block)
  unsigned int done_stats;
};

struct LayoutPartial {
  double sx, sy;
  float mnx, mxx, mny, mxy;
};

// ----------------------------------------------------------------------------
// small helpers
// ----------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Bilinear cell of a centred point: tsne_neighbors.cpp:1190-1206 (the CUDA
// twin is embedding_cuda_kernels.cpp:2634-2650).  Shared by splat and gather
// so both see identical weights.
struct Cell {
  int x0, y0;
  float tx, ty;
};

__device__ __forceinline__ Cell grid_cell(float x, float y, float lower_x, float lower_y,
                                          float inv_spacing, int g) {
  const float gx = fminf(static_cast<float>(g - 1), fmaxf(0.0f, (x - lower_x) * inv_spacing));
  const float gy = fminf(static_cast<float>(g - 1), fmaxf(0.0f, (y - lower_y) * inv_spacing));
  Cell c;
  c.x0 = min(g - 2, max(0, static_cast<int>(floorf(gx))));
  c.y0 = min(g - 2, max(0, static_cast<int>(floorf(gy))));
  c.tx = gx - static_cast<float>(c.x0);
  c.ty = gy - static_cast<float>(c.y0);
  return c;
}

__device__ __forceinline__ unsigned long long to_fixed(float v, double scale) {
  return static_cast<unsigned long long>(__double2ll_rn(static_cast<double>(v) * scale));
}

// Geometry from the centred bounds: tsne_neighbors.cpp:1176-1184 in FP64 as
// the CUDA finalize does (embedding_cuda_kernels.cpp:2567-2583), plus the
// fixed-point scales sized so that n contributions of the largest magnitude
// cannot overflow 62 bits.
__device__ inline void finalize_geometry(TsneStats* st, double sx, double sy, float mnx, float mxx,
                                         float mny, float mxy, int n, int g, bool centre) {
  const double mean_x = sx / static_cast<double>(n), mean_y = sy / static_cast<double>(n);
  // centre == false keeps the layout as given (force diagnostics)
  const float mxf = centre ? static_cast<float>(mean_x) : 0.0f;
  const float myf = centre ? static_cast<float>(mean_y) : 0.0f;
  // fl(a - m) is monotone in a, so the bounds of the centred layout are the
  // raw bounds shifted by the same float subtraction.
  const float cmnx = mnx - mxf, cmxx = mxx - mxf, cmny = mny - myf, cmxy = mxy - myf;
  const double cx = 0.5 * (static_cast<double>(cmnx) + static_cast<double>(cmxx));
  const double cy = 0.5 * (static_cast<double>(cmny) + static_cast<double>(cmxy));
  double span = fmax(static_cast<double>(cmxx) - cmnx, static_cast<double>(cmxy) - cmny);
  if (!isfinite(span) || span <= 0.0) span = 1.0;
  const double half = 0.55 * span + 1.0e-3;
  const double spacing = (2.0 * half) / static_cast<double>(g - 1);
  st->mean_x = mean_x;
  st->mean_y = mean_y;
  st->mean_xf = mxf;
  st->mean_yf = myf;
  st->min_x = cmnx;
  st->max_x = cmxx;
  st->min_y = cmny;
  st->max_y = cmxy;
  st->lower_x = static_cast<float>(cx - half);
  st->lower_y = static_cast<float>(cy - half);
  st->spacing = static_cast<float>(spacing);
  st->inv_spacing = static_cast<float>(1.0 / spacing);
  const double r = fmax(fmax(fabs(static_cast<double>(cmnx)), fabs(static_cast<double>(cmxx))),
                        fmax(fabs(static_cast<double>(cmny)), fabs(static_cast<double>(cmxy))));
  const int en = ilogb(static_cast<double>(n)) + 1;
  const double nr = fmax(static_cast<double>(n) * r, 1e-30);
  const int exy = ilogb(nr) + 1;
  st->fx_mass = ldexp(1.0, 61 - en);
  st->fx_xy = ldexp(1.0, min(61 - exy, 1000));
  const double rx = fmax(fabs(static_cast<double>(cmnx)), fabs(static_cast<double>(cmxx)));
  st->x_scale = rx > 0.0 && isfinite(rx) ? ldexp(1.0, ilogb(rx) + 1) : 1.0;
}

// Block reduction of LayoutPartial then last-block finalize over the nblk
// blocks that work on this layout (`partials` has nblk entries, this block is
// blk).  Returns true in thread 0 of the block that finalised.
__device__ inline bool reduce_layout_stats(LayoutPartial mine, LayoutPartial* partials,
                                           TsneStats* st, int n, int g, bool centre, int blk,
                                           int nblk) {
  __shared__ LayoutPartial red[kThreads / 64];
  __shared__ bool last;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  mine.sx = wave_sum(mine.sx);
  mine.sy = wave_sum(mine.sy);
  mine.mnx = static_cast<float>(wave_min(mine.mnx));
  mine.mxx = static_cast<float>(wave_max(mine.mxx));
  mine.mny = static_cast<float>(wave_min(mine.mny));
  mine.mxy = static_cast<float>(wave_max(mine.mxy));
  if (lane == 0) red[wid] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    LayoutPartial b = red[0];
    for (int w = 1; w < kThreads / 64; ++w) {
      b.sx += red[w].sx;
      b.sy += red[w].sy;
      b.mnx = fminf(b.mnx, red[w].mnx);
      b.mxx = fmaxf(b.mxx, red[w].mxx);
      b.mny = fminf(b.mny, red[w].mny);
      b.mxy = fmaxf(b.mxy, red[w].mxy);
    }
    partials[blk] = b;
    __threadfence();
    const unsigned prev = atomicAdd(&st->done_stats, 1u);
    last = (prev == static_cast<unsigned>(nblk - 1));
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  if (threadIdx.x == 0) {
    const volatile LayoutPartial* vp = partials;
    double sx = 0.0, sy = 0.0;
    float mnx = FLT_MAX, mxx = -FLT_MAX, mny = FLT_MAX, mxy = -FLT_MAX;
    for (int b = 0; b < nblk; ++b) {
      sx += vp[b].sx;
      sy += vp[b].sy;
      mnx = fminf(mnx, vp[b].mnx);
      mxx = fmaxf(mxx, vp[b].mxx);
      mny = fminf(mny, vp[b].mny);
      mxy = fmaxf(mxy, vp[b].mxy);
    }
    finalize_geometry(st, sx, sy, mnx, mxx, mny, mxy, n, g, centre);
    st->done_stats = 0u;
    return true;
  }
  return false;
}

// ----------------------------------------------------------------------------
// Layout statistics of a raw layout (initialisation; zero_mean_f and the
// bounds pass of the reference).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void layout_stats_kernel(const float2* __restrict__ y,
                                                                LayoutPartial* partials,
                                                                TsneStats* st, int n, int g,
                                                                int centre) {
  LayoutPartial m{0.0, 0.0, FLT_MAX, -FLT_MAX, FLT_MAX, -FLT_MAX};
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const float2 p = y[i];
    m.sx += p.x;
    m.sy += p.y;
    m.mnx = fminf(m.mnx, p.x);
    m.mxx = fmaxf(m.mxx, p.x);
    m.mny = fminf(m.mny, p.y);
    m.mxy = fmaxf(m.mxy, p.y);
  }
  reduce_layout_stats(m, partials, st, n, g, centre != 0, blockIdx.x, gridDim.x);
}

// ----------------------------------------------------------------------------
// Every stage below is a __device__ body taking its block index within the
// job (blk) and the job's block count (nblk); the single-job kernels pass
// blockIdx.x / gridDim.x and the job-table kernels of tsne_batch.hpp pass the
// job-local values, so a job gives the same bits alone and inside a batch.
//
// 1. splat: embedding_cuda_kernels.cpp:2622-2667 / tsne_neighbors.cpp:
//    1186-1226, accumulated in 64-bit fixed point (deterministic).
// grids: 3 channels of g*g int64, channel-major, row-major nodes.
// ----------------------------------------------------------------------------
// A block takes chunks of kSplatChunk consecutive points.  When the chunk's
// cells fit a kTileSide x kTileSide node window (neighbouring point ids are
// usually close in the layout, and early-exaggeration clusters are compact),
// the chunk accumulates in LDS and flushes each touched node once with global
// atomics; otherwise its points go straight to global atomics.  The sums are
// integers, so both routes give identical grids.
constexpr int kSplatPerThread = 4;
constexpr int kSplatChunk = kThreads * kSplatPerThread;
constexpr int kTileSide = 32;
constexpr int kTileNodes = kTileSide * kTileSide;

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ void splat_body(const float2* __restrict__ y,
                                           const TsneStats* __restrict__ st,
                                           unsigned long long* grids, int n, int g, int blk,
                                           int nblk) {
  __shared__ unsigned long long tile[3 * kTileNodes];
  __shared__ int box[4];
  const float mx = st->mean_xf, my = st->mean_yf;
  const float lx = st->lower_x, ly = st->lower_y, inv = st->inv_spacing;
  const double sm = st->fx_mass, sxy = st->fx_xy;
  const size_t gg = static_cast<size_t>(g) * g;
  const int n_chunks = (n + kSplatChunk - 1) / kSplatChunk;
  for (int ch = blk; ch < n_chunks; ch += nblk) {
    Cell cl[kSplatPerThread];
    float xc[kSplatPerThread], yc[kSplatPerThread];
    bool ok[kSplatPerThread];
    int bx0 = 1 << 30, by0 = 1 << 30, bx1 = -1, by1 = -1;
#pragma unroll
    for (int u = 0; u < kSplatPerThread; ++u) {
      const int i = ch * kSplatChunk + u * kThreads + threadIdx.x;
      ok[u] = i < n;
      if (ok[u]) {
        const float2 p = y[i];
        xc[u] = p.x - mx;
        yc[u] = p.y - my;
        cl[u] = grid_cell(xc[u], yc[u], lx, ly, inv, g);
        bx0 = min(bx0, cl[u].x0);
        by0 = min(by0, cl[u].y0);
        bx1 = max(bx1, cl[u].x0);
        by1 = max(by1, cl[u].y0);
      }
    }
    if (threadIdx.x == 0) {
      box[0] = 1 << 30;
      box[1] = 1 << 30;
      box[2] = -1;
      box[3] = -1;
    }
    __syncthreads();
    bx0 = wave_min_i(bx0);
    by0 = wave_min_i(by0);
    bx1 = wave_max_i(bx1);
    by1 = wave_max_i(by1);
    if ((threadIdx.x & 63) == 0) {
      atomicMin(&box[0], bx0);
      atomicMin(&box[1], by0);
      atomicMax(&box[2], bx1);
      atomicMax(&box[3], by1);
    }
    __syncthreads();
    const int ox = box[0], oy = box[1];
    const int bw = box[2] - ox + 2, bh = box[3] - oy + 2;  // node window
    const bool tiled = bw <= kTileSide && bh <= kTileSide;
    if (tiled) {
      for (int t = threadIdx.x; t < 3 * kTileNodes; t += kThreads) tile[t] = 0ull;
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < kSplatPerThread; ++u) {
      if (!ok[u]) continue;
      const Cell c = cl[u];
      const float w[4] = {(1.0f - c.tx) * (1.0f - c.ty), c.tx * (1.0f - c.ty),
                          (1.0f - c.tx) * c.ty, c.tx * c.ty};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int nx = c.x0 + (q & 1), ny = c.y0 + (q >> 1);
        const unsigned long long fm = to_fixed(w[q], sm);
        const unsigned long long fx = to_fixed(w[q] * xc[u], sxy);
        const unsigned long long fy = to_fixed(w[q] * yc[u], sxy);
        if (tiled) {
          const int t = (ny - oy) * kTileSide + (nx - ox);
          atomicAdd(tile + t, fm);
          atomicAdd(tile + kTileNodes + t, fx);
          atomicAdd(tile + 2 * kTileNodes + t, fy);
        } else {
          const size_t node = static_cast<size_t>(ny) * g + nx;
          atomicAdd(grids + node, fm);
          atomicAdd(grids + gg + node, fx);
          atomicAdd(grids + 2 * gg + node, fy);
        }
      }
    }
    if (tiled) {
      __syncthreads();
      for (int t = threadIdx.x; t < bw * bh; t += kThreads) {
        const int ty = t / bw, tx = t - ty * bw;
        const int lt = ty * kTileSide + tx;
        const size_t node = static_cast<size_t>(oy + ty) * g + (ox + tx);
#pragma unroll
        for (int c3 = 0; c3 < 3; ++c3) {
          const unsigned long long v = tile[c3 * kTileNodes + lt];
          if (v != 0ull) atomicAdd(grids + c3 * gg + node, v);
        }
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void splat_kernel(const float2* __restrict__ y,
                                                         const TsneStats* __restrict__ st,
                                                         unsigned long long* grids, int n, int g) {
  splat_body(y, st, grids, n, g, blockIdx.x, gridDim.x);
}

// ----------------------------------------------------------------------------
// 2. row FFTs.  Blocks [0, g): data row y.  Blocks [g, g + g/2): kernel rows
//    dy and dy + g/2.  Kernel row spectra are stored transposed,
//    kt[k2 * g + dy] for k2 in [0, g], so conv_cols reads them contiguously.
// ----------------------------------------------------------------------------
template <int F>
__device__ __forceinline__ void conv_rows_fwd_body(unsigned long long* grids,
                                                   const TsneStats* __restrict__ st,
                                                   const float2* __restrict__ tw,
                                                   float2* __restrict__ z1,
                                                   float2* __restrict__ z2,
                                                   float2* __restrict__ kt, int b) {
  constexpr int G = F / 2;
  __shared__ float2 buf[2 * F];
  if (b < G) {
    const int row = b;
    const double inv_m = 1.0 / st->fx_mass, inv_xy = 1.0 / st->fx_xy;
    const double inv_xs = inv_xy / st->x_scale;
    constexpr size_t gg = static_cast<size_t>(G) * G;
    for (int x = threadIdx.x; x < F; x += kThreads) {
      float2 a = make_float2(0.f, 0.f), c = make_float2(0.f, 0.f);
      if (x < G) {
        const size_t node = static_cast<size_t>(row) * G + x;
        const long long m = static_cast<long long>(grids[node]);
        const long long mxv = static_cast<long long>(grids[gg + node]);
        const long long myv = static_cast<long long>(grids[2 * gg + node]);
        grids[node] = 0ull;
        grids[gg + node] = 0ull;
        grids[2 * gg + node] = 0ull;
        a = make_float2(static_cast<float>(static_cast<double>(m) * inv_m),
                        static_cast<float>(static_cast<double>(mxv) * inv_xs));
        c = make_float2(static_cast<float>(static_cast<double>(myv) * inv_xy), 0.f);
      }
      buf[x] = a;
      buf[F + x] = c;
    }
    __syncthreads();
    fft_lds<F, 2, kThreads, false>(buf, tw);
    float2* o1 = z1 + static_cast<size_t>(row) * F;
    float2* o2 = z2 + static_cast<size_t>(row) * F;
    for (int x = threadIdx.x; x < F; x += kThreads) {
      o1[x] = buf[x];
      o2[x] = buf[F + x];
    }
  } else {
    const int dy0 = b - G;
    const float h = st->spacing;
    for (int t = threadIdx.x; t < 2 * F; t += kThreads) {
      const int which = t / F, xx = t - which * F;
      const int dy = dy0 + which * (G / 2);
      float2 v = make_float2(0.f, 0.f);
      if (xx != G) {
        const int dx = xx < G ? xx : xx - F;
        // tsne_neighbors.cpp:976-991: offsets in spacing units, q and q^2
        const float xo = static_cast<float>(dx) * h, yo = static_cast<float>(dy) * h;
        const float q = 1.0f / (1.0f + (xo * xo + yo * yo));
        v = make_float2(q, q * q);
      }
      buf[t] = v;
    }
    __syncthreads();
    fft_lds<F, 2, kThreads, false>(buf, tw);
    for (int t = threadIdx.x; t < 2 * (G + 1); t += kThreads) {
      const int which = t / (G + 1), k2 = t - which * (G + 1);
      const int dy = dy0 + which * (G / 2);
      kt[static_cast<size_t>(k2) * G + dy] = buf[which * F + k2];
    }
  }
}

template <int F>
__global__ __launch_bounds__(kThreads) void conv_rows_fwd(unsigned long long* grids,
                                                          const TsneStats* __restrict__ st,
                                                          const float2* __restrict__ tw,
                                                          float2* __restrict__ z1,
                                                          float2* __restrict__ z2,
                                                          float2* __restrict__ kt) {
  conv_rows_fwd_body<F>(grids, st, tw, z1, z2, kt, blockIdx.x);
}

// ----------------------------------------------------------------------------
// 3. column pass for the pair (c, F - c), c in [0, G].  LDS arrays:
//    0: z1 col c, 1: z2 col c, 2: z1 col F-c, 3: z2 col F-c, 4: kernel col.
// Output O1 = A*Kq + i*X*Kq2 (-> q, xq2) and O2 = A*Kq2 + i*Y*Kq2 (-> q2,
// yq2) for rows y < G, where A = FFT(mass), X = FFT(mass_x), Y = FFT(mass_y).
// The 1/F^2 of the inverse transform is folded into the kernel spectra.
// `lds` holds 5*F float2 (dynamic shared memory of the launching kernel).
// ----------------------------------------------------------------------------
template <int F>
__device__ __forceinline__ void conv_cols_body(const float2* __restrict__ z1,
                                               const float2* __restrict__ z2,
                                               const float2* __restrict__ kt,
                                               const float2* __restrict__ tw,
                                               float2* __restrict__ o1, float2* __restrict__ o2,
                                               double* partials, TsneStats* st, float2* lds,
                                               int c, int nblk) {
  constexpr int G = F / 2;
  float2* b0 = lds;
  float2* b1 = lds + F;
  float2* b2 = lds + 2 * F;
  float2* b3 = lds + 3 * F;
  float2* bk = lds + 4 * F;
  const int cp = (F - c) & (F - 1);
  const bool pair = (cp != c);
  for (int y = threadIdx.x; y < F; y += kThreads) {
    float2 a = make_float2(0.f, 0.f), bb = a, cc = a, dd = a;
    if (y < G) {
      const size_t r = static_cast<size_t>(y) * F;
      a = z1[r + c];
      bb = z2[r + c];
      cc = z1[r + cp];
      dd = z2[r + cp];
    }
    b0[y] = a;
    b1[y] = bb;
    b2[y] = cc;
    b3[y] = dd;
    const int d = y <= G ? y : F - y;
    bk[y] = d < G ? kt[static_cast<size_t>(c) * G + d] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<F, 5, kThreads, false>(lds, tw);
  constexpr float inv_ff = 1.0f / (static_cast<float>(F) * static_cast<float>(F));
  double zpart = 0.0;
  for (int t = threadIdx.x; t < F; t += kThreads) {
    const int k1 = t, k1m = (F - t) & (F - 1);
    const float2 z1a = b0[k1], z1b = b2[k1m];
    const float2 ya = b1[k1], yb = b3[k1m];
    const float kq = bk[k1].x * inv_ff, k2 = bk[k1].y * inv_ff;
    // A = (z + conj(zm))/2, X = (z - conj(zm))/(2i)
    const float2 A = make_float2(0.5f * (z1a.x + z1b.x), 0.5f * (z1a.y - z1b.y));
    const float2 X = make_float2(0.5f * (z1a.y + z1b.y), -0.5f * (z1a.x - z1b.x));
    // partner column values are the conjugates (Hermitian symmetry)
    const float2 Ab = cconj(A), Xb = cconj(X);
    const float2 o1a = make_float2(A.x * kq - X.y * k2, A.y * kq + X.x * k2);
    const float2 o2a = make_float2(A.x * k2 - ya.y * k2, A.y * k2 + ya.x * k2);
    const float2 o1b = make_float2(Ab.x * kq - Xb.y * k2, Ab.y * kq + Xb.x * k2);
    const float2 o2b = make_float2(Ab.x * k2 - yb.y * k2, Ab.y * k2 + yb.x * k2);
    zpart += 2.0 * (static_cast<double>(A.x) * A.x + static_cast<double>(A.y) * A.y) *
             static_cast<double>(kq);
    b0[k1] = o1a;
    b1[k1] = o2a;
    b2[k1m] = o1b;
    b3[k1m] = o2b;
  }
  // self-paired columns (c = 0, c = G) visit every element twice
  if (!pair) zpart *= 0.5;
  __syncthreads();
  fft_lds<F, 4, kThreads, true>(lds, tw);
  for (int y = threadIdx.x; y < G; y += kThreads) {
    const size_t r = static_cast<size_t>(y) * F;
    o1[r + c] = b0[y];
    o2[r + c] = b1[y];
    if (pair) {
      o1[r + cp] = b2[y];
      o2[r + cp] = b3[y];
    }
  }
  // Z partial: block reduce then last-block sum in column order
  __shared__ double red[kThreads / 64];
  __shared__ bool last;
  zpart = wave_sum(zpart);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = zpart;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) s += red[w];
    partials[c] = s;
    __threadfence();
    last = (atomicAdd(&st->done_cols, 1u) == static_cast<unsigned>(nblk - 1));
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    const volatile double* vp = partials;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += vp[b];
    // Parseval: sum_nodes mass*q = (1/F^2) sum_k |A(k)|^2 Kq_hat(k); kq above
    // already carries the 1/F^2.
    st->sum_q_total = s;
    st->done_cols = 0u;
  }
}

// XCD-aware column order: workgroups are dealt round-robin to the eight XCDs,
// so block b runs on XCD b % 8.  Giving each XCD a contiguous range of
// columns makes the blocks that share 128-byte lines of the strided column
// reads and writes run on one XCD, where they hit the same L2.  A launch has
// conv_cols_blocks(G) blocks; the surplus blocks of the last range exit.
__host__ __device__ constexpr int conv_cols_per_xcd(int g) { return (g + 1 + 7) / 8; }
__host__ __device__ constexpr int conv_cols_blocks(int g) { return 8 * conv_cols_per_xcd(g); }

__device__ __forceinline__ int xcd_column(int b, int g) {
  return (b & 7) * conv_cols_per_xcd(g) + (b >> 3);
}

template <int F>
__global__ __launch_bounds__(kThreads) void conv_cols(const float2* __restrict__ z1,
                                                      const float2* __restrict__ z2,
                                                      const float2* __restrict__ kt,
                                                      const float2* __restrict__ tw,
                                                      float2* __restrict__ o1,
                                                      float2* __restrict__ o2,
                                                      double* partials, TsneStats* st) {
  extern __shared__ float2 lds[];
  constexpr int G = F / 2;
  const int c = xcd_column(blockIdx.x, G);
  if (c > G) return;
  conv_cols_body<F>(z1, z2, kt, tw, o1, o2, partials, st, lds, c, G + 1);
}

// ----------------------------------------------------------------------------
// 4. inverse row FFTs for rows y < G; field(y, x) = (q2, xq2, yq2, q).
// ----------------------------------------------------------------------------
template <int F>
__device__ __forceinline__ void conv_rows_inv_body(const float2* __restrict__ o1,
                                                   const float2* __restrict__ o2,
                                                   const float2* __restrict__ tw,
                                                   const TsneStats* __restrict__ st,
                                                   float4* __restrict__ field, int row) {
  constexpr int G = F / 2;
  __shared__ float2 buf[2 * F];
  const float2* s1 = o1 + static_cast<size_t>(row) * F;
  const float2* s2 = o2 + static_cast<size_t>(row) * F;
  for (int x = threadIdx.x; x < F; x += kThreads) {
    buf[x] = s1[x];
    buf[F + x] = s2[x];
  }
  __syncthreads();
  fft_lds<F, 2, kThreads, true>(buf, tw);
  float4* out = field + static_cast<size_t>(row) * G;
  const double xs = st->x_scale;
  for (int x = threadIdx.x; x < G; x += kThreads) {
    const float2 a = buf[x], b = buf[F + x];
    out[x] = make_float4(b.x, static_cast<float>(static_cast<double>(a.y) * xs), b.y, a.x);
  }
}

template <int F>
__global__ __launch_bounds__(kThreads) void conv_rows_inv(const float2* __restrict__ o1,
                                                          const float2* __restrict__ o2,
                                                          const float2* __restrict__ tw,
                                                          const TsneStats* __restrict__ st,
                                                          float4* __restrict__ field) {
  conv_rows_inv_body<F>(o1, o2, tw, st, field, blockIdx.x);
}

// ----------------------------------------------------------------------------
// 5a. attraction (add_sparse_attractive_gradient_f, tsne_neighbors.cpp:
//     1061-1088; the CUDA directed-edge twin is embedding_cuda_kernels.cpp:
//     2788-2817).  TPR lanes own one point's CSR row; each lane keeps U
//     entries in flight (U independent col/val loads, then U independent
//     neighbour loads) so the row costs two dependent memory round trips per
//     U*TPR entries.  Lane-strided sums and a fixed butterfly reduction keep
//     it deterministic.  attr[i] = exaggeration * sum_j P_ij q_ij (y_i - y_j).
// ----------------------------------------------------------------------------
struct StepParams {
  float exaggeration, learning_rate, momentum, min_gain, max_step_norm;
};

__device__ __forceinline__ float sign_tsne(float x) { return x == 0.0f ? 0.0f : (x < 0.0f ? -1.0f : 1.0f); }

constexpr int kAttrUnroll = 4;

template <int TPR>
__device__ __forceinline__ void attract_body(const float2* __restrict__ y_in,
                                             const int* __restrict__ row_ptr,
                                             const int* __restrict__ col,
                                             const float* __restrict__ val,
                                             const TsneStats* __restrict__ st, float exaggeration,
                                             float2* __restrict__ attr, int n, int blk, int nblk) {
  constexpr int U = kAttrUnroll;
  const int lane = threadIdx.x % TPR;
  const int group = (blk * kThreads + static_cast<int>(threadIdx.x)) / TPR;
  const int n_groups = nblk * (kThreads / TPR);
  const float mxf = st->mean_xf, myf = st->mean_yf;
  for (int i = group; i < n; i += n_groups) {
    const float2 pi = y_in[i];
    const float xi = pi.x - mxf, yi = pi.y - myf;
    float ax = 0.0f, ay = 0.0f;
    const int beg = row_ptr[i], end = row_ptr[i + 1];
    for (int base = beg + lane; base < end; base += U * TPR) {
      int j[U];
      float v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int p = base + u * TPR;
        const bool ok = p < end;
        j[u] = ok ? col[p] : i;
        v[u] = ok ? val[p] : 0.0f;
      }
      float2 pj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) pj[u] = y_in[j[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float dx = xi - (pj[u].x - mxf), dy = yi - (pj[u].y - myf);
        const float q = 1.0f / (1.0f + dx * dx + dy * dy);
        const float cf = exaggeration * v[u] * q;
        ax += cf * dx;
        ay += cf * dy;
      }
    }
#pragma unroll
    for (int o = TPR / 2; o > 0; o >>= 1) {
      ax += __shfl_xor(ax, o, TPR);
      ay += __shfl_xor(ay, o, TPR);
    }
    if (lane == 0) attr[i] = make_float2(ax, ay);
  }
}

template <int TPR>
__global__ __launch_bounds__(kThreads) void attract_kernel(const float2* __restrict__ y_in,
                                                           const int* __restrict__ row_ptr,
                                                           const int* __restrict__ col,
                                                           const float* __restrict__ val,
                                                           const TsneStats* __restrict__ st,
                                                           float exaggeration,
                                                           float2* __restrict__ attr, int n) {
  attract_body<TPR>(y_in, row_ptr, col, val, st, exaggeration, attr, n, blockIdx.x, gridDim.x);
}

// ----------------------------------------------------------------------------
// 5b. gather + update, one thread per point (coalesced).  Back-interpolation
//     of (q2, xq2, yq2) at the point's cell (tsne_neighbors.cpp:1247-1258),
//     rep = -(y q2 - xq2, y q2 - yq2) / Z, grad = rep + attr.
// MODE 0: optimiser step (apply_open_tsne_update_f, tsne_neighbors.cpp:
//         1297-1336; CUDA twin embedding_cuda_kernels.cpp:2819-2886) into
//         y_out plus the layout statistics for the next iteration; returns
//         true in thread 0 of the block that finalised them.
// MODE 1: diagnostic; writes the repulsive force to diag_rep.
// ----------------------------------------------------------------------------
template <int MODE>
__device__ __forceinline__ bool update_body(const float2* __restrict__ y_in,
                                            float2* __restrict__ y_out,
                                            float2* __restrict__ gains, float2* __restrict__ upd,
                                            const float2* __restrict__ attr,
                                            const float4* __restrict__ field, TsneStats* st,
                                            LayoutPartial* partials, int n, int g, StepParams sp,
                                            float2* __restrict__ diag_rep, int blk, int nblk) {
  const float mxf = st->mean_xf, myf = st->mean_yf;
  const float lx = st->lower_x, ly = st->lower_y, inv = st->inv_spacing;
  const float inv_z = static_cast<float>(
      1.0 / fmax(st->sum_q_total - static_cast<double>(n), static_cast<double>(FLT_MIN)));
  LayoutPartial m{0.0, 0.0, FLT_MAX, -FLT_MAX, FLT_MAX, -FLT_MAX};
  for (int i = blk * kThreads + threadIdx.x; i < n; i += nblk * kThreads) {
    const float2 pi = y_in[i];
    const float xi = pi.x - mxf, yi = pi.y - myf;
    const Cell c = grid_cell(xi, yi, lx, ly, inv, g);
    const size_t p00 = static_cast<size_t>(c.y0) * g + c.x0;
    const float4 f00 = field[p00], f10 = field[p00 + 1];
    const float4 f01 = field[p00 + g], f11 = field[p00 + g + 1];
    const float w00 = (1.0f - c.tx) * (1.0f - c.ty), w10 = c.tx * (1.0f - c.ty);
    const float w01 = (1.0f - c.tx) * c.ty, w11 = c.tx * c.ty;
    const float q2 = w00 * f00.x + w10 * f10.x + w01 * f01.x + w11 * f11.x;
    const float xq2 = w00 * f00.y + w10 * f10.y + w01 * f01.y + w11 * f11.y;
    const float yq2 = w00 * f00.z + w10 * f10.z + w01 * f01.z + w11 * f11.z;
    const float rx = -(xi * q2 - xq2) * inv_z;
    const float ry = -(yi * q2 - yq2) * inv_z;
    if constexpr (MODE == 1) {
      diag_rep[i] = make_float2(rx, ry);
    } else {
      const float2 a = attr[i];
      const float gx = rx + a.x, gy = ry + a.y;
      float2 gn = gains[i];
      const float2 old = upd[i];
      gn.x = sign_tsne(old.x) != sign_tsne(gx) ? gn.x + 0.2f : gn.x * 0.8f + sp.min_gain;
      gn.y = sign_tsne(old.y) != sign_tsne(gy) ? gn.y + 0.2f : gn.y * 0.8f + sp.min_gain;
      gn.x = fmaxf(gn.x, sp.min_gain);
      gn.y = fmaxf(gn.y, sp.min_gain);
      float sx = sp.momentum * old.x - sp.learning_rate * gn.x * gx;
      float sy = sp.momentum * old.y - sp.learning_rate * gn.y * gy;
      if (isfinite(sp.max_step_norm) && sp.max_step_norm > 0.0f) {
        const float norm2 = sx * sx + sy * sy;
        if (norm2 > sp.max_step_norm * sp.max_step_norm) {
          const float scale = sp.max_step_norm / sqrtf(fmaxf(norm2, FLT_MIN));
          sx *= scale;
          sy *= scale;
        }
      }
      const float nx = xi + sx, ny = yi + sy;
      gains[i] = gn;
      upd[i] = make_float2(sx, sy);
      y_out[i] = make_float2(nx, ny);
      m.sx += nx;
      m.sy += ny;
      m.mnx = fminf(m.mnx, nx);
      m.mxx = fmaxf(m.mxx, nx);
      m.mny = fminf(m.mny, ny);
      m.mxy = fmaxf(m.mxy, ny);
    }
  }
  if constexpr (MODE == 0) return reduce_layout_stats(m, partials, st, n, g, true, blk, nblk);
  return false;
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void update_kernel(
    const float2* __restrict__ y_in, float2* __restrict__ y_out, float2* __restrict__ gains,
    float2* __restrict__ upd, const float2* __restrict__ attr, const float4* __restrict__ field,
    TsneStats* st, LayoutPartial* partials, int n, int g, StepParams sp,
    float2* __restrict__ diag_rep) {
  (void)update_body<MODE>(y_in, y_out, gains, upd, attr, field, st, partials, n, g, sp, diag_rep,
                          blockIdx.x, gridDim.x);
}

// Write the centred layout (final output): out_i = fl(y_i - mean).
__global__ void centred_copy_kernel(const float2* __restrict__ y, const TsneStats* __restrict__ st,
                                    float2* __restrict__ out, int n) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) {
    const float2 p = y[i];
    out[i] = make_float2(p.x - st->mean_xf, p.y - st->mean_yf);
  }
}

}  // namespace
}  // namespace fe
