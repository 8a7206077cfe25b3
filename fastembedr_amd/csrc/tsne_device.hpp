// t-SNE device kernels for gfx950.
//
// Per-iteration pipeline (all state resident in HBM):
//   1. splat            bilinear spread of every point onto the G x G grid
//                       (mass, mass*x, mass*y): points are binned by grid
//                       tile (count, column scan, scatter), then blocks sum
//                       every tile's points in 64-bit fixed point in LDS and
//                       write its nodes once, so the grid does not depend on
//                       arrival order;
//   2. conv_rows_fwd    row FFTs of z1 = mass + i*mass_x/x_scale and
//                       z2 = mass_y (zero padded to F = 2G), and of the G
//                       kernel rows Kq + i*Kq^2 (both real and even, so the
//                       spectra come out as real + i*real);
//   3. conv_cols        per column pair (c, F-c): column FFTs, Hermitian
//                       split of z1 into FFT(mass) and FFT(mass_x), the
//                       pointwise products with the kernel spectra, inverse
//                       column FFTs, and Z by Parseval in FP64;
//   4. conv_rows_inv    inverse row FFTs; writes the (q2, xq2, yq2, q) field;
//   5. attract          attraction over the symmetric P (tsne_pairs.hpp:
//                       every pair once; or the CSR kernels below), on a
//                       second stream beside 1-4;
//   6. update           per point: bilinear gather of the field, the
//                       gains/momentum/clip update, and the FP64 layout
//                       statistics (mean, bounds) for the next iteration.
// Centring is applied lazily: the layout buffer holds uncentred coordinates
// and every reader subtracts the float mean kept in TsneStats, which gives
// bit-for-bit the values an explicit centring pass would have written.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
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

// Workgroups are dealt round-robin to the eight XCDs, so block b runs on XCD
// b % 8.  xcd_remap gives the blocks of one XCD one contiguous range of
// logical block ids [x*per + min(x, rem), ...), so work on neighbouring
// points or rows (which share neighbours, grid cells and cache lines) stays
// in one XCD's L2.  It is a permutation of [0, nblk).
__host__ __device__ __forceinline__ int xcd_remap(int b, int nblk) {
  const int per = nblk / 8, rem = nblk % 8, x = b % 8;
  return x * per + (x < rem ? x : rem) + b / 8;
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

// Cross-block results are combined at kernel boundaries: a block writes its
// partial and a later launch folds the partials.  (A last-block reduction
// inside one launch would need a device-scope release fence per block, which
// on this GPU writes back the whole L2 of the block's XCD.)

__device__ __forceinline__ LayoutPartial combine(LayoutPartial a, const LayoutPartial& b) {
  a.sx += b.sx;
  a.sy += b.sy;
  a.mnx = fminf(a.mnx, b.mnx);
  a.mxx = fmaxf(a.mxx, b.mxx);
  a.mny = fminf(a.mny, b.mny);
  a.mxy = fmaxf(a.mxy, b.mxy);
  return a;
}

// Block reduction of LayoutPartial (fixed wave/LDS tree); thread 0 returns
// the block's value.
__device__ inline LayoutPartial block_layout(LayoutPartial mine) {
  __shared__ LayoutPartial red[kThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  mine.sx = wave_sum(mine.sx);
  mine.sy = wave_sum(mine.sy);
  mine.mnx = static_cast<float>(wave_min(mine.mnx));
  mine.mxx = static_cast<float>(wave_max(mine.mxx));
  mine.mny = static_cast<float>(wave_min(mine.mny));
  mine.mxy = static_cast<float>(wave_max(mine.mxy));
  __syncthreads();
  if (lane == 0) red[wid] = mine;
  __syncthreads();
  LayoutPartial b = red[0];
  for (int w = 1; w < kThreads / 64; ++w) b = combine(b, red[w]);
  return b;
}

// One block folds the nblk partials (thread t takes t, t + kThreads, ...,
// then the block tree) and sets the geometry of the next iteration.
__device__ inline void fold_layout(const LayoutPartial* __restrict__ partials, int nblk,
                                   TsneStats* st, int n, int g, bool centre) {
  LayoutPartial a{0.0, 0.0, FLT_MAX, -FLT_MAX, FLT_MAX, -FLT_MAX};
  for (int b = threadIdx.x; b < nblk; b += kThreads) a = combine(a, partials[b]);
  const LayoutPartial t = block_layout(a);
  if (threadIdx.x == 0)
    finalize_geometry(st, t.sx, t.sy, t.mnx, t.mxx, t.mny, t.mxy, n, g, centre);
}

__global__ __launch_bounds__(kThreads) void layout_finalize(const LayoutPartial* __restrict__ partials,
                                                            int nblk, TsneStats* st, int n, int g,
                                                            int centre) {
  fold_layout(partials, nblk, st, n, g, centre != 0);
}

// ----------------------------------------------------------------------------
// Layout statistics of a raw layout (initialisation; zero_mean_f and the
// bounds pass of the reference).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void layout_stats_kernel(const float2* __restrict__ y,
                                                                LayoutPartial* partials, int n) {
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
  const LayoutPartial b = block_layout(m);
  if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

// ----------------------------------------------------------------------------
// Every stage below is a __device__ body taking its block index within the
// job (blk) and the job's block count (nblk); the single-job kernels pass
// blockIdx.x / gridDim.x and the job-table kernels of tsne_batch.hpp pass the
// job-local values, so a job gives the same bits alone and inside a batch.
//
// 1. splat: embedding_cuda_kernels.cpp:2622-2667 / tsne_neighbors.cpp:
//    1186-1226, accumulated in 64-bit fixed point (deterministic).
//
// The node grid is cut into kTile x kTile tiles; tile (tx, ty) owns nodes
// [kTile*tx, kTile*tx + kTile) x [kTile*ty, kTile*ty + kTile).  A point in
// bilinear cell (x0, y0) touches nodes x0..x0+1 x y0..y0+1: its cell's tile,
// and also the tile to the right when x0 is that tile's last cell column, the
// tile below when y0 is its last cell row, or the diagonal one when both.
// Points are binned by (cell tile, edge class) with the classes ordered
//     0 interior, 1 right edge, 2 right and bottom edge, 3 bottom edge,
// so that what a tile owner needs from each neighbour is one contiguous range
// of bins: own tile [0, 4), left neighbour [1, 3), upper neighbour [2, 4),
// upper-left neighbour [2, 3).  Four launches, none with a global atomic on
// a shared address (those serialise across the XCDs):
//   a. splat_count    a block of kSplatThreads takes a range of kSplatChunk
//                     points (loads issued together), builds their bin
//                     histogram in LDS and stores it as its row of `rows`;
//   b. splat_colscan  per bin, the exclusive prefix of the rows down the
//                     blocks (in place) and the bin's total;
//   c. splat_scatter  every block scans the totals into bin starts in LDS
//                     (block 0 also stores them, with the first work item of
//                     every tile, for the tile pass); a point's slot is its
//                     bin start + its block's prefix + its rank among the
//                     block's points of that bin (an LDS atomic);
//   d. splat_tiles    a block takes work items of at most kSplatItem points
//                     of one tile in turn and accumulates each in an LDS window
//                     of the tile (a lane sums consecutive points of one cell
//                     in registers before touching LDS, which keeps collapsed
//                     clusters cheap).  A tile of one item writes its nodes
//                     with plain stores as the input of the row FFTs,
//                       sgrid(node) = (mass, mass*x / x_scale, mass*y, 0);
//                     the items of a larger tile add their windows to the
//                     64-bit `wide` grid, which the row FFT pass converts and
//                     clears.
// Ranks inside a bin follow LDS atomic order, but all sums are integer sums,
// so the grid is the same on every run and matches a direct splat.
// ----------------------------------------------------------------------------
constexpr int kTile = 32;
constexpr int kTileNodes = kTile * kTile;
constexpr int kMaxTilesSide = 32;  // grids up to 1024 x 1024 nodes
constexpr int kMaxTiles = kMaxTilesSide * kMaxTilesSide;
constexpr int kMaxBins = 4 * kMaxTiles;
// points per tile work item (1M points: 45 us for the four splat kernels
// against 50 with 1024-point items and 62 with 512; 4096 gave 50)
constexpr int kSplatItem = 2048;
constexpr int kItemPer = kSplatItem / kThreads;
// binning passes: blocks of kSplatThreads take kSplatChunk points; in the
// scans a thread owns kScanBins consecutive bins and kScanTiles tiles
constexpr int kSplatThreads = 1024;
constexpr int kSplatPerThread = 4;
constexpr int kSplatChunk = kSplatThreads * kSplatPerThread;
constexpr int kScanBins = kMaxBins / kSplatThreads;
constexpr int kScanTiles = kMaxTiles / kSplatThreads;
static_assert(kScanBins % 4 == 0 && kScanBins * kSplatThreads == kMaxBins &&
                  kScanTiles * kSplatThreads == kMaxTiles,
              "the binning scans split bins and tiles evenly over the threads");
static_assert(kSplatItem % kThreads == 0, "whole points per lane in the tile pass");
// column scan: a block takes kColCols bins, its waves a share of the rows
constexpr int kColCols = 64;
constexpr int kColWaves = kSplatThreads / 64;

struct SplatBuffers {
  int* rows;                 // splat_bin_blocks x bins: block histograms, then
                             // their exclusive prefixes down the blocks
  int* bin_count;            // bins: totals
  int* bin_start;            // bins + 1
  int* tile_items;           // tiles + 1: first work item of every tile
  float2* binned;            // n centred points in bin order
  unsigned long long* wide;  // 3 x g*g sums of multi-item tiles; zero between passes
  float4* sgrid;             // g*g input of the row FFTs
};

__host__ __device__ __forceinline__ int tiles_side(int g) { return (g + kTile - 1) / kTile; }
__host__ __device__ __forceinline__ int splat_bins(int g) { return 4 * tiles_side(g) * tiles_side(g); }
// launch size of splat_count and splat_scatter, and the rows of `rows`
__host__ __device__ __forceinline__ int splat_bin_blocks(int n) {
  return n > kSplatChunk ? (n + kSplatChunk - 1) / kSplatChunk : 1;
}
__host__ __device__ __forceinline__ int splat_colscan_blocks(int g) {
  return (splat_bins(g) + kColCols - 1) / kColCols;
}
// launch size of splat_tiles: a block per tile and kTileExtraBlocks more for
// tiles of several items, never more than the items can be (every tile has
// at most 1 + points/kSplatItem items and a point is counted by at most four
// tiles); blocks loop over further items
constexpr int kTileExtraBlocks = 256;
__host__ __device__ __forceinline__ int splat_item_blocks(int n, int g) {
  const long long tiles = static_cast<long long>(tiles_side(g)) * tiles_side(g);
  const long long bound = tiles + (4ll * n + kSplatItem - 1) / kSplatItem;
  return static_cast<int>(bound < tiles + kTileExtraBlocks ? bound : tiles + kTileExtraBlocks);
}

__device__ __forceinline__ int point_bin(const Cell& c, int ts) {
  const int tx = c.x0 / kTile, ty = c.y0 / kTile;
  const bool xe = (c.x0 % kTile) == kTile - 1, ye = (c.y0 % kTile) == kTile - 1;
  const int cls = xe ? (ye ? 2 : 1) : (ye ? 3 : 0);
  return 4 * (ty * ts + tx) + cls;
}

// Bin range r of the four that feed tile t (0 own, 1 left, 2 upper,
// 3 upper-left) as [x, y); a missing neighbour gives an empty range.
__device__ __forceinline__ int2 tile_range(const int* bs, int t, int ts, int r) {
  const int tx = t % ts, ty = t / ts;
  const int dx = r & 1, dy = r >> 1;
  if (tx < dx || ty < dy) return make_int2(0, 0);
  const int lo = r == 0 ? 0 : (r == 1 ? 1 : 2), hi = r == 0 ? 4 : (r == 2 ? 4 : 3);
  const int sb = 4 * ((ty - dy) * ts + (tx - dx));
  return make_int2(bs[sb + lo], bs[sb + hi]);
}

// points [beg, end) of block blk of nblk
__device__ __forceinline__ void point_range(int n, int blk, int nblk, int* beg, int* end) {
  const int chunk = (n + nblk - 1) / nblk;
  *beg = min(n, blk * chunk);
  *end = min(n, *beg + chunk);
}

// exclusive prefix sum of v over a block of NT threads in thread order
template <int NT>
__device__ __forceinline__ int block_exclusive_scan_nt(int v, int* total) {
  __shared__ int wsum[NT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    const int s = wsum[w];
    off += w < wid ? s : 0;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + incl - v;
}

// Centred coordinates and bins of the points i0 + threadIdx.x + u *
// kSplatThreads below `end` (bin -1 past it); all loads are issued first.
struct SplatLoad {
  float2 c[kSplatPerThread];
  int bin[kSplatPerThread];
};

__device__ __forceinline__ SplatLoad splat_load(const float2* __restrict__ y,
                                                const TsneStats* __restrict__ st, int i0, int end,
                                                int g) {
  const int ts = tiles_side(g);
  const float mx = st->mean_xf, my = st->mean_yf;
  const float lx = st->lower_x, ly = st->lower_y, inv = st->inv_spacing;
  SplatLoad L;
#pragma unroll
  for (int u = 0; u < kSplatPerThread; ++u) {
    const int i = i0 + threadIdx.x + u * kSplatThreads;
    L.c[u] = i < end ? y[i] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int u = 0; u < kSplatPerThread; ++u) {
    const int i = i0 + threadIdx.x + u * kSplatThreads;
    L.c[u] = make_float2(L.c[u].x - mx, L.c[u].y - my);
    L.bin[u] = i < end ? point_bin(grid_cell(L.c[u].x, L.c[u].y, lx, ly, inv, g), ts) : -1;
  }
  return L;
}

__device__ __forceinline__ void splat_count_body(const float2* __restrict__ y,
                                                 const TsneStats* __restrict__ st,
                                                 SplatBuffers sb, int n, int g, int blk,
                                                 int nblk) {
  __shared__ int hist[kMaxBins];
  const int nbins = splat_bins(g);
  for (int t = threadIdx.x; t < nbins; t += kSplatThreads) hist[t] = 0;
  __syncthreads();
  const int rb = xcd_remap(blk, nblk);
  int beg, end;
  point_range(n, rb, nblk, &beg, &end);
  for (int i0 = beg; i0 < end; i0 += kSplatChunk) {
    const SplatLoad L = splat_load(y, st, i0, end, g);
#pragma unroll
    for (int u = 0; u < kSplatPerThread; ++u)
      if (L.bin[u] >= 0) atomicAdd(&hist[L.bin[u]], 1);
  }
  __syncthreads();
  int* row = sb.rows + static_cast<size_t>(rb) * nbins;
  for (int t = threadIdx.x; t < nbins; t += kSplatThreads) row[t] = hist[t];
}

// Block blk owns bins [kColCols * blk, + kColCols); lane = bin, wave w =
// rows [w * per, (w + 1) * per) of the `nrows` block rows.
__device__ __forceinline__ void splat_colscan_body(SplatBuffers sb, int g, int nrows, int blk) {
  __shared__ int part[kColWaves][kColCols];
  constexpr int kU = 8;
  const int nbins = splat_bins(g);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int col = blk * kColCols + lane;
  const int per = (nrows + kColWaves - 1) / kColWaves;
  const int r0 = min(nrows, w * per), r1 = min(nrows, r0 + per);
  int* base = sb.rows + col;
  int s = 0;
  if (col < nbins) {
    for (int r = r0; r < r1; r += kU) {
      int v[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) v[u] = r + u < r1 ? base[static_cast<size_t>(r + u) * nbins] : 0;
#pragma unroll
      for (int u = 0; u < kU; ++u) s += v[u];
    }
  }
  part[w][lane] = s;
  __syncthreads();
  int run = 0, total = 0;
#pragma unroll
  for (int q = 0; q < kColWaves; ++q) {
    const int v = part[q][lane];
    run += q < w ? v : 0;
    total += v;
  }
  if (col >= nbins) return;
  if (w == 0) sb.bin_count[col] = total;
  for (int r = r0; r < r1; r += kU) {
    int v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) v[u] = r + u < r1 ? base[static_cast<size_t>(r + u) * nbins] : 0;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (r + u < r1) base[static_cast<size_t>(r + u) * nbins] = run;
      run += v[u];
    }
  }
}

__device__ __forceinline__ void splat_scatter_body(const float2* __restrict__ y,
                                                   const TsneStats* __restrict__ st,
                                                   SplatBuffers sb, int n, int g, int blk,
                                                   int nblk) {
  __shared__ int start[kMaxBins + 1];
  __shared__ int hist[kMaxBins];
  __shared__ int off[kMaxBins];
  const int ts = tiles_side(g), tiles = ts * ts, nbins = 4 * tiles;
  const int t = threadIdx.x, b0 = kScanBins * t;
  const int rb = xcd_remap(blk, nblk);
  const int* row = sb.rows + static_cast<size_t>(rb) * nbins;
  // bin starts: thread t scans bins b0 .. b0 + kScanBins - 1 (nbins is a
  // multiple of four, so a group of four is in or out as a whole); the
  // block's own prefixes are loaded with the totals
  int cnt[kScanBins], pre[kScanBins];
#pragma unroll
  for (int q = 0; q < kScanBins; q += 4) {
    int4 c4 = make_int4(0, 0, 0, 0), p4 = make_int4(0, 0, 0, 0);
    if (b0 + q < nbins) {
      c4 = *reinterpret_cast<const int4*>(sb.bin_count + b0 + q);
      p4 = *reinterpret_cast<const int4*>(row + b0 + q);
    }
    cnt[q] = c4.x;
    cnt[q + 1] = c4.y;
    cnt[q + 2] = c4.z;
    cnt[q + 3] = c4.w;
    pre[q] = p4.x;
    pre[q + 1] = p4.y;
    pre[q + 2] = p4.z;
    pre[q + 3] = p4.w;
  }
  int mine = 0;
#pragma unroll
  for (int q = 0; q < kScanBins; ++q) mine += cnt[q];
  int total;
  int run = block_exclusive_scan_nt<kSplatThreads>(mine, &total);
  int sv[kScanBins];
#pragma unroll
  for (int q = 0; q < kScanBins; ++q) {
    sv[q] = run;
    run += cnt[q];
    if (b0 + q < nbins) {
      start[b0 + q] = sv[q];
      off[b0 + q] = sv[q] + pre[q];
      hist[b0 + q] = 0;
    }
  }
  if (t == 0) start[nbins] = total;
  __syncthreads();
  if (blk == 0) {
    // the bin starts and the first work item of every tile, for the tile pass
#pragma unroll
    for (int q = 0; q < kScanBins; q += 4)
      if (b0 + q < nbins)
        *reinterpret_cast<int4*>(sb.bin_start + b0 + q) =
            make_int4(sv[q], sv[q + 1], sv[q + 2], sv[q + 3]);
    if (t == 0) sb.bin_start[nbins] = total;
    int items[kScanTiles];
    int my_items = 0;
#pragma unroll
    for (int q = 0; q < kScanTiles; ++q) {
      const int tt = kScanTiles * t + q;
      items[q] = 0;
      if (tt < tiles) {
        int pts = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int2 rg = tile_range(start, tt, ts, r);
          pts += rg.y - rg.x;
        }
        items[q] = max(1, (pts + kSplatItem - 1) / kSplatItem);
      }
      my_items += items[q];
    }
    int all;
    int first = block_exclusive_scan_nt<kSplatThreads>(my_items, &all);
#pragma unroll
    for (int q = 0; q < kScanTiles; ++q) {
      const int tt = kScanTiles * t + q;
      if (tt < tiles) sb.tile_items[tt] = first;
      first += items[q];
    }
    if (t == 0) sb.tile_items[tiles] = all;
  }
  int beg, end;
  point_range(n, rb, nblk, &beg, &end);
  for (int i0 = beg; i0 < end; i0 += kSplatChunk) {
    const SplatLoad L = splat_load(y, st, i0, end, g);
#pragma unroll
    for (int u = 0; u < kSplatPerThread; ++u)
      if (L.bin[u] >= 0) {
        const int r = atomicAdd(&hist[L.bin[u]], 1);
        sb.binned[off[L.bin[u]] + r] = L.c[u];
      }
  }
}

// adds a lane's pending cell sums (corner q: acc[3q .. 3q+2]) to the window
__device__ __forceinline__ void flush_cell(unsigned long long* win, int key, int g, int ox, int oy,
                                           const unsigned long long* acc) {
  const int x0 = key % g, y0 = key / g;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int nx = x0 + (q & 1) - ox, ny = y0 + (q >> 1) - oy;
    if (nx >= 0 && nx < kTile && ny >= 0 && ny < kTile) {
      const int t = ny * kTile + nx;
      atomicAdd(win + t, acc[3 * q]);
      atomicAdd(win + kTileNodes + t, acc[3 * q + 1]);
      atomicAdd(win + 2 * kTileNodes + t, acc[3 * q + 2]);
    }
  }
}

__device__ __forceinline__ float4 node_value(long long m, long long mxv, long long myv,
                                             double inv_m, double inv_xs, double inv_xy) {
  return make_float4(static_cast<float>(static_cast<double>(m) * inv_m),
                     static_cast<float>(static_cast<double>(mxv) * inv_xs),
                     static_cast<float>(static_cast<double>(myv) * inv_xy), 0.f);
}

// One work item: window cleared, points summed, nodes written.  Starts and
// ends with a barrier (the window and titems are shared with the block's
// previous item).
__device__ __forceinline__ void splat_tile_item(const TsneStats* __restrict__ st, SplatBuffers sb,
                                                int g, int item, unsigned long long* win,
                                                const int* titems) {
  const int ts = tiles_side(g), tiles = ts * ts;
  __syncthreads();
  for (int t = threadIdx.x; t < 3 * kTileNodes; t += kThreads) win[t] = 0ull;
  __syncthreads();
  int lo = 0, hi = tiles - 1;  // tile: last t with titems[t] <= item
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (titems[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  const int tile = lo;
  const int first = titems[tile], nitems = titems[tile + 1] - first;
  const int2 g0 = tile_range(sb.bin_start, tile, ts, 0), g1 = tile_range(sb.bin_start, tile, ts, 1);
  const int2 g2 = tile_range(sb.bin_start, tile, ts, 2), g3 = tile_range(sb.bin_start, tile, ts, 3);
  const int len0 = g0.y - g0.x, len1 = g1.y - g1.x, len2 = g2.y - g2.x;
  const int pts = len0 + len1 + len2 + (g3.y - g3.x);
  const int per = (pts + nitems - 1) / nitems;  // <= kSplatItem
  const int s0 = min(pts, (item - first) * per), s1 = min(pts, s0 + per);
  const int ox = (tile % ts) * kTile, oy = (tile / ts) * kTile;
  const float lx = st->lower_x, ly = st->lower_y, inv = st->inv_spacing;
  const double sm = st->fx_mass, sxy = st->fx_xy;
  // the item's points, all loads issued together
  float2 p[kItemPer];
#pragma unroll
  for (int u = 0; u < kItemPer; ++u) {
    // position v of the concatenated ranges own ++ left ++ upper ++ upper-left
    const int v = s0 + threadIdx.x + u * kThreads;
    int idx;
    if (v < len0) idx = g0.x + v;
    else if (v < len0 + len1) idx = g1.x + (v - len0);
    else if (v < len0 + len1 + len2) idx = g2.x + (v - len0 - len1);
    else idx = g3.x + (v - len0 - len1 - len2);
    p[u] = v < s1 ? sb.binned[idx] : make_float2(0.f, 0.f);
  }
  int pend = -1;
  unsigned long long acc[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) acc[q] = 0ull;
#pragma unroll
  for (int u = 0; u < kItemPer; ++u) {
    if (s0 + threadIdx.x + u * kThreads >= s1) break;
    const Cell c = grid_cell(p[u].x, p[u].y, lx, ly, inv, g);
    const int key = c.y0 * g + c.x0;
    if (key != pend) {
      if (pend >= 0) flush_cell(win, pend, g, ox, oy, acc);
      pend = key;
#pragma unroll
      for (int q = 0; q < 12; ++q) acc[q] = 0ull;
    }
    const float w[4] = {(1.0f - c.tx) * (1.0f - c.ty), c.tx * (1.0f - c.ty),
                        (1.0f - c.tx) * c.ty, c.tx * c.ty};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      acc[3 * q] += to_fixed(w[q], sm);
      acc[3 * q + 1] += to_fixed(w[q] * p[u].x, sxy);
      acc[3 * q + 2] += to_fixed(w[q] * p[u].y, sxy);
    }
  }
  if (pend >= 0) flush_cell(win, pend, g, ox, oy, acc);
  __syncthreads();
  const double inv_m = 1.0 / st->fx_mass, inv_xy = 1.0 / st->fx_xy;
  const double inv_xs = inv_xy / st->x_scale;
  const size_t gg = static_cast<size_t>(g) * g;
  if (nitems == 1) {
    for (int t = threadIdx.x; t < kTileNodes; t += kThreads) {
      const int nx = ox + (t % kTile), ny = oy + t / kTile;
      if (nx < g && ny < g)
        sb.sgrid[static_cast<size_t>(ny) * g + nx] =
            node_value(static_cast<long long>(win[t]), static_cast<long long>(win[kTileNodes + t]),
                       static_cast<long long>(win[2 * kTileNodes + t]), inv_m, inv_xs, inv_xy);
    }
    return;
  }
  // items of a larger tile add into the 64-bit grid; the row FFT pass
  // converts and clears it
  for (int t = threadIdx.x; t < kTileNodes; t += kThreads) {
    const int nx = ox + (t % kTile), ny = oy + t / kTile;
    if (nx < g && ny < g) {
      const size_t node = static_cast<size_t>(ny) * g + nx;
#pragma unroll
      for (int c3 = 0; c3 < 3; ++c3) {
        const unsigned long long v = win[c3 * kTileNodes + t];
        if (v) atomicAdd(sb.wide + c3 * gg + node, v);
      }
    }
  }
}

// Block blk of nblk takes the work items blk, blk + nblk, ...
__device__ __forceinline__ void splat_tiles_body(const TsneStats* __restrict__ st, SplatBuffers sb,
                                                 int g, int blk, int nblk) {
  __shared__ unsigned long long win[3 * kTileNodes];
  __shared__ int titems[kMaxTiles + 1];
  const int ts = tiles_side(g), tiles = ts * ts;
  const int total = sb.tile_items[tiles];
  if (blk >= total) return;
  for (int t = threadIdx.x; t <= tiles; t += kThreads) titems[t] = sb.tile_items[t];
  for (int item = blk; item < total; item += nblk) {
    splat_tile_item(st, sb, g, item, win, titems);
  }
}

// The splat and FFT kernels run beside the attraction in the overlapped
// iteration and are its critical path: their waves take a higher issue
// priority than the attraction's where both share a SIMD (1M points: 3780
// against 3700 iterations/s).
__device__ __forceinline__ void repulsion_wave_priority() { __builtin_amdgcn_s_setprio(1); }

__global__ __launch_bounds__(kSplatThreads) void splat_count(const float2* __restrict__ y,
                                                             const TsneStats* __restrict__ st,
                                                             SplatBuffers sb, int n, int g) {
  repulsion_wave_priority();
  splat_count_body(y, st, sb, n, g, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kSplatThreads) void splat_colscan(SplatBuffers sb, int g, int nrows) {
  repulsion_wave_priority();
  splat_colscan_body(sb, g, nrows, blockIdx.x);
}

__global__ __launch_bounds__(kSplatThreads) void splat_scatter(const float2* __restrict__ y,
                                                               const TsneStats* __restrict__ st,
                                                               SplatBuffers sb, int n, int g) {
  repulsion_wave_priority();
  splat_scatter_body(y, st, sb, n, g, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kThreads) void splat_tiles(const TsneStats* __restrict__ st,
                                                        SplatBuffers sb, int g) {
  repulsion_wave_priority();
  splat_tiles_body(st, sb, g, blockIdx.x, gridDim.x);
}

// ----------------------------------------------------------------------------
// 2. row FFTs.  Blocks [0, g): data row y.  Blocks [g, g + g/2): kernel rows
//    dy and dy + g/2.  Kernel row spectra are stored transposed,
//    kt[k2 * g + dy] for k2 in [0, g], so conv_cols reads them contiguously.
// ----------------------------------------------------------------------------
// threads of the FFT pass workgroups
constexpr int kFftThreads = 512;

template <int F>
__device__ __forceinline__ void conv_rows_fwd_body(SplatBuffers sb,
                                                   const TsneStats* __restrict__ st,
                                                   const float2* __restrict__ tw,
                                                   float2* __restrict__ z1,
                                                   float2* __restrict__ z2,
                                                   float2* __restrict__ kt, int b) {
  constexpr int G = F / 2;
  __shared__ float2 buf[2 * fft_span<F>()];
  if (b < G) {
    const int row = b;
    const int ts = tiles_side(G);
    const int* trow = sb.tile_items + (row / kTile) * ts;
    constexpr size_t gg = static_cast<size_t>(G) * G;
    // the row's nodes x < G (all loads issued before the LDS writes); the
    // zero-padded half x >= G
    constexpr int kIt = (G + kFftThreads - 1) / kFftThreads;
    float4 v[kIt];
    bool split[kIt];
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int x = threadIdx.x + it * kFftThreads;
      const int t = min(x, G - 1) / kTile;
      split[it] = x < G && trow[t + 1] - trow[t] > 1;
    }
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int x = threadIdx.x + it * kFftThreads;
      v[it] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (x < G && !split[it]) v[it] = sb.sgrid[static_cast<size_t>(row) * G + x];
    }
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      if (!split[it]) continue;
      // node of a tile split over several work items: its 64-bit sums
      const size_t node = static_cast<size_t>(row) * G + threadIdx.x + it * kFftThreads;
      const double inv_m = 1.0 / st->fx_mass, inv_xy = 1.0 / st->fx_xy;
      const long long m = static_cast<long long>(sb.wide[node]);
      const long long mxv = static_cast<long long>(sb.wide[gg + node]);
      const long long myv = static_cast<long long>(sb.wide[2 * gg + node]);
      sb.wide[node] = 0ull;
      sb.wide[gg + node] = 0ull;
      sb.wide[2 * gg + node] = 0ull;
      v[it] = node_value(m, mxv, myv, inv_m, inv_xy / st->x_scale, inv_xy);
    }
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int x = threadIdx.x + it * kFftThreads;
      if (x < G) {
        buf[fft_at<F>(0, x)] = make_float2(v[it].x, v[it].y);
        buf[fft_at<F>(1, x)] = make_float2(v[it].z, 0.f);
      }
    }
    for (int x = G + threadIdx.x; x < F; x += kFftThreads) {
      buf[fft_at<F>(0, x)] = make_float2(0.f, 0.f);
      buf[fft_at<F>(1, x)] = make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lds<F, 2, kFftThreads, false>(buf, tw);
    float2* o1 = z1 + static_cast<size_t>(row) * F;
    float2* o2 = z2 + static_cast<size_t>(row) * F;
    // z2 is the transform of a real row: columns G+1 .. F-1 are the
    // conjugates of columns G-1 .. 1 and are never read
    for (int x = threadIdx.x; x < F; x += kFftThreads) {
      o1[x] = buf[fft_at<F>(0, x)];
      if (x <= G) o2[x] = buf[fft_at<F>(1, x)];
    }
  } else {
    // kernel rows dy0 and dy0 + G/2; block G + j runs on XCD j % 8, which
    // gets a contiguous range of dy0, so the 8-byte stores of neighbouring
    // rows into one 128-byte line of the transposed spectrum meet in one L2
    const int j = b - G;
    const int dy0 = (j & 7) * (G / 16) + (j >> 3);
    const float h = st->spacing;
    for (int t = threadIdx.x; t < 2 * F; t += kFftThreads) {
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
      buf[fft_at<F>(which, xx)] = v;
    }
    __syncthreads();
    fft_lds<F, 2, kFftThreads, false>(buf, tw);
    for (int t = threadIdx.x; t < 2 * (G + 1); t += kFftThreads) {
      const int which = t / (G + 1), k2 = t - which * (G + 1);
      const int dy = dy0 + which * (G / 2);
      kt[static_cast<size_t>(k2) * G + dy] = buf[fft_at<F>(which, k2)];
    }
  }
}

template <int F>
__global__ __launch_bounds__(kFftThreads) void conv_rows_fwd(SplatBuffers sb,
                                                          const TsneStats* __restrict__ st,
                                                          const float2* __restrict__ tw,
                                                          float2* __restrict__ z1,
                                                          float2* __restrict__ z2,
                                                          float2* __restrict__ kt) {
  repulsion_wave_priority();
  conv_rows_fwd_body<F>(sb, st, tw, z1, z2, kt, blockIdx.x);
}

// ----------------------------------------------------------------------------
// 3. column pass for the pair (c, F - c), c in [0, G].  LDS arrays:
//    0: z1 col c, 1: z2 col c, 2: z1 col F-c, 3: kernel col.
// Output O1 = A*Kq + i*X*Kq2 (-> q, xq2) and O2 = A*Kq2 + i*Y*Kq2 (-> q2,
// yq2) for rows y < G, where A = FFT(mass), X = FFT(mass_x), Y = FFT(mass_y).
// mass and mass_y are real, so A(F-c, -k) = conj A(c, k) and
// Y(F-c, -k) = conj Y(c, k): column F-c of z2 is never transformed.  The
// 1/F^2 of the inverse transform is folded into the kernel spectra.
// `lds` holds 4 padded arrays, conv_cols_lds<F>() bytes (dynamic shared
// memory of the launching kernel).
// ----------------------------------------------------------------------------
template <int F>
__host__ __device__ constexpr size_t conv_cols_lds() {
  return 4 * static_cast<size_t>(fft_span<F>()) * sizeof(float2);
}

template <int F>
__device__ __forceinline__ void conv_cols_body(const float2* __restrict__ z1,
                                               const float2* __restrict__ z2,
                                               const float2* __restrict__ kt,
                                               const float2* __restrict__ tw,
                                               float2* __restrict__ o1, float2* __restrict__ o2,
                                               double* __restrict__ partials, float2* lds,
                                               int c) {
  constexpr int G = F / 2;
  constexpr int kPer = (F + kFftThreads - 1) / kFftThreads;  // spectrum entries per thread
  const int cp = (F - c) & (F - 1);
  const bool pair = (cp != c);
  // every load of the column is issued before the first LDS write
  constexpr int kIt = (F + kFftThreads - 1) / kFftThreads;
  float2 la[kIt], lb[kIt], lc[kIt], lk[kIt];
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int y = threadIdx.x + it * kFftThreads;
    la[it] = lb[it] = lc[it] = lk[it] = make_float2(0.f, 0.f);
    if (y < G) {
      const size_t r = static_cast<size_t>(y) * F;
      la[it] = z1[r + c];
      lb[it] = z2[r + c];
      lc[it] = z1[r + cp];
    }
    const int d = y <= G ? y : F - y;
    if (y < F && d < G) lk[it] = kt[static_cast<size_t>(c) * G + d];
  }
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int y = threadIdx.x + it * kFftThreads;
    if (y < F) {
      lds[fft_at<F>(0, y)] = la[it];
      lds[fft_at<F>(1, y)] = lb[it];
      lds[fft_at<F>(2, y)] = lc[it];
      lds[fft_at<F>(3, y)] = lk[it];
    }
  }
  __syncthreads();
  fft_lds<F, 4, kFftThreads, false>(lds, tw);
  constexpr float inv_ff = 1.0f / (static_cast<float>(F) * static_cast<float>(F));
  // the kernel spectrum is read before its array takes column F-c of O2
  float2 kv[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int t = threadIdx.x + j * kFftThreads;
    kv[j] = t < F ? lds[fft_at<F>(3, t)] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  double zpart = 0.0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int t = threadIdx.x + j * kFftThreads;
    if (t >= F) continue;
    const int k1 = t, k1m = (F - t) & (F - 1);
    const float2 z1a = lds[fft_at<F>(0, k1)], z1b = lds[fft_at<F>(2, k1m)];
    const float2 ya = lds[fft_at<F>(1, k1)], yb = cconj(ya);
    const float kq = kv[j].x * inv_ff, k2 = kv[j].y * inv_ff;
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
    lds[fft_at<F>(0, k1)] = o1a;
    lds[fft_at<F>(1, k1)] = o2a;
    lds[fft_at<F>(2, k1m)] = o1b;
    lds[fft_at<F>(3, k1m)] = o2b;
  }
  // self-paired columns (c = 0, c = G) visit every element twice
  if (!pair) zpart *= 0.5;
  __syncthreads();
  fft_lds<F, 4, kFftThreads, true>(lds, tw);
  for (int y = threadIdx.x; y < G; y += kFftThreads) {
    const size_t r = static_cast<size_t>(y) * F;
    o1[r + c] = lds[fft_at<F>(0, y)];
    o2[r + c] = lds[fft_at<F>(1, y)];
    if (pair) {
      o1[r + cp] = lds[fft_at<F>(2, y)];
      o2[r + cp] = lds[fft_at<F>(3, y)];
    }
  }
  // Z partial of the column pair; conv_rows_inv folds the partials
  __shared__ double red[kFftThreads / 64];
  zpart = wave_sum(zpart);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = zpart;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kFftThreads / 64; ++w) t += red[w];
    partials[c] = t;
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
__global__ __launch_bounds__(kFftThreads) void conv_cols(const float2* __restrict__ z1,
                                                      const float2* __restrict__ z2,
                                                      const float2* __restrict__ kt,
                                                      const float2* __restrict__ tw,
                                                      float2* __restrict__ o1,
                                                      float2* __restrict__ o2,
                                                      double* __restrict__ partials) {
  repulsion_wave_priority();
  extern __shared__ float2 lds[];
  constexpr int G = F / 2;
  const int c = xcd_column(blockIdx.x, G);
  if (c > G) return;
  conv_cols_body<F>(z1, z2, kt, tw, o1, o2, partials, lds, c);
}

// ----------------------------------------------------------------------------
// 4. inverse row FFTs for rows y < G; field(y, x) = (q2, xq2, yq2, q).
// ----------------------------------------------------------------------------
template <int F>
__device__ __forceinline__ void conv_rows_inv_body(const float2* __restrict__ o1,
                                                   const float2* __restrict__ o2,
                                                   const float2* __restrict__ tw,
                                                   const double* __restrict__ zpart,
                                                   TsneStats* st, float4* __restrict__ field,
                                                   int row) {
  constexpr int G = F / 2;
  __shared__ float2 buf[2 * fft_span<F>()];
  if (row == 0) {
    // Z: fold the G + 1 column-pair partials in a fixed order.  Parseval:
    // sum_nodes mass*q = (1/F^2) sum_k |A(k)|^2 Kq_hat(k) (the kernel
    // spectra already carry the 1/F^2)
    __shared__ double red[kFftThreads / 64];
    double a = 0.0;
    for (int c = threadIdx.x; c <= G; c += kFftThreads) a += zpart[c];
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < kFftThreads / 64; ++w) t += red[w];
      st->sum_q_total = t;
    }
  }
  const float2* s1 = o1 + static_cast<size_t>(row) * F;
  const float2* s2 = o2 + static_cast<size_t>(row) * F;
  {
    constexpr int kIt = (F + kFftThreads - 1) / kFftThreads;
    float2 a[kIt], b[kIt];
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int x = threadIdx.x + it * kFftThreads;
      if (x < F) {
        a[it] = s1[x];
        b[it] = s2[x];
      }
    }
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int x = threadIdx.x + it * kFftThreads;
      if (x < F) {
        buf[fft_at<F>(0, x)] = a[it];
        buf[fft_at<F>(1, x)] = b[it];
      }
    }
  }
  __syncthreads();
  fft_lds<F, 2, kFftThreads, true>(buf, tw);
  float4* out = field + static_cast<size_t>(row) * G;
  const double xs = st->x_scale;
  for (int x = threadIdx.x; x < G; x += kFftThreads) {
    const float2 a = buf[fft_at<F>(0, x)], b = buf[fft_at<F>(1, x)];
    out[x] = make_float4(b.x, static_cast<float>(static_cast<double>(a.y) * xs), b.y, a.x);
  }
}

template <int F>
__global__ __launch_bounds__(kFftThreads) void conv_rows_inv(const float2* __restrict__ o1,
                                                          const float2* __restrict__ o2,
                                                          const float2* __restrict__ tw,
                                                          const double* __restrict__ zpart,
                                                          TsneStats* st,
                                                          float4* __restrict__ field) {
  repulsion_wave_priority();
  conv_rows_inv_body<F>(o1, o2, tw, zpart, st, field, blockIdx.x);
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
  // a contiguous range of rows per block, the ranges of one XCD adjacent
  int beg, end;
  point_range(n, xcd_remap(blk, nblk), nblk, &beg, &end);
  const float mxf = st->mean_xf, myf = st->mean_yf;
  for (int i = beg + static_cast<int>(threadIdx.x) / TPR; i < end; i += kThreads / TPR) {
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
// 5a'. segment-blocked attraction (the terms of attract_body summed in another
//     fixed order).
//
//     Groups.  The rows of P are cut into groups of consecutive rows with
//     about the same number of entries; a workgroup of 16 waves takes one
//     group.  Each group has up to kMaxWindows windows: ranges of
//     kWindowPoints consecutive layout points chosen when the plan is built
//     to hold many of the group's columns (neighbours lie close in index
//     order when the input groups similar points).  The workgroup makes one
//     pass per window, copying the window into LDS, and a last pass for the
//     entries outside every window.
//
//     Segments.  Row r's entries whose columns fall in window w form the
//     segment (r, w); the rest form its last segment.  A segment keeps the
//     row's column order.  Segments are laid out group by group, pass by
//     pass, row by row, each padded to a multiple of 8 entries with entries
//     (column = the row itself, value 0) whose terms are exactly zero, so
//     that every aligned group of 8 entries (one lane's share of a step)
//     lies in one segment.
//
//     Waves.  A pass's entries are cut at 512-entry step boundaries into 16
//     ranges of about the same number of steps, one per wave; lane l of a
//     step takes entries 512 s + 8 l .. + 7 and their row lane_row[64 s + l],
//     sums their terms in order and scales the sum by the exaggeration.  Bit
//     l of starts[s] is set when lane l of step s begins a segment.  One
//     segmented scan over the 64 lanes per step (a segment of lanes begins at
//     a lane that begins a segment) carries each segment's sum to the lane
//     where the next one begins, which adds it to the force of the row of the
//     lane below it; lane 63 carries the open segment and its row into the
//     next step.  The segment open where a wave's range begins (its head) and
//     the one open where it ends (its tail) are left in LDS; after the pass
//     one wave joins them across the wave boundaries (a tail, then the heads
//     of the following waves up to the next segment start, in that order)
//     and adds each joined segment to its row.  A row's force is therefore
//     zero plus its segment sums in pass order, each formed in an order fixed
//     by the layout, whichever workgroups run them.
//
//     The padded arrays carry kLanePad entries past the end so that whole
//     steps and the prefetch of the next step load unconditionally.
// ----------------------------------------------------------------------------
constexpr int kLaneEntries = 8;
constexpr int kLaneStep = 64 * kLaneEntries;  // entries per wave step
constexpr int kLanePad = 2 * kLaneStep;       // padding entries after the segments
constexpr int kLanePadSteps = kLanePad / kLaneStep;
constexpr int kStagedWaves = 16;              // waves of a group's workgroup
constexpr int kStagedThreads = 64 * kStagedWaves;
constexpr int kWindowPoints = 20432;          // + the wave edges: 160 KB of LDS
constexpr size_t kStagedLds = kWindowPoints * sizeof(float2);
constexpr int kMaxWindows = 4;
constexpr int kPasses = kMaxWindows + 1;      // the last pass: entries outside the windows
constexpr int kWindowMinEntries = 2048;       // fewer entries do not pay for a window
constexpr int kWindowBins = 1024;             // candidate windows counted per group

// attraction kernels: one lane group per row of P (any P), the
// segment-blocked kernel reading the layout directly or through the windows,
// and the pair-once kernel of tsne_pairs.hpp
enum AttractionKind { kAttrRows = 0, kAttrLanes = 1, kAttrStaged = 2, kAttrPairs = 3 };

// the segments and their tables
struct BlockTables {
  const int* col;                    // padded entries (+ kLanePad)
  const unsigned short* col16;       // window passes: column - window start
  const float* val;
  const int* lane_row;               // row of every 8 padded entries
  const unsigned long long* starts;  // per step: lanes that begin a segment
  const int* group_row;              // first row of every group (groups + 1)
  const int2* window;                // per group and window: (lo, points); 0 points: none
  const int2* pass_range;            // per group and pass: its entries [begin, end)
};

// The segments are read once per launch: their loads carry the non-temporal
// hint so they do not push the layout windows out of L2.
__device__ __forceinline__ int4 stream_load(const int4* p) {
  typedef int v4 __attribute__((ext_vector_type(4)));
  const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
  return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 stream_load(const float4* p) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = __builtin_nontemporal_load(reinterpret_cast<const v4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

struct LaneCols {
  int j[kLaneEntries];
  int row;
  unsigned long long starts;
};

// LOCAL: the window passes' 16-bit columns (one 16-byte load per lane)
template <bool LOCAL>
__device__ __forceinline__ LaneCols lane_cols(const BlockTables& t, int base, int lane) {
  LaneCols c;
  if constexpr (LOCAL) {
    const int4 a = stream_load(reinterpret_cast<const int4*>(t.col16 + base + kLaneEntries * lane));
    const unsigned int w[4] = {static_cast<unsigned int>(a.x), static_cast<unsigned int>(a.y),
                               static_cast<unsigned int>(a.z), static_cast<unsigned int>(a.w)};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      c.j[2 * q] = static_cast<int>(w[q] & 0xffffu);
      c.j[2 * q + 1] = static_cast<int>(w[q] >> 16);
    }
  } else {
    const int* p = t.col + base + kLaneEntries * lane;
    const int4 a = stream_load(reinterpret_cast<const int4*>(p));
    const int4 b = stream_load(reinterpret_cast<const int4*>(p + 4));
    c.j[0] = a.x;
    c.j[1] = a.y;
    c.j[2] = a.z;
    c.j[3] = a.w;
    c.j[4] = b.x;
    c.j[5] = b.y;
    c.j[6] = b.z;
    c.j[7] = b.w;
  }
  c.row = __builtin_nontemporal_load(t.lane_row + base / kLaneEntries + lane);
  c.starts = t.starts[base / kLaneStep];
  return c;
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, true));
}
__device__ __forceinline__ float dpp_wave_shr1(float v) { return dpp_move<0x138, 0xF>(v); }

// Segmented inclusive scan of (sx, sy) over the wave: lane l sums lanes
// lim .. l, where lim (<= l) is the first lane of its segment.  The steps add
// the value of a source lane below only while the source lies in the segment.
__device__ __forceinline__ void segmented_wave_scan(float& sx, float& sy, int lane, int lim) {
#define FE_SEG_STEP(CTRL, MASK, SRC)                   \
  {                                                    \
    const float ux = dpp_move<CTRL, MASK>(sx);          \
    const float uy = dpp_move<CTRL, MASK>(sy);          \
    if ((SRC) >= lim) {                                \
      sx += ux;                                        \
      sy += uy;                                        \
    }                                                  \
  }
  FE_SEG_STEP(0x111, 0xF, lane - 1)
  FE_SEG_STEP(0x112, 0xF, lane - 2)
  FE_SEG_STEP(0x114, 0xF, lane - 4)
  FE_SEG_STEP(0x118, 0xF, lane - 8)
  FE_SEG_STEP(0x142, 0xA, (lane & ~15) - 1)
  FE_SEG_STEP(0x143, 0xC, 31)
#undef FE_SEG_STEP
}

// Sources of layout points for the attraction: straight from the layout, or
// from a window of the layout staged in LDS.  fetch() issues the loads of a
// point and value() forms the centred point fl(y - mean) from what they
// returned; a step issues every fetch before it uses any value, so all its
// loads are in flight together.
struct GlobalPoints {
  static constexpr bool kLocal = false;
  const float2* __restrict__ y;
  float mx, my;
  using Fetch = float2;
  __device__ __forceinline__ Fetch fetch(int j) const { return y[j]; }
  __device__ __forceinline__ float2 value(const Fetch& f) const {
    return make_float2(f.x - mx, f.y - my);
  }
};

// A window pass: every column of the pass lies in the window staged in LDS
// (already centred, the same fl(y - mean)) and is stored as its offset from
// the window's start.
struct WindowPoints {
  static constexpr bool kLocal = true;
  const float2* __restrict__ y;
  const float2* lds;
  float mx, my;
  using Fetch = float2;
  __device__ __forceinline__ Fetch fetch(int o) const { return lds[o]; }
  __device__ __forceinline__ float2 value(const Fetch& f) const { return f; }
};

// What a wave leaves for the pass's join: the sum of its entries before its
// first segment start (head), whether it has a start (-1: empty range), the
// segment open at its end (tail) and that segment's row.
struct WaveEdge {
  float2 head;
  float2 tail;
  int tail_row;
  int has_start;
};

// Entries e0 .. e1-1 of one wave (e0 < e1): segments that begin and end in
// the range are added to their rows' forces in attr; the head and tail go to
// *edge.
template <class SRC>
__device__ __forceinline__ void attract_range(const SRC& src, const BlockTables& t, int e0, int e1,
                                              float exaggeration, float2* __restrict__ attr,
                                              WaveEdge* edge) {
  const int lane = threadIdx.x & 63;
  const unsigned long long upto_lane = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const float mx = src.mx, my = src.my;
  float cx = 0.0f, cy = 0.0f;  // the segment open at the start of the step
  int crow = 0;                // and its row
  bool head_open = true;       // no segment start seen yet
  int base = e0 & ~(kLaneStep - 1);
  LaneCols cur = lane_cols<SRC::kLocal>(t, base, lane);
  // the step's values, loaded one step ahead (the segments are padded by
  // kLanePad entries past their end, so the read past a range stays inside)
  float4 va = stream_load(reinterpret_cast<const float4*>(t.val + base + kLaneEntries * lane));
  float4 vb = stream_load(reinterpret_cast<const float4*>(t.val + base + kLaneEntries * lane + 4));
  for (; base < e1; base += kLaneStep) {
    const int eb = base + kLaneEntries * lane;
    const bool valid = eb >= e0 && eb < e1;
    // a lane that begins a segment closes the one below it
    const bool seen = valid && ((cur.starts >> lane) & 1ull);
    const int row = cur.row;
    const float v[kLaneEntries] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
    // the row's own point (adjacent rows: through the caches), the gathers,
    // then the next step's columns, rows and values
    const float2 pr = src.y[row];
    typename SRC::Fetch fj[kLaneEntries];
#pragma unroll
    for (int u = 0; u < kLaneEntries; ++u) fj[u] = src.fetch(cur.j[u]);
    cur = lane_cols<SRC::kLocal>(t, base + kLaneStep, lane);
    va = stream_load(reinterpret_cast<const float4*>(t.val + eb + kLaneStep));
    vb = stream_load(reinterpret_cast<const float4*>(t.val + eb + kLaneStep + 4));
    // keep every load of the step above the first wait
    __builtin_amdgcn_sched_barrier(0);
    const float xi = pr.x - mx, yi = pr.y - my;
    float sx = 0.0f, sy = 0.0f;
#pragma unroll
    for (int u = 0; u < kLaneEntries; ++u) {
      const float2 pj = src.value(fj[u]);
      const float dx = xi - pj.x, dy = yi - pj.y;
      const float q = __builtin_amdgcn_rcpf(1.0f + dx * dx + dy * dy);
      const float c = v[u] * q;
      sx += c * dx;
      sy += c * dy;
    }
    float tx = valid ? exaggeration * sx : 0.0f, ty = valid ? exaggeration * sy : 0.0f;
    const unsigned long long seen_mask = __ballot(seen);
    const unsigned long long upto = seen_mask & upto_lane;
    const int lim = upto ? 63 - __clzll(upto) : 0;  // first lane of this lane's segment
    if (lane == 0 && !seen) {
      tx += cx;
      ty += cy;
    }
    segmented_wave_scan(tx, ty, lane, lim);
    // the open segment arriving at this lane and its row: the carry for lane
    // 0, else the lane below
    float ix = dpp_wave_shr1(tx), iy = dpp_wave_shr1(ty);
    int irow = __builtin_amdgcn_update_dpp(0, row, 0x138, 0xF, 0xF, true);
    if (lane == 0) {
      ix = cx;
      iy = cy;
      irow = crow;
    }
    if (seen) {
      // the wave's first start closes its head
      if (head_open && lane == __builtin_ctzll(seen_mask)) {
        edge->head = make_float2(ix, iy);
      } else {
        const float2 a = attr[irow];
        attr[irow] = make_float2(a.x + ix, a.y + iy);
      }
    }
    if (seen_mask) head_open = false;
    cx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tx), 63));
    cy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ty), 63));
    crow = __builtin_amdgcn_readlane(row, 63);
  }
  if (lane == 0) {
    if (head_open) {  // one segment runs through the whole range
      edge->head = make_float2(cx, cy);
      edge->tail = make_float2(0.f, 0.f);
      edge->tail_row = 0;
      edge->has_start = 0;
    } else {
      edge->tail = make_float2(cx, cy);
      edge->tail_row = t.lane_row[(e1 - 1) / kLaneEntries];
      edge->has_start = 1;
    }
  }
}

// The join of a pass (lanes 0 .. kStagedWaves of one wave): lane w >= 1 with
// a segment start closes the segment open before it, whose sum is the tail
// of the last wave with a start before w plus the heads of the waves after
// that one up to w; lane kStagedWaves closes the segment open at the pass's
// end the same way.
__device__ __forceinline__ void join_waves(const WaveEdge* edges, float2* __restrict__ attr) {
  const int lane = threadIdx.x & 63;
  if (lane > kStagedWaves) return;
  if (lane < kStagedWaves && edges[lane].has_start != 1) return;
  int j = lane - 1;
  while (j >= 0 && edges[j].has_start != 1) --j;
  if (j < 0) return;  // nothing open before: the pass begins with a segment start
  float2 sum = edges[j].tail;
  for (int i = j + 1; i < lane; ++i) {
    if (edges[i].has_start == 0) {
      sum.x += edges[i].head.x;
      sum.y += edges[i].head.y;
    }
  }
  if (lane < kStagedWaves) {
    sum.x += edges[lane].head.x;
    sum.y += edges[lane].head.y;
  }
  const int r = edges[j].tail_row;
  const float2 a = attr[r];
  attr[r] = make_float2(a.x + sum.x, a.y + sum.y);
}

// The window's points, centred (fl(y - mean), as every reader forms them).
// Every load of a thread's share is issued before the first LDS write, so
// the window arrives in one round trip to memory rather than one per point a
// thread copies.
constexpr int kStageLoads = (kWindowPoints + kStagedThreads - 1) / kStagedThreads;
__device__ __forceinline__ void stage_window(const float2* __restrict__ y, int lo, int m,
                                             float mx, float my, float2* sy) {
  float2 r[kStageLoads];
#pragma unroll
  for (int u = 0; u < kStageLoads; ++u) {
    const int i = static_cast<int>(threadIdx.x) + u * kStagedThreads;
    if (i < m) r[u] = y[lo + i];
  }
#pragma unroll
  for (int u = 0; u < kStageLoads; ++u) {
    const int i = static_cast<int>(threadIdx.x) + u * kStagedThreads;
    if (i < m) sy[i] = make_float2(r[u].x - mx, r[u].y - my);
  }
}

// One group: zero its rows' forces, then the passes in order.  STAGE: the
// window passes read the staged windows (the same values as the layout).
template <bool STAGE>
__device__ __forceinline__ void attract_group(const float2* __restrict__ y, const BlockTables& t,
                                              float mx, float my, float exaggeration,
                                              float2* __restrict__ attr, int g, float2* sy) {
  __shared__ WaveEdge edges[kStagedWaves];
  const int rb = t.group_row[g], re = t.group_row[g + 1];
  for (int r = rb + static_cast<int>(threadIdx.x); r < re; r += kStagedThreads)
    attr[r] = make_float2(0.f, 0.f);
  const int w = static_cast<int>(threadIdx.x >> 6);
  for (int k = 0; k < kPasses; ++k) {
    const int2 pr = t.pass_range[g * kPasses + k];
    if (pr.x >= pr.y) continue;  // the same for the whole workgroup
    const bool staged = STAGE && k < kMaxWindows;
    int2 win = make_int2(0, 0);
    if (staged) {
      win = t.window[g * kMaxWindows + k];
      stage_window(y, win.x, win.y, mx, my, sy);
    }
    // the window, the zeroed forces and the previous pass's join are done
    __syncthreads();
    const int s0 = pr.x / kLaneStep, steps = (pr.y - 1) / kLaneStep - s0 + 1;
    const int e0 = max(pr.x, (s0 + w * steps / kStagedWaves) * kLaneStep);
    const int e1 = min(pr.y, (s0 + (w + 1) * steps / kStagedWaves) * kLaneStep);
    if (e0 < e1) {
      if (staged) {
        const WindowPoints src{y, sy, mx, my};
        attract_range(src, t, e0, e1, exaggeration, attr, &edges[w]);
      } else {
        const GlobalPoints src{y, mx, my};
        attract_range(src, t, e0, e1, exaggeration, attr, &edges[w]);
      }
    } else if ((threadIdx.x & 63) == 0) {
      edges[w].has_start = -1;
    }
    __syncthreads();
    if (w == 0) join_waves(edges, attr);
  }
}

template <bool STAGE>
__global__ __launch_bounds__(kStagedThreads) void attract_blocked(
    const float2* __restrict__ y_in, BlockTables t, const TsneStats* __restrict__ st,
    float exaggeration, float2* __restrict__ attr, int n) {
  extern __shared__ float2 sy[];
  attract_group<STAGE>(y_in, t, st->mean_xf, st->mean_yf, exaggeration, attr,
                       xcd_remap(blockIdx.x, gridDim.x), sy);
}

// ---- set-up of the segments ----

// lengths of the rows rounded up to whole lanes (at least one), to cut the
// groups by padded entries
__global__ __launch_bounds__(kThreads) void lane_lengths(const int* __restrict__ row_ptr, int n,
                                                         int* __restrict__ len) {
  const int r = blockIdx.x * kThreads + threadIdx.x;
  if (r > n) return;
  if (r == n) {
    len[n] = 0;
    return;
  }
  const int l = row_ptr[r + 1] - row_ptr[r];
  len[r] = max(kLaneEntries, (l + kLaneEntries - 1) / kLaneEntries * kLaneEntries);
}

// first row starting at or after entry w * nnz / parts (parts + 1 values)
__global__ __launch_bounds__(kThreads) void cut_rows(const int* __restrict__ row_ptr, int n,
                                                     long long nnz, int parts,
                                                     int* __restrict__ first_row) {
  const int w = blockIdx.x * kThreads + threadIdx.x;
  if (w > parts) return;
  const long long target = static_cast<long long>(w) * nnz / parts;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (row_ptr[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  first_row[w] = lo;
}

// Windows of one group (one block): the column bounds; when they span at most
// kWindowPoints one window covers them, else the columns are counted in
// candidate windows lo = min + b * kWindowPoints and the (at most
// kMaxWindows) fullest with kWindowMinEntries or more are kept, in order.
__global__ __launch_bounds__(256) void blk_windows(const int* __restrict__ row_ptr,
                                                   const int* __restrict__ col,
                                                   const int* __restrict__ group_row, int n,
                                                   int2* __restrict__ window) {
  __shared__ int hist[kWindowBins];
  __shared__ int red[2][256 / 64];
  const int g = blockIdx.x;
  const int e0 = row_ptr[group_row[g]], e1 = row_ptr[group_row[g + 1]];
  int mn = INT_MAX, mx = -1;
  for (int e = e0 + static_cast<int>(threadIdx.x); e < e1; e += 256) {
    const int j = col[e];
    mn = min(mn, j);
    mx = max(mx, j);
  }
  mn = wave_min_i(mn);
  mx = wave_max_i(mx);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = mn;
    red[1][threadIdx.x >> 6] = mx;
  }
  for (int b = threadIdx.x; b < kWindowBins; b += 256) hist[b] = 0;
  __syncthreads();
  mn = min(min(red[0][0], red[0][1]), min(red[0][2], red[0][3]));
  mx = max(max(red[1][0], red[1][1]), max(red[1][2], red[1][3]));
  int2* out = window + g * kMaxWindows;
  if (mx < mn) {  // no entries
    if (threadIdx.x < kMaxWindows) out[threadIdx.x] = make_int2(0, 0);
    return;
  }
  if (mx - mn + 1 <= kWindowPoints) {
    if (threadIdx.x < kMaxWindows)
      out[threadIdx.x] = threadIdx.x == 0 ? make_int2(mn, mx - mn + 1) : make_int2(0, 0);
    return;
  }
  for (int e = e0 + static_cast<int>(threadIdx.x); e < e1; e += 256) {
    const int b = (col[e] - mn) / kWindowPoints;
    if (b < kWindowBins) atomicAdd(&hist[b], 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int pick[kMaxWindows];
    int np = 0;
    for (int q = 0; q < kMaxWindows; ++q) {
      int best = -1, bc = kWindowMinEntries - 1;
      for (int b = 0; b < kWindowBins; ++b) {
        if (hist[b] > bc) {
          bc = hist[b];
          best = b;
        }
      }
      if (best < 0) break;
      pick[np++] = best;
      hist[best] = -1;
    }
    // ascending window order
    for (int a = 1; a < np; ++a)
      for (int c = a; c > 0 && pick[c - 1] > pick[c]; --c) {
        const int tmp = pick[c];
        pick[c] = pick[c - 1];
        pick[c - 1] = tmp;
      }
    for (int q = 0; q < kMaxWindows; ++q) {
      if (q < np) {
        const int lo = mn + pick[q] * kWindowPoints;
        out[q] = make_int2(lo, min(kWindowPoints, n - lo));
      } else {
        out[q] = make_int2(0, 0);
      }
    }
  }
}

__device__ __forceinline__ int lower_col(const int* __restrict__ col, int lo, int hi, int v) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (col[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int group_of_row(const int* __restrict__ group_row, int groups, int r) {
  int lo = 0, hi = groups - 1;  // last g with group_row[g] <= r
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (group_row[mid] <= r) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// slot of segment (row r, pass k): groups in order, passes in order, rows in order
__device__ __forceinline__ int seg_slot(const int* __restrict__ group_row, int g, int k, int r) {
  const int rb = group_row[g], rows = group_row[g + 1] - rb;
  return kPasses * rb + k * rows + (r - rb);
}

// entries [x, y) of a row's column range [s, e) that fall in window w
// (empty when the window is absent)
__device__ __forceinline__ int2 row_window(const int* __restrict__ col, int s, int e, int2 w) {
  if (w.y <= 0) return make_int2(s, s);
  const int a = lower_col(col, s, e, w.x);
  return make_int2(a, lower_col(col, a, e, w.x + w.y));
}

// padded length of every segment (one thread per row)
__global__ __launch_bounds__(kThreads) void blk_seg_lengths(const int* __restrict__ row_ptr,
                                                            const int* __restrict__ col,
                                                            const int* __restrict__ group_row,
                                                            int groups,
                                                            const int2* __restrict__ window, int n,
                                                            int* __restrict__ len) {
  const int r = blockIdx.x * kThreads + threadIdx.x;
  if (r > n) return;
  if (r == n) {  // the scans' last element
    len[kPasses * n] = 0;
    return;
  }
  const int g = group_of_row(group_row, groups, r);
  const int s = row_ptr[r], e = row_ptr[r + 1];
  int2 ab[kMaxWindows];
#pragma unroll
  for (int k = 0; k < kMaxWindows; ++k) ab[k] = row_window(col, s, e, window[g * kMaxWindows + k]);
  int rest = e - s;
#pragma unroll
  for (int k = 0; k < kPasses; ++k) {
    int l;
    if (k < kMaxWindows) {
      l = ab[k].y - ab[k].x;
      rest -= l;
    } else {
      l = rest;
    }
    const int slot = seg_slot(group_row, g, k, r);
    len[slot] = (l + kLaneEntries - 1) / kLaneEntries * kLaneEntries;
  }
}

// the padded segments (one wave per row): entries, padding, segment starts
__global__ __launch_bounds__(kThreads) void blk_fill(const int* __restrict__ row_ptr,
                                                     const int* __restrict__ col,
                                                     const float* __restrict__ val,
                                                     const int* __restrict__ group_row, int groups,
                                                     const int2* __restrict__ window, int n,
                                                     const int* __restrict__ off,
                                                     int* __restrict__ lane_row,
                                                     int* __restrict__ pcol,
                                                     unsigned short* __restrict__ pcol16,
                                                     float* __restrict__ pval,
                                                     unsigned long long* starts) {
  const int r = blockIdx.x * (kThreads / 64) + static_cast<int>(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= n) return;
  const int g = group_of_row(group_row, groups, r);
  const int s = row_ptr[r], e = row_ptr[r + 1];
  int2 ab[kMaxWindows];
#pragma unroll
  for (int k = 0; k < kMaxWindows; ++k) ab[k] = row_window(col, s, e, window[g * kMaxWindows + k]);
  int rest_out = 0;  // entries of the last segment written so far
#pragma unroll
  for (int k = 0; k < kPasses; ++k) {
    const int slot = seg_slot(group_row, g, k, r);
    const int o = off[slot], plen = off[slot + 1] - o;
    if (plen == 0) continue;
    if (lane == 0) atomicOr(starts + o / kLaneStep, 1ull << ((o / kLaneEntries) & 63));
    for (int q = lane; q < plen / kLaneEntries; q += 64) lane_row[o / kLaneEntries + q] = r;
    if (k < kMaxWindows) {
      const int l = ab[k].y - ab[k].x;
      const int lo = window[g * kMaxWindows + k].x;
      for (int q = lane; q < plen; q += 64) {
        const bool real = q < l;
        const int j = real ? col[ab[k].x + q] : r;
        pcol[o + q] = j;
        pcol16[o + q] = static_cast<unsigned short>(real ? j - lo : 0);
        pval[o + q] = real ? val[ab[k].x + q] : 0.0f;
      }
    } else {
      // the entries outside every window, in column order
      for (int c0 = s; c0 < e; c0 += 64) {
        const int q = c0 + lane;
        bool out = q < e;
        if (out) {
#pragma unroll
          for (int w = 0; w < kMaxWindows; ++w) out = out && (q < ab[w].x || q >= ab[w].y);
        }
        const unsigned long long m = __ballot(out);
        if (out) {
          const int at = rest_out + __popcll(m & ((1ull << lane) - 1ull));
          pcol[o + at] = col[q];
          pval[o + at] = val[q];
        }
        rest_out += __popcll(m);
      }
      for (int q = rest_out + lane; q < plen; q += 64) {
        pcol[o + q] = r;
        pval[o + q] = 0.0f;
      }
    }
  }
}

// per group and pass, its entries [begin, end)
__global__ __launch_bounds__(kThreads) void blk_pass_ranges(const int* __restrict__ group_row,
                                                            int groups,
                                                            const int* __restrict__ off,
                                                            int2* __restrict__ pass_range) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= groups * kPasses) return;
  const int g = i / kPasses, k = i % kPasses;
  const int rb = group_row[g], rows = group_row[g + 1] - rb;
  const int slot = kPasses * rb + k * rows;
  pass_range[i] = make_int2(off[slot], off[slot + rows]);
}


// ----------------------------------------------------------------------------
// 5b. gather + update, one thread per point (coalesced).  Back-interpolation
//     of (q2, xq2, yq2) at the point's cell (tsne_neighbors.cpp:1247-1258),
//     rep = -(y q2 - xq2, y q2 - yq2) / Z, grad = rep + attr.
// MODE 0: optimiser step (apply_open_tsne_update_f, tsne_neighbors.cpp:
//         1297-1336; CUDA twin embedding_cuda_kernels.cpp:2819-2886) into
//         y_out plus this block's partial of the layout statistics
//         (layout_finalize folds them into the next iteration's geometry).
// MODE 1: diagnostic; writes the repulsive force to diag_rep.
// ----------------------------------------------------------------------------
// The optimiser step of one point from its centred position and repulsive
// force: gains, momentum, step clip (apply_open_tsne_update_f,
// tsne_neighbors.cpp:1297-1336), the new position into y_out and into this
// thread's layout statistics.
__device__ __forceinline__ void opt_step(int i, float xi, float yi, float rx, float ry, float2 a,
                                         float2* __restrict__ gains, float2* __restrict__ upd,
                                         float2* __restrict__ y_out, const StepParams& sp,
                                         LayoutPartial& m) {
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

// The attraction of point i as the update reads it: from the array the
// attraction kernel wrote.
struct AttrArray {
  const float2* __restrict__ a;
  __device__ __forceinline__ float2 operator()(int i) const { return a[i]; }
};

template <int MODE, class AttrAt>
__device__ __forceinline__ void update_body(const float2* __restrict__ y_in,
                                            float2* __restrict__ y_out,
                                            float2* __restrict__ gains, float2* __restrict__ upd,
                                            AttrAt attr,
                                            const float4* __restrict__ field, TsneStats* st,
                                            LayoutPartial* partials, int n, int g, StepParams sp,
                                            float2* __restrict__ diag_rep, int blk, int nblk) {
  const float mxf = st->mean_xf, myf = st->mean_yf;
  const float lx = st->lower_x, ly = st->lower_y, inv = st->inv_spacing;
  const float inv_z = static_cast<float>(
      1.0 / fmax(st->sum_q_total - static_cast<double>(n), static_cast<double>(FLT_MIN)));
  LayoutPartial m{0.0, 0.0, FLT_MAX, -FLT_MAX, FLT_MAX, -FLT_MAX};
  // a contiguous range of points per block, the ranges of one XCD adjacent
  int beg, end;
  point_range(n, xcd_remap(blk, nblk), nblk, &beg, &end);
  for (int i = beg + threadIdx.x; i < end; i += kThreads) {
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
      opt_step(i, xi, yi, rx, ry, attr(i), gains, upd, y_out, sp, m);
    }
  }
  if constexpr (MODE == 0) {
    const LayoutPartial b = block_layout(m);
    if (threadIdx.x == 0) partials[blk] = b;
  }
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void update_kernel(
    const float2* __restrict__ y_in, float2* __restrict__ y_out, float2* __restrict__ gains,
    float2* __restrict__ upd, const float2* __restrict__ attr, const float4* __restrict__ field,
    TsneStats* st, LayoutPartial* partials, int n, int g, StepParams sp,
    float2* __restrict__ diag_rep) {
  update_body<MODE>(y_in, y_out, gains, upd, AttrArray{attr}, field, st, partials, n, g, sp,
                    diag_rep, blockIdx.x, gridDim.x);
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

// ----------------------------------------------------------------------------
// 6. Exact repulsion: the pair form of compute_gradient_pair_symmetric_f
//    (tsne_neighbors.cpp:1090-1150), the reference CPU route for n <= 3000.
//    For every point i the unnormalised force sum_{j != i} -q_ij^2 (y_i - y_j)
//    and Z = sum_{i != j} q_ij with q_ij = 1 / (1 + |y_i - y_j|^2).
//    One thread per point; blockIdx.y splits the j range into `splits`
//    contiguous slices so small layouts still fill the GPU.  A slice streams
//    the centred layout through LDS in tiles of kExactTile points; each tile's
//    float partial joins FP64 running sums, so every sum has a fixed order.
//    Partials: part[s * n + i] (force of slice s) and zpart[s * gridDim.x + b]
//    (Z of block b's points over slice s); the readers fold them in that order.
// ----------------------------------------------------------------------------
constexpr int kExactTile = 512;

template <bool FORCES>
__global__ __launch_bounds__(kThreads) void exact_rep_kernel(const float2* __restrict__ y,
                                                             const TsneStats* __restrict__ st,
                                                             double2* __restrict__ part,
                                                             double* __restrict__ zpart, int n,
                                                             int chunk) {
  __shared__ float2 tile[kExactTile];
  __shared__ double zred[kThreads / 64];
  const float mxf = st->mean_xf, myf = st->mean_yf;
  const int i = blockIdx.x * kThreads + threadIdx.x;
  const int j0 = blockIdx.y * chunk, j1 = min(n, j0 + chunk);
  float xi = 0.0f, yi = 0.0f;
  if (i < n) {
    const float2 p = y[i];
    xi = p.x - mxf;
    yi = p.y - myf;
  }
  double fx = 0.0, fy = 0.0, z = 0.0;
  for (int base = j0; base < j1; base += kExactTile) {
    const int m = min(kExactTile, j1 - base);
    __syncthreads();
    for (int t = threadIdx.x; t < m; t += kThreads) {
      const float2 p = y[base + t];
      tile[t] = make_float2(p.x - mxf, p.y - myf);
    }
    __syncthreads();
    float tx = 0.0f, ty = 0.0f, tz = 0.0f;
    for (int t = 0; t < m; ++t) {
      const float dx = xi - tile[t].x, dy = yi - tile[t].y;
      const float q = base + t == i ? 0.0f : 1.0f / (1.0f + dx * dx + dy * dy);
      tz += q;
      if constexpr (FORCES) {
        const float c = -(q * q);
        tx += c * dx;
        ty += c * dy;
      }
    }
    z += tz;
    if constexpr (FORCES) {
      fx += tx;
      fy += ty;
    }
  }
  if constexpr (FORCES) {
    if (i < n) part[static_cast<size_t>(blockIdx.y) * n + i] = make_double2(fx, fy);
  }
  if (i >= n) z = 0.0;
  z = wave_sum(z);
  if ((threadIdx.x & 63) == 0) zred[threadIdx.x >> 6] = z;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += zred[w];
    zpart[static_cast<size_t>(blockIdx.y) * gridDim.x + blockIdx.x] = t;
  }
}

// Z from the exact partials, folded in slice-major block order.
__device__ __forceinline__ double fold_exact_z(const double* __restrict__ zpart, int nz) {
  __shared__ double zs;
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int b = 0; b < nz; ++b) t += zpart[b];
    zs = t;
  }
  __syncthreads();
  return zs;
}

// Update of the exact route.  MODE 0: optimiser step from the normalised exact
// repulsion (value * 1 / max(Z, FLT_MIN) in float, as the reference scales its
// summed float force) plus the attraction, the layout partials; block 0 also
// keeps Z + n in st->sum_q_total (the quantity the FFT route keeps there).
// MODE 1: diagnostic; the normalised force into diag_rep, Z into st.
template <int MODE>
__global__ __launch_bounds__(kThreads) void update_exact_kernel(
    const float2* __restrict__ y_in, float2* __restrict__ y_out, float2* __restrict__ gains,
    float2* __restrict__ upd, const float2* __restrict__ attr, const double2* __restrict__ part,
    int splits, const double* __restrict__ zpart, int nz, TsneStats* st,
    LayoutPartial* partials, int n, StepParams sp, float2* __restrict__ diag_rep) {
  const double zsum = fold_exact_z(zpart, nz);
  const float inv_z = static_cast<float>(1.0 / fmax(zsum, static_cast<double>(FLT_MIN)));
  if (blockIdx.x == 0 && threadIdx.x == 0) st->sum_q_total = zsum + static_cast<double>(n);
  const float mxf = st->mean_xf, myf = st->mean_yf;
  LayoutPartial m{0.0, 0.0, FLT_MAX, -FLT_MAX, FLT_MAX, -FLT_MAX};
  int beg, end;
  point_range(n, xcd_remap(blockIdx.x, gridDim.x), gridDim.x, &beg, &end);
  for (int i = beg + threadIdx.x; i < end; i += kThreads) {
    double sx = 0.0, sy = 0.0;
    for (int s = 0; s < splits; ++s) {
      const double2 v = part[static_cast<size_t>(s) * n + i];
      sx += v.x;
      sy += v.y;
    }
    const float rx = static_cast<float>(sx) * inv_z, ry = static_cast<float>(sy) * inv_z;
    if constexpr (MODE == 1) {
      diag_rep[i] = make_float2(rx, ry);
    } else {
      const float2 pi = y_in[i];
      opt_step(i, pi.x - mxf, pi.y - myf, rx, ry, attr[i], gains, upd, y_out, sp, m);
    }
  }
  if constexpr (MODE == 0) {
    const LayoutPartial b = block_layout(m);
    if (threadIdx.x == 0) partials[blockIdx.x] = b;
  }
}

// ----------------------------------------------------------------------------
// 7. KL divergence of P against Q = q / Z (evaluate_kl_f, tsne_neighbors.cpp:
//    1338-1365): per CSR row the FP64 sum of P_ij log((P_ij + 1e-9) /
//    (q_ij / Z + 1e-9)), q_ij = 1 / (1 + d2) in FP64 from the float squared
//    distance of the centred layout, Z the exact sum rounded to float as
//    compute_sum_q_f returns it.  One thread per row; block partials in row
//    order (the host adds them in block order).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void kl_rows_kernel(const float2* __restrict__ y,
                                                           const TsneStats* __restrict__ st,
                                                           const int* __restrict__ row_ptr,
                                                           const int* __restrict__ col,
                                                           const float* __restrict__ val, int n,
                                                           double zf, double* __restrict__ part) {
  __shared__ double red[kThreads / 64];
  const float mxf = st->mean_xf, myf = st->mean_yf;
  const int i = blockIdx.x * kThreads + threadIdx.x;
  double row = 0.0;
  if (i < n) {
    const float2 pi = y[i];
    const float xi = pi.x - mxf, yi = pi.y - myf;
    for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
      const float2 pj = y[col[e]];
      const float dx = xi - (pj.x - mxf), dy = yi - (pj.y - myf);
      const float d2 = dx * dx + dy * dy;
      const double p = static_cast<double>(val[e]);
      const double q = (1.0 / (1.0 + static_cast<double>(d2))) / zf;
      row += p * log((p + 1e-9) / (q + 1e-9));
    }
  }
  row = wave_sum(row);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = row;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[w];
    part[blockIdx.x] = t;
  }
}

}  // namespace
}  // namespace fe
