// The UMAP fuzzy graph on gfx950 (umap_graph_csr_float_cpp,
// src/fast_knn_umap.cpp:3995-4048, at the default edge budget n_cols, which
// builds with build_graph_csr_native_direct_values :1562-1780).
//
//   ug_mean_partials / ug_mean_fold   the global mean of the finite
//                     non-negative distances (smooth_knn_dist :876-923),
//                     summed as double-double per block, folded in block order
//   ug_smooth_rows    one thread per row: rho = the smallest positive finite
//                     distance, sigma by the reference's 64-step FP64 search
//                     for sum exp(-(d - rho)/sigma) = log2(k) (tolerance 1e-5)
//                     and its floor (:925-1003)
//   ug_count / ug_fill
//                     every kept directed membership w = exp(-(d - rho)/sigma)
//                     (1 at d <= rho; float arithmetic, :1608-1634) goes into
//                     both endpoint rows, tagged forward or reverse
//   ug_sort_union_wave / ug_sort_union
//                     one wave per row of at most 512 entries (a bitonic sort
//                     across lanes and registers), one block per longer row
//                     (LDS bitonic sort) of (neighbour,
//                     direction, weight bits); per neighbour the largest
//                     forward and reverse weights (the reference's max-dedupe
//                     of :1291-1364 and its incoming lists) give
//                     f + r - f*r, kept above 1e-6 (:1702-1768)
//   ug_compact / ug_epochs
//                     CSR rows in ascending neighbour order, COO heads for the
//                     epoch kernel, epochs_per_sample = max_w / max(w, 1e-6)
//                     (:1264-1282)
//
// Float operations that the reference does in float are done with explicit
// rounding (no fused multiply-add); exp of a float argument is the double exp
// rounded to float.  Every sum has a fixed order, so the graph does not depend
// on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/fastembedr_hip.h"
#include "fe_common.hpp"
#include "tsne_device.hpp"
#include "tsne_affinity.hpp"

namespace fe {
namespace {

constexpr int kUgThreads = 256;
constexpr int kUgMeanBlocks = 1024;
constexpr unsigned long long kUgReverse = 1ull << 31;

// double-double accumulation (Knuth two-sum)
struct DD {
  double hi, lo;
};
__device__ __forceinline__ DD dd_add(DD a, double b) {
  const double s = __dadd_rn(a.hi, b);
  const double bb = __dsub_rn(s, a.hi);
  const double err = __dadd_rn(__dsub_rn(a.hi, __dsub_rn(s, bb)), __dsub_rn(b, bb));
  const double lo = __dadd_rn(a.lo, err);
  const double hi = __dadd_rn(s, lo);
  return DD{hi, __dsub_rn(lo, __dsub_rn(hi, s))};
}
__device__ __forceinline__ DD dd_add_dd(DD a, DD b) { return dd_add(dd_add(a, b.hi), b.lo); }

__device__ __forceinline__ float dist_at(const float* __restrict__ d, int n, int col, int i) {
  return d[static_cast<size_t>(col) * n + i];
}

// min / max of the used index columns (the reference's offset detection)
__global__ __launch_bounds__(kUgThreads) void ug_index_range(const int* __restrict__ idx, int n,
                                                             int col_start, int n_cols,
                                                             int* range) {
  const long long total = static_cast<long long>(n) * n_cols;
  int mn = INT_MAX, mx = INT_MIN;
  for (long long e = blockIdx.x * static_cast<long long>(kUgThreads) + threadIdx.x; e < total;
       e += static_cast<long long>(gridDim.x) * kUgThreads) {
    const int v = idx[static_cast<size_t>(col_start) * n + e];
    mn = min(mn, v);
    mx = max(mx, v);
  }
  mn = wave_min_i(mn);
  mx = wave_max_i(mx);
  if ((threadIdx.x & 63) == 0) {
    atomicMin(range, mn);
    atomicMax(range + 1, mx);
  }
}

__global__ __launch_bounds__(kUgThreads) void ug_mean_partials(const float* __restrict__ dist,
                                                               int n, int col_start, int n_cols,
                                                               DD* partials,
                                                               unsigned long long* counts) {
  __shared__ DD sd[kUgThreads];
  __shared__ unsigned long long sc[kUgThreads];
  const long long total = static_cast<long long>(n) * n_cols;
  const long long per = (total + gridDim.x - 1) / gridDim.x;
  const long long b0 = blockIdx.x * per, b1 = min(total, b0 + per);
  DD acc{0.0, 0.0};
  unsigned long long cnt = 0;
  for (long long e = b0 + threadIdx.x; e < b1; e += kUgThreads) {
    const float d = dist[static_cast<size_t>(col_start) * n + e];
    if (isfinite(d) && d >= 0.0f) {
      acc = dd_add(acc, static_cast<double>(d));
      ++cnt;
    }
  }
  sd[threadIdx.x] = acc;
  sc[threadIdx.x] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    DD t{0.0, 0.0};
    unsigned long long c = 0;
    for (int q = 0; q < kUgThreads; ++q) {
      t = dd_add_dd(t, sd[q]);
      c += sc[q];
    }
    partials[blockIdx.x] = t;
    counts[blockIdx.x] = c;
  }
}

__global__ void ug_mean_fold(const DD* __restrict__ partials,
                             const unsigned long long* __restrict__ counts, int nb,
                             double* mean) {
  if (threadIdx.x != 0) return;
  DD t{0.0, 0.0};
  unsigned long long c = 0;
  for (int q = 0; q < nb; ++q) {
    t = dd_add_dd(t, partials[q]);
    c += counts[q];
  }
  *mean = c > 0 ? __ddiv_rn(__dadd_rn(t.hi, t.lo), static_cast<double>(c)) : 1.0;
}

__global__ __launch_bounds__(kUgThreads) void ug_smooth_rows(const float* __restrict__ dist, int n,
                                                             int col_start, int k,
                                                             const double* __restrict__ gmean_p,
                                                             float* __restrict__ sigmas,
                                                             float* __restrict__ rhos) {
  const int i = blockIdx.x * kUgThreads + threadIdx.x;
  if (i >= n) return;
  const double target = log2(static_cast<double>(k));
  const double gmean = *gmean_p;
  double rho = INFINITY, rsum = 0.0;
  int rcount = 0;
  for (int j = 0; j < k; ++j) {
    const float d = dist_at(dist, n, col_start + j, i);
    if (!isfinite(d)) continue;
    if (d >= 0.0f) {
      rsum = __dadd_rn(rsum, static_cast<double>(d));
      ++rcount;
    }
    if (d > 0.0f && static_cast<double>(d) < rho) rho = static_cast<double>(d);
  }
  if (!isfinite(rho)) rho = 0.0;
  rhos[i] = static_cast<float>(rho);
  const double big = DBL_MAX;
  double sigma = 1.0, best = 1.0, best_diff = big, lo = 0.0, hi = big;
  for (int it = 0; it < 64; ++it) {
    const double safe = fmax(sigma, 1.0e-12);
    double psum = 0.0;
    for (int j = 0; j < k; ++j) {
      const float raw = dist_at(dist, n, col_start + j, i);
      if (!isfinite(raw)) continue;
      const double d = __dsub_rn(static_cast<double>(raw), rho);
      psum = __dadd_rn(psum, d <= 0.0 ? 1.0 : exp(__ddiv_rn(-d, safe)));
    }
    const double diff = fabs(__dsub_rn(psum, target));
    if (diff < best_diff) {
      best_diff = diff;
      best = sigma;
    }
    if (psum > target) {
      hi = sigma;
      sigma = __dmul_rn(0.5, __dadd_rn(lo, hi));
    } else {
      lo = sigma;
      sigma = hi == big ? __dmul_rn(sigma, 2.0) : __dmul_rn(0.5, __dadd_rn(lo, hi));
    }
    if (diff < 1.0e-5) break;
  }
  const double rmean = rcount > 0 ? __ddiv_rn(rsum, static_cast<double>(rcount)) : gmean;
  const double floor_s = __dmul_rn(1.0e-3, rho > 0.0 ? rmean : gmean);
  best = fmax(best, floor_s);
  sigmas[i] = static_cast<float>(fmax(best, 1.0e-12));
}

// the kept membership of entry (i, column), or 0 when the entry is dropped
__device__ __forceinline__ float ug_membership(const int* __restrict__ idx,
                                               const float* __restrict__ dist, int n,
                                               int col, int i, int offset,
                                               const float* __restrict__ sigmas,
                                               const float* __restrict__ rhos, int* nb_out) {
  const int nb = idx[static_cast<size_t>(col) * n + i] - offset;
  const float d = dist_at(dist, n, col, i);
  *nb_out = nb;
  if (nb < 0 || nb >= n || nb == i || !isfinite(d)) return 0.0f;
  const float rho = rhos[i];
  float w = 1.0f;
  if (!(d <= rho)) {
    const float x = __fdiv_rn(-__fsub_rn(d, rho), sigmas[i]);
    w = static_cast<float>(exp(static_cast<double>(x)));
  }
  return isfinite(w) && w > 0.0f ? w : 0.0f;
}

__global__ __launch_bounds__(kUgThreads) void ug_count(const int* __restrict__ idx,
                                                       const float* __restrict__ dist, int n,
                                                       int col_start, int k, int offset,
                                                       const float* __restrict__ sigmas,
                                                       const float* __restrict__ rhos,
                                                       int* __restrict__ deg) {
  const long long total = static_cast<long long>(n) * k;
  for (long long e = blockIdx.x * static_cast<long long>(kUgThreads) + threadIdx.x; e < total;
       e += static_cast<long long>(gridDim.x) * kUgThreads) {
    const int c = static_cast<int>(e / n), i = static_cast<int>(e - static_cast<long long>(c) * n);
    int nb;
    if (ug_membership(idx, dist, n, col_start + c, i, offset, sigmas, rhos, &nb) > 0.0f) {
      atomicAdd(deg + i, 1);
      atomicAdd(deg + nb, 1);
    }
  }
}

__global__ __launch_bounds__(kUgThreads) void ug_fill(const int* __restrict__ idx,
                                                      const float* __restrict__ dist, int n,
                                                      int col_start, int k, int offset,
                                                      const float* __restrict__ sigmas,
                                                      const float* __restrict__ rhos,
                                                      const int* __restrict__ raw_ptr,
                                                      int* __restrict__ cursor,
                                                      unsigned long long* __restrict__ raw) {
  const long long total = static_cast<long long>(n) * k;
  for (long long e = blockIdx.x * static_cast<long long>(kUgThreads) + threadIdx.x; e < total;
       e += static_cast<long long>(gridDim.x) * kUgThreads) {
    const int c = static_cast<int>(e / n), i = static_cast<int>(e - static_cast<long long>(c) * n);
    int nb;
    const float w = ug_membership(idx, dist, n, col_start + c, i, offset, sigmas, rhos, &nb);
    if (!(w > 0.0f)) continue;
    const unsigned long long bits = __float_as_uint(w);  // positive: 31 bits
    int pos = raw_ptr[i] + atomicAdd(cursor + i, 1);
    raw[pos] = (static_cast<unsigned long long>(static_cast<unsigned>(nb)) << 32) | bits;
    pos = raw_ptr[nb] + atomicAdd(cursor + nb, 1);
    raw[pos] = (static_cast<unsigned long long>(static_cast<unsigned>(i)) << 32) | kUgReverse | bits;
  }
}

__device__ __forceinline__ bool ug_kept(float w) { return isfinite(w) && w > 1.0e-6f; }

// the union weight f + r - f*r of the neighbour group that starts at sorted
// entry t (f, r: the largest forward and reverse weights of the group)
__device__ __forceinline__ float ug_group_weight(const unsigned long long* keys, int t, int len) {
  const unsigned long long nb = keys[t] >> 32;
  float f = 0.0f, r = 0.0f;
  for (int u = t; u < len && (keys[u] >> 32) == nb; ++u) {
    const float w = __uint_as_float(static_cast<unsigned>(keys[u] & 0x7fffffffu));
    if (keys[u] & kUgReverse) r = fmaxf(r, w);
    else f = fmaxf(f, w);
  }
  return __fsub_rn(__fadd_rn(f, r), __fmul_rn(f, r));
}

// One row by one block: LDS bitonic sort of up to kSortCap entries, then
// one union weight per neighbour, written to the front of the row's raw range
// as (neighbour << 32 | weight bits).
__device__ __forceinline__ void sort_union_long(unsigned long long* __restrict__ raw,
                                                const int* __restrict__ raw_ptr, int row,
                                                int* __restrict__ union_len,
                                                unsigned* error_flags) {
  __shared__ unsigned long long keys[kSortCap];
  const int beg = raw_ptr[row], len = raw_ptr[row + 1] - beg;
  if (len > kSortCap) {
    // ug_sort_union_global takes it
    if (threadIdx.x == 0) atomicOr(error_flags, kLongRow);
    return;
  }
  int pw = 1;
  while (pw < len) pw <<= 1;
  for (int t = threadIdx.x; t < pw; t += kUgThreads) keys[t] = t < len ? raw[beg + t] : ~0ull;
  __syncthreads();
  for (int size = 2; size <= pw; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < pw / 2; t += kUgThreads) {
        const int a = 2 * stride * (t / stride) + (t % stride), b = a + stride;
        const bool up = (a & size) == 0;
        const unsigned long long ka = keys[a], kb = keys[b];
        if ((ka > kb) == up) {
          keys[a] = kb;
          keys[b] = ka;
        }
      }
      __syncthreads();
    }
  }
  // a thread owns a chunk of consecutive sorted entries and the neighbour
  // groups that start in it
  constexpr int CH = kSortCap / kUgThreads;
  const int c0 = threadIdx.x * CH;
  auto starts_group = [&](int t) { return t == 0 || (keys[t] >> 32) != (keys[t - 1] >> 32); };
  int mine = 0;
  for (int t = c0; t < min(c0 + CH, len); ++t)
    if (starts_group(t) && ug_kept(ug_group_weight(keys, t, len))) ++mine;
  int kept_total;
  int out = block_exclusive_scan_nt<kUgThreads>(mine, &kept_total);
  if (threadIdx.x == 0) union_len[row] = kept_total;
  for (int t = c0; t < min(c0 + CH, len); ++t) {
    if (!starts_group(t)) continue;
    const float w = ug_group_weight(keys, t, len);
    if (!ug_kept(w)) continue;
    raw[beg + out] = (keys[t] >> 32 << 32) | __float_as_uint(w);
    ++out;
  }
}

// Union rows of (64 R/2, 64 R] raw entries (R = 1: at most 64), one wave
// each: entry e = 64 r + lane sits in register r of its lane; a bitonic sort
// exchanges across registers for strides of 64 and more and across lanes
// (shuffles) below.  The groups' union weights come from the sorted keys in
// LDS; the kept ones are written to the front of the row's raw range as
// (neighbour << 32 | weight bits).  The widest class queues the rows longer
// than 64 R for ug_sort_union.
template <int R>
__global__ __launch_bounds__(kUgThreads) void ug_sort_union_wave(
    unsigned long long* __restrict__ raw, const int* __restrict__ raw_ptr, int n,
    int* __restrict__ union_len, int* __restrict__ long_rows, int* __restrict__ long_count) {
  constexpr int kCap = 64 * R;
  __shared__ unsigned long long sk[(kUgThreads / 64) * kCap];
  const int lane = threadIdx.x & 63, wave = static_cast<int>(threadIdx.x >> 6);
  const int row = blockIdx.x * (kUgThreads / 64) + wave;
  int beg = 0, len = 0;
  bool mine = false;  // this wave writes the row
  if (row < n) {
    beg = raw_ptr[row];
    len = raw_ptr[row + 1] - beg;
    if (len > kCap && R == kWaveSortMaxR) {
      if (lane == 0) long_rows[atomicAdd(long_count, 1)] = row;
    }
    mine = wave_sort_class<R>(len);
  }
  if (!mine) return;  // wave-uniform: the wave owns no row of this class
  unsigned long long key[R];
#pragma unroll
  for (int r = 0; r < R; ++r) key[r] = 64 * r + lane < len ? raw[beg + 64 * r + lane] : ~0ull;
  wave_sort_keys<R>(key, lane);
  unsigned long long* wk = sk + wave * kCap;
#pragma unroll
  for (int r = 0; r < R; ++r) wk[64 * r + lane] = key[r];
  wave_lds_barrier();
  int base = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int e = 64 * r + lane;
    bool keep = false;
    unsigned long long out = 0;
    if (e < len && (e == 0 || (key[r] >> 32) != (wk[e - 1] >> 32))) {
      const float w = ug_group_weight(wk, e, len);
      keep = ug_kept(w);
      out = (key[r] >> 32 << 32) | __float_as_uint(w);
    }
    const unsigned long long m = __ballot(keep);
    if (keep) raw[beg + base + __popcll(m & ((1ull << lane) - 1ull))] = out;
    base += __popcll(m);
  }
  if (lane == 0) union_len[row] = base;
}

// The queued rows of more than 64 kWaveSortMaxR entries, one block per row: the same
// result through an LDS bitonic sort of up to kSortCap entries.
__global__ __launch_bounds__(kUgThreads) void ug_sort_union(unsigned long long* __restrict__ raw,
                                                            const int* __restrict__ raw_ptr,
                                                            const int* __restrict__ long_rows,
                                                            const int* __restrict__ long_count,
                                                            int* __restrict__ union_len,
                                                            unsigned* error_flags) {
  const int rows = *long_count;
  for (int q = blockIdx.x; q < rows; q += gridDim.x) {
    __syncthreads();  // the previous row's keys are read
    sort_union_long(raw, raw_ptr, long_rows[q], union_len, error_flags);
  }
}

// The queued rows longer than kSortCap: global sort (sort_row_global), then
// one union weight per neighbour written over the row.
__global__ __launch_bounds__(kLongThreads) void ug_sort_union_global(
    unsigned long long* __restrict__ raw, unsigned long long* __restrict__ scratch,
    const int* __restrict__ raw_ptr, const int* __restrict__ long_rows,
    const int* __restrict__ long_count, int* __restrict__ union_len) {
  __shared__ unsigned long long keys[kSortCap];
  const int rows = *long_count;
  for (int q = blockIdx.x; q < rows; q += gridDim.x) {
    const int row = long_rows[q];
    const int beg = raw_ptr[row], len = raw_ptr[row + 1] - beg;
    if (len <= kSortCap) continue;  // block-uniform
    __syncthreads();
    sort_row_global(raw + beg, scratch + beg, len, keys);
    const unsigned long long* sorted = scratch + beg;
    auto starts_group = [&](int t) { return t == 0 || (sorted[t] >> 32) != (sorted[t - 1] >> 32); };
    const int ch = (len + kLongThreads - 1) / kLongThreads;
    const int c0 = threadIdx.x * ch, c1 = min(c0 + ch, len);
    int mine = 0;
    for (int t = c0; t < c1; ++t)
      if (starts_group(t) && ug_kept(ug_group_weight(sorted, t, len))) ++mine;
    int kept_total;
    int out = block_exclusive_scan_nt<kLongThreads>(mine, &kept_total);
    if (threadIdx.x == 0) union_len[row] = kept_total;
    for (int t = c0; t < c1; ++t) {
      if (!starts_group(t)) continue;
      const float w = ug_group_weight(sorted, t, len);
      if (!ug_kept(w)) continue;
      raw[beg + out] = (sorted[t] >> 32 << 32) | __float_as_uint(w);
      ++out;
    }
  }
}

__global__ __launch_bounds__(kUgThreads) void ug_compact(const unsigned long long* __restrict__ raw,
                                                         const int* __restrict__ raw_ptr,
                                                         const int* __restrict__ offsets, int n,
                                                         int* __restrict__ heads,
                                                         int* __restrict__ nbrs,
                                                         float* __restrict__ weights,
                                                         unsigned* max_bits) {
  __shared__ unsigned wave_max[kUgThreads / 64];
  const int lane = threadIdx.x & 63;
  unsigned mb = 0u;
  // rows strided over the grid, so the largest weight takes one atomic per
  // block (same-address atomics serialise)
  for (int row = blockIdx.x * (kUgThreads / 64) + static_cast<int>(threadIdx.x >> 6); row < n;
       row += gridDim.x * (kUgThreads / 64)) {
    const int src = raw_ptr[row], dst = offsets[row], len = offsets[row + 1] - dst;
    for (int t = lane; t < len; t += 64) {
      const unsigned long long e = raw[src + t];
      const unsigned wb = static_cast<unsigned>(e & 0xffffffffu);
      heads[dst + t] = row;
      nbrs[dst + t] = static_cast<int>(e >> 32);
      weights[dst + t] = __uint_as_float(wb);
      mb = max(mb, wb);  // positive finite floats order as their bits
    }
  }
  for (int o = 32; o > 0; o >>= 1) mb = max(mb, static_cast<unsigned>(__shfl_xor(static_cast<int>(mb), o, 64)));
  if (lane == 0) wave_max[threadIdx.x >> 6] = mb;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned m = 0u;
    for (int w = 0; w < kUgThreads / 64; ++w) m = max(m, wave_max[w]);
    if (m) atomicMax(max_bits, m);
  }
}

__global__ __launch_bounds__(kUgThreads) void ug_epochs(const float* __restrict__ weights,
                                                        long long nnz,
                                                        const unsigned* __restrict__ max_bits,
                                                        float* __restrict__ eps) {
  const float mw = __uint_as_float(*max_bits);
  for (long long e = blockIdx.x * static_cast<long long>(kUgThreads) + threadIdx.x; e < nnz;
       e += static_cast<long long>(gridDim.x) * kUgThreads)
    eps[e] = mw > 0.0f ? __fdiv_rn(mw, fmaxf(weights[e], 1.0e-6f)) : INFINITY;
}

int ug_scan(const int* in, int n, int* out, int* sums, int* total, hipStream_t s) {
  const int nb = (n + kScanBlock - 1) / kScanBlock;
  scan_blocks<<<nb, 256, 0, s>>>(in, n, out, sums);
  FE_LAUNCH_CHECK("scan_blocks");
  scan_sums<<<1, 256, 0, s>>>(sums, nb, total);
  FE_LAUNCH_CHECK("scan_sums");
  scan_add<<<(n + 255) / 256, 256, 0, s>>>(out, n, sums);
  FE_LAUNCH_CHECK("scan_add");
  return 0;
}

}  // namespace

class UmapGraph {
 public:
  int n = 0;
  long long nnz = 0;
  float max_weight = 0.0f;
  Arena store;  // offsets, heads, neighbours, weights, epochs_per_sample
  int* offsets = nullptr;
  int* heads = nullptr;
  int* nbrs = nullptr;
  float* weights = nullptr;
  float* eps = nullptr;

  int build(const int* indices, const float* distances, int n_, int k_total, int col_start,
            int n_cols, int index_offset, bool device_inputs);
};

int UmapGraph::build(const int* indices, const float* distances, int n_, int k_total,
                     int col_start, int n_cols, int index_offset, bool device_inputs) {
  n = n_;
  Stream stream;
  if (stream.init()) return 1;
  hipStream_t s = stream.get();
  const size_t total = static_cast<size_t>(n) * k_total;
  const int k = n_cols;
  // inputs, per-row smoothing, both-endpoint degrees and raw entries
  Arena work;
  ArenaPlan wp;
  if (!device_inputs) {
    wp.add<int>(total);
    wp.add<float>(total);
  }
  const int nbs = (n + 1 + kScanBlock - 1) / kScanBlock;
  wp.add<int>(2);
  wp.add<DD>(kUgMeanBlocks);
  wp.add<unsigned long long>(kUgMeanBlocks);
  wp.add<double>(1);
  wp.add<float>(n);
  wp.add<float>(n);
  wp.add<int>(n + 1);   // degrees
  wp.add<int>(n + 1);   // raw row offsets
  wp.add<int>(n + 1);   // cursors / union lengths
  wp.add<int>(n + 1);   // final offsets
  wp.add<int>(nbs);
  wp.add<int>(1);
  wp.add<int>(1);       // rows queued for the block sort
  wp.add<unsigned>(2);  // error flags, max weight bits
  if (work.init(wp.bytes, "UMAP fuzzy graph workspace")) return 1;
  const int* d_idx = indices;
  const float* d_dist = distances;
  if (!device_inputs) {
    int* hi = work.take<int>(total);
    float* hd = work.take<float>(total);
    FE_CHECK(hipMemcpyAsync(hi, indices, total * sizeof(int), hipMemcpyHostToDevice, s));
    FE_CHECK(hipMemcpyAsync(hd, distances, total * sizeof(float), hipMemcpyHostToDevice, s));
    d_idx = hi;
    d_dist = hd;
  }
  int* range = work.take<int>(2);
  DD* mpart = work.take<DD>(kUgMeanBlocks);
  unsigned long long* mcount = work.take<unsigned long long>(kUgMeanBlocks);
  double* gmean = work.take<double>(1);
  float* sigmas = work.take<float>(n);
  float* rhos = work.take<float>(n);
  int* deg = work.take<int>(n + 1);
  int* raw_ptr = work.take<int>(n + 1);
  int* cursor = work.take<int>(n + 1);
  int* offs = work.take<int>(n + 1);
  int* sums = work.take<int>(nbs);
  int* tot = work.take<int>(1);
  int* long_count = work.take<int>(1);
  unsigned* flags = work.take<unsigned>(2);
  FE_CHECK(hipMemsetAsync(flags, 0, 2 * sizeof(unsigned), s));
  if (index_offset < 0) {
    const int init[2] = {INT_MAX, INT_MIN};
    FE_CHECK(hipMemcpyAsync(range, init, sizeof(init), hipMemcpyHostToDevice, s));
    const int blocks = static_cast<int>(std::min<long long>((static_cast<long long>(n) * k + kUgThreads - 1) / kUgThreads, 4096));
    ug_index_range<<<blocks, kUgThreads, 0, s>>>(d_idx, n, col_start, k, range);
    FE_LAUNCH_CHECK("ug_index_range");
    int h[2];
    FE_CHECK(hipMemcpyAsync(h, range, sizeof(h), hipMemcpyDeviceToHost, s));
    FE_CHECK(hipStreamSynchronize(s));
    // fast_knn_umap.cpp:4013-4022
    index_offset = (h[0] >= 1 && h[1] <= n) ? 1 : 0;
  }
  ug_mean_partials<<<kUgMeanBlocks, kUgThreads, 0, s>>>(d_dist, n, col_start, k, mpart, mcount);
  FE_LAUNCH_CHECK("ug_mean_partials");
  ug_mean_fold<<<1, 64, 0, s>>>(mpart, mcount, kUgMeanBlocks, gmean);
  FE_LAUNCH_CHECK("ug_mean_fold");
  ug_smooth_rows<<<(n + kUgThreads - 1) / kUgThreads, kUgThreads, 0, s>>>(d_dist, n, col_start, k,
                                                                           gmean, sigmas, rhos);
  FE_LAUNCH_CHECK("ug_smooth_rows");
  FE_CHECK(hipMemsetAsync(deg, 0, (n + 1) * sizeof(int), s));
  const int eblocks = static_cast<int>(std::min<long long>((static_cast<long long>(n) * k + kUgThreads - 1) / kUgThreads, 8192));
  ug_count<<<eblocks, kUgThreads, 0, s>>>(d_idx, d_dist, n, col_start, k, index_offset, sigmas,
                                          rhos, deg);
  FE_LAUNCH_CHECK("ug_count");
  if (ug_scan(deg, n + 1, raw_ptr, sums, tot, s)) return 1;
  int h_raw = 0;
  FE_CHECK(hipMemcpyAsync(&h_raw, raw_ptr + n, sizeof(int), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  Arena rawa;
  ArenaPlan rp;
  rp.add<unsigned long long>(std::max(h_raw, 1));
  if (rawa.init(rp.bytes, "UMAP fuzzy graph entries")) return 1;
  unsigned long long* raw = rawa.take<unsigned long long>(std::max(h_raw, 1));
  FE_CHECK(hipMemsetAsync(cursor, 0, (n + 1) * sizeof(int), s));
  ug_fill<<<eblocks, kUgThreads, 0, s>>>(d_idx, d_dist, n, col_start, k, index_offset, sigmas,
                                         rhos, raw_ptr, cursor, raw);
  FE_LAUNCH_CHECK("ug_fill");
  // cursors become the union lengths (the last slot stays 0 for the scan)
  FE_CHECK(hipMemsetAsync(cursor, 0, (n + 1) * sizeof(int), s));
  // rows of up to 512 entries by one wave each, one launch per length class;
  // the degrees are no longer needed and hold the queue of longer rows
  FE_CHECK(hipMemsetAsync(long_count, 0, sizeof(int), s));
  constexpr int kRowsPerBlock = kUgThreads / 64;
  const int wblocks = (n + kRowsPerBlock - 1) / kRowsPerBlock;
  ug_sort_union_wave<1><<<wblocks, kUgThreads, 0, s>>>(raw, raw_ptr, n, cursor, deg, long_count);
  ug_sort_union_wave<2><<<wblocks, kUgThreads, 0, s>>>(raw, raw_ptr, n, cursor, deg, long_count);
  ug_sort_union_wave<4><<<wblocks, kUgThreads, 0, s>>>(raw, raw_ptr, n, cursor, deg, long_count);
  ug_sort_union_wave<kWaveSortMaxR><<<wblocks, kUgThreads, 0, s>>>(raw, raw_ptr, n, cursor, deg,
                                                                 long_count);
  FE_LAUNCH_CHECK("ug_sort_union_wave");
  ug_sort_union<<<std::min(n, 2048), kUgThreads, 0, s>>>(raw, raw_ptr, deg, long_count, cursor,
                                                         flags);
  FE_LAUNCH_CHECK("ug_sort_union");
  if (ug_scan(cursor, n + 1, offs, sums, tot, s)) return 1;
  int h_nnz = 0;
  unsigned h_flags = 0;
  FE_CHECK(hipMemcpyAsync(&h_nnz, offs + n, sizeof(int), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipMemcpyAsync(&h_flags, flags, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  Arena long_scratch;
  if (h_flags & kLongRow) {
    // rows longer than kSortCap: sorted in global memory, one block each,
    // then the union lengths scanned again
    ArenaPlan lp;
    lp.add<unsigned long long>(std::max(h_raw, 1));
    if (long_scratch.init(lp.bytes, "UMAP fuzzy graph long-row sort")) return 1;
    unsigned long long* scratch = long_scratch.take<unsigned long long>(std::max(h_raw, 1));
    ug_sort_union_global<<<64, kLongThreads, 0, s>>>(raw, scratch, raw_ptr, deg, long_count,
                                                    cursor);
    FE_LAUNCH_CHECK("ug_sort_union_global");
    if (ug_scan(cursor, n + 1, offs, sums, tot, s)) return 1;
    FE_CHECK(hipMemcpyAsync(&h_nnz, offs + n, sizeof(int), hipMemcpyDeviceToHost, s));
    FE_CHECK(hipStreamSynchronize(s));
  }
  nnz = h_nnz;
  ArenaPlan sp;
  sp.add<int>(n + 1);
  sp.add<int>(std::max<long long>(nnz, 1));
  sp.add<int>(std::max<long long>(nnz, 1));
  sp.add<float>(std::max<long long>(nnz, 1));
  sp.add<float>(std::max<long long>(nnz, 1));
  if (store.init(sp.bytes, "UMAP fuzzy graph")) return 1;
  offsets = store.take<int>(n + 1);
  heads = store.take<int>(std::max<long long>(nnz, 1));
  nbrs = store.take<int>(std::max<long long>(nnz, 1));
  weights = store.take<float>(std::max<long long>(nnz, 1));
  eps = store.take<float>(std::max<long long>(nnz, 1));
  FE_CHECK(hipMemcpyAsync(offsets, offs, (n + 1) * sizeof(int), hipMemcpyDeviceToDevice, s));
  ug_compact<<<std::min((n + 3) / 4, 2048), kUgThreads, 0, s>>>(raw, raw_ptr, offs, n, heads,
                                                                nbrs, weights, flags + 1);
  FE_LAUNCH_CHECK("ug_compact");
  if (nnz > 0) {
    const int pblocks = static_cast<int>(std::min<long long>((nnz + kUgThreads - 1) / kUgThreads, 8192));
    ug_epochs<<<pblocks, kUgThreads, 0, s>>>(weights, nnz, flags + 1, eps);
    FE_LAUNCH_CHECK("ug_epochs");
  }
  unsigned mb = 0;
  FE_CHECK(hipMemcpyAsync(&mb, flags + 1, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  std::memcpy(&max_weight, &mb, sizeof(float));
  return 0;
}

}  // namespace fe

namespace fe {
int umap_run(const int* heads, const int* tails, const float* weights,
             const float* epochs_per_sample, const float* init, int n, int n_edges, int n_epochs,
             int nsr, float lr, float a, float b, float gamma, unsigned seed, int mode, float* out,
             bool device_inputs, bool symmetric);  // umap.hip
}  // namespace fe

struct fastembedr_hip_umap_graph {
  fe::UmapGraph g;
};

extern "C" {

int fastembedr_hip_umap_graph_create(const int* indices, const float* distances, int n,
                                     int k_total, int col_start, int n_cols, int edge_budget,
                                     int index_offset, int device_inputs,
                                     fastembedr_hip_umap_graph** graph) {
  using namespace fe;
  clear_error();
  if (!indices || !distances || !graph) return fail("null pointer");
  *graph = nullptr;
  // fast_knn_umap.cpp:4001-4009
  if (n < 2 || k_total < 1) return fail("indices and distances must have the same dimensions");
  if (col_start < 0 || n_cols < 1 || col_start + n_cols > k_total)
    return fail("invalid KNN column range");
  if (edge_budget < 1) return fail("edge_budget must be positive");
  if (edge_budget < n_cols)
    return fail("the HIP fuzzy graph supports the default edge budget (edge_budget >= n_cols) only");
  if (n_cols > kMaxK) return fail("the HIP fuzzy graph supports at most 256 neighbours");
  auto h = std::make_unique<fastembedr_hip_umap_graph>();
  if (h->g.build(indices, distances, n, k_total, col_start, n_cols, index_offset,
                 device_inputs != 0))
    return 1;
  *graph = h.release();
  return 0;
}

int fastembedr_hip_umap_graph_info(fastembedr_hip_umap_graph* graph, long long* nnz,
                                   float* max_weight) {
  fe::clear_error();
  if (!graph) return fe::fail("null graph");
  if (nnz) *nnz = graph->g.nnz;
  if (max_weight) *max_weight = graph->g.max_weight;
  return 0;
}

int fastembedr_hip_umap_graph_export(fastembedr_hip_umap_graph* graph, int* offsets,
                                     int* neighbors, float* weights, float* epochs_per_sample) {
  using namespace fe;
  clear_error();
  if (!graph) return fail("null graph");
  const UmapGraph& g = graph->g;
  if (offsets)
    FE_CHECK(hipMemcpy(offsets, g.offsets, (g.n + 1) * sizeof(int), hipMemcpyDeviceToHost));
  if (g.nnz > 0) {
    if (neighbors)
      FE_CHECK(hipMemcpy(neighbors, g.nbrs, g.nnz * sizeof(int), hipMemcpyDeviceToHost));
    if (weights)
      FE_CHECK(hipMemcpy(weights, g.weights, g.nnz * sizeof(float), hipMemcpyDeviceToHost));
    if (epochs_per_sample)
      FE_CHECK(hipMemcpy(epochs_per_sample, g.eps, g.nnz * sizeof(float), hipMemcpyDeviceToHost));
  }
  return 0;
}

int fastembedr_hip_umap_graph_device(fastembedr_hip_umap_graph* graph, const int** heads,
                                     const int** tails, const float** weights,
                                     const float** epochs_per_sample) {
  fe::clear_error();
  if (!graph) return fe::fail("null graph");
  if (heads) *heads = graph->g.heads;
  if (tails) *tails = graph->g.nbrs;
  if (weights) *weights = graph->g.weights;
  if (epochs_per_sample) *epochs_per_sample = graph->g.eps;
  return 0;
}

int fastembedr_hip_umap_graph_optimize(fastembedr_hip_umap_graph* graph, const float* init,
                                       int n_epochs, int negative_sample_rate,
                                       float learning_rate, float a, float b,
                                       float repulsion_strength, unsigned int seed,
                                       int optimizer_mode, float* out) {
  fe::clear_error();
  if (!graph) return fe::fail("null graph");
  const auto& g = graph->g;
  if (g.nnz < 1) return fe::fail("the fuzzy graph has no edges");
  // the graph is symmetric with equal weights both ways (f + r - f*r of a
  // row's forward and reverse memberships is the same float either way), so
  // the deterministic mode applies each pair's attraction through its heads
  return fe::umap_run(g.heads, g.nbrs, g.weights, g.eps, init, g.n, static_cast<int>(g.nnz),
                      n_epochs, negative_sample_rate, learning_rate, a, b, repulsion_strength,
                      seed, optimizer_mode, out, true, true);
}

void fastembedr_hip_umap_graph_destroy(fastembedr_hip_umap_graph* graph) { delete graph; }

}  // extern "C"
