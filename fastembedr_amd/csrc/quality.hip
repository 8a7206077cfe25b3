// Exact structure metrics on gfx950 (exact_structure_metrics_cpp,
// src/embedding_utils.cpp:1582-1752): for a sample of n <= 5000 observations
// with high-dimensional and low-dimensional coordinates, every row ranks all
// other rows by FP64 squared distance in both spaces (ties by index), and
// the k nearest in each space give trustworthiness, continuity, kNN
// preservation and mean neighbour-rank error for every requested k.
//
// One 1024-thread workgroup per row: the squared distances are formed in
// FP64 in the reference's order (sum over columns of delta * delta, no fused
// multiply-add), sorted in LDS as (distance bits, index) pairs by a bitonic
// network, turned into ranks, and the row's penalties are summed as 64-bit
// integers, so the result equals the reference's bit for bit whatever the
// order of the rows.  A second launch adds the rows' sums in row order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fastembedr_hip.h"
#include "fe_common.hpp"

namespace fe {
namespace {

constexpr int kRankThreads = 1024;
constexpr int kRankMaxN = 5000;
constexpr int kRankSums = 4;  // shared, trust penalty, continuity penalty, rank error

__host__ __device__ constexpr int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// squared distance of rows i and j of a column-major n x p matrix
__device__ __forceinline__ double col_major_d2(const double* __restrict__ m, int n, int p, int i,
                                               int j) {
  double a = 0.0;
  for (int c = 0; c < p; ++c) {
    const double d = __dsub_rn(m[static_cast<size_t>(c) * n + i], m[static_cast<size_t>(c) * n + j]);
    a = __dadd_rn(a, __dmul_rn(d, d));
  }
  return a;
}

// (key, index) pairs, ascending by key then index
__device__ __forceinline__ bool pair_greater(unsigned long long ka, int ia, unsigned long long kb,
                                             int ib) {
  return ka > kb || (ka == kb && ia > ib);
}

__device__ void bitonic_sort(unsigned long long* key, int* idx, int m) {
  for (int size = 2; size <= m; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < (m >> 1); t += kRankThreads) {
        const int a = 2 * stride * (t / stride) + (t % stride), b = a + stride;
        const bool up = (a & size) == 0;
        const unsigned long long ka = key[a], kb = key[b];
        const int ia = idx[a], ib = idx[b];
        if (pair_greater(ka, ia, kb, ib) == up) {
          key[a] = kb;
          key[b] = ka;
          idx[a] = ib;
          idx[b] = ia;
        }
      }
      __syncthreads();
    }
  }
}

// the row's distances to every row in one space, sorted; the row itself and
// the padding sort last (key all ones; the row before the padding indices)
__device__ void sort_row(const double* __restrict__ m, int n, int p, int i, int mm,
                         unsigned long long* key, int* idx) {
  for (int j = threadIdx.x; j < mm; j += kRankThreads) {
    unsigned long long k = ~0ull;
    if (j < n && j != i) k = static_cast<unsigned long long>(__double_as_longlong(col_major_d2(m, n, p, i, j)));
    key[j] = k;
    idx[j] = j;
  }
  __syncthreads();
  bitonic_sort(key, idx, mm);
}

__device__ __forceinline__ long long block_sum_ll(long long v, long long* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  long long t = 0;
  if (threadIdx.x == 0)
    for (int q = 0; q < kRankThreads / 64; ++q) t += red[q];
  __syncthreads();
  return t;
}

// One row: sums[(row * nk + q) * 4 + s] for every requested k.
// LDS: key[mm], idx[mm], high ranks[n], low ranks[n], high order[kmax].
__global__ __launch_bounds__(kRankThreads) void rank_rows(const double* __restrict__ high, int ph,
                                                          const double* __restrict__ low, int pl,
                                                          int n, int mm,
                                                          const int* __restrict__ ks, int nk,
                                                          int kmax, long long* __restrict__ sums) {
  extern __shared__ unsigned long long smem[];
  __shared__ long long red[kRankThreads / 64];
  unsigned long long* key = smem;
  int* idx = reinterpret_cast<int*>(key + mm);
  int* hr = idx + mm;
  int* lr = hr + n;
  int* ho = lr + n;
  const int i = blockIdx.x;
  sort_row(high, n, ph, i, mm, key, idx);
  for (int r = threadIdx.x; r < n - 1; r += kRankThreads) hr[idx[r]] = r + 1;
  for (int r = threadIdx.x; r < kmax; r += kRankThreads) ho[r] = idx[r];
  __syncthreads();
  sort_row(low, n, pl, i, mm, key, idx);
  for (int r = threadIdx.x; r < n - 1; r += kRankThreads) lr[idx[r]] = r + 1;
  __syncthreads();
  for (int q = 0; q < nk; ++q) {
    const int k = ks[q];
    long long shared = 0, trust = 0, cont = 0, rerr = 0;
    for (int r = threadIdx.x; r < k; r += kRankThreads) {
      const int rank_in_high = hr[idx[r]];
      const int rank_in_low = lr[ho[r]];
      if (rank_in_high <= k) ++shared;
      else trust += rank_in_high - k;
      if (rank_in_low > k) cont += rank_in_low - k;
      rerr += rank_in_high > r + 1 ? rank_in_high - (r + 1) : (r + 1) - rank_in_high;
    }
    shared = block_sum_ll(shared, red);
    trust = block_sum_ll(trust, red);
    cont = block_sum_ll(cont, red);
    rerr = block_sum_ll(rerr, red);
    if (threadIdx.x == 0) {
      long long* o = sums + (static_cast<size_t>(i) * nk + q) * kRankSums;
      o[0] = shared;
      o[1] = trust;
      o[2] = cont;
      o[3] = rerr;
    }
  }
}

// totals over the rows, in row order (integers: exact)
__global__ void rank_totals(const long long* __restrict__ sums, int n, int width,
                            long long* __restrict__ totals) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= width) return;
  long long t = 0;
  for (int i = 0; i < n; ++i) t += sums[static_cast<size_t>(i) * width + s];
  totals[s] = t;
}

}  // namespace
}  // namespace fe

extern "C" {

int fastembedr_hip_exact_structure_metrics(const double* high, int n, int high_cols,
                                           const double* low, int low_cols,
                                           const int* requested_k, int nk, double* out) {
  using namespace fe;
  clear_error();
  if (high == nullptr || low == nullptr || requested_k == nullptr || out == nullptr)
    return fail("null host pointer");
  // embedding_utils.cpp:1587-1605
  if (n < 3) return fail("at least three observations are required");
  if (high_cols < 1 || low_cols < 1) return fail("input matrices must have columns");
  if (n > kRankMaxN) return fail("exact rank metrics support at most 5000 sampled observations");
  if (nk < 1) return fail("requested_k must not be empty");
  int kmax = 0;
  for (int q = 0; q < nk; ++q) {
    const int k = requested_k[q];
    if (k < 1 || k >= n || 2 * n - 3 * k - 1 <= 0)
      return fail("each k must satisfy 1 <= k < (2 * n - 1) / 3");
    kmax = std::max(kmax, k);
  }
  const int mm = next_pow2(n);
  const size_t lds = static_cast<size_t>(mm) * (sizeof(unsigned long long) + sizeof(int)) +
                     2 * static_cast<size_t>(n) * sizeof(int) + static_cast<size_t>(kmax) * sizeof(int);
  Stream stream;
  if (stream.init()) return 1;
  hipStream_t s = stream.get();
  const size_t hb = static_cast<size_t>(n) * high_cols, lb = static_cast<size_t>(n) * low_cols;
  const int width = nk * kRankSums;
  ArenaPlan ap;
  ap.add<double>(hb);
  ap.add<double>(lb);
  ap.add<int>(nk);
  ap.add<long long>(static_cast<size_t>(n) * width);
  ap.add<long long>(width);
  Arena mem;
  if (mem.init(ap.bytes, "structure metrics")) return 1;
  double* d_high = mem.take<double>(hb);
  double* d_low = mem.take<double>(lb);
  int* d_ks = mem.take<int>(nk);
  long long* d_sums = mem.take<long long>(static_cast<size_t>(n) * width);
  long long* d_tot = mem.take<long long>(width);
  FE_CHECK(hipMemcpyAsync(d_high, high, hb * sizeof(double), hipMemcpyHostToDevice, s));
  FE_CHECK(hipMemcpyAsync(d_low, low, lb * sizeof(double), hipMemcpyHostToDevice, s));
  FE_CHECK(hipMemcpyAsync(d_ks, requested_k, nk * sizeof(int), hipMemcpyHostToDevice, s));
  FE_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&rank_rows),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
  rank_rows<<<n, kRankThreads, lds, s>>>(d_high, high_cols, d_low, low_cols, n, mm, d_ks, nk, kmax,
                                         d_sums);
  FE_LAUNCH_CHECK("rank_rows");
  rank_totals<<<(width + 63) / 64, 64, 0, s>>>(d_sums, n, width, d_tot);
  FE_LAUNCH_CHECK("rank_totals");
  std::vector<long long> tot(width);
  FE_CHECK(hipMemcpyAsync(tot.data(), d_tot, width * sizeof(long long), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  // embedding_utils.cpp:1732-1743
  for (int q = 0; q < nk; ++q) {
    const double shared = static_cast<double>(tot[q * kRankSums]);
    const double trust_penalty = static_cast<double>(tot[q * kRankSums + 1]);
    const double continuity_penalty = static_cast<double>(tot[q * kRankSums + 2]);
    const double rank_error = static_cast<double>(tot[q * kRankSums + 3]);
    const double k = static_cast<double>(requested_k[q]);
    const double normalizer = 2.0 / (static_cast<double>(n) * k * (2.0 * n - 3.0 * k - 1.0));
    out[q] = std::max(0.0, std::min(1.0, 1.0 - normalizer * trust_penalty));
    out[nk + q] = std::max(0.0, std::min(1.0, 1.0 - normalizer * continuity_penalty));
    out[2 * nk + q] = shared / (static_cast<double>(n) * k);
    out[3 * nk + q] = rank_error / (static_cast<double>(n) * k);
  }
  return 0;
}

}  // extern "C"
