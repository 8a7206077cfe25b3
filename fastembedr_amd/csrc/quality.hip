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

__host__ __device__ constexpr int pow2_at_least_i(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

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

// ---- KNN structure score ----
//
// knn_structure_score (knn_structure_score_cpp, src/embedding_utils.cpp:
// 1386-1580; structure_score_rows_kernel, src/embedding_cuda_kernels.cpp:
// 1827-1944): one workgroup per kept row scans every row of the layout once
// -- FP64 squared distances, each counted against the row's sorted
// high-dimensional neighbours (continuity ranks) and offered to the row's
// list of nearest layout rows (LDS, merged by a bitonic sort) -- then the
// preservation, local trustworthiness, continuity and label-majority terms;
// the rows' terms are added in row order.

constexpr int kScoreThreads = 256;
constexpr int kScoreMaxK = 256;
constexpr int kScoreSlots = 512;  // list (<= 256) + one chunk of candidates
constexpr int kScoreWidth = 6;

__device__ __forceinline__ double layout_d2(const double* __restrict__ y, int n, int dim, int i,
                                            int j) {
  double d = 0.0;
  for (int c = 0; c < dim; ++c) {
    const double diff = __dsub_rn(y[static_cast<size_t>(c) * n + i], y[static_cast<size_t>(c) * n + j]);
    d = __dadd_rn(d, __dmul_rn(diff, diff));
  }
  return d;
}

__device__ __forceinline__ bool pair_lt(double a, int ia, double b, int ib) {
  return a < b || (a == b && ia < ib);
}

// ascending bitonic sort of the first m (power of two) pairs, whole block
__device__ void block_sort_pairs(double* d, int* j, int m) {
  for (int size = 2; size <= m; size <<= 1) {
    for (int h = size >> 1; h > 0; h >>= 1) {
      for (int t = threadIdx.x; t < m / 2; t += kScoreThreads) {
        const int a = 2 * h * (t / h) + (t % h), b = a + h;
        const bool up = (a & size) == 0;
        if (pair_lt(d[b], j[b], d[a], j[a]) == up) {
          const double td = d[a];
          const int tj = j[a];
          d[a] = d[b];
          j[a] = j[b];
          d[b] = td;
          j[b] = tj;
        }
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(kScoreThreads) void score_rows(
    const double* __restrict__ y, int n, int dim, const int* __restrict__ ind, int index_rows,
    int limit, const int* __restrict__ keep, int compact, int k, const int* __restrict__ labels,
    int n_levels, double* __restrict__ rows) {
  __shared__ double td[kScoreMaxK];
  __shared__ int tj[kScoreMaxK];
  __shared__ int diff[kScoreMaxK + 1];
  __shared__ double sd[kScoreSlots];
  __shared__ int sj[kScoreSlots];
  __shared__ int ia[kScoreThreads], ib[kScoreThreads];
  __shared__ int pending, n_targets, repeated;
  const int kk = blockIdx.x, tid = threadIdx.x;
  double* out = rows + static_cast<size_t>(kk) * kScoreWidth;
  const int q = keep[kk] - 1;
  const int row = compact ? kk : q;
  if (q < 0 || q >= n || row < 0 || row >= index_rows) {  // not scored
    if (tid < kScoreWidth) out[tid] = 0.0;
    return;
  }
  // the targets: the first k high-dimensional neighbours, by (layout d2, index)
  {
    double d = INFINITY;
    int j = INT_MAX;
    if (tid < k) {
      const int nb = ind[static_cast<size_t>(tid) * index_rows + row] - 1;
      if (nb >= 0 && nb < n) {
        d = layout_d2(y, n, dim, q, nb);
        j = nb;
      }
    }
    td[tid] = d;
    tj[tid] = j;
  }
  diff[tid] = 0;
  for (int s = tid; s < kScoreSlots; s += kScoreThreads) {
    sd[s] = INFINITY;
    sj[s] = INT_MAX;
  }
  if (tid == 0) {
    diff[kScoreMaxK] = 0;
    pending = 0;
    n_targets = 0;
    repeated = 0;
  }
  __syncthreads();
  block_sort_pairs(td, tj, kScoreMaxK);
  if (tj[tid] != INT_MAX && (tid == kScoreMaxK - 1 || tj[tid + 1] == INT_MAX)) n_targets = tid + 1;
  // a target listed twice (adjacent after the sort) or the row itself
  if (tj[tid] != INT_MAX && (tj[tid] == q || (tid > 0 && tj[tid] == tj[tid - 1]))) repeated = 1;
  __syncthreads();
  const int T = n_targets;
  const bool bound = T == k && repeated == 0;
  const int KF = pow2_at_least_i(k);
  const int N = pow2_at_least_i(KF + kScoreThreads);
  // a key at or below tau may enter the list: at first the farthest target
  // (k distinct targets other than the row are k candidates, so the k-th
  // nearest is no farther)
  double tau_d = bound ? td[T - 1] : INFINITY;
  int tau_j = bound ? tj[T - 1] : INT_MAX;
  for (int b0 = 0; b0 < n; b0 += kScoreThreads) {
    const int c = b0 + tid;
    if (c < n && c != q) {
      const double d = layout_d2(y, n, dim, q, c);
      int lo = 0, hi = T;  // the first target after (d, c)
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (pair_lt(d, c, td[mid], tj[mid])) hi = mid;
        else lo = mid + 1;
      }
      if (lo < T) atomicAdd(&diff[lo], 1);
      if (d < tau_d || (d == tau_d && c <= tau_j)) {
        const int pos = KF + atomicAdd(&pending, 1);
        sd[pos] = d;
        sj[pos] = c;
      }
    }
    __syncthreads();
    if (pending > 0) {
      block_sort_pairs(sd, sj, N);
      for (int s = KF + tid; s < N; s += kScoreThreads) {
        sd[s] = INFINITY;
        sj[s] = INT_MAX;
      }
      if (tid == 0) pending = 0;
      __syncthreads();
      if (sj[k - 1] != INT_MAX) {
        tau_d = sd[k - 1];
        tau_j = sj[k - 1];
      }
    }
  }
  if (sj[k - 1] == INT_MAX) {  // fewer than k other rows
    if (tid < kScoreWidth) out[tid] = 0.0;
    return;
  }
  // high ranks of the k nearest layout rows
  if (tid < k) {
    const int c = sj[tid];
    int rank = limit + 1;
    for (int s = 0; s < limit; ++s)
      if (ind[static_cast<size_t>(s) * index_rows + row] - 1 == c) {
        rank = s + 1;
        break;
      }
    ia[tid] = rank;
    int count = 0, lab = 0;
    if (labels != nullptr) {
      lab = labels[c];
      if (lab >= 1 && lab <= n_levels)
        for (int s = 0; s < k; ++s) count += labels[sj[s]] == lab ? 1 : 0;
    }
    ib[tid] = count;
  }
  __syncthreads();
  if (tid == 0) {
    int shared = 0;
    double trust_penalty = 0.0;
    for (int r = 0; r < k; ++r) {
      if (ia[r] <= k) ++shared;
      else trust_penalty += static_cast<double>(ia[r] - k);
    }
    double cont_penalty = 0.0;
    int before = 0;
    for (int t = 0; t < T; ++t) {
      before += diff[t];
      const int low_rank = 1 + before;
      if (low_rank > k) cont_penalty += static_cast<double>(low_rank - k);
    }
    double hit = 0.0, hit_n = 0.0;
    if (labels != nullptr && n_levels > 0) {
      const int truth = labels[q];
      if (truth >= 1 && truth <= n_levels) {
        int best = 0, best_count = 0;  // the largest count, then the smallest label
        for (int r = 0; r < k; ++r) {
          const int lab = labels[sj[r]];
          if (ib[r] > best_count || (ib[r] == best_count && ib[r] > 0 && lab < best)) {
            best_count = ib[r];
            best = lab;
          }
        }
        if (best_count > 0) {
          hit = best == truth ? 1.0 : 0.0;
          hit_n = 1.0;
        }
      }
    }
    const double kd = static_cast<double>(k);
    const double trust_denom = kd * static_cast<double>(max(1, limit + 1 - k));
    const double cont_denom = kd * static_cast<double>(max(1, n - k));
    out[0] = static_cast<double>(shared) / kd;
    out[1] = fmax(0.0, fmin(1.0, 1.0 - trust_penalty / trust_denom));
    out[2] = fmax(0.0, fmin(1.0, 1.0 - cont_penalty / cont_denom));
    out[3] = hit;
    out[4] = hit_n;
    out[5] = 1.0;
  }
}

// totals of the rows, in row order
__global__ void score_totals(const double* __restrict__ rows, int keep_n,
                             double* __restrict__ totals) {
  const int m = threadIdx.x;
  if (m >= kScoreWidth) return;
  double t = 0.0;
  for (int r = 0; r < keep_n; ++r) t = __dadd_rn(t, rows[static_cast<size_t>(r) * kScoreWidth + m]);
  totals[m] = t;
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

int fastembedr_hip_knn_structure_score(const double* layout, const int* indices, const int* keep,
                                       const int* labels, int n, int index_rows,
                                       int high_rank_limit, int preserve_k, int keep_n,
                                       int compact_indices, int n_label_levels, double* totals) {
  using namespace fe;
  clear_error();
  // fastembedr_cuda_knn_structure_score's checks (embedding_cuda_kernels.cpp:4583-4596)
  if (layout == nullptr || indices == nullptr || keep == nullptr || totals == nullptr)
    return fail("null host pointer");
  if (n < 2 || index_rows < 1 || high_rank_limit < 1 || preserve_k < 1 ||
      preserve_k > high_rank_limit || preserve_k > kScoreMaxK || keep_n < 1)
    return fail("invalid structure-score dimensions");
  if (n_label_levels > 0 && labels == nullptr) return fail("labels pointer is null");
  const int dim = 2;
  Stream stream;
  if (stream.init()) return 1;
  hipStream_t s = stream.get();
  const size_t yb = static_cast<size_t>(n) * dim;
  const size_t ib = static_cast<size_t>(index_rows) * high_rank_limit;
  const bool with_labels = n_label_levels > 0;
  ArenaPlan ap;
  ap.add<double>(yb);
  ap.add<int>(ib);
  ap.add<int>(keep_n);
  if (with_labels) ap.add<int>(n);
  ap.add<double>(static_cast<size_t>(keep_n) * kScoreWidth);
  ap.add<double>(kScoreWidth);
  Arena mem;
  if (mem.init(ap.bytes, "structure score")) return 1;
  double* d_y = mem.take<double>(yb);
  int* d_ind = mem.take<int>(ib);
  int* d_keep = mem.take<int>(keep_n);
  int* d_lab = with_labels ? mem.take<int>(n) : nullptr;
  double* d_rows = mem.take<double>(static_cast<size_t>(keep_n) * kScoreWidth);
  double* d_tot = mem.take<double>(kScoreWidth);
  FE_CHECK(hipMemcpyAsync(d_y, layout, yb * sizeof(double), hipMemcpyHostToDevice, s));
  FE_CHECK(hipMemcpyAsync(d_ind, indices, ib * sizeof(int), hipMemcpyHostToDevice, s));
  FE_CHECK(hipMemcpyAsync(d_keep, keep, keep_n * sizeof(int), hipMemcpyHostToDevice, s));
  if (with_labels)
    FE_CHECK(hipMemcpyAsync(d_lab, labels, static_cast<size_t>(n) * sizeof(int),
                            hipMemcpyHostToDevice, s));
  score_rows<<<keep_n, kScoreThreads, 0, s>>>(d_y, n, dim, d_ind, index_rows, high_rank_limit,
                                              d_keep, compact_indices ? 1 : 0, preserve_k, d_lab,
                                              with_labels ? n_label_levels : 0, d_rows);
  FE_LAUNCH_CHECK("score_rows");
  score_totals<<<1, 64, 0, s>>>(d_rows, keep_n, d_tot);
  FE_LAUNCH_CHECK("score_totals");
  FE_CHECK(hipMemcpyAsync(totals, d_tot, kScoreWidth * sizeof(double), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  return 0;
}

}  // extern "C"
