// Device construction of the symmetric t-SNE affinity matrix P (setup path,
// run once per job).
//
//   perplexity_rows   one wavefront per KNN row: FP64 bisection on beta to the
//                     target entropy log(perplexity), with the min-d^2 shift,
//                     tolerance 1e-5 and 200 iterations of
//                     compute_row_probabilities_flat (tsne_neighbors.cpp:
//                     452-530).  Reads the reference's column-major k x n
//                     KNN layout through an LDS transpose; writes row-major
//                     0-based neighbour ids and float p_{j|i}.
//   sym_row_counts / sym_count / sym_fill / sym_sort_merge_wave /
//   sym_sort_merge / sym_compact
//                     the symmetrisation of build_tsne_probabilities_float
//                     (tsne_neighbors.cpp:616-724): every directed entry is
//                     placed in both endpoint rows (the forward ones per row
//                     without atomics, the reverse ones with one atomic
//                     each), each row is sorted by (column, value bits) —
//                     by one wave for rows of up to 512 entries, by one
//                     block in LDS up to kSortCap, by one block in global
//                     memory (LDS-sorted chunks merged pairwise) beyond —
//                     so the result does not
//                     depend on atomic arrival order, equal columns are
//                     summed in FP64, and P_ij = float(sum) *
//                     float(0.5 / total_mass).  Rows come out with ascending
//                     columns, exactly as the reference's key-sorted fill
//                     lays them out.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace fe {
// kernels have internal linkage: every translation unit that includes this
// header gets its own copy of the device code
namespace {

constexpr int kAffRowsPerBlock = 64;  // rows staged per 256-thread block
constexpr int kMaxK = 256;            // embedding_cuda_kernels.cpp:5661
constexpr int kSortCap = 8192;        // longest symmetric row sorted in LDS

constexpr int kWaveSortMaxR = 8;      // rows of up to 64 * 8 entries are sorted by one wave

// Ascending bitonic sort of the 64 R keys of one wave: entry e = 64 r + lane
// is register r of its lane.  Strides of 64 and more exchange registers
// within a lane; shorter strides exchange across lanes with shuffles.
template <int R>
__device__ __forceinline__ void wave_sort_keys(unsigned long long (&key)[R], int lane) {
#pragma unroll
  for (int size = 2; size <= 64 * R; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= 64) {
        const int rs = stride / 64;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (r & rs) continue;
          const bool up = ((64 * r) & size) == 0;
          const unsigned long long a = key[r], b = key[r | rs];
          const bool swap = (a > b) == up;
          key[r] = swap ? b : a;
          key[r | rs] = swap ? a : b;
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const unsigned long long other = __shfl_xor(key[r], stride, 64);
          const bool up = ((64 * r + lane) & size) == 0;
          const bool keep_min = ((lane & stride) == 0) == up;
          key[r] = keep_min ? (other < key[r] ? other : key[r]) : (other > key[r] ? other : key[r]);
        }
      }
    }
  }
}

// Orders a wave's LDS stores before its lanes read each other's entries
// (the wave touches no other wave's LDS, so no block barrier is needed).
__device__ __forceinline__ void wave_lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the length class of a row sorted by one wave with R keys per lane:
// (32 R, 64 R] entries, and every row of at most 64 for R = 1
template <int R>
__device__ __forceinline__ bool wave_sort_class(int len) {
  return len <= 64 * R && (R == 1 || len > 32 * R);
}

// kLongRow: some row is longer than kSortCap and is left to the global sort
enum AffinityError : unsigned { kErrIndexRange = 1u, kErrDistance = 2u, kLongRow = 4u };

// ---------------------------------------------------------------------------
// Rows longer than kSortCap (a point in very many KNN lists), one block of
// kLongThreads per row: chunks of kSortCap keys are sorted in LDS into
// `scratch`, then sorted runs are merged pairwise between `scratch` and the
// row itself (merge path: every thread writes an equal share of the output,
// its start found by a binary search on its diagonal; ties take the left
// run first) until one run remains, which ends in `scratch`.  The caller's
// group pass then reads `scratch` and writes its compacted result over the
// row.  The sorted order is a function of the keys only.
// ---------------------------------------------------------------------------
constexpr int kLongThreads = 256;

// ascending bitonic sort of keys[0, pw) (pw a power of two) by the block
__device__ __forceinline__ void block_bitonic(unsigned long long* keys, int pw) {
  for (int size = 2; size <= pw; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < pw / 2; t += kLongThreads) {
        const int lo = 2 * stride * (t / stride) + (t % stride);
        const int hi = lo + stride;
        const bool up = ((lo & size) == 0);
        const unsigned long long a = keys[lo], b = keys[hi];
        if ((a > b) == up) {
          keys[lo] = b;
          keys[hi] = a;
        }
      }
      __syncthreads();
    }
  }
}

// entries of a among the first d outputs of the merge of a and b
__device__ __forceinline__ int merge_split(const unsigned long long* a, int la,
                                           const unsigned long long* b, int lb, int d) {
  int lo = max(0, d - lb), hi = min(d, la);
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (a[m] <= b[d - 1 - m]) lo = m + 1;
    else hi = m;
  }
  return lo;
}

__device__ __forceinline__ void merge_runs(const unsigned long long* a, int la,
                                           const unsigned long long* b, int lb,
                                           unsigned long long* out) {
  const long long total = static_cast<long long>(la) + lb;
  const int d0 = static_cast<int>(total * threadIdx.x / kLongThreads);
  const int d1 = static_cast<int>(total * (threadIdx.x + 1) / kLongThreads);
  int i = merge_split(a, la, b, lb, d0), j = d0 - i;
  for (int o = d0; o < d1; ++o) {
    const bool take_a = j >= lb || (i < la && a[i] <= b[j]);
    out[o] = take_a ? a[i++] : b[j++];
  }
}

// Sorts row[0, len) into scratch[0, len); keys: kSortCap entries of LDS.
__device__ inline void sort_row_global(unsigned long long* row, unsigned long long* scratch,
                                       int len, unsigned long long* keys) {
  for (int c0 = 0; c0 < len; c0 += kSortCap) {
    const int cl = min(kSortCap, len - c0);
    int pw = 1;
    while (pw < cl) pw <<= 1;
    __syncthreads();  // the previous chunk's keys are stored
    for (int t = threadIdx.x; t < pw; t += kLongThreads) keys[t] = t < cl ? row[c0 + t] : ~0ull;
    __syncthreads();
    block_bitonic(keys, pw);
    for (int t = threadIdx.x; t < cl; t += kLongThreads) scratch[c0 + t] = keys[t];
  }
  __syncthreads();
  unsigned long long* src = scratch;
  unsigned long long* dst = row;
  for (int w = kSortCap; w < len; w *= 2) {
    for (int lo = 0; lo < len; lo += 2 * w) {
      const int mid = min(lo + w, len), hi = min(lo + 2 * w, len);
      merge_runs(src + lo, mid - lo, src + mid, hi - mid, dst + lo);
    }
    __syncthreads();
    unsigned long long* t = src;
    src = dst;
    dst = t;
  }
  if (src != scratch) {
    for (int t = threadIdx.x; t < len; t += kLongThreads) scratch[t] = src[t];
    __syncthreads();
  }
}

template <typename DistT>
constexpr int aff_rows_per_block() {
  return sizeof(DistT) == 8 ? kAffRowsPerBlock / 2 : kAffRowsPerBlock;
}

template <typename DistT>
__global__ __launch_bounds__(256) void perplexity_rows(const int* __restrict__ indices,
                                                       const DistT* __restrict__ distances,
                                                       int n, int n_targets, int k,
                                                       int index_offset, double perplexity,
                                                       int* __restrict__ nbr,
                                                       float* __restrict__ pdir,
                                                       unsigned* error_flags) {
  // n rows of k neighbours whose ids must lie in [0, n_targets) after the
  // offset (n_targets = n for t-SNE, the reference size for the transform)
  constexpr int RPB = aff_rows_per_block<DistT>();
  extern __shared__ unsigned char smem_raw[];
  DistT* s_d = reinterpret_cast<DistT*>(smem_raw);
  int* s_idx = reinterpret_cast<int*>(s_d + RPB * k);
  const int row0 = blockIdx.x * RPB;
  const int rows = min(RPB, n - row0);
  // coalesced column-major loads: for each neighbour column j, rows are
  // contiguous in memory
  for (int t = threadIdx.x; t < rows * k; t += blockDim.x) {
    const int j = t / rows, r = t - j * rows;
    const size_t src = static_cast<size_t>(j) * n + row0 + r;
    const int nb = indices[src] - index_offset;
    const DistT d = distances[src];
    if (nb < 0 || nb >= n_targets) atomicOr(error_flags, kErrIndexRange);
    if (!isfinite(static_cast<double>(d)) || d < DistT(0)) atomicOr(error_flags, kErrDistance);
    s_idx[r * k + j] = nb;
    s_d[r * k + j] = d;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int PER = kMaxK / 64;
  const double target = log(perplexity);
  for (int r = wid; r < rows; r += blockDim.x / 64) {
    const int row = row0 + r;
    double d2[PER];
    double mn = DBL_MAX, mx = 0.0;
#pragma unroll
    for (int s = 0; s < PER; ++s) {
      const int j = lane + 64 * s;
      d2[s] = 0.0;
      if (j < k) {
        const double d = static_cast<double>(s_d[r * k + j]);
        d2[s] = d * d;
        mn = fmin(mn, d2[s]);
        mx = fmax(mx, d2[s]);
      }
    }
    mn = -wave_max(-mn);
    mx = wave_max(mx);
    double p[PER];
    const double spread = mx - mn;
    if (spread <= DBL_EPSILON * fmax(1.0, mx)) {
#pragma unroll
      for (int s = 0; s < PER; ++s) p[s] = 1.0 / static_cast<double>(k);
    } else {
      bool found = false;
      double beta = 1.0, min_beta = -DBL_MAX, max_beta = DBL_MAX, sum_p = DBL_MIN;
      for (int iter = 0; !found && iter < 200; ++iter) {
        double sp = 0.0, sh = 0.0;
#pragma unroll
        for (int s = 0; s < PER; ++s) {
          const int j = lane + 64 * s;
          p[s] = 0.0;
          if (j < k) {
            const double sh2 = d2[s] - mn;
            p[s] = exp(-beta * sh2);
            sp += p[s];
            sh += sh2 * p[s];
          }
        }
        sum_p = DBL_MIN + wave_sum(sp);
        sh = wave_sum(sh);
        const double entropy = beta * sh / sum_p + log(sum_p);
        const double diff = entropy - target;
        if (fabs(diff) < 1e-5) {
          found = true;
        } else if (diff > 0.0) {
          min_beta = beta;
          beta = (max_beta == DBL_MAX || max_beta == -DBL_MAX) ? beta * 2.0 : (beta + max_beta) / 2.0;
        } else {
          max_beta = beta;
          beta = (min_beta == -DBL_MAX || min_beta == DBL_MAX) ? beta / 2.0 : (beta + min_beta) / 2.0;
        }
        if (!isfinite(beta)) break;
      }
      if (!isfinite(sum_p) || sum_p <= DBL_MIN) {
        double tied = 0.0;
#pragma unroll
        for (int s = 0; s < PER; ++s) {
          const int j = lane + 64 * s;
          if (j < k && fabs(d2[s] - mn) <= DBL_EPSILON * fmax(1.0, mn)) tied += 1.0;
        }
        tied = wave_sum(tied);
        const double mass = 1.0 / fmax(tied, 1.0);
#pragma unroll
        for (int s = 0; s < PER; ++s) p[s] = fabs(d2[s] - mn) <= DBL_EPSILON * fmax(1.0, mn) ? mass : 0.0;
      } else {
        const double inv = 1.0 / sum_p;
#pragma unroll
        for (int s = 0; s < PER; ++s) p[s] *= inv;
      }
    }
#pragma unroll
    for (int s = 0; s < PER; ++s) {
      const int j = lane + 64 * s;
      if (j < k) {
        const size_t o = static_cast<size_t>(row) * k + j;
        nbr[o] = s_idx[r * k + j];
        pdir[o] = static_cast<float>(p[s]);
      }
    }
  }
}

// One wave per row: its forward (non-self) entries, added to its degree with
// one atomic and kept in `cursor` as the start of its reverse entries.
__global__ __launch_bounds__(256) void sym_row_counts(const int* __restrict__ nbr, int n, int k,
                                                      int* __restrict__ deg,
                                                      int* __restrict__ cursor) {
  const int row = blockIdx.x * 4 + static_cast<int>(threadIdx.x >> 6);
  if (row >= n) return;
  const int lane = threadIdx.x & 63;
  int cnt = 0;
  for (int c0 = 0; c0 < k; c0 += 64) {
    const int c = c0 + lane;
    cnt += __popcll(__ballot(c < k && nbr[static_cast<long long>(row) * k + c] != row));
  }
  if (lane == 0) {
    cursor[row] = cnt;
    if (cnt) atomicAdd(deg + row, cnt);
  }
}

// Count the reverse endpoint of every non-self directed entry (the forward
// ones are counted per row); block partial sums of the directed mass (FP64)
// for the global normaliser.
__global__ __launch_bounds__(256) void sym_count(const int* __restrict__ nbr,
                                                 const float* __restrict__ pdir, long long total,
                                                 int k, int* __restrict__ deg,
                                                 double* __restrict__ mass_partials) {
  double local = 0.0;
  for (long long e = blockIdx.x * 256ll + threadIdx.x; e < total; e += gridDim.x * 256ll) {
    const int i = static_cast<int>(e / k);
    const int j = nbr[e];
    if (j == i) continue;
    atomicAdd(deg + j, 1);
    local += static_cast<double>(pdir[e]);
  }
  __shared__ double red[4];
  local = wave_sum(local);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) mass_partials[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// One wave per row: its forward entries go to the front of its raw range in
// KNN order (no atomics), each one's reverse entry to the neighbour's range
// after that row's forward entries (one atomic on its cursor).
__global__ __launch_bounds__(256) void sym_fill(const int* __restrict__ nbr,
                                                const float* __restrict__ pdir, int n, int k,
                                                const int* __restrict__ raw_ptr,
                                                int* __restrict__ cursor,
                                                unsigned long long* __restrict__ raw) {
  const int row = blockIdx.x * 4 + static_cast<int>(threadIdx.x >> 6);
  if (row >= n) return;
  const int lane = threadIdx.x & 63;
  const int base = raw_ptr[row];
  int fwd = 0;
  for (int c0 = 0; c0 < k; c0 += 64) {
    const int c = c0 + lane;
    int j = row;
    float p = 0.0f;
    if (c < k) {
      const long long e = static_cast<long long>(row) * k + c;
      j = nbr[e];
      p = pdir[e];
    }
    const bool fw = j != row;
    const unsigned long long m = __ballot(fw);
    if (fw) {
      const unsigned long long bits = __float_as_uint(p);
      raw[base + fwd + __popcll(m & ((1ull << lane) - 1ull))] =
          (static_cast<unsigned long long>(static_cast<unsigned>(j)) << 32) | bits;
      const int pos = raw_ptr[j] + atomicAdd(cursor + j, 1);
      raw[pos] = (static_cast<unsigned long long>(static_cast<unsigned>(row)) << 32) | bits;
    }
    fwd += __popcll(m);
  }
}

// The sorted row's merge: equal columns summed in FP64 in sorted order,
// written back to the front of the row's raw range as
// (col << 32 | float(sum) bits).
__device__ __forceinline__ unsigned long long merged_entry(const unsigned long long* keys, int t,
                                                           int len) {
  const unsigned long long cidx = keys[t] >> 32;
  double s = 0.0;
  for (int u = t; u < len && (keys[u] >> 32) == cidx; ++u)
    s += static_cast<double>(__uint_as_float(static_cast<unsigned>(keys[u] & 0xffffffffu)));
  return (cidx << 32) | __float_as_uint(static_cast<float>(s));
}

// Symmetric rows of length class R (wave_sort_class), one wave each: the
// keys (col << 32 | value bits) sorted across the wave, then merged.  The
// widest class queues the rows longer than 64 kWaveSortMaxR entries for
// sym_sort_merge.
template <int R>
__global__ __launch_bounds__(256) void sym_sort_merge_wave(unsigned long long* __restrict__ raw,
                                                           const int* __restrict__ raw_ptr, int n,
                                                           int* __restrict__ merged_len,
                                                           int* __restrict__ long_rows,
                                                           int* __restrict__ long_count) {
  constexpr int kCap = 64 * R;
  __shared__ unsigned long long sk[4 * kCap];
  const int lane = threadIdx.x & 63, wave = static_cast<int>(threadIdx.x >> 6);
  const int row = blockIdx.x * 4 + wave;
  if (row >= n) return;
  const int beg = raw_ptr[row], len = raw_ptr[row + 1] - beg;
  if (R == kWaveSortMaxR && len > kCap && lane == 0) long_rows[atomicAdd(long_count, 1)] = row;
  if (!wave_sort_class<R>(len)) return;  // wave-uniform
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
    const bool head = e < len && (e == 0 || (key[r] >> 32) != (wk[e - 1] >> 32));
    const unsigned long long out = head ? merged_entry(wk, e, len) : 0ull;
    const unsigned long long m = __ballot(head);
    if (head) raw[beg + base + __popcll(m & ((1ull << lane) - 1ull))] = out;
    base += __popcll(m);
  }
  if (lane == 0) merged_len[row] = base;
}

// One queued long row by one block: LDS bitonic sort of up to kSortCap
// entries, then the merge.
__device__ __forceinline__ void sort_merge_long(unsigned long long* __restrict__ raw,
                                                const int* __restrict__ raw_ptr, int row,
                                                int* __restrict__ merged_len,
                                                unsigned* error_flags) {
  __shared__ unsigned long long keys[kSortCap];
  const int beg = raw_ptr[row], len = raw_ptr[row + 1] - beg;
  if (len > kSortCap) {
    // sym_sort_merge_global takes it
    if (threadIdx.x == 0) atomicOr(error_flags, kLongRow);
    return;
  }
  int pw = 1;
  while (pw < len) pw <<= 1;
  for (int t = threadIdx.x; t < pw; t += 256) keys[t] = t < len ? raw[beg + t] : ~0ull;
  __syncthreads();
  for (int size = 2; size <= pw; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < pw / 2; t += 256) {
        const int lo = 2 * stride * (t / stride) + (t % stride);
        const int hi = lo + stride;
        const bool up = ((lo & size) == 0);
        const unsigned long long a = keys[lo], b = keys[hi];
        if ((a > b) == up) {
          keys[lo] = b;
          keys[hi] = a;
        }
      }
      __syncthreads();
    }
  }
  // merge: a thread owns a chunk of consecutive sorted entries
  constexpr int CH = kSortCap / 256;
  const int c0 = threadIdx.x * CH;
  int heads = 0;
  for (int t = c0; t < min(c0 + CH, len); ++t)
    if (t == 0 || (keys[t] >> 32) != (keys[t - 1] >> 32)) ++heads;
  int total;
  int out = block_exclusive_scan_nt<256>(heads, &total);
  if (threadIdx.x == 0) merged_len[row] = total;
  for (int t = c0; t < min(c0 + CH, len); ++t) {
    if (t == 0 || (keys[t] >> 32) != (keys[t - 1] >> 32)) {
      raw[beg + out] = merged_entry(keys, t, len);
      ++out;
    }
  }
}

// The rows queued by sym_sort_merge_wave, strided over the grid.
__global__ __launch_bounds__(256) void sym_sort_merge(unsigned long long* __restrict__ raw,
                                                      const int* __restrict__ raw_ptr,
                                                      const int* __restrict__ long_rows,
                                                      const int* __restrict__ long_count,
                                                      int* __restrict__ merged_len,
                                                      unsigned* error_flags) {
  const int rows = *long_count;
  for (int q = blockIdx.x; q < rows; q += gridDim.x) {
    __syncthreads();  // the previous row's keys are read
    sort_merge_long(raw, raw_ptr, long_rows[q], merged_len, error_flags);
  }
}

// The queued rows longer than kSortCap: global sort, then the merge of
// equal columns written over the row.
__global__ __launch_bounds__(kLongThreads) void sym_sort_merge_global(
    unsigned long long* __restrict__ raw, unsigned long long* __restrict__ scratch,
    const int* __restrict__ raw_ptr, const int* __restrict__ long_rows,
    const int* __restrict__ long_count, int* __restrict__ merged_len) {
  __shared__ unsigned long long keys[kSortCap];
  const int rows = *long_count;
  for (int q = blockIdx.x; q < rows; q += gridDim.x) {
    const int row = long_rows[q];
    const int beg = raw_ptr[row], len = raw_ptr[row + 1] - beg;
    if (len <= kSortCap) continue;  // block-uniform
    __syncthreads();
    sort_row_global(raw + beg, scratch + beg, len, keys);
    const unsigned long long* sorted = scratch + beg;
    const int ch = (len + kLongThreads - 1) / kLongThreads;
    const int c0 = threadIdx.x * ch, c1 = min(c0 + ch, len);
    int heads = 0;
    for (int t = c0; t < c1; ++t)
      if (t == 0 || (sorted[t] >> 32) != (sorted[t - 1] >> 32)) ++heads;
    int total;
    int out = block_exclusive_scan_nt<kLongThreads>(heads, &total);
    if (threadIdx.x == 0) merged_len[row] = total;
    for (int t = c0; t < c1; ++t) {
      if (t == 0 || (sorted[t] >> 32) != (sorted[t - 1] >> 32)) {
        raw[beg + out] = merged_entry(sorted, t, len);
        ++out;
      }
    }
  }
}

__global__ __launch_bounds__(256) void sym_compact(const unsigned long long* __restrict__ raw,
                                                   const int* __restrict__ raw_ptr,
                                                   const int* __restrict__ row_ptr, int n,
                                                   float norm, int* __restrict__ col,
                                                   float* __restrict__ val) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  const int lane = threadIdx.x & 63;
  const int src = raw_ptr[row], dst = row_ptr[row], len = row_ptr[row + 1] - dst;
  for (int t = lane; t < len; t += 64) {
    const unsigned long long e = raw[src + t];
    col[dst + t] = static_cast<int>(e >> 32);
    val[dst + t] = __uint_as_float(static_cast<unsigned>(e & 0xffffffffu)) * norm;
  }
}

// Exclusive scan of n ints into out[0..n] (out[n] = total), three launches:
// per-block scans, a single-block scan of the block totals, offsets.
constexpr int kScanBlock = 1024;

__global__ __launch_bounds__(256) void scan_blocks(const int* __restrict__ in, int n,
                                                   int* __restrict__ out, int* __restrict__ sums) {
  __shared__ int s[kScanBlock];
  const int base = blockIdx.x * kScanBlock;
  for (int t = threadIdx.x; t < kScanBlock; t += 256) s[t] = base + t < n ? in[base + t] : 0;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    int v[4];
    for (int q = 0; q < 4; ++q) {
      const int t = threadIdx.x + q * 256;
      v[q] = t >= off ? s[t - off] : 0;
    }
    __syncthreads();
    for (int q = 0; q < 4; ++q) s[threadIdx.x + q * 256] += v[q];
    __syncthreads();
  }
  for (int t = threadIdx.x; t < kScanBlock; t += 256) {
    if (base + t < n) out[base + t] = t == 0 ? 0 : s[t - 1];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = s[kScanBlock - 1];
}

__global__ __launch_bounds__(256) void scan_sums(int* sums, int nb, int* total_out) {
  if (threadIdx.x == 0) {
    long long run = 0;
    for (int b = 0; b < nb; ++b) {
      const int v = sums[b];
      sums[b] = static_cast<int>(run);
      run += v;
    }
    *total_out = static_cast<int>(run);
  }
}

__global__ __launch_bounds__(256) void scan_add(int* __restrict__ out, int n,
                                                const int* __restrict__ sums) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] += sums[i / kScanBlock];
}

}  // namespace
}  // namespace fe
