// Exact Euclidean k-nearest neighbours on gfx950: the producer behind the
// reference's exact native route (native_cuda_knn_impl,
// src/native_knn_cuda_impl.cpp:520-700, which hands FAISS bfKnn or cuVS
// brute force the k + 1 nearest rows and finalize_cuvs_knn_kernel,
// src/embedding_cuda_kernels.cpp:1428-1480, drops the row itself).
//
// The result is exact: for every row q, the k rows j != q with the smallest
// squared distance sum_c (x[q][c] - x[j][c])^2 formed in FP64 from the
// float32 inputs in column order, ties by the smaller index, distances
// (float) sqrt(d2).
//
//   knn_col_partials / knn_col_fold   column means (FP64, fixed order)
//   knn_pack          row-major copies: centred float32 rows padded to a
//                     multiple of 8 columns (the MFMA operands) with their
//                     float32 norms, the original rows, and FP64 norms
//   knn_select        MFMA 16x16x4 f32 tiles of q . c over all candidates;
//                     key fl(|c|^2 - 2 q . c) with the index in the low bits;
//                     each query keeps its KL smallest keys in LDS (pending
//                     keys below the list's last one are merged in by a
//                     wave-wide bitonic sort), so the list is the exact
//                     KL-smallest set of the keys, whatever the schedule
//   knn_refine        one wave per query: FP64 squared distances of the KL
//                     candidates, sorted by (d2, index), the row itself
//                     dropped, the first k written; a bound on the key's
//                     rounding error proves that no candidate outside the
//                     list can come before the k-th, else the row goes to
//   knn_scan          one block per unproved row: FP64 distances to every
//                     row, the k best kept through LDS merges
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/fastembedr_hip.h"
#include "fe_common.hpp"

namespace fe {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kKnnMaxK = 256;       // the reference's kMaxNativeCudaK
constexpr int kKnnMargin = 16;      // list entries beyond k + 1
constexpr int kKnnChunk = 64;       // columns per LDS chunk (16 MFMA k-steps)
constexpr int kKnnTile = 64;        // candidates per tile (4 MFMA row blocks)
constexpr int kKnnQ = 16;           // queries per wave (the MFMA's N)
constexpr int kKnnBuf = 64;         // pending keys per query
constexpr int kKnnThreads = 256;
constexpr int kKnnPartBlocks = 256; // row blocks of the column sums
constexpr unsigned long long kNoKey = ~0ull;

__host__ __device__ constexpr int pow2_at_least(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// float -> unsigned with the same order
__device__ __forceinline__ unsigned int order_bits(float f) {
  const unsigned int u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float order_float(unsigned int o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m) {
  const int lo = __shfl_xor(static_cast<int>(static_cast<unsigned int>(v)), m, 64);
  const int hi = __shfl_xor(static_cast<int>(static_cast<unsigned int>(v >> 32)), m, 64);
  return (static_cast<unsigned long long>(static_cast<unsigned int>(hi)) << 32) |
         static_cast<unsigned int>(lo);
}

// Ascending bitonic sort of 64 * R values held R per lane (element
// lane * R + r), across the wave.
template <int R>
__device__ __forceinline__ void wave_sort(unsigned long long (&e)[R]) {
  constexpr int N = 64 * R;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
    for (int j = size >> 1; j > 0; j >>= 1) {
      if (j >= R) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const unsigned long long o = shfl_xor_u64(e[r], j / R);
          const int s = lane * R + r;
          const bool up = (s & size) == 0, low = (s & j) == 0;
          e[r] = (low == up) ? (o < e[r] ? o : e[r]) : (o > e[r] ? o : e[r]);
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if ((r & j) == 0) {
            const int s = lane * R + r;
            const bool up = (s & size) == 0;
            const unsigned long long a = e[r], b = e[r | j];
            if ((a > b) == up) {
              e[r] = b;
              e[r | j] = a;
            }
          }
        }
      }
    }
  }
}

// ---- preparation ----

// partial FP64 sums of column c over the row block blockIdx.x
__global__ __launch_bounds__(kKnnThreads) void knn_col_partials(const float* __restrict__ x, int n,
                                                                double* __restrict__ part) {
  __shared__ double red[kKnnThreads / 64];
  const int c = blockIdx.y, b = blockIdx.x;
  const long long rows = (static_cast<long long>(n) + kKnnPartBlocks - 1) / kKnnPartBlocks;
  const long long r0 = b * rows, r1 = std::min<long long>(n, r0 + rows);
  double s = 0.0;
  for (long long i = r0 + threadIdx.x; i < r1; i += kKnnThreads)
    s = __dadd_rn(s, static_cast<double>(x[static_cast<size_t>(c) * n + i]));
  for (int o = 32; o > 0; o >>= 1) s = __dadd_rn(s, __shfl_xor(s, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kKnnThreads / 64; ++w) t = __dadd_rn(t, red[w]);
    part[static_cast<size_t>(c) * kKnnPartBlocks + b] = t;
  }
}

__global__ void knn_col_fold(const double* __restrict__ part, int p, int n,
                             double* __restrict__ mean) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= p) return;
  double t = 0.0;
  for (int b = 0; b < kKnnPartBlocks; ++b) t = __dadd_rn(t, part[static_cast<size_t>(c) * kKnnPartBlocks + b]);
  mean[c] = t / n;
}

// One thread per row (n_pad rows; rows >= n are zero with norm +inf).
__global__ __launch_bounds__(kKnnThreads) void knn_pack(const float* __restrict__ x, int n, int n_pad,
                                                        int p, int P, const double* __restrict__ mean,
                                                        float* __restrict__ xc, float* __restrict__ xr,
                                                        float* __restrict__ nrm,
                                                        double* __restrict__ qn,
                                                        double* __restrict__ norm_part) {
  __shared__ double red[kKnnThreads / 64];
  const int i = blockIdx.x * kKnnThreads + threadIdx.x;
  double nd = 0.0;
  if (i < n_pad) {
    float* row = xc + static_cast<size_t>(i) * P;
    if (i < n) {
      float nf = 0.0f;
      for (int c = 0; c < p; ++c) {
        const float v = x[static_cast<size_t>(c) * n + i];
        xr[static_cast<size_t>(i) * p + c] = v;
        const float w = static_cast<float>(__dsub_rn(static_cast<double>(v), mean[c]));
        row[c] = w;
        nf = __fadd_rn(nf, __fmul_rn(w, w));
        nd = __dadd_rn(nd, __dmul_rn(static_cast<double>(w), static_cast<double>(w)));
      }
      for (int c = p; c < P; ++c) row[c] = 0.0f;
      nrm[i] = nf;
      qn[i] = nd;
    } else {
      for (int c = 0; c < P; ++c) row[c] = 0.0f;
      nrm[i] = __int_as_float(0x7f800000);
    }
  }
  // the largest FP64 norm of the block
  for (int o = 32; o > 0; o >>= 1) nd = fmax(nd, __shfl_xor(nd, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nd;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kKnnThreads / 64; ++w) t = fmax(t, red[w]);
    norm_part[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kKnnThreads) void knn_norm_fold(const double* __restrict__ part, int m,
                                                             double* __restrict__ cmax2) {
  __shared__ double red[kKnnThreads / 64];
  double t = 0.0;
  for (int b = threadIdx.x; b < m; b += kKnnThreads) t = fmax(t, part[b]);
  for (int o = 32; o > 0; o >>= 1) t = fmax(t, __shfl_xor(t, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kKnnThreads / 64; ++w) s = fmax(s, red[w]);
    *cmax2 = s;
  }
}

// ---- candidate lists ----

constexpr int kSelWaves = 8;                       // two waves per SIMD
constexpr int kSelThreads = 64 * kSelWaves;
constexpr int kSelQueries = kKnnQ * kSelWaves;     // queries per workgroup
// float4 loads per thread for one candidate chunk (64 rows x 64 columns)
constexpr int kSelFetch = kKnnTile * kKnnChunk / 4 / kSelThreads;
// LDS: two candidate chunks, the query chunk, two norm tiles, the pending
// keys and their counts (the lists themselves live in the output array)
constexpr size_t kSelCand = kKnnChunk * kKnnTile;  // floats per candidate chunk
constexpr size_t kSelLds = sizeof(float) * (2 * kSelCand + kKnnChunk * kSelQueries + 2 * kKnnTile) +
                           sizeof(unsigned long long) * kKnnBuf * kSelQueries +
                           sizeof(int) * kSelQueries;
static_assert(kSelFetch * kSelThreads == kKnnTile * kKnnChunk / 4, "whole chunk per fetch");

__device__ __forceinline__ unsigned long long shfl_u64(unsigned long long v, int src) {
  const int lo = __shfl(static_cast<int>(static_cast<unsigned int>(v)), src, 64);
  const int hi = __shfl(static_cast<int>(static_cast<unsigned int>(v >> 32)), src, 64);
  return (static_cast<unsigned long long>(static_cast<unsigned int>(hi)) << 32) |
         static_cast<unsigned int>(lo);
}

// Merge the pending keys of one query into its sorted list (whole wave):
// the (at most 64) pending keys are sorted one per lane; the list (KL
// ascending keys, KL / 64 per lane, element lane * KL / 64 + r) against the
// pending keys in reverse order, element by element, keeps the smaller of
// each pair -- the KL smallest keys of both, as a bitonic sequence -- which
// a bitonic merge puts in order.  Every lane reads and writes only its own
// list elements.  Returns the list's new last key (every lane).
template <int KL>
__device__ __forceinline__ unsigned long long merge_pending(unsigned long long* __restrict__ list,
                                                            const unsigned long long* pend,
                                                            int m) {
  constexpr int RL = KL / 64;
  const int lane = threadIdx.x & 63;
  unsigned long long pk[1] = {lane < m ? pend[lane] : kNoKey};
  wave_sort<1>(pk);
  unsigned long long e[RL];
#pragma unroll
  for (int r = 0; r < RL; ++r) {
    const int i = lane * RL + r;
    const int j = KL - 1 - i;  // the pending key paired with list[i]
    const unsigned long long o = shfl_u64(pk[0], j < 64 ? j : 0);
    const unsigned long long l = list[i];
    e[r] = (j < 64 && o < l) ? o : l;
  }
  constexpr int kStages = __builtin_ctz(KL);
#pragma unroll
  for (int t = 0; t < kStages; ++t) {
    const int h = (KL / 2) >> t;
    if (h >= RL) {
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const unsigned long long o = shfl_xor_u64(e[r], h / RL);
        const bool low = ((lane * RL + r) & h) == 0;
        e[r] = low ? (o < e[r] ? o : e[r]) : (o > e[r] ? o : e[r]);
      }
    } else {
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        if ((r & h) == 0 && e[r] > e[r | h]) {
          const unsigned long long t2 = e[r];
          e[r] = e[r | h];
          e[r | h] = t2;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RL; ++r) list[lane * RL + r] = e[r];
  return shfl_u64(e[RL - 1], 63);
}

// The keys of one MFMA block (candidates t0 + 16 b + 4 (lane / 16) + r)
// below the query's current last key go to its pending buffer.  Candidates
// come in increasing index order, so a key is below the last key exactly when
// its value is below the last value (tau_v; -inf for a padding query slot);
// padding candidates have norm +inf and never pass.
__device__ __forceinline__ void offer_keys(const v4f& acc, int b, int t0, int lane,
                                           const float* snrm, float tau_v, int* count,
                                           unsigned long long* pending) {
  const int cl = 16 * b + 4 * (lane >> 4);
  const v4f nr = *reinterpret_cast<const v4f*>(snrm + cl);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const float v = __fsub_rn(nr[r], __fmul_rn(2.0f, acc[r]));
    if (v < tau_v) {
      const unsigned long long key = (static_cast<unsigned long long>(order_bits(v)) << 32) |
                                     static_cast<unsigned int>(t0 + cl + r);
      pending[atomicAdd(count, 1)] = key;
    }
  }
}

// the value part of a list's last key (+inf while the list is not full;
// -inf for a padding query slot, which takes no keys)
__device__ __forceinline__ float last_value(unsigned long long key, bool qvalid) {
  if (!qvalid) return -__int_as_float(0x7f800000);
  return key == kNoKey ? __int_as_float(0x7f800000)
                       : order_float(static_cast<unsigned int>(key >> 32));
}

// Merge every query of wave w with more than `limit` pending keys; the lane's
// tau_v follows its query's list.
template <int KL>
__device__ __forceinline__ void merge_full(unsigned long long* __restrict__ lists,
                                           const unsigned long long* buf, int* cnt, int w,
                                           int lane, int gbase, int limit, bool qvalid,
                                           float& tau_v) {
  const bool need = lane < kKnnQ && cnt[w * kKnnQ + lane] > limit;
  unsigned long long m = __ballot(need);
  while (m) {
    const int j = __builtin_ctzll(m);
    m &= m - 1;
    const int qs = w * kKnnQ + j;
    const unsigned long long last = merge_pending<KL>(
        lists + static_cast<size_t>(gbase + j) * KL, buf + qs * kKnnBuf, cnt[qs]);
    if (lane == 0) cnt[qs] = 0;
    if ((lane & 15) == j) tau_v = last_value(last, qvalid);
  }
}

// One step of the walk = (candidate tile, column chunk).  While the MFMAs
// of a step run, the next step's candidate chunk is already on its way into
// registers; it goes to the other LDS buffer afterwards, so one barrier per
// step suffices.  Slot e of a chunk: candidate e / 16, float4 e % 16.
__device__ __forceinline__ void fetch_chunk(const float* __restrict__ xc,
                                            const float* __restrict__ nrm, int P, int t0, int c0,
                                            int tid, v4f (&f)[kSelFetch], float& fn) {
  const int kq = min(kKnnChunk, P - c0) / 4;
#pragma unroll
  for (int u = 0; u < kSelFetch; ++u) {
    const int e = tid + u * kSelThreads;
    const int c = e >> 4, q = e & 15;
    if (q < kq)
      f[u] = *reinterpret_cast<const v4f*>(xc + static_cast<size_t>(t0 + c) * P + c0 + 4 * q);
  }
  if (c0 == 0 && tid < kKnnTile) fn = nrm[t0 + tid];
}

__device__ __forceinline__ void store_chunk(int P, int c0, int tid, const v4f (&f)[kSelFetch],
                                            float fn, float* sc, float* snrm) {
  const int kq = min(kKnnChunk, P - c0) / 4;
#pragma unroll
  for (int u = 0; u < kSelFetch; ++u) {
    const int e = tid + u * kSelThreads;
    const int c = e >> 4, q = e & 15;
    if (q < kq) *reinterpret_cast<v4f*>(sc + (q * kKnnTile + c) * 4) = f[u];
  }
  if (c0 == 0 && tid < kKnnTile) snrm[tid] = fn;
}

// The workgroup's query rows (centred, row-major xc), columns c0 .. c0 + 63,
// into LDS as [k-step][query][4] (the MFMA's B operand layout).
__device__ __forceinline__ void stage_queries(const float* __restrict__ xc, int P, int c0,
                                              const int* __restrict__ qlist, int nq, int tid,
                                              float* sq) {
  const int kq = min(kKnnChunk, P - c0) / 4;
  for (int e = tid; e < kSelQueries * kq; e += kSelThreads) {
    const int q = e / kq, s = e - q * kq;
    const int g = blockIdx.x * kSelQueries + q;
    const int row = g < nq ? (qlist ? qlist[g] : g) : 0;
    *reinterpret_cast<v4f*>(sq + (s * kSelQueries + q) * 4) =
        *reinterpret_cast<const v4f*>(xc + static_cast<size_t>(row) * P + c0 + 4 * s);
  }
}

// Stage one: each wave owns kKnnQ queries; the workgroup walks every
// candidate tile.  lists: nq lists of KL keys (ascending), kept in place.
// ONE: the rows fit one 64-column chunk.
template <int KL, bool ONE>
__global__ __launch_bounds__(kSelThreads) void knn_select(
    const float* __restrict__ xc, const float* __restrict__ nrm, int n_pad, int P,
    const float* __restrict__ xqc, const int* __restrict__ qlist, int nq,
    unsigned long long* __restrict__ lists) {
  constexpr int RL = KL / 64;
  extern __shared__ unsigned long long knn_smem[];
  // the MFMA operands first (small LDS offsets), then the pending keys
  float* sc0 = reinterpret_cast<float*>(knn_smem);          // [step][candidate][4], x2
  float* sq = sc0 + 2 * kSelCand;                            // [step][query][4]
  float* snrm0 = sq + kKnnChunk * kSelQueries;               // x2
  unsigned long long* buf = reinterpret_cast<unsigned long long*>(snrm0 + 2 * kKnnTile);
  int* cnt = reinterpret_cast<int*>(buf + kKnnBuf * kSelQueries);

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int qs = w * kKnnQ + (lane & 15);  // this lane's query slot
  const int gbase = blockIdx.x * kSelQueries + w * kKnnQ;
  const bool qvalid = gbase + (lane & 15) < nq;

  // the wave's lists start empty (each lane writes the elements it merges)
  for (int j = 0; j < kKnnQ && gbase + j < nq; ++j) {
#pragma unroll
    for (int r = 0; r < RL; ++r) lists[static_cast<size_t>(gbase + j) * KL + lane * RL + r] = kNoKey;
  }
  if (tid < kSelQueries) cnt[tid] = 0;
  float tau_v = qvalid ? __int_as_float(0x7f800000) : -__int_as_float(0x7f800000);

  const int nchunks = ONE ? 1 : (P + kKnnChunk - 1) / kKnnChunk;
  const int nsteps = (n_pad / kKnnTile) * nchunks;
  v4f pf[kSelFetch];
  float pfn = 0.0f;
  stage_queries(xqc, P, 0, qlist, nq, tid, sq);
  fetch_chunk(xc, nrm, P, 0, 0, tid, pf, pfn);
  store_chunk(P, 0, tid, pf, pfn, sc0, snrm0);
  __syncthreads();

  const float* qb = sq + (w * kKnnQ + (lane & 15)) * 4 + (lane >> 4);
  v4f acc[4];
  for (int step = 0; step < nsteps; ++step) {
    const int tile = ONE ? step : step / nchunks;
    const int chunk = ONE ? 0 : step - tile * nchunks;
    const int t0 = tile * kKnnTile;
    const float* sc = sc0 + (step & 1) * kSelCand;
    const float* snrm = snrm0 + (tile & 1) * kKnnTile;
    const bool more = step + 1 < nsteps;
    const int nt0 = ONE ? t0 + kKnnTile : ((step + 1) / nchunks) * kKnnTile;
    const int nc0 = ONE ? 0 : ((step + 1) - ((step + 1) / nchunks) * nchunks) * kKnnChunk;
    if (more) fetch_chunk(xc, nrm, P, nt0, nc0, tid, pf, pfn);
    if (chunk == 0) {
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[b] = v4f{0.f, 0.f, 0.f, 0.f};
    }
    // P is a multiple of 8: the chunk has an even number of k-steps, taken
    // two at a time (no branch between the MFMAs of a chunk)
    const int kq = ONE ? P / 4 : min(kKnnChunk, P - chunk * kKnnChunk) / 4;
    const float* cb = sc + (lane & 15) * 4 + (lane >> 4);
    for (int s = 0; s < kq; s += 2) {
      const float b0 = qb[s * kSelQueries * 4], b1 = qb[(s + 1) * kSelQueries * 4];
      float a0[4], a1[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        a0[b] = cb[(s * kKnnTile + 16 * b) * 4];
        a1[b] = cb[((s + 1) * kKnnTile + 16 * b) * 4];
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[b], b0, acc[b], 0, 0, 0);
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[b], b1, acc[b], 0, 0, 0);
    }
    if (chunk == nchunks - 1) {
      // two blocks add at most 32 keys to a query, so a query with more than
      // kKnnBuf - 32 pending keys is merged after each pair of blocks
      unsigned long long* pend = buf + qs * kKnnBuf;
      offer_keys(acc[0], 0, t0, lane, snrm, tau_v, cnt + qs, pend);
      offer_keys(acc[1], 1, t0, lane, snrm, tau_v, cnt + qs, pend);
      merge_full<KL>(lists, buf, cnt, w, lane, gbase, kKnnBuf - 32, qvalid, tau_v);
      offer_keys(acc[2], 2, t0, lane, snrm, tau_v, cnt + qs, pend);
      offer_keys(acc[3], 3, t0, lane, snrm, tau_v, cnt + qs, pend);
      merge_full<KL>(lists, buf, cnt, w, lane, gbase, kKnnBuf - 32, qvalid, tau_v);
    }
    if (more)
      store_chunk(P, nc0, tid, pf, pfn, sc0 + ((step + 1) & 1) * kSelCand,
                  snrm0 + ((nt0 / kKnnTile) & 1) * kKnnTile);
    __syncthreads();
    if (!ONE && more && nchunks > 1) {
      // past 64 columns the query chunk changes with the step
      stage_queries(xqc, P, nc0, qlist, nq, tid, sq);
      __syncthreads();
    }
  }
  // the rest of the pending keys
  merge_full<KL>(lists, buf, cnt, w, lane, gbase, 0, qvalid, tau_v);
}

template <int KL>
int run_select_kl(const float* xc, const float* nrm, int n_pad, int P, const float* xqc, int nq,
                  unsigned long long* lists, hipStream_t s) {
  const bool one = P <= kKnnChunk;
  const void* fn = one ? reinterpret_cast<const void*>(&knn_select<KL, true>)
                       : reinterpret_cast<const void*>(&knn_select<KL, false>);
  static bool attr_set[2] = {false, false};
  if (!attr_set[one]) {
    FE_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(kSelLds)));
    attr_set[one] = true;
  }
  const int blocks = (nq + kSelQueries - 1) / kSelQueries;
  if (one)
    knn_select<KL, true><<<blocks, kSelThreads, kSelLds, s>>>(xc, nrm, n_pad, P, xqc, nullptr, nq,
                                                              lists);
  else
    knn_select<KL, false><<<blocks, kSelThreads, kSelLds, s>>>(xc, nrm, n_pad, P, xqc, nullptr, nq,
                                                               lists);
  FE_LAUNCH_CHECK("knn_select");
  return 0;
}

// ---- exact distances ----

// FP64 squared distance of query row q (row-major xq) and candidate row j
// (row-major xr), summed in column order
__device__ __forceinline__ double exact_d2(const float* __restrict__ xq, const float* __restrict__ xr,
                                          int p, int q, int j) {
  const float* a = xq + static_cast<size_t>(q) * p;
  const float* b = xr + static_cast<size_t>(j) * p;
  double s = 0.0;
  for (int c = 0; c < p; ++c) {
    const double d = __dsub_rn(static_cast<double>(b[c]), static_cast<double>(a[c]));
    s = __dadd_rn(s, __dmul_rn(d, d));
  }
  return s;
}

// (d2, index) as one orderable 96-bit pair, packed in two words
struct DistKey {
  unsigned long long d;  // bits of a non-negative double (+inf: none)
  int j;
};
__device__ __forceinline__ bool key_less(unsigned long long da, int ja, unsigned long long db, int jb) {
  return da < db || (da == db && ja < jb);
}

// Ascending bitonic sort of 64 * R (d, j) pairs across the wave.
template <int R>
__device__ __forceinline__ void wave_sort_pairs(unsigned long long (&d)[R], int (&j)[R]) {
  constexpr int N = 64 * R;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
    for (int h = size >> 1; h > 0; h >>= 1) {
      if (h >= R) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const unsigned long long od = shfl_xor_u64(d[r], h / R);
          const int oj = __shfl_xor(j[r], h / R, 64);
          const int s = lane * R + r;
          const bool up = (s & size) == 0, low = (s & h) == 0;
          const bool other_first = key_less(od, oj, d[r], j[r]);
          const bool take = (low == up) ? other_first : !other_first;
          if (take) {
            d[r] = od;
            j[r] = oj;
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if ((r & h) == 0) {
            const int s = lane * R + r;
            const bool up = (s & size) == 0;
            const bool gt = key_less(d[r | h], j[r | h], d[r], j[r]);
            if (gt == up) {
              const unsigned long long td = d[r];
              const int tj = j[r];
              d[r] = d[r | h];
              j[r] = j[r | h];
              d[r | h] = td;
              j[r | h] = tj;
            }
          }
        }
      }
    }
  }
}

// Stage two: one wave per query.  xq / xr: original query / candidate rows;
// self: the queries are the candidates and each drops itself; outputs are
// column-major nq x k.
template <int KL>
__global__ __launch_bounds__(kKnnThreads) void knn_refine(
    const unsigned long long* __restrict__ lists, const float* __restrict__ xq,
    const float* __restrict__ xr, int p, int P, const double* __restrict__ qn,
    const double* __restrict__ cmax2, const int* __restrict__ qlist, int nq, int self, int k,
    int index_base, int force_scan, int* __restrict__ out_idx, float* __restrict__ out_dist,
    int* __restrict__ failed, int* __restrict__ failed_count) {
  constexpr int R = KL / 64;
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * (kKnnThreads / 64) + (threadIdx.x >> 6);
  if (g >= nq) return;
  const int q = qlist ? qlist[g] : g;
  const unsigned long long* L = lists + static_cast<size_t>(g) * KL;
  unsigned long long d[R];
  int j[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const unsigned long long key = L[lane * R + r];
    if (key == kNoKey) {
      d[r] = 0x7ff0000000000000ull;
      j[r] = INT_MAX;
    } else {
      j[r] = static_cast<int>(static_cast<unsigned int>(key));
      d[r] = static_cast<unsigned long long>(__double_as_longlong(exact_d2(xq, xr, p, q, j[r])));
    }
  }
  wave_sort_pairs<R>(d, j);
  // position of the row itself (at most once)
  int self_at = INT_MAX;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (self && j[r] == q) self_at = lane * R + r;
  for (int o = 32; o > 0; o >>= 1) self_at = min(self_at, __shfl_xor(self_at, o, 64));
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int s = lane * R + r;
    if (s == self_at || j[r] == INT_MAX) continue;
    const int o = s - (self_at < s ? 1 : 0);
    if (o < k) {
      out_idx[static_cast<size_t>(o) * nq + q] = j[r] + index_base;
      out_dist[static_cast<size_t>(o) * nq + q] =
          static_cast<float>(__dsqrt_rn(__longlong_as_double(static_cast<long long>(d[r]))));
    }
  }
  // the k-th answer's squared distance: sorted position k - 1, or k when
  // the row itself comes before it
  const int sk = self_at <= k - 1 ? k : k - 1;
  unsigned long long mine = 0;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (r == sk % R) mine = d[r];
  const int src = sk / R;
  const unsigned long long dkb =
      (static_cast<unsigned long long>(static_cast<unsigned int>(
           __shfl(static_cast<int>(static_cast<unsigned int>(mine >> 32)), src, 64)))
       << 32) |
      static_cast<unsigned int>(__shfl(static_cast<int>(static_cast<unsigned int>(mine)), src, 64));
  if (lane != 0) return;
  const double dk = __longlong_as_double(static_cast<long long>(dkb));
  const unsigned long long last = L[KL - 1];
  bool proved = last == kNoKey;  // every other row is in the list
  if (!proved) {
    // any row c outside the list has key >= the list's last key v:
    //   |q~ - c~|^2 = key_exact + |q~|^2 >= v + |q~|^2 - E,
    //   E = 4 (P + 4) u (Cmax^2 + 2 |q~| Cmax)   (u = 2^-24; the dot product,
    //       the float norm and the subtraction, with room to spare),
    // and the centred rows differ from x - mean by at most u (|q~| + |c~|)
    // each, so |q - c| >= sqrt(v + |q~|^2 - E) - eta.
    const double u = 5.9604644775390625e-08;
    const double v = static_cast<double>(order_float(static_cast<unsigned int>(last >> 32)));
    const double qq = qn[q], cm2 = *cmax2;
    const double qnorm = sqrt(qq), cmax = sqrt(cm2);
    const double E = 4.0 * (P + 4) * u * (cm2 + 2.0 * qnorm * cmax);
    const double eta = 1.01 * u * (qnorm + cmax);
    const double t = v + qq - E;
    if (t > 0.0) {
      const double a = sqrt(t) * (1.0 - 1e-12) - eta;
      proved = a > 0.0 && a * a * (1.0 - 1e-12) > dk;
    }
  }
  if (!proved || force_scan) failed[atomicAdd(failed_count, 1)] = q;
}

// ---- full scan of the rows the bound did not settle ----

constexpr int kScanSlots = 1024;  // list (<= 256) + one chunk (256), padded

__global__ __launch_bounds__(kKnnThreads) void knn_scan(const float* __restrict__ xq,
                                                        const float* __restrict__ xr, int n, int p,
                                                        const int* __restrict__ rows, int nq,
                                                        int self, int k, int index_base,
                                                        int* __restrict__ out_idx,
                                                        float* __restrict__ out_dist) {
  __shared__ double sd[kScanSlots];
  __shared__ int sj[kScanSlots];
  __shared__ int pending;
  const int q = rows[blockIdx.x];
  const int tid = threadIdx.x;
  const int KF = pow2_at_least(k);          // kept entries (>= k)
  const int N = pow2_at_least(KF + kKnnThreads);
  for (int s = tid; s < kScanSlots; s += kKnnThreads) {
    sd[s] = INFINITY;
    sj[s] = INT_MAX;
  }
  if (tid == 0) pending = 0;
  __syncthreads();
  double tau_d = INFINITY;
  int tau_j = INT_MAX;
  for (int b0 = 0; b0 < n; b0 += kKnnThreads) {
    const int c = b0 + tid;
    if (c < n && !(self && c == q)) {
      const double d = exact_d2(xq, xr, p, q, c);
      if (d < tau_d || (d == tau_d && c < tau_j)) {
        const int pos = KF + atomicAdd(&pending, 1);
        sd[pos] = d;
        sj[pos] = c;
      }
    }
    __syncthreads();
    if (pending > 0) {
      for (int size = 2; size <= N; size <<= 1) {
        for (int h = size >> 1; h > 0; h >>= 1) {
          for (int t = tid; t < N / 2; t += kKnnThreads) {
            const int a = 2 * h * (t / h) + (t % h), bb = a + h;
            const bool up = (a & size) == 0;
            const bool gt = sd[bb] < sd[a] || (sd[bb] == sd[a] && sj[bb] < sj[a]);
            if (gt == up) {
              const double td = sd[a];
              const int tj = sj[a];
              sd[a] = sd[bb];
              sj[a] = sj[bb];
              sd[bb] = td;
              sj[bb] = tj;
            }
          }
          __syncthreads();
        }
      }
      for (int s = KF + tid; s < N; s += kKnnThreads) {
        sd[s] = INFINITY;
        sj[s] = INT_MAX;
      }
      __syncthreads();
      if (tid == 0) pending = 0;
    }
    tau_d = sd[k - 1];
    tau_j = sj[k - 1];
    __syncthreads();
  }
  for (int o = tid; o < k; o += kKnnThreads) {
    out_idx[static_cast<size_t>(o) * nq + q] = sj[o] + index_base;
    out_dist[static_cast<size_t>(o) * nq + q] = static_cast<float>(__dsqrt_rn(sd[o]));
  }
}

template <int KL>
int run_refine(const unsigned long long* lists, const float* xq, const float* xr, int p, int P,
               const double* qn, const double* cmax2, int nq, int self, int k, int index_base,
               int force_scan, int* out_idx, float* out_dist, int* failed, int* failed_count,
               hipStream_t s) {
  const int per = kKnnThreads / 64;
  knn_refine<KL><<<(nq + per - 1) / per, kKnnThreads, 0, s>>>(
      lists, xq, xr, p, P, qn, cmax2, nullptr, nq, self, k, index_base, force_scan, out_idx,
      out_dist, failed, failed_count);
  FE_LAUNCH_CHECK("knn_refine");
  return 0;
}

int list_length(int k) {
  const int need = k + 1 + kKnnMargin;
  return need <= 64 ? 64 : need <= 128 ? 128 : need <= 256 ? 256 : 512;
}

// Exact k-NN of the query rows among the candidate rows; self: the query
// rows are the candidates and each drops itself.  Both matrices column-major
// (rows x p) float32; outputs column-major nq x k.
int exact_knn_run(const float* data, int n, const float* query, int nq, int p, int k, bool self,
                  int device_inputs, int index_base, int* indices, float* distances,
                  int device_outputs, int* scanned_rows) {
  const int P = (p + 7) / 8 * 8;
  const int n_pad = (n + kKnnTile - 1) / kKnnTile * kKnnTile;
  const int KL = list_length(k);
  const bool force_scan = [] {
    const char* e = std::getenv("FASTEMBEDR_HIP_KNN_SCAN_ALL");
    return e != nullptr && std::strcmp(e, "1") == 0;
  }();
  Stream stream;
  if (stream.init()) return 1;
  hipStream_t s = stream.get();
  const size_t np = static_cast<size_t>(n) * p, nqp = static_cast<size_t>(nq) * p;
  const size_t nk = static_cast<size_t>(nq) * k;
  const int pack_blocks = (n_pad + kKnnThreads - 1) / kKnnThreads;
  const int qpack_blocks = (nq + kKnnThreads - 1) / kKnnThreads;
  ArenaPlan ap;
  if (!device_inputs) {
    ap.add<float>(np);
    if (!self) ap.add<float>(nqp);
  }
  if (!device_outputs) {
    ap.add<int>(nk);
    ap.add<float>(nk);
  }
  ap.add<double>(static_cast<size_t>(p) * kKnnPartBlocks);
  ap.add<double>(p);
  ap.add<float>(static_cast<size_t>(n_pad) * P);  // candidates: centred, padded
  ap.add<float>(np);                              // candidates: original
  ap.add<float>(n_pad);
  ap.add<double>(n);
  ap.add<double>(pack_blocks);
  ap.add<double>(1);
  if (!self) {
    ap.add<float>(static_cast<size_t>(nq) * P);   // queries: centred
    ap.add<float>(nqp);                           // queries: original
    ap.add<float>(nq);
    ap.add<double>(nq);
    ap.add<double>(qpack_blocks);
  }
  ap.add<unsigned long long>(static_cast<size_t>(nq) * KL);
  ap.add<int>(nq);
  ap.add<int>(1);
  Arena mem;
  if (mem.init(ap.bytes, "exact KNN")) return 1;
  const float* x = data;
  const float* xq_in = self ? data : query;
  if (!device_inputs) {
    float* up = mem.take<float>(np);
    FE_CHECK(hipMemcpyAsync(up, data, np * sizeof(float), hipMemcpyHostToDevice, s));
    x = up;
    xq_in = up;
    if (!self) {
      float* uq = mem.take<float>(nqp);
      FE_CHECK(hipMemcpyAsync(uq, query, nqp * sizeof(float), hipMemcpyHostToDevice, s));
      xq_in = uq;
    }
  }
  int* o_idx = indices;
  float* o_dist = distances;
  if (!device_outputs) {
    o_idx = mem.take<int>(nk);
    o_dist = mem.take<float>(nk);
  }
  double* part = mem.take<double>(static_cast<size_t>(p) * kKnnPartBlocks);
  double* mean = mem.take<double>(p);
  float* xc = mem.take<float>(static_cast<size_t>(n_pad) * P);
  float* xr = mem.take<float>(np);
  float* nrm = mem.take<float>(n_pad);
  double* qn = mem.take<double>(n);
  double* norm_part = mem.take<double>(pack_blocks);
  double* cmax2 = mem.take<double>(1);
  float* xqc = xc;
  float* xqr = xr;
  double* qqn = qn;
  float* q_nrm = nullptr;
  double* q_part = nullptr;
  if (!self) {
    xqc = mem.take<float>(static_cast<size_t>(nq) * P);
    xqr = mem.take<float>(nqp);
    q_nrm = mem.take<float>(nq);
    qqn = mem.take<double>(nq);
    q_part = mem.take<double>(qpack_blocks);
  }
  unsigned long long* lists = mem.take<unsigned long long>(static_cast<size_t>(nq) * KL);
  int* failed = mem.take<int>(nq);
  int* failed_count = mem.take<int>(1);

  // the candidates' column means centre both matrices
  knn_col_partials<<<dim3(kKnnPartBlocks, p), kKnnThreads, 0, s>>>(x, n, part);
  FE_LAUNCH_CHECK("knn_col_partials");
  knn_col_fold<<<(p + 63) / 64, 64, 0, s>>>(part, p, n, mean);
  FE_LAUNCH_CHECK("knn_col_fold");
  knn_pack<<<pack_blocks, kKnnThreads, 0, s>>>(x, n, n_pad, p, P, mean, xc, xr, nrm, qn, norm_part);
  FE_LAUNCH_CHECK("knn_pack");
  knn_norm_fold<<<1, kKnnThreads, 0, s>>>(norm_part, pack_blocks, cmax2);
  FE_LAUNCH_CHECK("knn_norm_fold");
  if (!self) {
    knn_pack<<<qpack_blocks, kKnnThreads, 0, s>>>(xq_in, nq, nq, p, P, mean, xqc, xqr, q_nrm, qqn,
                                                  q_part);
    FE_LAUNCH_CHECK("knn_pack (queries)");
  }
  FE_CHECK(hipMemsetAsync(failed_count, 0, sizeof(int), s));
  const int sf = self ? 1 : 0, fs = force_scan ? 1 : 0;
  int rc = 0;
  switch (KL) {
    case 64:
      rc = run_select_kl<64>(xc, nrm, n_pad, P, xqc, nq, lists, s) ||
           run_refine<64>(lists, xqr, xr, p, P, qqn, cmax2, nq, sf, k, index_base, fs, o_idx,
                          o_dist, failed, failed_count, s);
      break;
    case 128:
      rc = run_select_kl<128>(xc, nrm, n_pad, P, xqc, nq, lists, s) ||
           run_refine<128>(lists, xqr, xr, p, P, qqn, cmax2, nq, sf, k, index_base, fs, o_idx,
                           o_dist, failed, failed_count, s);
      break;
    case 256:
      rc = run_select_kl<256>(xc, nrm, n_pad, P, xqc, nq, lists, s) ||
           run_refine<256>(lists, xqr, xr, p, P, qqn, cmax2, nq, sf, k, index_base, fs, o_idx,
                           o_dist, failed, failed_count, s);
      break;
    default:
      rc = run_select_kl<512>(xc, nrm, n_pad, P, xqc, nq, lists, s) ||
           run_refine<512>(lists, xqr, xr, p, P, qqn, cmax2, nq, sf, k, index_base, fs, o_idx,
                           o_dist, failed, failed_count, s);
      break;
  }
  if (rc) return 1;
  int h_failed = 0;
  FE_CHECK(hipMemcpyAsync(&h_failed, failed_count, sizeof(int), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  if (h_failed > 0) {
    knn_scan<<<h_failed, kKnnThreads, 0, s>>>(xqr, xr, n, p, failed, nq, sf, k, index_base, o_idx,
                                              o_dist);
    FE_LAUNCH_CHECK("knn_scan");
  }
  if (!device_outputs) {
    FE_CHECK(hipMemcpyAsync(indices, o_idx, nk * sizeof(int), hipMemcpyDeviceToHost, s));
    FE_CHECK(hipMemcpyAsync(distances, o_dist, nk * sizeof(float), hipMemcpyDeviceToHost, s));
  }
  FE_CHECK(hipStreamSynchronize(s));
  if (scanned_rows) *scanned_rows = h_failed;
  return 0;
}

}  // namespace
}  // namespace fe

extern "C" {

int fastembedr_hip_exact_knn(const float* data, int n, int p, int k, int device_inputs,
                             int index_base, int* indices, float* distances, int device_outputs,
                             int* scanned_rows) {
  using namespace fe;
  clear_error();
  if (data == nullptr || indices == nullptr || distances == nullptr)
    return fail("null data or output pointer");
  // native_knn_cuda_impl checks (src/native_knn_cuda_impl.cpp:536-543)
  if (n < 2 || p < 1) return fail("`data` must contain at least two rows and one column.");
  if (k < 1 || k >= n || k > kKnnMaxK)
    return fail("exact KNN requires k in [1, min(n - 1, 256)]");
  if (index_base != 0 && index_base != 1) return fail("index_base must be 0 or 1");
  if (p > 65535) return fail("exact KNN supports at most 65535 columns");
  if (static_cast<long long>(n) * std::max(p, k) > (1ll << 31) - 1)
    return fail("exact KNN supports at most 2^31 - 1 elements per array");
  return exact_knn_run(data, n, data, n, p, k, true, device_inputs, index_base, indices, distances,
                       device_outputs, scanned_rows);
}

int fastembedr_hip_exact_query_knn(const float* data, int n, const float* query, int n_query,
                                   int p, int k, int device_inputs, int index_base, int* indices,
                                   float* distances, int device_outputs, int* scanned_rows) {
  using namespace fe;
  clear_error();
  if (data == nullptr || query == nullptr || indices == nullptr || distances == nullptr)
    return fail("null data or output pointer");
  // native_cuda_query_knn_impl checks (src/native_knn_cuda_impl.cpp:894-906)
  if (n < 1 || n_query < 1 || p < 1)
    return fail("`data` and `query` must have at least one row and matching columns.");
  if (k < 1 || k > n || k > kKnnMaxK)
    return fail("exact query KNN requires k in [1, min(nrow(data), 256)]");
  if (index_base != 0 && index_base != 1) return fail("index_base must be 0 or 1");
  if (p > 65535) return fail("exact KNN supports at most 65535 columns");
  if (static_cast<long long>(std::max(n, n_query)) * std::max(p, k) > (1ll << 31) - 1)
    return fail("exact KNN supports at most 2^31 - 1 elements per array");
  return exact_knn_run(data, n, query, n_query, p, k, false, device_inputs, index_base, indices,
                       distances, device_outputs, scanned_rows);
}

}  // extern "C"
