// Fixed-reference t-SNE transform on gfx950.
//
// Semantics of transform_tsne_cpp (src/tsne_neighbors.cpp:2124-2373) for two
// output dimensions: per-query conditional affinities p_{j|i} (FP64
// bisection, rows sum to 1), median / weighted / random / explicit start,
// and per iteration: repulsion from every reference point (n_ref <= threshold
// or n_negatives >= n_ref) or from n_negatives reference points drawn with
// the per-(iteration, query) xorshift32 stream of :1566-1600, attraction to
// the k reference neighbours scaled by the phase exaggeration, gradient clip,
// gains (+0.2 / x0.8 + 0.01, floor 0.01), momentum, step clip.
//
// The whole optimisation is ONE launch: a group of 16 lanes owns a query and
// keeps its state, its k neighbour coordinates and its p row in registers
// for every iteration; the only memory traffic inside the loop is the
// repulsion samples' reference coordinates (resident in L2 / MALL, or in LDS
// when every reference point is used).  Each lane draws a contiguous slice
// of the query's negative samples; it starts its xorshift32 stream by a
// GF(2) jump-ahead matrix (host-built), so the sample ids equal the
// sequential reference stream exactly.  Arithmetic is float32 with the
// per-query sums reduced in a fixed butterfly order (deterministic).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "../../include/fastembedr_hip.h"
#include "fe_common.hpp"
#include "serve_iface.hpp"
#include "tsne_device.hpp"
#include "tsne_affinity.hpp"

namespace fe {
namespace {

constexpr int kQL = 16;             // lanes per query
constexpr int kQPerBlock = 256 / kQL;
constexpr int kMaxPerLane = kMaxK / kQL;
constexpr int kLdsRef = 4096;       // exact-repulsion reference staged in LDS

struct TransformParams {
  int n_ref, n_query, k, total_iter, early_iter, n_neg, exact, seed;
  float lr, early_exag, exag, mom0, mom1, max_grad, max_step;
  int neg_chunk;
};

// src/tsne_neighbors.cpp:67-88
__host__ __device__ inline uint32_t mix_uint32(uint32_t v) {
  v ^= v >> 16u;
  v *= 0x7feb352du;
  v ^= v >> 15u;
  v *= 0x846ca68bu;
  v ^= v >> 16u;
  return v == 0u ? 0x6d2b79f5u : v;
}

__host__ __device__ inline uint32_t xorshift32_step(uint32_t s) {
  s ^= s << 13u;
  s ^= s >> 17u;
  s ^= s << 5u;
  return s;
}

__device__ __forceinline__ uint32_t apply_matrix(const uint32_t* cols, uint32_t s) {
  uint32_t v = 0u;
#pragma unroll 8
  for (int b = 0; b < 32; ++b) v ^= ((s >> b) & 1u) ? cols[b] : 0u;
  return v;
}

__device__ __forceinline__ float sign_f(float x) { return x == 0.0f ? 0.0f : (x < 0.0f ? -1.0f : 1.0f); }

template <int MODE>  // MODE 1: reference staged in LDS (exact repulsion)
__global__ __launch_bounds__(256) void transform_kernel(const float2* __restrict__ ref,
                                                        const int* __restrict__ nbr,
                                                        const float* __restrict__ prow,
                                                        const float2* __restrict__ y0,
                                                        const uint32_t* __restrict__ jump,
                                                        TransformParams tp,
                                                        float2* __restrict__ out) {
  __shared__ uint32_t s_jump[kQL * 32];
  extern __shared__ float2 s_ref[];
  for (int t = threadIdx.x; t < kQL * 32; t += blockDim.x) s_jump[t] = jump[t];
  if constexpr (MODE == 1)
    for (int t = threadIdx.x; t < tp.n_ref; t += blockDim.x) s_ref[t] = ref[t];
  __syncthreads();
  const int lane = threadIdx.x % kQL;
  const int q = blockIdx.x * kQPerBlock + threadIdx.x / kQL;
  const bool active = q < tp.n_query;
  const int qq = active ? q : 0;
  // neighbours and affinities in registers
  float2 rn[kMaxPerLane];
  float pn[kMaxPerLane];
#pragma unroll
  for (int s = 0; s < kMaxPerLane; ++s) {
    const int j = lane + s * kQL;
    pn[s] = 0.0f;
    rn[s] = make_float2(0.f, 0.f);
    if (j < tp.k) {
      const size_t o = static_cast<size_t>(qq) * tp.k + j;
      rn[s] = ref[nbr[o]];
      pn[s] = prow[o];
    }
  }
  float2 y = y0[qq];
  float ux = 0.f, uy = 0.f, gnx = 1.f, gny = 1.f;
  const uint32_t* my_jump = s_jump + lane * 32;
  const int s_beg = lane * tp.neg_chunk;
  const int s_end = min(tp.n_neg, s_beg + tp.neg_chunk);
  for (int it = 0; it < tp.total_iter; ++it) {
    const bool early = it < tp.early_iter;
    const float ex = early ? tp.early_exag : tp.exag;
    const float mom = early ? tp.mom0 : tp.mom1;
    // repulsion: single pass of sum q and sum q^2 (y - r); scaled by 1/sum_q
    float sq = 0.f, rx = 0.f, ry = 0.f;
    if (tp.exact) {
      for (int r = lane; r < tp.n_ref; r += kQL) {
        const float2 p = MODE == 1 ? s_ref[r] : ref[r];
        const float dx = y.x - p.x, dy = y.y - p.y;
        const float qv = 1.0f / (1.0f + dx * dx + dy * dy);
        sq += qv;
        rx += qv * qv * dx;
        ry += qv * qv * dy;
      }
    } else {
      const int iter_seed = tp.seed + 65537 * (it + 1);
      const uint32_t st0 = mix_uint32(static_cast<uint32_t>(iter_seed) ^
                                      (static_cast<uint32_t>(qq + 1) * 0x9e3779b9u));
      uint32_t st = apply_matrix(my_jump, st0);
      for (int m = s_beg; m < s_end; ++m) {
        st = xorshift32_step(st);
        const int r = static_cast<int>((static_cast<uint64_t>(st) * static_cast<uint64_t>(tp.n_ref)) >> 32u);
        const float2 p = ref[r];
        const float dx = y.x - p.x, dy = y.y - p.y;
        const float qv = 1.0f / (1.0f + dx * dx + dy * dy);
        sq += qv;
        rx += qv * qv * dx;
        ry += qv * qv * dy;
      }
    }
    // attraction
    float ax = 0.f, ay = 0.f;
#pragma unroll
    for (int s = 0; s < kMaxPerLane; ++s) {
      const float dx = y.x - rn[s].x, dy = y.y - rn[s].y;
      const float qv = 1.0f / (1.0f + dx * dx + dy * dy);
      const float c = ex * pn[s] * qv;
      ax += c * dx;
      ay += c * dy;
    }
#pragma unroll
    for (int o = kQL / 2; o > 0; o >>= 1) {
      sq += __shfl_xor(sq, o, kQL);
      rx += __shfl_xor(rx, o, kQL);
      ry += __shfl_xor(ry, o, kQL);
      ax += __shfl_xor(ax, o, kQL);
      ay += __shfl_xor(ay, o, kQL);
    }
    const float inv = 1.0f / (sq + FLT_MIN);
    float gx = ax - rx * inv, gy = ay - ry * inv;
    const float gn2 = gx * gx + gy * gy;
    if (gn2 > tp.max_grad * tp.max_grad) {
      const float sc = tp.max_grad / (sqrtf(gn2) + 1e-12f);
      gx *= sc;
      gy *= sc;
    }
    gnx = sign_f(ux) != sign_f(gx) ? gnx + 0.2f : gnx * 0.8f + 0.01f;
    gny = sign_f(uy) != sign_f(gy) ? gny + 0.2f : gny * 0.8f + 0.01f;
    gnx = fmaxf(gnx, 0.01f);
    gny = fmaxf(gny, 0.01f);
    ux = mom * ux - tp.lr * gnx * gx;
    uy = mom * uy - tp.lr * gny * gy;
    const float sn2 = ux * ux + uy * uy;
    if (sn2 > tp.max_step * tp.max_step) {
      const float sc = tp.max_step / (sqrtf(sn2) + 1e-12f);
      ux *= sc;
      uy *= sc;
    }
    y.x += ux;
    y.y += uy;
  }
  if (active && lane == 0) out[q] = y;
}

// median / weighted starts (initialize_tsne_transform, :1417-1451): one
// group of kQL lanes per query; the median uses exact ranks with index
// tie-breaks, which selects the same order statistics as nth_element.
__global__ __launch_bounds__(256) void transform_init_kernel(const float2* __restrict__ ref,
                                                             const int* __restrict__ nbr,
                                                             const float* __restrict__ dist_f,
                                                             int n_query, int k, int weighted,
                                                             float2* __restrict__ y0) {
  const int lane = threadIdx.x % kQL;
  const int q = blockIdx.x * kQPerBlock + threadIdx.x / kQL;
  if (q >= n_query) return;
  const int* nb = nbr + static_cast<size_t>(q) * k;
  if (weighted) {
    double num_x = 0.0, num_y = 0.0, den = 0.0;
    for (int j = lane; j < k; j += kQL) {
      const float2 r = ref[nb[j]];
      const double w = 1.0 / (fmax(0.0, static_cast<double>(dist_f[static_cast<size_t>(q) * k + j])) + 1e-6);
      num_x += w * r.x;
      num_y += w * r.y;
      den += w;
    }
#pragma unroll
    for (int o = kQL / 2; o > 0; o >>= 1) {
      num_x += __shfl_xor(num_x, o, kQL);
      num_y += __shfl_xor(num_y, o, kQL);
      den += __shfl_xor(den, o, kQL);
    }
    if (lane == 0)
      y0[q] = make_float2(static_cast<float>(num_x / (den + DBL_MIN)),
                          static_cast<float>(num_y / (den + DBL_MIN)));
    return;
  }
  const int lo_rank = (k - 1) / 2, hi_rank = k / 2;  // equal for odd k
  float med[2];
  for (int d = 0; d < 2; ++d) {
    float lo = 0.f, hi = 0.f;
    for (int j = lane; j < k; j += kQL) {
      const float2 rj = ref[nb[j]];
      const float vj = d == 0 ? rj.x : rj.y;
      int rank = 0;
      for (int u = 0; u < k; ++u) {
        const float2 ru = ref[nb[u]];
        const float vu = d == 0 ? ru.x : ru.y;
        rank += (vu < vj || (vu == vj && u < j)) ? 1 : 0;
      }
      if (rank == lo_rank) lo = vj;
      if (rank == hi_rank) hi = vj;
    }
#pragma unroll
    for (int o = kQL / 2; o > 0; o >>= 1) {
      lo += __shfl_xor(lo, o, kQL);
      hi += __shfl_xor(hi, o, kQL);
    }
    med[d] = (k & 1) ? hi : static_cast<float>(0.5 * (static_cast<double>(lo) + hi));
  }
  if (lane == 0) y0[q] = make_float2(med[0], med[1]);
}

// 32x32 GF(2) matrices as 32 column words: cols[b] = image of bit b.
struct Gf2 {
  uint32_t c[32];
};

static uint32_t gf2_apply(const Gf2& m, uint32_t s) {
  uint32_t v = 0u;
  for (int b = 0; b < 32; ++b)
    if ((s >> b) & 1u) v ^= m.c[b];
  return v;
}

static Gf2 gf2_mul(const Gf2& a, const Gf2& b) {  // a after b
  Gf2 r;
  for (int j = 0; j < 32; ++j) r.c[j] = gf2_apply(a, b.c[j]);
  return r;
}

static Gf2 gf2_pow_xorshift(long long e) {
  Gf2 base, acc;
  for (int b = 0; b < 32; ++b) {
    base.c[b] = xorshift32_step(1u << b);
    acc.c[b] = 1u << b;
  }
  while (e > 0) {
    if (e & 1) acc = gf2_mul(base, acc);
    base = gf2_mul(base, base);
    e >>= 1;
  }
  return acc;
}

}  // namespace

// Resident reference layout (the served "fixed reference").
struct TransformReference {
  int n_ref = 0;
  Arena mem;
  float2* layout = nullptr;
};

// A transform request with its arguments checked and its host-side inputs
// built: the start layout when the host makes it (explicit or random), the
// weighted start's row-major distances, the jump matrices.
struct TransformTask {
  TransformParams tp{};
  int index_offset = 0;
  double perplexity = 0.0;
  int start = 0;  // 0 median, 1 weighted (both by the init kernel), 2 host start
  bool lds_ref = false;
  std::size_t dist_size = 0;
  std::vector<float> host_start;
  std::vector<float> dist_rm;
  std::vector<uint32_t> jump;
};

template <typename DistT>
static int transform_prepare(int n_ref, const int* indices, const DistT* distances, int n_query,
                             int k, int index_offset, const float* init, int initialization,
                             double perplexity, int n_iter, int early_iter, double lr,
                             double early_exag, double exag, double mom0, double mom1,
                             double max_grad, double max_step, int n_negatives, int threshold,
                             int seed, TransformTask* t) {
  if (!indices || !distances) return fail("null transform input pointer");
  if (n_query < 1 || k < 1) return fail("KNN input must have at least one query row and one neighbor column.");
  if (k > kMaxK) return fail("HIP transform supports at most 256 neighbors.");
  if (!(perplexity > 0.0)) return fail("`perplexity` must be positive.");
  if (n_iter < 0 || early_iter < 0) return fail("iteration counts must be non-negative.");
  if (n_iter + early_iter < 1) return fail("at least one transform iteration is required.");
  if (!(lr > 0.0)) return fail("`learning_rate` must be positive.");
  if (!(early_exag > 0.0) || !(exag > 0.0)) return fail("exaggeration values must be positive.");
  if (mom0 < 0.0 || mom1 < 0.0) return fail("momentum values must be non-negative.");
  if (n_negatives < 1) return fail("`n_negatives` must be positive.");
  if (initialization < 0 || initialization > 2) return fail("`initialization` must be 'median', 'weighted', or 'random'.");
  if (index_offset != 0 && index_offset != 1) return fail("index_offset must be 0 or 1");
  if (threshold < 1) threshold = 1;
  if (!(max_grad > 0.0) || !std::isfinite(max_grad)) max_grad = FLT_MAX;
  if (!(max_step > 0.0) || !std::isfinite(max_step)) max_step = FLT_MAX;
  if (n_negatives > n_ref) n_negatives = n_ref;
  const bool exact = n_ref <= threshold || n_negatives >= n_ref;

  // the KNN checks of the P-row kernel, on the host before anything is
  // enqueued (the kernel still flags them)
  const std::size_t total = static_cast<std::size_t>(n_query) * k;
  for (std::size_t e = 0; e < total; ++e) {
    const long long nb = static_cast<long long>(indices[e]) - index_offset;
    if (nb < 0 || nb >= n_ref) return fail("KNN indices are out of range for `reference_layout`.");
  }
  for (std::size_t e = 0; e < total; ++e) {
    const double d = static_cast<double>(distances[e]);
    if (!std::isfinite(d) || d < 0.0) return fail("KNN distances must be finite and non-negative.");
  }

  t->index_offset = index_offset;
  t->perplexity = perplexity;
  t->dist_size = sizeof(DistT);
  if (init) {
    for (std::size_t e = 0; e < 2 * static_cast<std::size_t>(n_query); ++e)
      if (!std::isfinite(init[e])) return fail("`Y_init` must contain only finite coordinates.");
    t->start = 2;
    t->host_start.assign(init, init + 2 * static_cast<std::size_t>(n_query));
  } else if (initialization == 2) {
    // tsne_neighbors.cpp:1397-1414: per-query std::mt19937 and N(0, 1e-4)
    t->start = 2;
    t->host_start.resize(2 * static_cast<std::size_t>(n_query));
    const uint32_t rs = static_cast<uint32_t>(seed);
    for (int i = 0; i < n_query; ++i) {
      std::mt19937 rng(mix_uint32(rs ^ (static_cast<uint32_t>(i + 1) * 0x9e3779b9u)));
      std::normal_distribution<double> normal(0.0, 1.0e-4);
      t->host_start[2 * i] = static_cast<float>(normal(rng));
      t->host_start[2 * i + 1] = static_cast<float>(normal(rng));
    }
  } else {
    t->start = initialization;
    if (initialization == 1) {
      // column-major input -> row-major float copy for the init kernel
      t->dist_rm.resize(total);
      for (std::size_t e = 0; e < total; ++e) {
        const std::size_t i = e / k, j = e % k;
        t->dist_rm[e] = static_cast<float>(distances[j * n_query + i]);
      }
    }
  }

  TransformParams& tp = t->tp;
  tp.n_ref = n_ref;
  tp.n_query = n_query;
  tp.k = k;
  tp.total_iter = n_iter + early_iter;
  tp.early_iter = early_iter;
  tp.n_neg = exact ? n_ref : n_negatives;
  tp.exact = exact ? 1 : 0;
  tp.seed = seed;
  tp.lr = static_cast<float>(lr);
  tp.early_exag = static_cast<float>(early_exag);
  tp.exag = static_cast<float>(exag);
  tp.mom0 = static_cast<float>(mom0);
  tp.mom1 = static_cast<float>(mom1);
  tp.max_grad = static_cast<float>(max_grad);
  tp.max_step = static_cast<float>(max_step);
  tp.neg_chunk = (n_negatives + kQL - 1) / kQL;
  t->lds_ref = exact && n_ref <= kLdsRef;
  t->jump.resize(kQL * 32);
  for (int l = 0; l < kQL; ++l) {
    const Gf2 m = gf2_pow_xorshift(static_cast<long long>(l) * tp.neg_chunk);
    std::memcpy(t->jump.data() + 32 * l, m.c, sizeof(m.c));
  }
  return 0;
}

// Device workspace of a request, carved in this order by transform_enqueue.
static std::size_t transform_device_bytes(const TransformTask& t) {
  const std::size_t q = t.tp.n_query, total = q * t.tp.k;
  ArenaPlan plan;
  plan.add<int>(total);                              // ids upload
  plan.add<unsigned char>(total * t.dist_size);      // distances upload
  plan.add<int>(total);                              // nbr
  plan.add<float>(total);                            // p rows
  plan.add<float>(t.start == 1 ? total : 0);         // float distances (weighted start)
  plan.add<float2>(q);                               // start
  plan.add<float2>(q);                               // result
  plan.add<uint32_t>(kQL * 32);                      // jump matrices
  plan.add<unsigned>(4);
  return plan.bytes;
}

// Host-side pointers of one enqueue: where the uploads come from and the
// downloads go.
struct TransformIO {
  const int* idx;
  const void* dist;
  const float* start;    // host start (start == 2)
  const float* dist_rm;  // weighted start
  const uint32_t* jump;
  float* out;
  unsigned* flags;
};

// Enqueues the request on s without waiting: uploads, P rows, start, the
// optimisation, and the downloads of the result and the error flags.
template <typename DistT>
static int transform_enqueue(const TransformReference& R, const TransformTask& t,
                             const TransformIO& io, void* dev, hipStream_t s) {
  const int n_query = t.tp.n_query, k = t.tp.k;
  const std::size_t total = static_cast<std::size_t>(n_query) * k;
  unsigned char* base = static_cast<unsigned char*>(dev);
  std::size_t off = 0;
  auto take = [&](std::size_t bytes) -> void* {
    if (!bytes) return nullptr;
    void* p = base + off;
    off += align_up(bytes);
    return p;
  };
  int* d_idx = static_cast<int*>(take(total * sizeof(int)));
  DistT* d_dist = static_cast<DistT*>(take(total * sizeof(DistT)));
  int* nbr = static_cast<int*>(take(total * sizeof(int)));
  float* prow = static_cast<float*>(take(total * sizeof(float)));
  float* dist_f = static_cast<float*>(take(t.start == 1 ? total * sizeof(float) : 0));
  float2* y0 = static_cast<float2*>(take(n_query * sizeof(float2)));
  float2* yo = static_cast<float2*>(take(n_query * sizeof(float2)));
  uint32_t* jump = static_cast<uint32_t*>(take(kQL * 32 * sizeof(uint32_t)));
  unsigned* flags = static_cast<unsigned*>(take(4 * sizeof(unsigned)));
  FE_CHECK(hipMemcpyAsync(d_idx, io.idx, total * sizeof(int), hipMemcpyHostToDevice, s));
  FE_CHECK(hipMemcpyAsync(d_dist, io.dist, total * sizeof(DistT), hipMemcpyHostToDevice, s));
  FE_CHECK(hipMemsetAsync(flags, 0, 4 * sizeof(unsigned), s));
  constexpr int rpb = aff_rows_per_block<DistT>();
  const std::size_t lds = static_cast<std::size_t>(rpb) * k * (sizeof(DistT) + sizeof(int));
  if (lds > 65536)
    FE_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&perplexity_rows<DistT>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
  perplexity_rows<DistT><<<(n_query + rpb - 1) / rpb, 256, lds, s>>>(
      d_idx, d_dist, n_query, R.n_ref, k, t.index_offset, t.perplexity, nbr, prow, flags);
  FE_LAUNCH_CHECK("perplexity_rows(transform)");
  if (t.start == 2) {
    FE_CHECK(hipMemcpyAsync(y0, io.start, n_query * sizeof(float2), hipMemcpyHostToDevice, s));
  } else {
    if (t.start == 1)
      FE_CHECK(hipMemcpyAsync(dist_f, io.dist_rm, total * sizeof(float), hipMemcpyHostToDevice, s));
    transform_init_kernel<<<(n_query + kQPerBlock - 1) / kQPerBlock, 256, 0, s>>>(
        R.layout, nbr, dist_f, n_query, k, t.start == 1 ? 1 : 0, y0);
    FE_LAUNCH_CHECK("transform_init_kernel");
  }
  FE_CHECK(hipMemcpyAsync(jump, io.jump, kQL * 32 * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  const int blocks = (n_query + kQPerBlock - 1) / kQPerBlock;
  if (t.lds_ref) {
    transform_kernel<1><<<blocks, 256, R.n_ref * sizeof(float2), s>>>(R.layout, nbr, prow, y0,
                                                                      jump, t.tp, yo);
  } else {
    transform_kernel<0><<<blocks, 256, 0, s>>>(R.layout, nbr, prow, y0, jump, t.tp, yo);
  }
  FE_LAUNCH_CHECK("transform_kernel");
  FE_CHECK(hipMemcpyAsync(io.out, yo, n_query * sizeof(float2), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipMemcpyAsync(io.flags, flags, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  return 0;
}

static int transform_flags_error(unsigned flags) {
  if (flags & kErrIndexRange) return fail("KNN indices are out of range for `reference_layout`.");
  if (flags & kErrDistance) return fail("KNN distances must be finite and non-negative.");
  return 0;
}

// One request, synchronously: its own workspace, enqueue, wait.
template <typename DistT>
static int transform_run(const TransformReference& R, const int* indices, const DistT* distances,
                         int n_query, int k, int index_offset, const float* init,
                         int initialization, double perplexity, int n_iter, int early_iter,
                         double lr, double early_exag, double exag, double mom0, double mom1,
                         double max_grad, double max_step, int n_negatives, int threshold,
                         int seed, float* out, hipStream_t s) {
  if (!out) return fail("null transform input pointer");
  TransformTask t;
  if (transform_prepare<DistT>(R.n_ref, indices, distances, n_query, k, index_offset, init,
                               initialization, perplexity, n_iter, early_iter, lr, early_exag,
                               exag, mom0, mom1, max_grad, max_step, n_negatives, threshold, seed,
                               &t))
    return 1;
  Arena a;
  if (a.init(transform_device_bytes(t), "HIP transform workspace")) return 1;
  unsigned h_flags = 0;
  const TransformIO io{indices, distances, t.host_start.data(), t.dist_rm.data(), t.jump.data(),
                       out, &h_flags};
  void* dev = a.take<unsigned char>(a.capacity());
  if (transform_enqueue<DistT>(R, t, io, dev, s)) {
    (void)hipStreamSynchronize(s);
    return 1;
  }
  FE_CHECK(hipStreamSynchronize(s));
  return transform_flags_error(h_flags);
}

static int make_reference(const float* layout, int n_ref, TransformReference* R, hipStream_t s) {
  if (!layout) return fail("null reference layout");
  if (n_ref < 1) return fail("`reference_layout` must be a non-empty matrix.");
  for (size_t t = 0; t < 2 * static_cast<size_t>(n_ref); ++t)
    if (!std::isfinite(layout[t])) return fail("`reference_layout` must contain only finite coordinates.");
  R->n_ref = n_ref;
  ArenaPlan p;
  p.add<float2>(n_ref);
  if (R->mem.init(p.bytes, "HIP transform reference")) return 1;
  R->layout = R->mem.take<float2>(n_ref);
  FE_CHECK(hipMemcpyAsync(R->layout, layout, n_ref * sizeof(float2), hipMemcpyHostToDevice, s));
  return check(hipStreamSynchronize(s), "reference upload");
}

// ---- the serving session's interface (serve_iface.hpp) ----

int transform_task_create(const TransformReference& R, const int* indices, const float* distances,
                          int n_query, int k, int index_offset, const float* init,
                          int initialization, double perplexity, int n_iter,
                          int early_exaggeration_iter, double learning_rate,
                          double early_exaggeration, double exaggeration,
                          double initial_momentum, double final_momentum, double max_grad_norm,
                          double max_step_norm, int n_negatives, int exact_repulsion_threshold,
                          int seed, TransformTaskPtr* out) {
  TransformTaskPtr t(new TransformTask);
  if (transform_prepare<float>(R.n_ref, indices, distances, n_query, k, index_offset, init,
                               initialization, perplexity, n_iter, early_exaggeration_iter,
                               learning_rate, early_exaggeration, exaggeration, initial_momentum,
                               final_momentum, max_grad_norm, max_step_norm, n_negatives,
                               exact_repulsion_threshold, seed, t.get()))
    return 1;
  *out = std::move(t);
  return 0;
}

void transform_task_destroy(TransformTask* t) { delete t; }

std::size_t transform_task_device_bytes(const TransformTask& t) { return transform_device_bytes(t); }

int transform_task_queries(const TransformTask& t) { return t.tp.n_query; }

// pinned staging, in this order: ids, distances, host start, weighted-start
// distances, jump matrices, result, flags
namespace {
struct PinnedLayout {
  std::size_t idx, dist, start, dist_rm, jump, out, flags, bytes;
};
PinnedLayout pinned_layout(const TransformTask& t) {
  const std::size_t q = t.tp.n_query, total = q * t.tp.k;
  PinnedLayout L{};
  std::size_t off = 0;
  auto put = [&](std::size_t bytes) {
    const std::size_t at = off;
    off += align_up(bytes);
    return at;
  };
  L.idx = put(total * sizeof(int));
  L.dist = put(total * sizeof(float));
  L.start = put(t.host_start.size() * sizeof(float));
  L.dist_rm = put(t.dist_rm.size() * sizeof(float));
  L.jump = put(t.jump.size() * sizeof(uint32_t));
  L.out = put(q * sizeof(float2));
  L.flags = put(4 * sizeof(unsigned));
  L.bytes = off;
  return L;
}
}  // namespace

std::size_t transform_task_pinned_bytes(const TransformTask& t) { return pinned_layout(t).bytes; }

int transform_task_enqueue(const TransformReference& R, const TransformTask& t, const int* indices,
                           const float* distances, void* dev, void* pinned, hipStream_t s) {
  const PinnedLayout L = pinned_layout(t);
  unsigned char* h = static_cast<unsigned char*>(pinned);
  const std::size_t total = static_cast<std::size_t>(t.tp.n_query) * t.tp.k;
  std::memcpy(h + L.idx, indices, total * sizeof(int));
  std::memcpy(h + L.dist, distances, total * sizeof(float));
  if (!t.host_start.empty())
    std::memcpy(h + L.start, t.host_start.data(), t.host_start.size() * sizeof(float));
  if (!t.dist_rm.empty())
    std::memcpy(h + L.dist_rm, t.dist_rm.data(), t.dist_rm.size() * sizeof(float));
  std::memcpy(h + L.jump, t.jump.data(), t.jump.size() * sizeof(uint32_t));
  const TransformIO io{reinterpret_cast<const int*>(h + L.idx), h + L.dist,
                       reinterpret_cast<const float*>(h + L.start),
                       reinterpret_cast<const float*>(h + L.dist_rm),
                       reinterpret_cast<const uint32_t*>(h + L.jump),
                       reinterpret_cast<float*>(h + L.out), reinterpret_cast<unsigned*>(h + L.flags)};
  return transform_enqueue<float>(R, t, io, dev, s);
}

int transform_task_result(const TransformTask& t, const void* pinned, const float** layout) {
  const PinnedLayout L = pinned_layout(t);
  const unsigned char* h = static_cast<const unsigned char*>(pinned);
  if (transform_flags_error(*reinterpret_cast<const unsigned*>(h + L.flags))) return 1;
  *layout = reinterpret_cast<const float*>(h + L.out);
  return 0;
}

}  // namespace fe

struct fastembedr_hip_transform_reference {
  fe::TransformReference impl;
  fe::Stream stream;
};

const fe::TransformReference* fe::transform_reference_of(
    const fastembedr_hip_transform_reference* ref) {
  return ref ? &ref->impl : nullptr;
}

extern "C" {

int fastembedr_hip_tsne_transform(const float* reference_layout, int n_reference,
                                  const int* indices, const double* distances, int n_query, int k,
                                  int index_offset, const float* init, int initialization,
                                  double perplexity, int n_iter, int early_exaggeration_iter,
                                  double learning_rate, double early_exaggeration,
                                  double exaggeration, double initial_momentum,
                                  double final_momentum, double max_grad_norm,
                                  double max_step_norm, int n_negatives,
                                  int exact_repulsion_threshold, int seed, float* out) {
  fe::clear_error();
  fe::Stream stream;
  if (stream.init()) return 1;
  fe::TransformReference R;
  if (fe::make_reference(reference_layout, n_reference, &R, stream.get())) return 1;
  return fe::transform_run<double>(R, indices, distances, n_query, k, index_offset, init,
                                   initialization, perplexity, n_iter, early_exaggeration_iter,
                                   learning_rate, early_exaggeration, exaggeration,
                                   initial_momentum, final_momentum, max_grad_norm, max_step_norm,
                                   n_negatives, exact_repulsion_threshold, seed, out, stream.get());
}

int fastembedr_hip_transform_reference_create(const float* reference_layout, int n_reference,
                                              fastembedr_hip_transform_reference** ref) {
  fe::clear_error();
  if (!ref) return fe::fail("null reference handle");
  *ref = nullptr;
  auto* r = new (std::nothrow) fastembedr_hip_transform_reference;
  if (!r) return fe::fail("out of host memory");
  if (r->stream.init() || fe::make_reference(reference_layout, n_reference, &r->impl, r->stream.get())) {
    delete r;
    return 1;
  }
  *ref = r;
  return 0;
}

int fastembedr_hip_transform_reference_run(fastembedr_hip_transform_reference* ref,
                                           const int* indices, const float* distances,
                                           int n_query, int k, int index_offset, const float* init,
                                           int initialization, double perplexity, int n_iter,
                                           int early_exaggeration_iter, double learning_rate,
                                           double early_exaggeration, double exaggeration,
                                           double initial_momentum, double final_momentum,
                                           double max_grad_norm, double max_step_norm,
                                           int n_negatives, int exact_repulsion_threshold,
                                           int seed, float* out) {
  fe::clear_error();
  if (!ref) return fe::fail("null reference handle");
  return fe::transform_run<float>(ref->impl, indices, distances, n_query, k, index_offset, init,
                                  initialization, perplexity, n_iter, early_exaggeration_iter,
                                  learning_rate, early_exaggeration, exaggeration, initial_momentum,
                                  final_momentum, max_grad_norm, max_step_norm, n_negatives,
                                  exact_repulsion_threshold, seed, out, ref->stream.get());
}

void fastembedr_hip_transform_reference_destroy(fastembedr_hip_transform_reference* ref) {
  delete ref;
}

}  // extern "C"
