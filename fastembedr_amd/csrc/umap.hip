// UMAP edge-sampled SGD epoch on gfx950.
//
// Semantics of fastembedr_cuda_umap_optimize_coo and its epoch kernels
// (src/embedding_cuda_kernels.cpp:3044-3258, 5483-5608): per directed COO
// edge, the closed-form positive-sample count of this epoch, attraction
// -2ab d^(2(b-1)) / (1 + a d^(2b)) clipped to +-4 on both endpoints, the
// closed-form negative-sample count with hash-chosen negative vertices,
// repulsion 2 gamma b / ((0.001 + d^2)(1 + a d^(2b))) clipped to +-4 on the
// head, alpha = lr (1 - e / E), then coordinates sanitised to [-16, 16].
//
// optimizer_mode 0 updates the layout in place with float atomics (the
// reference's hogwild form; not bit-reproducible; see umap_epoch_hogwild for
// how the atomics are grouped).  optimizer_mode 1 reads the
// epoch-start layout and accumulates every update into a 64-bit fixed-point
// delta (scale 2^40), so the epoch result does not depend on arrival order;
// the apply kernel adds, sanitises and clears the delta in one pass, which
// keeps the epoch loop free of memsets and capturable in a single hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fastembedr_hip.h"
#include "fe_common.hpp"
#include "serve_iface.hpp"

namespace fe {

namespace {

constexpr int kUThreads = 256;
constexpr double kDeltaScale = 1099511627776.0;  // 2^40

struct UmapParams {
  int n, n_epochs, nsr;
  unsigned seed;
  float lr, a, b, gamma;
  unsigned long long nmod;  // umap_nmod(n): x mod n by multiplication
};

// c = ceil(2^64 / n) for n >= 1: for every 32-bit x, x mod n is the high 64
// bits of (c * x mod 2^64) * n (Lemire, Kaser and Kurz, "Faster remainder by
// direct computation", 2019), the same value as the % operator.
inline unsigned long long umap_nmod(int n) {
  return ~0ull / static_cast<unsigned long long>(n > 0 ? n : 1) + 1ull;
}

__device__ __forceinline__ unsigned fast_mod(unsigned x, unsigned long long c, unsigned n) {
  const unsigned long long low = c * x;
  return static_cast<unsigned>(__umul64hi(low, static_cast<unsigned long long>(n)));
}

__device__ __forceinline__ unsigned mix_uint(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// deterministic_vertex (embedding_cuda_kernels.cpp:262-274)
__device__ __forceinline__ unsigned negative_vertex(unsigned n, unsigned long long nmod,
                                                    unsigned seed, unsigned epoch, unsigned head,
                                                    unsigned tail, unsigned s) {
  unsigned x = seed;
  x ^= epoch * 0x9e3779b9u;
  x ^= (head + 1u) * 0x85ebca6bu;
  x ^= (tail + 1u) * 0xc2b2ae35u;
  x ^= (s + 1u) * 0x27d4eb2du;
  return fast_mod(mix_uint(x), nmod, n);  // mix_uint(x) % n
}

__device__ __forceinline__ float clip4(float x) { return fminf(4.0f, fmaxf(-4.0f, x)); }

// fast_positive_pow (:364-383): exponent-interpolated x^b for x > 0
__device__ __forceinline__ float fast_pow(float x, float b) {
  if (x <= 0.0f) return 0.0f;
  const unsigned xb = __float_as_uint(x);
  constexpr float bias = 1064866805.0f;
  const int whole = static_cast<int>(b);
  const float frac = b - static_cast<float>(whole);
  const unsigned ib = static_cast<unsigned>(__fadd_rn(__fmul_rn(frac, __fsub_rn(static_cast<float>(xb), bias)), bias));
  const float fp = __uint_as_float(ib);
  float ip = 1.0f, base = x;
  int e = whole;
  while (e > 0) {
    if (e & 1) ip *= base;
    base *= base;
    e >>= 1;
  }
  return ip * fp;
}

__device__ __forceinline__ float attract(float d2, const UmapParams& p) {
  if (d2 <= 0.0f) return 0.0f;
  const float d2b = fast_pow(d2, p.b);
  return -2.0f * p.a * p.b * (d2b / d2) / (p.a * d2b + 1.0f);
}

__device__ __forceinline__ float repel(float d2, const UmapParams& p) {
  if (d2 <= 0.0f) return 0.0f;
  const float d2b = fast_pow(d2, p.b);
  return p.gamma * 2.0f * p.b / ((0.001f + d2) * (p.a * d2b + 1.0f));
}

// schedules (:3044-3084); IEEE division and floor, no fast-math
__device__ __forceinline__ int n_positive(float period, unsigned epoch) {
  if (period <= 0.0f || !isfinite(period) || epoch == 0u) return 0;
  const float now = static_cast<float>(epoch), prev = static_cast<float>(epoch - 1u);
  const int cur = static_cast<int>(floorf(__fdiv_rn(now, period)));
  const int pre = static_cast<int>(floorf(__fdiv_rn(prev, period)));
  const int s = cur - pre;
  return s > 0 ? s : 0;
}

__device__ __forceinline__ int n_negative(float period, int nsr, unsigned epoch) {
  if (nsr <= 0 || period <= 0.0f || !isfinite(period) || epoch == 0u) return 0;
  const float now = static_cast<float>(epoch), prev = static_cast<float>(epoch - 1u);
  const int cur = static_cast<int>(floorf(__fdiv_rn(now, period)));
  const int pre = static_cast<int>(floorf(__fdiv_rn(prev, period)));
  if (cur <= pre) return 0;
  const float np = __fdiv_rn(period, static_cast<float>(nsr));
  if (np <= 0.0f || !isfinite(np)) return 0;
  const int cur_total = static_cast<int>(floorf(__fdiv_rn(__fsub_rn(now, np), np)));
  int prev_total = 0;
  if (cur > 1) {
    const float pa = ceilf(__fmul_rn(static_cast<float>(cur - 1), period));
    prev_total = static_cast<int>(floorf(__fdiv_rn(__fsub_rn(pa, np), np)));
  }
  const int s = cur_total - prev_total;
  return s > 0 ? s : 0;
}

__device__ __forceinline__ long long fixed(float v) { return __double2ll_rn(static_cast<double>(v) * kDeltaScale); }

__device__ __forceinline__ unsigned long long shfl_up_u64(unsigned long long v, int d) {
  const int lo = __shfl_up(static_cast<int>(static_cast<unsigned>(v)), d, 64);
  const int hi = __shfl_up(static_cast<int>(static_cast<unsigned>(v >> 32)), d, 64);
  return (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
         static_cast<unsigned>(lo);
}

// Epoch-start coordinates of point i from the layout in memory or in LDS.
struct GlobalPos {
  const float* __restrict__ y;
  __device__ __forceinline__ float2 operator()(size_t i) const {
    return make_float2(y[2 * i], y[2 * i + 1]);
  }
};
struct LdsPos {
  const float* y;
  __device__ __forceinline__ float2 operator()(size_t i) const {
    return make_float2(y[2 * i], y[2 * i + 1]);
  }
};

// The head's share of symmetric-graph edge (h, t) in a mode 1 epoch, as
// integers of the fixed-point delta: twice its attraction (the reverse edge's
// update of t is exactly this edge's tail update, see umap_epoch_grouped)
// times the positive samples, plus its repulsions.  Nothing when the edge
// takes no positive sample this epoch.
template <class Pos>
__device__ __forceinline__ void sym_edge_head(Pos pos, int h, int t, float per, const UmapParams& p,
                                              unsigned epoch, float alpha, long long* hx,
                                              long long* hy) {
  constexpr float eps = 1.1920928955078125e-7f;
  const int npos = n_positive(per, epoch);
  if (npos <= 0) return;
  const float2 hp = pos(static_cast<size_t>(h));
  {
    const float2 tp = pos(static_cast<size_t>(t));
    const float dx = hp.x - tp.x;
    const float dy = hp.y - tp.y;
    const float d2 = fmaxf(eps, dx * dx + dy * dy);
    const float c = attract(d2, p);
    const float gx = clip4(c * dx) * alpha, gy = clip4(c * dy) * alpha;
    *hx += 2 * (npos * fixed(gx));
    *hy += 2 * (npos * fixed(gy));
  }
  const int nneg = n_negative(per, p.nsr, epoch);
  for (int s = 0; s < nneg; ++s) {
    const unsigned v = negative_vertex(static_cast<unsigned>(p.n), p.nmod, p.seed, epoch,
                                       static_cast<unsigned>(h), static_cast<unsigned>(t),
                                       static_cast<unsigned>(s));
    if (static_cast<int>(v) == h || static_cast<int>(v) == t) continue;
    const float2 vp = pos(static_cast<size_t>(v));
    const float dx = hp.x - vp.x;
    const float dy = hp.y - vp.y;
    const float d2 = fmaxf(eps, dx * dx + dy * dy);
    const float c = repel(d2, p);
    *hx += fixed(clip4(c * dx) * alpha);
    *hy += fixed(clip4(c * dy) * alpha);
  }
}

// Inclusive sums over the runs of equal heads in a wave (COO heads come in
// row order); `last` marks the lane that ends its run.
__device__ __forceinline__ void head_run_sums(int h, long long hx, long long hy,
                                              unsigned long long* ux, unsigned long long* uy,
                                              bool* last) {
  const int lane = threadIdx.x & 63;
  const unsigned long long upto_lane = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int hprev = __shfl_up(h, 1, 64);
  const bool start = lane == 0 || hprev != h;
  const unsigned long long starts = __ballot(start) & upto_lane;
  const int lim = 63 - __clzll(starts);  // first lane of this lane's run
  unsigned long long x = static_cast<unsigned long long>(hx);
  unsigned long long y = static_cast<unsigned long long>(hy);
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long ox = shfl_up_u64(x, d), oy = shfl_up_u64(y, d);
    if (lane - d >= lim) {
      x += ox;
      y += oy;
    }
  }
  const int hnext = __shfl_down(h, 1, 64);
  *last = lane == 63 || hnext != h;
  *ux = x;
  *uy = y;
}

// optimizer_mode 1 with the head updates summed before they reach memory.
// The epoch reads its start layout, so an edge's positive samples are all the
// same update (npos times the fixed-point value); the head's share of an edge
// (positives and negatives) is summed in a register, and the lanes of a wave
// holding consecutive edges of one head (COO heads in row order) add theirs
// by a segmented scan, one atomic pair per head run.  The tail's share goes
// straight to memory.  Integer sums: the same delta as one atomic per sample.
// SYM: every edge (h, t) has its reverse (t, h) with the same period (the
// symmetric fuzzy graph); the reverse edge's positive update of t is exactly
// this edge's tail update (the same d2, the difference negated, a symmetric
// clip and rounding), so each edge adds twice its head update and no tail's.
// (blk, nblk: this block's index among the nblk blocks that share the edges)
template <bool SYM>
__device__ __forceinline__ void epoch_grouped_body(unsigned long long* delta,
                                                   const float* __restrict__ layout,
                                                   const int* __restrict__ heads,
                                                   const int* __restrict__ tails,
                                                   const float* __restrict__ period, int n_edges,
                                                   const UmapParams& p, unsigned epoch, int blk,
                                                   int nblk) {
  // per wave: the sums of the negative samples of each lane's edge
  __shared__ unsigned long long neg_sum[kUThreads / 64][2][64];
  const float alpha = p.lr * (1.0f - __fdiv_rn(static_cast<float>(epoch),
                                               fmaxf(1.0f, static_cast<float>(p.n_epochs))));
  constexpr float eps = 1.1920928955078125e-7f;
  const int lane = threadIdx.x & 63;
  unsigned long long* nsx = neg_sum[threadIdx.x >> 6][0];
  unsigned long long* nsy = neg_sum[threadIdx.x >> 6][1];
  const GlobalPos pos{layout};
  for (int base = blk * kUThreads + static_cast<int>(threadIdx.x & ~63); base < n_edges;
       base += nblk * kUThreads) {
    const int e = base + lane;
    int h = -1, t = -1, nneg = 0;
    float2 hp = make_float2(0.0f, 0.0f);
    long long hx = 0, hy = 0;
    if (e < n_edges) {
      h = heads[e];
      t = tails[e];
      if (h >= 0 && h < p.n && t >= 0 && t < p.n && h != t) {
        const float per = period[e];
        const int npos = n_positive(per, epoch);
        if (npos > 0) {
          hp = pos(static_cast<size_t>(h));
          const float2 tp = pos(static_cast<size_t>(t));
          const float dx = hp.x - tp.x;
          const float dy = hp.y - tp.y;
          const float d2 = fmaxf(eps, dx * dx + dy * dy);
          const float c = attract(d2, p);
          const float gx = clip4(c * dx) * alpha, gy = clip4(c * dy) * alpha;
          const long long fx = npos * fixed(gx), fy = npos * fixed(gy);
          if constexpr (SYM) {
            hx += 2 * fx;
            hy += 2 * fy;
          } else {
            hx += fx;
            hy += fy;
            const size_t tb = static_cast<size_t>(t) * 2;
            atomicAdd(delta + tb, static_cast<unsigned long long>(-fx));
            atomicAdd(delta + tb + 1, static_cast<unsigned long long>(-fy));
          }
          nneg = n_negative(per, p.nsr, epoch);
        }
      }
    }
    // The wave's negative samples, flattened: sample i of the wave belongs to
    // the lane o whose prefix range [incl(o) - nneg(o), incl(o)) holds it, and
    // the wave's lanes take the samples 64 at a time, whatever lane owns them
    // (only the sampled edges of an epoch have any, so a loop over each
    // lane's own samples would leave most lanes idle).  A sample adds into
    // its owner's LDS sums; integer sums, so the same delta in any order.
    nsx[lane] = 0ull;
    nsy[lane] = 0ull;
    int incl = nneg;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int o = __shfl_up(incl, d, 64);
      if (lane >= d) incl += o;
    }
    // the wave's sample count, in a scalar register: the rounds below are
    // wave-uniform, so every shuffle in them reads lanes that are active
    const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
    for (int r = 0; r < total; r += 64) {
      const int i = r + lane;
      // owner: the number of lanes whose inclusive count is <= i
      int o = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (__shfl(incl, o + step - 1, 64) <= i) o += step;
      o = min(o, 63);
      const int oh = __shfl(h, o, 64), ot = __shfl(t, o, 64);
      const int first = __shfl(incl, o, 64) - __shfl(nneg, o, 64);
      const float ox = __shfl(hp.x, o, 64), oy = __shfl(hp.y, o, 64);
      if (i < total) {
        const unsigned v = negative_vertex(static_cast<unsigned>(p.n), p.nmod, p.seed, epoch,
                                           static_cast<unsigned>(oh), static_cast<unsigned>(ot),
                                           static_cast<unsigned>(i - first));
        if (static_cast<int>(v) != oh && static_cast<int>(v) != ot) {
          const float2 vp = pos(static_cast<size_t>(v));
          const float dx = ox - vp.x;
          const float dy = oy - vp.y;
          const float d2 = fmaxf(eps, dx * dx + dy * dy);
          const float c = repel(d2, p);
          atomicAdd(nsx + o, static_cast<unsigned long long>(fixed(clip4(c * dx) * alpha)));
          atomicAdd(nsy + o, static_cast<unsigned long long>(fixed(clip4(c * dy) * alpha)));
        }
      }
    }
    // the wave's LDS operations complete in order: the sums are whole here
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    hx += static_cast<long long>(nsx[lane]);
    hy += static_cast<long long>(nsy[lane]);
    unsigned long long ux, uy;
    bool last;
    head_run_sums(h, hx, hy, &ux, &uy, &last);
    if (last && h >= 0 && h < p.n && (ux != 0ull || uy != 0ull)) {
      atomicAdd(delta + static_cast<size_t>(h) * 2, ux);
      atomicAdd(delta + static_cast<size_t>(h) * 2 + 1, uy);
    }
  }
}

template <bool SYM>
__global__ __launch_bounds__(kUThreads) void umap_epoch_grouped(unsigned long long* delta,
                                                                const float* __restrict__ layout,
                                                                const int* __restrict__ heads,
                                                                const int* __restrict__ tails,
                                                                const float* __restrict__ period,
                                                                int n_edges, UmapParams p,
                                                                unsigned epoch) {
  epoch_grouped_body<SYM>(delta, layout, heads, tails, period, n_edges, p, epoch,
                          static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
}

// optimizer_mode 0 (hogwild): every read sees the live layout and every
// update is a float atomic on it, as in the reference's in-place kernel; an
// edge's positive samples run one after another in registers (each seeing the
// previous ones), its tail receives their sum in one atomic pair, and the
// head's share of the edges of one head run (positives and negatives) is
// summed across the lanes holding the run, one atomic pair per run.  Which
// concurrent update a read sees is a race, in the reference too.  SYM (the
// symmetric fuzzy graph, where the reverse edge (t, h) moves t as this edge
// would): the head takes both endpoints' share of a positive sample, twice its
// own, and the tail none, so the pair's distance shrinks by the same amount
// per sample and no tail atomics are issued.
template <bool SYM>
__device__ __forceinline__ void epoch_hogwild_body(float* layout, const int* __restrict__ heads,
                                                   const int* __restrict__ tails,
                                                   const float* __restrict__ period, int n_edges,
                                                   const UmapParams& p, unsigned epoch, int blk,
                                                   int nblk) {
  const float alpha = p.lr * (1.0f - __fdiv_rn(static_cast<float>(epoch),
                                               fmaxf(1.0f, static_cast<float>(p.n_epochs))));
  constexpr float eps = 1.1920928955078125e-7f;
  const int lane = threadIdx.x & 63;
  const unsigned long long upto_lane = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  for (int base = blk * kUThreads + static_cast<int>(threadIdx.x & ~63); base < n_edges;
       base += nblk * kUThreads) {
    const int e = base + lane;
    int h = -1;
    float ux = 0.0f, uy = 0.0f;
    if (e < n_edges) {
      h = heads[e];
      const int t = tails[e];
      if (h >= 0 && h < p.n && t >= 0 && t < p.n && h != t) {
        const float per = period[e];
        const int npos = n_positive(per, epoch);
        if (npos > 0) {
          const size_t hb = static_cast<size_t>(h) * 2, tb = static_cast<size_t>(t) * 2;
          float hx = layout[hb], hy = layout[hb + 1];
          float tx = layout[tb], ty = layout[tb + 1];
          float sx = 0.0f, sy = 0.0f;
          for (int s = 0; s < npos; ++s) {
            const float dx = hx - tx, dy = hy - ty;
            const float d2 = fmaxf(eps, dx * dx + dy * dy);
            const float c = attract(d2, p);
            const float gx = clip4(c * dx) * alpha, gy = clip4(c * dy) * alpha;
            if constexpr (SYM) {
              hx += 2.0f * gx;
              hy += 2.0f * gy;
              sx += 2.0f * gx;
              sy += 2.0f * gy;
            } else {
              hx += gx;
              hy += gy;
              tx -= gx;
              ty -= gy;
              sx += gx;
              sy += gy;
            }
          }
          if constexpr (!SYM) {
            atomicAdd(layout + tb, -sx);
            atomicAdd(layout + tb + 1, -sy);
          }
          ux = sx;
          uy = sy;
          const int nneg = n_negative(per, p.nsr, epoch);
          for (int s = 0; s < nneg; ++s) {
            const unsigned v = negative_vertex(static_cast<unsigned>(p.n), p.nmod, p.seed, epoch,
                                               static_cast<unsigned>(h), static_cast<unsigned>(t),
                                               static_cast<unsigned>(s));
            if (static_cast<int>(v) == h || static_cast<int>(v) == t) continue;
            const size_t vb = static_cast<size_t>(v) * 2;
            const float dx = hx - layout[vb];
            const float dy = hy - layout[vb + 1];
            const float d2 = fmaxf(eps, dx * dx + dy * dy);
            const float c = repel(d2, p);
            const float gx = clip4(c * dx) * alpha, gy = clip4(c * dy) * alpha;
            hx += gx;
            hy += gy;
            ux += gx;
            uy += gy;
          }
        }
      }
    }
    // segmented inclusive scan over the runs of equal heads
    const int hprev = __shfl_up(h, 1, 64);
    const bool start = lane == 0 || hprev != h;
    const unsigned long long starts = __ballot(start) & upto_lane;
    const int lim = 63 - __clzll(starts);  // first lane of this lane's run
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const float ox = __shfl_up(ux, d, 64), oy = __shfl_up(uy, d, 64);
      if (lane - d >= lim) {
        ux += ox;
        uy += oy;
      }
    }
    const int hnext = __shfl_down(h, 1, 64);
    const bool last = lane == 63 || hnext != h;
    if (last && h >= 0 && h < p.n && (ux != 0.0f || uy != 0.0f)) {
      atomicAdd(layout + static_cast<size_t>(h) * 2, ux);
      atomicAdd(layout + static_cast<size_t>(h) * 2 + 1, uy);
    }
  }
}

template <bool SYM>
__global__ __launch_bounds__(kUThreads) void umap_epoch_hogwild(float* layout,
                                                                const int* __restrict__ heads,
                                                                const int* __restrict__ tails,
                                                                const float* __restrict__ period,
                                                                int n_edges, UmapParams p,
                                                                unsigned epoch) {
  epoch_hogwild_body<SYM>(layout, heads, tails, period, n_edges, p, epoch,
                          static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
}

// sanitise (:3227-3258); MODE 1 first adds and clears the fixed-point delta
template <int MODE>
__device__ __forceinline__ void apply_body(float* layout, unsigned long long* delta, int i) {
  float v = layout[i];
  if constexpr (MODE == 1) {
    const long long d = static_cast<long long>(delta[i]);
    delta[i] = 0ull;
    v = static_cast<float>(static_cast<double>(v) + static_cast<double>(d) / kDeltaScale);
  }
  if (!isfinite(v)) v = 0.0f;
  layout[i] = fminf(16.0f, fmaxf(-16.0f, v));
}

template <int MODE>
__global__ __launch_bounds__(kUThreads) void umap_apply_kernel(float* layout,
                                                               unsigned long long* delta, int m) {
  const int i = blockIdx.x * kUThreads + threadIdx.x;
  if (i < m) apply_body<MODE>(layout, delta, i);
}

__global__ void umap_schedule_kernel(const int* __restrict__ heads, const int* __restrict__ tails,
                                     const float* __restrict__ period, int n_edges, int n,
                                     unsigned epoch, int nsr, unsigned seed,
                                     unsigned long long nmod, int max_neg, int* n_pos,
                                     int* n_neg, int* neg_ids) {
  const int e = blockIdx.x * kUThreads + threadIdx.x;
  if (e >= n_edges) return;
  const int h = heads[e], t = tails[e];
  const int np = n_positive(period[e], epoch);
  const int nn = np > 0 ? n_negative(period[e], nsr, epoch) : 0;
  n_pos[e] = np;
  n_neg[e] = nn;
  for (int s = 0; s < max_neg; ++s) {
    int v = -1;
    if (s < nn) {
      v = static_cast<int>(negative_vertex(static_cast<unsigned>(n), nmod, seed, epoch,
                                           static_cast<unsigned>(h), static_cast<unsigned>(t),
                                           static_cast<unsigned>(s)));
      if (v == h || v == t) v = -1;
    }
    neg_ids[static_cast<size_t>(e) * max_neg + s] = v;
  }
}

}  // namespace

// symmetric: the COO is a symmetric graph (see umap_epoch_grouped); only
// the resident fuzzy graph passes it
int umap_run(const int* heads, const int* tails, const float* weights,
             const float* epochs_per_sample, const float* init, int n, int n_edges, int n_epochs,
             int nsr, float lr, float a, float b, float gamma, unsigned seed, int mode, float* out,
             bool device_inputs, bool symmetric) {
  clear_error();
  if (!heads || !tails || !weights || !epochs_per_sample || !init || !out)
    return fail("null host pointer");
  if (n < 2 || n_edges < 1 || n_epochs < 1 || nsr < 0 || !(lr > 0.0f) || !(gamma > 0.0f))
    return fail("invalid HIP COO UMAP optimization parameters");
  if (mode != 0 && mode != 1) return fail("optimizer_mode must be 0 (atomic) or 1 (delta)");
  Stream stream;
  if (stream.init()) return 1;
  hipStream_t s = stream.get();
  const size_t m = 2 * static_cast<size_t>(n);
  ArenaPlan plan;
  if (!device_inputs) {
    plan.add<int>(n_edges);
    plan.add<int>(n_edges);
    plan.add<float>(n_edges);
  }
  plan.add<float>(m);
  plan.add<unsigned long long>(mode == 1 ? m : 1);
  Arena arena;
  if (arena.init(plan.bytes, "HIP COO UMAP optimizer allocation")) return 1;
  const int* dh = heads;
  const int* dt = tails;
  const float* dp = epochs_per_sample;
  if (!device_inputs) {
    int* h = arena.take<int>(n_edges);
    int* t = arena.take<int>(n_edges);
    float* pp = arena.take<float>(n_edges);
    FE_CHECK(hipMemcpyAsync(h, heads, n_edges * sizeof(int), hipMemcpyHostToDevice, s));
    FE_CHECK(hipMemcpyAsync(t, tails, n_edges * sizeof(int), hipMemcpyHostToDevice, s));
    FE_CHECK(hipMemcpyAsync(pp, epochs_per_sample, n_edges * sizeof(float), hipMemcpyHostToDevice, s));
    dh = h;
    dt = t;
    dp = pp;
  }
  float* layout = arena.take<float>(m);
  unsigned long long* delta = arena.take<unsigned long long>(mode == 1 ? m : 1);
  FE_CHECK(hipMemcpyAsync(layout, init, m * sizeof(float), hipMemcpyHostToDevice, s));
  FE_CHECK(hipMemsetAsync(delta, 0, (mode == 1 ? m : 1) * sizeof(unsigned long long), s));
  const UmapParams p{n, n_epochs, nsr, seed, lr, a, b, gamma, umap_nmod(n)};
  const int eblocks = static_cast<int>(std::min<long long>((n_edges + kUThreads - 1) / kUThreads, 8192));
  const int rblocks = static_cast<int>((m + kUThreads - 1) / kUThreads);
  auto enqueue = [&](int epoch) -> int {
    if (mode == 0) {
      if (symmetric)
        umap_epoch_hogwild<true><<<eblocks, kUThreads, 0, s>>>(layout, dh, dt, dp, n_edges, p,
                                                               static_cast<unsigned>(epoch));
      else
        umap_epoch_hogwild<false><<<eblocks, kUThreads, 0, s>>>(layout, dh, dt, dp, n_edges, p,
                                                                static_cast<unsigned>(epoch));
      FE_LAUNCH_CHECK("umap_epoch_hogwild");
      umap_apply_kernel<0><<<rblocks, kUThreads, 0, s>>>(layout, delta, static_cast<int>(m));
    } else {
      if (symmetric)
        umap_epoch_grouped<true><<<eblocks, kUThreads, 0, s>>>(delta, layout, dh, dt, dp, n_edges,
                                                               p, static_cast<unsigned>(epoch));
      else
        umap_epoch_grouped<false><<<eblocks, kUThreads, 0, s>>>(delta, layout, dh, dt, dp, n_edges,
                                                                p, static_cast<unsigned>(epoch));
      FE_LAUNCH_CHECK("umap_epoch_grouped");
      umap_apply_kernel<1><<<rblocks, kUThreads, 0, s>>>(layout, delta, static_cast<int>(m));
    }
    FE_LAUNCH_CHECK("umap_apply_kernel");
    return 0;
  };
  // the whole epoch schedule as one graph (epoch numbers differ per node)
  bool done = false;
  const char* env = std::getenv("FASTEMBEDR_HIP_GRAPH_CAPTURE");
  if (!(env && env[0] == '0')) {
    hipGraph_t graph = nullptr;
    bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
    for (int e = 0; ok && e < n_epochs; ++e) ok = enqueue(e) == 0;
    const hipError_t end = hipStreamEndCapture(s, &graph);
    if (ok && end == hipSuccess && graph) {
      GraphExec exec;
      if (exec.instantiate(graph) == 0 && hipGraphLaunch(exec.get(), s) == hipSuccess) {
        FE_CHECK(hipStreamSynchronize(s));
        done = true;
      }
    } else if (graph) {
      (void)hipGraphDestroy(graph);
    }
    (void)hipGetLastError();
    clear_error();
  }
  if (!done)
    for (int e = 0; e < n_epochs; ++e)
      if (enqueue(e)) return 1;
  FE_CHECK(hipMemcpyAsync(out, layout, m * sizeof(float), hipMemcpyDeviceToHost, s));
  FE_CHECK(hipStreamSynchronize(s));
  (void)weights;
  return 0;
}

// ---------------------------------------------------------------------------
// Epochs that allocate nothing: the fixed-point delta of optimizer_mode 1
// lives in a file-scope device array instead of a per-call allocation.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxPoints = 1 << 21;  // 2,097,152 points: 32 MiB per device
constexpr int kMaxDevices = 64;

// Every GPU the library runs on holds its own copy until the process ends.
// Host code must never pass `scratch` itself to a kernel (in host code the
// name is a stand-in, not a device address): it uses the per-device address
// from hipGetSymbolAddress below.
__device__ __align__(256) unsigned long long scratch[2 * static_cast<size_t>(kNoAllocMaxPoints)];

namespace {
std::mutex g_scratch_mutex;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};

__global__ void clear_u64_kernel(unsigned long long* p, int m) {
  const int i = blockIdx.x * kUThreads + threadIdx.x;
  if (i < m) p[i] = 0ull;
}
}  // namespace

// The current device's address of `scratch`, looked up once per device.
static int scratch_address(unsigned long long** out) {
  int dev = 0;
  FE_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail("device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !p)
      return check(rc != hipSuccess ? rc : hipErrorInvalidSymbol,
                   "hipGetSymbolAddress(UMAP scratch)");
    g_scratch_ptr[dev] = static_cast<unsigned long long*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int umap_run_noalloc(const int* heads, const int* tails, const float* epochs_per_sample,
                     float* layout, int n, int n_edges, int n_epochs, int nsr, float lr, float a,
                     float b, float gamma, unsigned seed, int mode, hipStream_t s) {
  clear_error();
  if (!heads || !tails || !epochs_per_sample || !layout) return fail("null device pointer");
  if (n < 2 || n_edges < 1 || n_epochs < 1 || nsr < 0 || !(lr > 0.0f) || !(gamma > 0.0f))
    return fail("invalid HIP COO UMAP optimization parameters");
  if (mode != 0 && mode != 1) return fail("optimizer_mode must be 0 (atomic) or 1 (delta)");
  if (mode == 1 && n > kNoAllocMaxPoints)
    return fail("the allocation-free UMAP entry point takes at most " +
                std::to_string(kNoAllocMaxPoints) + " points in optimizer_mode 1 (got " +
                std::to_string(n) + ")");
  unsigned long long* delta = nullptr;
  if (mode == 1 && scratch_address(&delta)) return 1;
  const int m = 2 * n;
  const int rblocks = (m + kUThreads - 1) / kUThreads;
  if (mode == 1) {
    // the previous call on this device may have stopped part way
    clear_u64_kernel<<<rblocks, kUThreads, 0, s>>>(delta, m);
    FE_LAUNCH_CHECK("clear_u64_kernel");
  }
  const UmapParams p{n, n_epochs, nsr, seed, lr, a, b, gamma, umap_nmod(n)};
  const int eblocks = static_cast<int>(std::min<long long>((n_edges + kUThreads - 1) / kUThreads, 8192));
  for (int e = 0; e < n_epochs; ++e) {
    if (mode == 0) {
      umap_epoch_hogwild<false><<<eblocks, kUThreads, 0, s>>>(layout, heads, tails,
                                                              epochs_per_sample, n_edges, p,
                                                              static_cast<unsigned>(e));
      FE_LAUNCH_CHECK("umap_epoch_hogwild");
      umap_apply_kernel<0><<<rblocks, kUThreads, 0, s>>>(layout, delta, m);
    } else {
      umap_epoch_grouped<false><<<eblocks, kUThreads, 0, s>>>(delta, layout, heads, tails,
                                                              epochs_per_sample, n_edges, p,
                                                              static_cast<unsigned>(e));
      FE_LAUNCH_CHECK("umap_epoch_grouped");
      umap_apply_kernel<1><<<rblocks, kUThreads, 0, s>>>(layout, delta, m);
    }
    FE_LAUNCH_CHECK("umap_apply_kernel");
  }
  return 0;
}


// ---------------------------------------------------------------------------
// Batched UMAP jobs: many resident fuzzy graphs advance together through
// job-table launches (per epoch: one epoch launch per optimizer mode, one
// apply launch per mode, one advance launch), whatever the number of jobs.
// A job's blocks find it by a binary search of the launch's prefix table and
// run the epoch body with job-local block indices; each job reads its own
// epoch counter from its descriptor, so a captured chunk of epochs stays
// valid while jobs sit at different epochs, and a finished job's blocks
// exit at once.  Mode 1 sums are integers, so a job gives the bits of its
// resident graph optimised alone (fastembedr_hip_umap_graph_optimize).
// ---------------------------------------------------------------------------
namespace {

struct UmapJobDesc {
  const int* heads;
  const int* tails;
  const float* period;
  float* layout;
  unsigned long long* delta;
  int n, n_edges, epoch, active, mode;
  UmapParams p;
};

__device__ __forceinline__ int umap_find_entry(const int* __restrict__ first, int m) {
  const int bid = static_cast<int>(blockIdx.x);
  int lo = 0, hi = m - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= bid) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int MODE>
__global__ __launch_bounds__(kUThreads) void umap_batch_epoch(UmapJobDesc* jobs,
                                                              const int* __restrict__ slot,
                                                              const int* __restrict__ first,
                                                              int m) {
  const int e = umap_find_entry(first, m);
  const UmapJobDesc* J = jobs + slot[e];
  if (!J->active) return;
  const int blk = static_cast<int>(blockIdx.x) - first[e], nblk = first[e + 1] - first[e];
  const unsigned epoch = static_cast<unsigned>(J->epoch);
  if constexpr (MODE == 1)
    epoch_grouped_body<true>(J->delta, J->layout, J->heads, J->tails, J->period, J->n_edges,
                             J->p, epoch, blk, nblk);
  else
    epoch_hogwild_body<true>(J->layout, J->heads, J->tails, J->period, J->n_edges, J->p, epoch,
                             blk, nblk);
}

template <int MODE>
__global__ __launch_bounds__(kUThreads) void umap_batch_apply(UmapJobDesc* jobs,
                                                              const int* __restrict__ slot,
                                                              const int* __restrict__ first,
                                                              int m) {
  const int e = umap_find_entry(first, m);
  const UmapJobDesc* J = jobs + slot[e];
  if (!J->active) return;
  const int i = (static_cast<int>(blockIdx.x) - first[e]) * kUThreads + static_cast<int>(threadIdx.x);
  if (i < 2 * J->n) apply_body<MODE>(J->layout, J->delta, i);
}

// Persistent form for small mode 1 jobs: a fixed grid of workgroups pulls
// jobs from a queue (one atomic per job), and the workgroup that takes a job
// runs `epochs` of its epochs start to end with the layout and the delta in
// LDS: the negative samples' gathers and the head sums stay on the CU and no
// launch separates the epochs.  The arithmetic is sym_edge_head's and
// apply_body's, with integer sums, so a job gives the bits of the job-table
// path and of its graph optimised alone.  Every workgroup exits once the
// queue is empty.
constexpr int kPersistThreads = 1024;
constexpr int kPersistMaxPoints = 6144;  // 16 + 8 bytes per point: 144 KB of LDS
constexpr size_t kPersistLds = 24ull * kPersistMaxPoints;

__global__ __launch_bounds__(kPersistThreads) void umap_batch_persistent(
    UmapJobDesc* jobs, const int* __restrict__ queue, int queue_len, unsigned* head, int epochs) {
  extern __shared__ unsigned long long pers_lds[];
  unsigned long long* d = pers_lds;                                          // 2 per point
  float* y = reinterpret_cast<float*>(pers_lds + 2 * kPersistMaxPoints);    // 2 per point
  __shared__ int s_q;
  for (;;) {
    __syncthreads();  // the previous job is written back and s_q read
    if (threadIdx.x == 0) s_q = static_cast<int>(atomicAdd(head, 1u));
    __syncthreads();
    const int q = s_q;
    if (q >= queue_len) return;
    UmapJobDesc* J = jobs + queue[q];
    if (!J->active) continue;
    const int n = J->n, m = 2 * n, n_edges = J->n_edges;
    const UmapParams p = J->p;
    for (int i = static_cast<int>(threadIdx.x); i < m; i += kPersistThreads) {
      y[i] = J->layout[i];
      d[i] = 0ull;
    }
    const int e0 = J->epoch, e1 = min(p.n_epochs, e0 + epochs);
    __syncthreads();
    for (int ep = e0; ep < e1; ++ep) {
      const unsigned epoch = static_cast<unsigned>(ep);
      const float alpha = p.lr * (1.0f - __fdiv_rn(static_cast<float>(epoch),
                                                   fmaxf(1.0f, static_cast<float>(p.n_epochs))));
      for (int base = static_cast<int>(threadIdx.x & ~63); base < n_edges;
           base += kPersistThreads) {
        const int e = base + static_cast<int>(threadIdx.x & 63);
        int h = -1;
        long long hx = 0, hy = 0;
        if (e < n_edges) {
          h = J->heads[e];
          const int t = J->tails[e];
          if (h >= 0 && h < n && t >= 0 && t < n && h != t)
            sym_edge_head(LdsPos{y}, h, t, J->period[e], p, epoch, alpha, &hx, &hy);
        }
        unsigned long long ux, uy;
        bool last;
        head_run_sums(h, hx, hy, &ux, &uy, &last);
        if (last && h >= 0 && h < n && (ux != 0ull || uy != 0ull)) {
          atomicAdd(d + 2 * h, ux);
          atomicAdd(d + 2 * h + 1, uy);
        }
      }
      __syncthreads();
      for (int i = static_cast<int>(threadIdx.x); i < m; i += kPersistThreads) apply_body<1>(y, d, i);
      __syncthreads();
    }
    for (int i = static_cast<int>(threadIdx.x); i < m; i += kPersistThreads) J->layout[i] = y[i];
    if (threadIdx.x == 0) {
      J->epoch = e1;
      if (e1 >= p.n_epochs) J->active = 0;
    }
  }
}

__global__ void umap_queue_reset(unsigned* head) { *head = 0u; }

// compute units of the current device (256 on MI355X)
int compute_units() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
      prop.multiProcessorCount > 0)
    return prop.multiProcessorCount;
  return 256;
}

// one thread per job: the epoch counter, then the job's activity
__global__ void umap_batch_advance(UmapJobDesc* jobs, const int* __restrict__ slot, int m) {
  const int e = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= m) return;
  UmapJobDesc* J = jobs + slot[e];
  if (!J->active) return;
  J->epoch += 1;
  if (J->epoch >= J->p.n_epochs) J->active = 0;
}

}  // namespace

int umap_job_validate(const fastembedr_hip_umap_job& j) {
  if (!j.indices || !j.distances || !j.init) return fail("null UMAP job pointer");
  if (j.n < 2 || j.k_total < 1 || j.n_neighbors < 1 || j.n_neighbors > j.k_total ||
      j.n_epochs < 1 || j.negative_sample_rate < 0 || !(j.learning_rate > 0.0f) ||
      !(j.repulsion_strength > 0.0f))
    return fail("invalid HIP UMAP job parameters");
  if (j.optimizer_mode != 0 && j.optimizer_mode != 1)
    return fail("optimizer_mode must be 0 (atomic) or 1 (delta)");
  for (size_t t = 0; t < 2 * static_cast<size_t>(j.n); ++t)
    if (!std::isfinite(j.init[t])) return fail("UMAP init must contain only finite coordinates.");
  return 0;
}

// Device bytes a job of this shape holds while resident: the graph at its
// worst case of 2 n k entries (COO heads, tails, weights, periods and the
// CSR), the layout and mode 1's delta; setup adds the graph build's peak.
static long long umap_job_bytes(int n, int k, int mode, bool setup) {
  const long long nk2 = 2ll * n * k;
  long long b = 20ll * nk2 + 4ll * (n + 1) + 8ll * n + (mode == 1 ? 16ll * n : 0) + 4096;
  if (setup) b += 24ll * nk2 + 32ll * n;
  return b;
}

class UmapBatch {
 public:
  struct Slot {
    fastembedr_hip_umap_graph* graph = nullptr;
    Arena mem;
    long long bytes = 0;
    int n = 0;
  };
  int max_jobs = 0;
  long long budget = 0, used = 0;
  Stream stream;
  std::vector<std::unique_ptr<Slot>> slots;
  std::vector<UmapJobDesc> desc;
  Arena dev;
  UmapJobDesc* d_desc = nullptr;
  int* d_tab = nullptr;
  int tab_cap = 0;
  bool dirty = true;
  struct Launch {
    int kind;  // 0 epoch, 1 apply, 2 advance
    int mode;
    int off, m, blocks;
  };
  std::vector<Launch> launches;
  std::unique_ptr<GraphExec> graph;
  // FASTEMBEDR_HIP_UMAP_PERSISTENT=1 runs the small mode 1 jobs in the
  // persistent kernel, on `side` beside the launches of the other jobs.  Off
  // by default: one workgroup per job is compute-bound on its CU, and on
  // MI355X the captured job-table launches were faster even when the small
  // jobs fill the GPU (512 jobs of 1k-6k points, 200 epochs: 0.50 s against
  // 0.59 s, the same bits); =auto uses it when there are at least as many
  // small jobs as compute units.
  int persistent_mode = 0;  // 0 never, 1 always, -1 when the small jobs fill the GPU
  bool use_persistent = false;
  int pers_off = 0, pers_len = 0;
  unsigned* d_head = nullptr;
  Stream side;
  hipEvent_t fork = nullptr, join = nullptr;

  static bool small_job(const UmapJobDesc& d) { return d.mode == 1 && d.n <= kPersistMaxPoints; }
  bool is_persistent(int q) const { return use_persistent && small_job(desc[q]); }

  ~UmapBatch() {
    if (stream.get()) (void)hipStreamSynchronize(stream.get());
    if (side.get()) (void)hipStreamSynchronize(side.get());
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    for (auto& sl : slots)
      if (sl && sl->graph) fastembedr_hip_umap_graph_destroy(sl->graph);
  }

  int init(int max_jobs_, long long budget_) {
    if (max_jobs_ < 1) return fail("max_jobs must be positive");
    max_jobs = max_jobs_;
    budget = budget_;
    if (stream.init()) return 1;
    slots.resize(max_jobs);
    desc.resize(max_jobs);
    std::memset(desc.data(), 0, desc.size() * sizeof(UmapJobDesc));
    tab_cap = 7 * (2 * max_jobs + 1) + 16;
    const char* env = std::getenv("FASTEMBEDR_HIP_UMAP_PERSISTENT");
    if (env && env[0] == '1') persistent_mode = 1;
    if (env && std::strcmp(env, "auto") == 0) persistent_mode = -1;
    if (side.init()) return 1;
    FE_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    FE_CHECK(hipEventCreateWithFlags(&join, hipEventDisableTiming));
    ArenaPlan p;
    p.add<UmapJobDesc>(max_jobs);
    p.add<int>(tab_cap);
    p.add<unsigned>(1);
    if (dev.init(p.bytes, "UMAP batch tables")) return 1;
    d_desc = dev.take<UmapJobDesc>(max_jobs);
    d_tab = dev.take<int>(tab_cap);
    d_head = dev.take<unsigned>(1);
    FE_CHECK(hipMemsetAsync(d_desc, 0, max_jobs * sizeof(UmapJobDesc), stream.get()));
    return check(hipStreamSynchronize(stream.get()), "UMAP batch init");
  }

  // 0 admitted, 1 error, 2 not admitted now (budget or no free slot)
  int add(const fastembedr_hip_umap_job& j, bool device_inputs, int* slot_out) {
    if (umap_job_validate(j)) return 1;
    int slot = -1;
    for (int q = 0; q < max_jobs; ++q)
      if (!slots[q]) {
        slot = q;
        break;
      }
    if (slot < 0) {
      set_error("no free batch slot");
      return 2;
    }
    if (budget > 0 &&
        used + umap_job_bytes(j.n, j.n_neighbors, j.optimizer_mode, true) > budget) {
      set_error("HBM budget does not admit this job yet");
      return 2;
    }
    auto sl = std::make_unique<Slot>();
    if (fastembedr_hip_umap_graph_create(j.indices, j.distances, j.n, j.k_total, 0, j.n_neighbors,
                                         j.n_neighbors, j.index_offset, device_inputs ? 1 : 0,
                                         &sl->graph))
      return 1;
    long long nnz = 0;
    float max_w = 0.0f;
    const int* heads = nullptr;
    const int* tails = nullptr;
    const float* weights = nullptr;
    const float* period = nullptr;
    if (fastembedr_hip_umap_graph_info(sl->graph, &nnz, &max_w) ||
        fastembedr_hip_umap_graph_device(sl->graph, &heads, &tails, &weights, &period)) {
      fastembedr_hip_umap_graph_destroy(sl->graph);
      return 1;
    }
    if (nnz < 1 || nnz > 2147483647ll) {
      fastembedr_hip_umap_graph_destroy(sl->graph);
      return fail("the UMAP graph has no edges or too many for 32-bit edge ids");
    }
    const size_t m = 2 * static_cast<size_t>(j.n);
    ArenaPlan plan;
    plan.add<float>(m);
    plan.add<unsigned long long>(j.optimizer_mode == 1 ? m : 1);
    hipStream_t s = stream.get();
    if (sl->mem.init(plan.bytes, "UMAP batch layout")) {
      fastembedr_hip_umap_graph_destroy(sl->graph);
      return 1;
    }
    float* layout = sl->mem.take<float>(m);
    unsigned long long* delta = sl->mem.take<unsigned long long>(j.optimizer_mode == 1 ? m : 1);
    const int mm = j.optimizer_mode == 1 ? static_cast<int>(m) : 1;
    FE_CHECK(hipMemcpyAsync(layout, j.init, m * sizeof(float), hipMemcpyHostToDevice, s));
    FE_CHECK(hipMemsetAsync(delta, 0, mm * sizeof(unsigned long long), s));
    UmapJobDesc d;
    std::memset(&d, 0, sizeof(d));
    d.heads = heads;
    d.tails = tails;
    d.period = period;
    d.layout = layout;
    d.delta = delta;
    d.n = j.n;
    d.n_edges = static_cast<int>(nnz);
    d.epoch = 0;
    d.active = 1;
    d.mode = j.optimizer_mode;
    d.p = UmapParams{j.n, j.n_epochs, j.negative_sample_rate, j.seed, j.learning_rate, j.a, j.b,
                     j.repulsion_strength, umap_nmod(j.n)};
    desc[slot] = d;
    FE_CHECK(hipMemcpyAsync(d_desc + slot, &desc[slot], sizeof(UmapJobDesc),
                            hipMemcpyHostToDevice, s));
    FE_CHECK(hipStreamSynchronize(s));
    sl->n = j.n;
    sl->bytes = 20ll * nnz + 4ll * (j.n + 1) + static_cast<long long>(sl->mem.capacity());
    used += sl->bytes;
    slots[slot] = std::move(sl);
    dirty = true;
    *slot_out = slot;
    return 0;
  }

  int remove(int slot) {
    if (slot < 0 || slot >= max_jobs || !slots[slot]) return fail("invalid batch slot");
    FE_CHECK(hipStreamSynchronize(stream.get()));
    UmapJobDesc z;
    std::memset(&z, 0, sizeof(z));
    desc[slot] = z;
    FE_CHECK(hipMemcpy(d_desc + slot, &z, sizeof(UmapJobDesc), hipMemcpyHostToDevice));
    used -= slots[slot]->bytes;
    fastembedr_hip_umap_graph_destroy(slots[slot]->graph);
    slots[slot].reset();
    dirty = true;
    return 0;
  }

  int rebuild() {
    launches.clear();
    std::vector<int> tab;
    int small = 0;
    for (int q = 0; q < max_jobs; ++q)
      if (slots[q] && small_job(desc[q])) ++small;
    use_persistent = persistent_mode == 1 || (persistent_mode < 0 && small >= compute_units());
    auto add_launch = [&](int kind, int mode, const std::vector<std::pair<int, int>>& ent) {
      if (ent.empty()) return;
      Launch L{kind, mode, static_cast<int>(tab.size()), static_cast<int>(ent.size()), 0};
      for (auto& e : ent) tab.push_back(e.first);
      int acc = 0;
      for (auto& e : ent) {
        tab.push_back(acc);
        acc += e.second;
      }
      tab.push_back(acc);
      L.blocks = acc;
      launches.push_back(L);
    };
    for (int mode = 0; mode <= 1; ++mode) {
      std::vector<std::pair<int, int>> ent;
      for (int q = 0; q < max_jobs; ++q)
        if (slots[q] && desc[q].mode == mode && !is_persistent(q))
          ent.push_back({q, static_cast<int>(std::min<long long>(
                                (desc[q].n_edges + kUThreads - 1) / kUThreads, 8192))});
      add_launch(0, mode, ent);
    }
    for (int mode = 0; mode <= 1; ++mode) {
      std::vector<std::pair<int, int>> ent;
      for (int q = 0; q < max_jobs; ++q)
        if (slots[q] && desc[q].mode == mode && !is_persistent(q))
          ent.push_back({q, (2 * desc[q].n + kUThreads - 1) / kUThreads});
      add_launch(1, mode, ent);
    }
    std::vector<std::pair<int, int>> all;
    for (int q = 0; q < max_jobs; ++q)
      if (slots[q] && !is_persistent(q)) all.push_back({q, 1});
    add_launch(2, 0, all);
    // the persistent kernel's queue: small jobs first in slot order
    pers_off = static_cast<int>(tab.size());
    pers_len = 0;
    for (int q = 0; q < max_jobs; ++q)
      if (slots[q] && is_persistent(q)) {
        tab.push_back(q);
        ++pers_len;
      }
    if (static_cast<int>(tab.size()) > tab_cap) return fail("UMAP batch table overflow");
    if (!tab.empty())
      FE_CHECK(hipMemcpy(d_tab, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice));
    dirty = false;
    return 0;
  }

  int enqueue_epoch() {
    hipStream_t s = stream.get();
    for (const Launch& L : launches) {
      const int* sl = d_tab + L.off;
      const int* fi = sl + L.m;
      if (L.kind == 0) {
        if (L.mode == 1) umap_batch_epoch<1><<<L.blocks, kUThreads, 0, s>>>(d_desc, sl, fi, L.m);
        else umap_batch_epoch<0><<<L.blocks, kUThreads, 0, s>>>(d_desc, sl, fi, L.m);
      } else if (L.kind == 1) {
        if (L.mode == 1) umap_batch_apply<1><<<L.blocks, kUThreads, 0, s>>>(d_desc, sl, fi, L.m);
        else umap_batch_apply<0><<<L.blocks, kUThreads, 0, s>>>(d_desc, sl, fi, L.m);
      } else {
        umap_batch_advance<<<(L.m + kUThreads - 1) / kUThreads, kUThreads, 0, s>>>(d_desc, sl, L.m);
      }
      FE_LAUNCH_CHECK("UMAP batch kernel");
    }
    return 0;
  }

  int step(int epochs) {
    if (epochs < 0) return fail("epochs must be non-negative");
    if (dirty) {
      graph.reset();
      if (rebuild()) return 1;
    }
    hipStream_t s = stream.get();
    if (pers_len > 0 && epochs > 0) {
      static bool attr_done = false;
      if (!attr_done) {
        FE_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&umap_batch_persistent),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     static_cast<int>(kPersistLds)));
        attr_done = true;
      }
      hipStream_t a = side.get();
      FE_CHECK(hipEventRecord(fork, s));
      FE_CHECK(hipStreamWaitEvent(a, fork, 0));
      umap_queue_reset<<<1, 1, 0, a>>>(d_head);
      FE_LAUNCH_CHECK("umap_queue_reset");
      umap_batch_persistent<<<std::min(pers_len, compute_units()), kPersistThreads, kPersistLds,
                              a>>>(d_desc, d_tab + pers_off, pers_len, d_head, epochs);
      FE_LAUNCH_CHECK("umap_batch_persistent");
      FE_CHECK(hipEventRecord(join, a));
    }
    if (!launches.empty() && enqueue_table(epochs)) return 1;
    // the batch's stream passes only once the persistent jobs are done too
    if (pers_len > 0 && epochs > 0) FE_CHECK(hipStreamWaitEvent(s, join, 0));
    return 0;
  }

  // the job-table launches of `epochs` epochs: captured chunks, then the rest
  int enqueue_table(int epochs) {
    hipStream_t s = stream.get();
    constexpr int kChunk = 25;
    int left = epochs;
    const char* env = std::getenv("FASTEMBEDR_HIP_GRAPH_CAPTURE");
    if (!(env && env[0] == '0') && left >= kChunk) {
      if (!graph) {
        hipGraph_t g = nullptr;
        bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
        for (int e = 0; ok && e < kChunk; ++e) ok = enqueue_epoch() == 0;
        const hipError_t end = hipStreamEndCapture(s, &g);
        if (ok && end == hipSuccess && g) {
          auto exec = std::make_unique<GraphExec>();
          if (exec->instantiate(g) == 0) graph = std::move(exec);
        } else if (g) {
          (void)hipGraphDestroy(g);
        }
        (void)hipGetLastError();
        clear_error();
      }
      if (graph)
        for (; left >= kChunk; left -= kChunk) FE_CHECK(hipGraphLaunch(graph->get(), s));
    }
    for (; left > 0; --left)
      if (enqueue_epoch()) return 1;
    return 0;
  }

  int status(int* epochs, int* active) {
    std::vector<UmapJobDesc> h(max_jobs);
    FE_CHECK(hipMemcpyAsync(h.data(), d_desc, max_jobs * sizeof(UmapJobDesc),
                            hipMemcpyDeviceToHost, stream.get()));
    FE_CHECK(hipStreamSynchronize(stream.get()));
    for (int q = 0; q < max_jobs; ++q) {
      if (epochs) epochs[q] = slots[q] ? h[q].epoch : -1;
      if (active) active[q] = slots[q] ? h[q].active : 0;
    }
    return 0;
  }

  int layout(int slot, float* out) {
    if (slot < 0 || slot >= max_jobs || !slots[slot]) return fail("invalid batch slot");
    if (!out) return fail("null output pointer");
    FE_CHECK(hipMemcpyAsync(out, desc[slot].layout, 2 * static_cast<size_t>(slots[slot]->n) *
                                                         sizeof(float),
                            hipMemcpyDeviceToHost, stream.get()));
    return check(hipStreamSynchronize(stream.get()), "UMAP batch layout");
  }
};
}  // namespace fe

extern "C" {

int fastembedr_hip_umap_optimize_coo(const int* heads, const int* tails, const float* weights,
                                     const float* epochs_per_sample, const float* init, int n,
                                     int n_edges, int n_epochs, int negative_sample_rate,
                                     float learning_rate, float a, float b,
                                     float repulsion_strength, unsigned int seed,
                                     int optimizer_mode, float* out) {
  return fe::umap_run(heads, tails, weights, epochs_per_sample, init, n, n_edges, n_epochs,
                      negative_sample_rate, learning_rate, a, b, repulsion_strength, seed,
                      optimizer_mode, out, false, false);
}

int fastembedr_hip_umap_optimize_coo_device(const int* heads, const int* tails,
                                            const float* weights, const float* epochs_per_sample,
                                            const float* init, int n, int n_edges, int n_epochs,
                                            int negative_sample_rate, float learning_rate, float a,
                                            float b, float repulsion_strength, unsigned int seed,
                                            int optimizer_mode, float* out) {
  return fe::umap_run(heads, tails, weights, epochs_per_sample, init, n, n_edges, n_epochs,
                      negative_sample_rate, learning_rate, a, b, repulsion_strength, seed,
                      optimizer_mode, out, true, false);
}

int fastembedr_hip_umap_optimize_coo_device_noalloc(
    const int* heads, const int* tails, const float* weights, const float* epochs_per_sample,
    float* layout, int n, int n_edges, int n_epochs, int negative_sample_rate,
    float learning_rate, float a, float b, float repulsion_strength, unsigned int seed,
    int optimizer_mode, void* stream) {
  (void)weights;
  return fe::umap_run_noalloc(heads, tails, epochs_per_sample, layout, n, n_edges, n_epochs,
                              negative_sample_rate, learning_rate, a, b, repulsion_strength, seed,
                              optimizer_mode, static_cast<hipStream_t>(stream));
}

int fastembedr_hip_umap_schedule(const int* heads, const int* tails,
                                 const float* epochs_per_sample, int n_edges, int n, int epoch,
                                 int negative_sample_rate, unsigned int seed, int max_neg,
                                 int* n_pos, int* n_neg, int* neg_ids) {
  fe::clear_error();
  if (!heads || !tails || !epochs_per_sample || !n_pos || !n_neg || (max_neg > 0 && !neg_ids))
    return fe::fail("null host pointer");
  if (n < 2 || n_edges < 1 || epoch < 0 || max_neg < 0 || negative_sample_rate < 0)
    return fe::fail("invalid UMAP schedule arguments");
  fe::ArenaPlan plan;
  plan.add<int>(n_edges);
  plan.add<int>(n_edges);
  plan.add<float>(n_edges);
  plan.add<int>(n_edges);
  plan.add<int>(n_edges);
  plan.add<int>(static_cast<size_t>(n_edges) * (max_neg > 0 ? max_neg : 1));
  fe::Arena a;
  if (a.init(plan.bytes, "UMAP schedule")) return 1;
  int* h = a.take<int>(n_edges);
  int* t = a.take<int>(n_edges);
  float* p = a.take<float>(n_edges);
  int* dp = a.take<int>(n_edges);
  int* dn = a.take<int>(n_edges);
  int* di = a.take<int>(static_cast<size_t>(n_edges) * (max_neg > 0 ? max_neg : 1));
  FE_CHECK(hipMemcpy(h, heads, n_edges * sizeof(int), hipMemcpyHostToDevice));
  FE_CHECK(hipMemcpy(t, tails, n_edges * sizeof(int), hipMemcpyHostToDevice));
  FE_CHECK(hipMemcpy(p, epochs_per_sample, n_edges * sizeof(float), hipMemcpyHostToDevice));
  fe::umap_schedule_kernel<<<(n_edges + 255) / 256, 256>>>(h, t, p, n_edges, n,
                                                           static_cast<unsigned>(epoch),
                                                           negative_sample_rate, seed,
                                                           fe::umap_nmod(n), max_neg, dp, dn, di);
  FE_LAUNCH_CHECK("umap_schedule_kernel");
  FE_CHECK(hipDeviceSynchronize());
  FE_CHECK(hipMemcpy(n_pos, dp, n_edges * sizeof(int), hipMemcpyDeviceToHost));
  FE_CHECK(hipMemcpy(n_neg, dn, n_edges * sizeof(int), hipMemcpyDeviceToHost));
  if (max_neg > 0)
    FE_CHECK(hipMemcpy(neg_ids, di, static_cast<size_t>(n_edges) * max_neg * sizeof(int),
                       hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"

struct fastembedr_hip_umap_batch {
  fe::UmapBatch impl;
};

hipStream_t fe::umap_batch_stream(const fastembedr_hip_umap_batch* batch) {
  return batch ? batch->impl.stream.get() : nullptr;
}

void fe::umap_batch_set_budget(fastembedr_hip_umap_batch* batch, long long budget) {
  if (batch) batch->impl.budget = budget;
}

long long fe::umap_batch_used(const fastembedr_hip_umap_batch* batch) {
  return batch ? batch->impl.used : 0;
}

extern "C" {

long long fastembedr_hip_umap_job_bytes(int n, int n_neighbors, int optimizer_mode,
                                        int setup_peak) {
  return fe::umap_job_bytes(n, n_neighbors, optimizer_mode, setup_peak != 0);
}

int fastembedr_hip_umap_batch_create(int max_jobs, long long hbm_budget_bytes,
                                     fastembedr_hip_umap_batch** batch) {
  fe::clear_error();
  if (!batch) return fe::fail("null batch pointer");
  *batch = nullptr;
  auto* b = new (std::nothrow) fastembedr_hip_umap_batch;
  if (!b) return fe::fail("out of host memory");
  if (fe::guarded("UMAP batch create", [&] { return b->impl.init(max_jobs, hbm_budget_bytes); })) {
    delete b;
    return 1;
  }
  *batch = b;
  return 0;
}

int fastembedr_hip_umap_batch_add(fastembedr_hip_umap_batch* batch,
                                  const fastembedr_hip_umap_job* job, int device_inputs,
                                  int* slot) {
  fe::clear_error();
  if (!batch || !job || !slot) return fe::fail("null batch, job or slot pointer");
  return fe::guarded("UMAP batch add",
                     [&] { return batch->impl.add(*job, device_inputs != 0, slot); });
}

int fastembedr_hip_umap_batch_step(fastembedr_hip_umap_batch* batch, int n_epochs) {
  fe::clear_error();
  if (!batch) return fe::fail("null batch pointer");
  return fe::guarded("UMAP batch step", [&] { return batch->impl.step(n_epochs); });
}

int fastembedr_hip_umap_batch_status(fastembedr_hip_umap_batch* batch, int* epochs, int* active) {
  fe::clear_error();
  if (!batch) return fe::fail("null batch pointer");
  return fe::guarded("UMAP batch status", [&] { return batch->impl.status(epochs, active); });
}

int fastembedr_hip_umap_batch_layout(fastembedr_hip_umap_batch* batch, int slot, float* out) {
  fe::clear_error();
  if (!batch) return fe::fail("null batch pointer");
  return batch->impl.layout(slot, out);
}

int fastembedr_hip_umap_batch_remove(fastembedr_hip_umap_batch* batch, int slot) {
  fe::clear_error();
  if (!batch) return fe::fail("null batch pointer");
  return batch->impl.remove(slot);
}

void fastembedr_hip_umap_batch_destroy(fastembedr_hip_umap_batch* batch) { delete batch; }

}  // extern "C"
