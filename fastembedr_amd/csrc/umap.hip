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

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fastembedr_hip.h"
#include "fe_common.hpp"

namespace fe {

namespace {

constexpr int kUThreads = 256;
constexpr double kDeltaScale = 1099511627776.0;  // 2^40

struct UmapParams {
  int n, n_epochs, nsr;
  unsigned seed;
  float lr, a, b, gamma;
};

__device__ __forceinline__ unsigned mix_uint(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// deterministic_vertex (embedding_cuda_kernels.cpp:262-274)
__device__ __forceinline__ unsigned negative_vertex(unsigned n, unsigned seed, unsigned epoch,
                                                    unsigned head, unsigned tail, unsigned s) {
  unsigned x = seed;
  x ^= epoch * 0x9e3779b9u;
  x ^= (head + 1u) * 0x85ebca6bu;
  x ^= (tail + 1u) * 0xc2b2ae35u;
  x ^= (s + 1u) * 0x27d4eb2du;
  return mix_uint(x) % n;
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
template <bool SYM>
__global__ __launch_bounds__(kUThreads) void umap_epoch_grouped(unsigned long long* delta,
                                                                const float* __restrict__ layout,
                                                                const int* __restrict__ heads,
                                                                const int* __restrict__ tails,
                                                                const float* __restrict__ period,
                                                                int n_edges, UmapParams p,
                                                                unsigned epoch) {
  const float alpha = p.lr * (1.0f - __fdiv_rn(static_cast<float>(epoch),
                                               fmaxf(1.0f, static_cast<float>(p.n_epochs))));
  constexpr float eps = 1.1920928955078125e-7f;
  const int lane = threadIdx.x & 63;
  const unsigned long long upto_lane = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  for (int base = blockIdx.x * kUThreads + (threadIdx.x & ~63); base < n_edges;
       base += gridDim.x * kUThreads) {
    const int e = base + lane;
    int h = -1;
    long long hx = 0, hy = 0;
    if (e < n_edges) {
      h = heads[e];
      const int t = tails[e];
      if (h >= 0 && h < p.n && t >= 0 && t < p.n && h != t) {
        const float per = period[e];
        const int npos = n_positive(per, epoch);
        if (npos > 0) {
          const size_t hb = static_cast<size_t>(h) * 2, tb = static_cast<size_t>(t) * 2;
          const float hxv = layout[hb], hyv = layout[hb + 1];
          {
            const float dx = hxv - layout[tb];
            const float dy = hyv - layout[tb + 1];
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
              atomicAdd(delta + tb, static_cast<unsigned long long>(-fx));
              atomicAdd(delta + tb + 1, static_cast<unsigned long long>(-fy));
            }
          }
          const int nneg = n_negative(per, p.nsr, epoch);
          for (int s = 0; s < nneg; ++s) {
            const unsigned v = negative_vertex(static_cast<unsigned>(p.n), p.seed, epoch,
                                               static_cast<unsigned>(h), static_cast<unsigned>(t),
                                               static_cast<unsigned>(s));
            if (static_cast<int>(v) == h || static_cast<int>(v) == t) continue;
            const size_t vb = static_cast<size_t>(v) * 2;
            const float dx = hxv - layout[vb];
            const float dy = hyv - layout[vb + 1];
            const float d2 = fmaxf(eps, dx * dx + dy * dy);
            const float c = repel(d2, p);
            hx += fixed(clip4(c * dx) * alpha);
            hy += fixed(clip4(c * dy) * alpha);
          }
        }
      }
    }
    // segmented inclusive scan over the runs of equal heads
    const int hprev = __shfl_up(h, 1, 64);
    const bool start = lane == 0 || hprev != h;
    const unsigned long long starts = __ballot(start) & upto_lane;
    const int lim = 63 - __clzll(starts);  // first lane of this lane's run
    unsigned long long ux = static_cast<unsigned long long>(hx);
    unsigned long long uy = static_cast<unsigned long long>(hy);
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const unsigned long long ox = shfl_up_u64(ux, d), oy = shfl_up_u64(uy, d);
      if (lane - d >= lim) {
        ux += ox;
        uy += oy;
      }
    }
    const int hnext = __shfl_down(h, 1, 64);
    const bool last = lane == 63 || hnext != h;
    if (last && h >= 0 && h < p.n && (ux != 0ull || uy != 0ull)) {
      atomicAdd(delta + static_cast<size_t>(h) * 2, ux);
      atomicAdd(delta + static_cast<size_t>(h) * 2 + 1, uy);
    }
  }
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
__global__ __launch_bounds__(kUThreads) void umap_epoch_hogwild(float* layout,
                                                                const int* __restrict__ heads,
                                                                const int* __restrict__ tails,
                                                                const float* __restrict__ period,
                                                                int n_edges, UmapParams p,
                                                                unsigned epoch) {
  const float alpha = p.lr * (1.0f - __fdiv_rn(static_cast<float>(epoch),
                                               fmaxf(1.0f, static_cast<float>(p.n_epochs))));
  constexpr float eps = 1.1920928955078125e-7f;
  const int lane = threadIdx.x & 63;
  const unsigned long long upto_lane = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  for (int base = blockIdx.x * kUThreads + (threadIdx.x & ~63); base < n_edges;
       base += gridDim.x * kUThreads) {
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
            const unsigned v = negative_vertex(static_cast<unsigned>(p.n), p.seed, epoch,
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

// sanitise (:3227-3258); MODE 1 first adds and clears the fixed-point delta
template <int MODE>
__global__ __launch_bounds__(kUThreads) void umap_apply_kernel(float* layout,
                                                               unsigned long long* delta, int m) {
  const int i = blockIdx.x * kUThreads + threadIdx.x;
  if (i >= m) return;
  float v = layout[i];
  if constexpr (MODE == 1) {
    const long long d = static_cast<long long>(delta[i]);
    delta[i] = 0ull;
    v = static_cast<float>(static_cast<double>(v) + static_cast<double>(d) / kDeltaScale);
  }
  if (!isfinite(v)) v = 0.0f;
  layout[i] = fminf(16.0f, fmaxf(-16.0f, v));
}

__global__ void umap_schedule_kernel(const int* __restrict__ heads, const int* __restrict__ tails,
                                     const float* __restrict__ period, int n_edges, int n,
                                     unsigned epoch, int nsr, unsigned seed, int max_neg,
                                     int* n_pos, int* n_neg, int* neg_ids) {
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
      v = static_cast<int>(negative_vertex(static_cast<unsigned>(n), seed, epoch,
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
  const UmapParams p{n, n_epochs, nsr, seed, lr, a, b, gamma};
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
  const UmapParams p{n, n_epochs, nsr, seed, lr, a, b, gamma};
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
                                                           negative_sample_rate, seed, max_neg,
                                                           dp, dn, di);
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
