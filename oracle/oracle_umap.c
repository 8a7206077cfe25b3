/*
 * oracle_umap.c -- CPU restatement of the UMAP edge-sampled SGD epoch of the
 * reference GPU path (TEST INFRASTRUCTURE ONLY; the product never links it).
 *
 * Follows src/embedding_cuda_kernels.cpp:
 *   mix_uint / deterministic_vertex       :253-274
 *   clip4                                  :290-292
 *   fast_positive_pow                      :364-383
 *   attractive_coeff / repulsive_coeff     :385-407 (objective 0, UMAP)
 *   positive/negative sample schedules     :3044-3084
 *   embed_epoch_coo_{atomic,delta}_kernel  :3086-3225
 *   sanitize (+-16, non-finite -> 0)       :3227-3258
 *   host epoch loop                        :5565-5599
 * in the same float32 arithmetic.  The epoch is evaluated sequentially:
 * mode 1 ("delta") computes every update from the epoch-start layout, sums
 * them in FP64 and applies them once, which is the deterministic form; mode 0
 * applies each update immediately in edge order, one serialisation of the
 * reference's hogwild atomics.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

unsigned int or_umap_mix_uint(unsigned int x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

int or_umap_negative_vertex(int n, unsigned int seed, int epoch, int head, int tail, int s) {
  unsigned int x = seed;
  x ^= (unsigned int)epoch * 0x9e3779b9u;
  x ^= ((unsigned int)head + 1u) * 0x85ebca6bu;
  x ^= ((unsigned int)tail + 1u) * 0xc2b2ae35u;
  x ^= ((unsigned int)s + 1u) * 0x27d4eb2du;
  return (int)(or_umap_mix_uint(x) % (unsigned int)n);
}

static float clip4(float x) { return fminf(4.0f, fmaxf(-4.0f, x)); }

static float as_float(unsigned int u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static unsigned int as_uint(float f) {
  unsigned int u;
  memcpy(&u, &f, 4);
  return u;
}

float or_umap_fast_pow(float x, float b) {
  if (x <= 0.0f) return 0.0f;
  const unsigned int xb = as_uint(x);
  const float bias = 1064866805.0f;
  const int whole = (int)b;
  const float frac = b - (float)whole;
  const unsigned int ib = (unsigned int)(frac * ((float)xb - bias) + bias);
  const float fp = as_float(ib);
  float ip = 1.0f, base = x;
  int e = whole;
  while (e > 0) {
    if (e & 1) ip *= base;
    base *= base;
    e >>= 1;
  }
  return ip * fp;
}

static float attractive_coeff(float d2, float a, float b) {
  if (d2 <= 0.0f) return 0.0f;
  const float d2b = or_umap_fast_pow(d2, b);
  return -2.0f * a * b * (d2b / d2) / (a * d2b + 1.0f);
}

static float repulsive_coeff(float d2, float a, float b, float gamma) {
  if (d2 <= 0.0f) return 0.0f;
  const float d2b = or_umap_fast_pow(d2, b);
  return gamma * 2.0f * b / ((0.001f + d2) * (a * d2b + 1.0f));
}

int or_umap_n_positive(int epoch, float period) {
  if (period <= 0.0f || !isfinite(period) || epoch == 0) return 0;
  const float now = (float)epoch, prev = (float)(epoch - 1);
  const int cur = (int)floorf(now / period), pre = (int)floorf(prev / period);
  const int s = cur - pre;
  return s > 0 ? s : 0;
}

int or_umap_n_negative(int epoch, float period, int nsr) {
  if (nsr <= 0 || period <= 0.0f || !isfinite(period) || epoch == 0) return 0;
  const float now = (float)epoch, prev = (float)(epoch - 1);
  const int cur = (int)floorf(now / period), pre = (int)floorf(prev / period);
  if (cur <= pre) return 0;
  const float np = period / (float)nsr;
  if (np <= 0.0f || !isfinite(np)) return 0;
  const int cur_total = (int)floorf((now - np) / np);
  int prev_total = 0;
  if (cur > 1) {
    const float pa = ceilf((float)(cur - 1) * period);
    prev_total = (int)floorf((pa - np) / np);
  }
  const int s = cur_total - prev_total;
  return s > 0 ? s : 0;
}

/* One epoch over the COO edges; y is n x 2 float (in place). */
void or_umap_epoch(const int* heads, const int* tails, const float* period, int n_edges,
                   double* delta_scratch, int n, int epoch, int n_epochs, int nsr, double lr,
                   double a, double b, double gamma, unsigned int seed, int mode, float* y) {
  const float alpha = (float)lr * (1.0f - (float)epoch / fmaxf(1.0f, (float)n_epochs));
  const float eps = 1.1920928955078125e-7f;
  const float fa = (float)a, fb = (float)b, fg = (float)gamma;
  double* delta = delta_scratch;
  if (mode == 1) memset(delta, 0, sizeof(double) * 2 * (size_t)n);
  for (int e = 0; e < n_edges; ++e) {
    const int h = heads[e], t = tails[e];
    if (h < 0 || h >= n || t < 0 || t >= n || h == t) continue;
    const int npos = or_umap_n_positive(epoch, period[e]);
    if (npos <= 0) continue;
    for (int s = 0; s < npos; ++s) {
      const float dx = y[2 * h] - y[2 * t], dy = y[2 * h + 1] - y[2 * t + 1];
      const float d2 = fmaxf(eps, dx * dx + dy * dy);
      const float c = attractive_coeff(d2, fa, fb);
      const float gx = clip4(c * dx) * alpha, gy = clip4(c * dy) * alpha;
      if (mode == 1) {
        delta[2 * h] += gx; delta[2 * h + 1] += gy;
        delta[2 * t] -= gx; delta[2 * t + 1] -= gy;
      } else {
        y[2 * h] += gx; y[2 * h + 1] += gy;
        y[2 * t] -= gx; y[2 * t + 1] -= gy;
      }
    }
    const int nneg = or_umap_n_negative(epoch, period[e], nsr);
    for (int s = 0; s < nneg; ++s) {
      const int v = or_umap_negative_vertex(n, seed, epoch, h, t, s);
      if (v == h || v == t) continue;
      const float dx = y[2 * h] - y[2 * v], dy = y[2 * h + 1] - y[2 * v + 1];
      const float d2 = fmaxf(eps, dx * dx + dy * dy);
      const float c = repulsive_coeff(d2, fa, fb, fg);
      const float gx = clip4(c * dx) * alpha, gy = clip4(c * dy) * alpha;
      if (mode == 1) {
        delta[2 * h] += gx; delta[2 * h + 1] += gy;
      } else {
        y[2 * h] += gx; y[2 * h + 1] += gy;
      }
    }
  }
  for (int i = 0; i < 2 * n; ++i) {
    float v = mode == 1 ? (float)((double)y[i] + delta[i]) : y[i];
    if (!isfinite(v)) v = 0.0f;
    y[i] = fminf(16.0f, fmaxf(-16.0f, v));
  }
}

/* The full epoch loop of fastembedr_cuda_umap_optimize_coo (:5565-5599). */
void or_umap_optimize(const int* heads, const int* tails, const float* period, int n_edges,
                      int n, int n_epochs, int nsr, double lr, double a, double b, double gamma,
                      unsigned int seed, int mode, float* y) {
  double* delta = (double*)malloc(sizeof(double) * 2 * (size_t)n);
  for (int e = 0; e < n_epochs; ++e)
    or_umap_epoch(heads, tails, period, n_edges, delta, n, e, n_epochs, nsr, lr, a, b, gamma, seed,
                  mode, y);
  free(delta);
}

/* Per-edge schedule of one epoch: n_pos, n_neg and the first max_neg
 * negative vertex ids (-1 where skipped because it equals head or tail). */
void or_umap_schedule(const int* heads, const int* tails, const float* period, int n_edges, int n,
                      int epoch, int nsr, unsigned int seed, int max_neg, int* n_pos, int* n_neg,
                      int* neg_ids) {
  for (int e = 0; e < n_edges; ++e) {
    const int h = heads[e], t = tails[e];
    n_pos[e] = or_umap_n_positive(epoch, period[e]);
    n_neg[e] = n_pos[e] > 0 ? or_umap_n_negative(epoch, period[e], nsr) : 0;
    for (int s = 0; s < max_neg; ++s) {
      int v = -1;
      if (s < n_neg[e]) {
        v = or_umap_negative_vertex(n, seed, epoch, h, t, s);
        if (v == h || v == t) v = -1;
      }
      neg_ids[(size_t)e * max_neg + s] = v;
    }
  }
}
