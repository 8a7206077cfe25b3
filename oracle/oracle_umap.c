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

/* ---- fuzzy graph (umap_graph_csr_float_cpp, src/fast_knn_umap.cpp:3995-4048,
 * building with build_graph_csr_native_direct_values :1562-1780 at the default
 * edge budget = n_cols) ------------------------------------------------------
 * smooth_knn_dist :876-1003 (one thread: the global mean in long double, row
 * order); directed memberships :1608-1634; per-row sort by neighbour and
 * max-dedupe (compact_directed_rows_inplace :1291-1364); incoming lists in
 * source order :1641-1700; union f + r - f*r with the 1e-6 floor :1702-1768;
 * epochs_per_sample :1264-1282.  indices / distances: column-major n x k_total
 * (R matrices), columns col_start .. col_start+n_cols-1 used.  Returns nnz, or
 * -1 when cap is too small. */
static void or_smooth_knn_rows(const float* dist, int n, int k, float* sigmas, float* rhos) {
  const double target = log2((double)k);
  long double gsum = 0.0L;
  size_t gcount = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < k; ++j) {
      const float d = dist[(size_t)i * k + j];
      if (isfinite(d) && d >= 0.0f) { gsum += (long double)d; ++gcount; }
    }
  const double gmean = gcount > 0 ? (double)(gsum / (long double)gcount) : 1.0;
  for (int i = 0; i < n; ++i) {
    const float* row = dist + (size_t)i * k;
    double rho = INFINITY, rsum = 0.0;
    int rcount = 0;
    for (int j = 0; j < k; ++j) {
      const float d = row[j];
      if (!isfinite(d)) continue;
      if (d >= 0.0f) { rsum += (double)d; ++rcount; }
      if (d > 0.0f && (double)d < rho) rho = (double)d;
    }
    if (!isfinite(rho)) rho = 0.0;
    rhos[i] = (float)rho;
    const double big = 1.79769313486231570815e+308;
    double sigma = 1.0, best = 1.0, best_diff = big, lo = 0.0, hi = big;
    for (int it = 0; it < 64; ++it) {
      const double safe = sigma > 1.0e-12 ? sigma : 1.0e-12;
      double psum = 0.0;
      for (int j = 0; j < k; ++j) {
        if (!isfinite(row[j])) continue;
        const double d = (double)row[j] - rho;
        psum += d <= 0.0 ? 1.0 : exp(-d / safe);
      }
      const double diff = fabs(psum - target);
      if (diff < best_diff) { best_diff = diff; best = sigma; }
      if (psum > target) {
        hi = sigma;
        sigma = 0.5 * (lo + hi);
      } else {
        lo = sigma;
        sigma = hi == big ? sigma * 2.0 : 0.5 * (lo + hi);
      }
      if (diff < 1.0e-5) break;
    }
    const double rmean = rcount > 0 ? rsum / (double)rcount : gmean;
    const double floor_s = 1.0e-3 * (rho > 0.0 ? rmean : gmean);
    if (best < floor_s) best = floor_s;
    sigmas[i] = (float)(best > 1.0e-12 ? best : 1.0e-12);
  }
}

int or_umap_fuzzy_graph(const int* indices, const float* distances, int n, int k_total,
                        int col_start, int n_cols, int index_offset, int* offsets,
                        int* neighbors, float* weights, float* eps, float* max_weight,
                        float* sigmas, float* rhos, long long cap) {
  const int k = n_cols;
  float* dist = (float*)malloc(sizeof(float) * (size_t)n * k);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < k; ++j) dist[(size_t)i * k + j] = distances[(size_t)(j + col_start) * n + i];
  or_smooth_knn_rows(dist, n, k, sigmas, rhos);
  /* directed rows, sorted by neighbour (insertion, stable), equal neighbours
   * merged by max */
  int* dn = (int*)malloc(sizeof(int) * (size_t)n * k);
  float* dw = (float*)malloc(sizeof(float) * (size_t)n * k);
  int* dc = (int*)calloc((size_t)n, sizeof(int));
  int* incnt = (int*)calloc((size_t)n + 1, sizeof(int));
  for (int i = 0; i < n; ++i) {
    int* rn = dn + (size_t)i * k;
    float* rw = dw + (size_t)i * k;
    int c = 0;
    for (int j = 0; j < k; ++j) {
      const int nb = indices[(size_t)(j + col_start) * n + i] - index_offset;
      const float d = dist[(size_t)i * k + j];
      if (nb < 0 || nb >= n || nb == i || !isfinite(d)) continue;
      const float w = d <= rhos[i] ? 1.0f : expf(-(d - rhos[i]) / sigmas[i]);
      if (!isfinite(w) || w <= 0.0f) continue;
      int at = c++;
      while (at > 0 && rn[at - 1] > nb) { rn[at] = rn[at - 1]; rw[at] = rw[at - 1]; --at; }
      rn[at] = nb;
      rw[at] = w;
    }
    int u = 0;
    for (int r = 0; r < c;) {
      const int nb = rn[r];
      float w = rw[r++];
      while (r < c && rn[r] == nb) { if (rw[r] > w) w = rw[r]; ++r; }
      rn[u] = nb;
      rw[u] = w;
      ++u;
    }
    dc[i] = u;
    for (int r = 0; r < u; ++r) incnt[rn[r] + 1]++;
  }
  for (int i = 0; i < n; ++i) incnt[i + 1] += incnt[i];
  long long tot_in = incnt[n];
  int* in_src = (int*)malloc(sizeof(int) * (size_t)(tot_in > 0 ? tot_in : 1));
  float* in_w = (float*)malloc(sizeof(float) * (size_t)(tot_in > 0 ? tot_in : 1));
  int* cur = (int*)malloc(sizeof(int) * (size_t)n);
  memcpy(cur, incnt, sizeof(int) * (size_t)n);
  for (int i = 0; i < n; ++i) /* sources in ascending order */
    for (int r = 0; r < dc[i]; ++r) {
      const int t = dn[(size_t)i * k + r];
      in_src[cur[t]] = i;
      in_w[cur[t]] = dw[(size_t)i * k + r];
      cur[t]++;
    }
  long long nnz = 0;
  offsets[0] = 0;
  for (int i = 0; i < n; ++i) {
    int o = 0, p = incnt[i];
    const int oe = dc[i], pe = incnt[i + 1];
    const int* rn = dn + (size_t)i * k;
    const float* rw = dw + (size_t)i * k;
    while (o < oe || p < pe) {
      const int a = o < oe ? rn[o] : 2147483647, b = p < pe ? in_src[p] : 2147483647;
      int nb;
      float f = 0.0f, r = 0.0f;
      if (a == b) { nb = a; f = rw[o++]; r = in_w[p++]; }
      else if (a < b) { nb = a; f = rw[o++]; }
      else { nb = b; r = in_w[p++]; }
      if (nb == i) continue;
      const float s = f + r, pr = f * r;
      const float w = s - pr;
      if (!isfinite(w) || w <= 1.0e-6f) continue;
      if (nnz >= cap) { nnz = -1; goto done; }
      neighbors[nnz] = nb;
      weights[nnz] = w;
      ++nnz;
    }
    offsets[i + 1] = (int)nnz;
  }
  {
    float mw = 0.0f;
    for (long long e = 0; e < nnz; ++e) if (isfinite(weights[e]) && weights[e] > mw) mw = weights[e];
    *max_weight = mw;
    for (long long e = 0; e < nnz; ++e) {
      const float w = weights[e] > 1.0e-6f ? weights[e] : 1.0e-6f;
      eps[e] = mw > 0.0f ? mw / w : INFINITY;
    }
  }
done:
  free(dist); free(dn); free(dw); free(dc); free(incnt); free(in_src); free(in_w); free(cur);
  return (int)nnz;
}
