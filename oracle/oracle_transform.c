/*
 * oracle_transform.c -- FP64 CPU restatement of the fixed-reference t-SNE
 * transform (TEST INFRASTRUCTURE ONLY; the product never links it).
 *
 * Follows src/tsne_neighbors.cpp:
 *   mix_uint32 / xorshift32 / uniform_index          :67-88
 *   compute_row_probabilities_flat                   :452-530 (or_row_probabilities)
 *   initialize_tsne_transform (median / weighted)    :1384-1452
 *   compute_tsne_transform_gradient_2d_flat          :1552-1648
 *   transform_tsne_cpp iteration loop                :2258-2349
 * for two output dimensions.  The "random" initialisation (std::mt19937 +
 * std::normal_distribution) is not restated here; tests pass it explicitly.
 * Layouts are row-major n x 2; indices row-major n_query x k, 0-based.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void or_row_probabilities(const double* dist, int k, double perplexity, double* row_p);

uint32_t or_mix_uint32(uint32_t v) {
  v ^= v >> 16u;
  v *= 0x7feb352du;
  v ^= v >> 15u;
  v *= 0x846ca68bu;
  v ^= v >> 16u;
  return v == 0u ? 0x6d2b79f5u : v;
}

static uint32_t xorshift32(uint32_t* s) {
  *s ^= *s << 13u;
  *s ^= *s >> 17u;
  *s ^= *s << 5u;
  return *s;
}

int or_uniform_index(uint32_t* s, int n) {
  return (int)(((uint64_t)xorshift32(s) * (uint64_t)n) >> 32u);
}

static int cmp_double(const void* a, const void* b) {
  const double x = *(const double*)a, y = *(const double*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/*
 * Returns 0 on success.  init: NULL -> initialization (0 median, 1 weighted);
 * otherwise the explicit n_query x 2 start.  out: n_query x 2.
 */
int or_transform(int n_ref, const double* ref, int n_query, int k, const int* idx,
                 const double* dist, const double* init, int initialization, double perplexity,
                 int n_iter, int early_iter, double lr, double early_exag, double exag,
                 double mom0, double mom1, double max_grad, double max_step, int n_negatives,
                 int threshold, int seed, double* out) {
  if (n_ref < 1 || n_query < 1 || k < 1 || perplexity <= 0.0 || n_iter + early_iter < 1) return 5;
  if (threshold < 1) threshold = 1;
  if (max_grad <= 0.0 || !isfinite(max_grad)) max_grad = DBL_MAX;
  if (max_step <= 0.0 || !isfinite(max_step)) max_step = DBL_MAX;
  if (n_negatives > n_ref) n_negatives = n_ref;
  for (size_t e = 0; e < (size_t)n_query * k; ++e) {
    if (idx[e] < 0 || idx[e] >= n_ref) return 1;
    if (!isfinite(dist[e]) || dist[e] < 0.0) return 2;
  }
  const int exact = n_ref <= threshold || n_negatives >= n_ref;
  double* p = (double*)malloc(sizeof(double) * (size_t)n_query * k);
  double* qv = (double*)malloc(sizeof(double) * (size_t)(exact ? n_ref : n_negatives));
  int* sampled = (int*)malloc(sizeof(int) * (size_t)(n_negatives > 0 ? n_negatives : 1));
  double* vals = (double*)malloc(sizeof(double) * (size_t)k);
  for (int i = 0; i < n_query; ++i) or_row_probabilities(dist + (size_t)i * k, k, perplexity, p + (size_t)i * k);
  for (int i = 0; i < n_query; ++i) {
    double y0, y1;
    if (init) {
      y0 = init[2 * i];
      y1 = init[2 * i + 1];
    } else if (initialization == 1) {
      double num[2] = {0, 0}, den = DBL_MIN;
      for (int j = 0; j < k; ++j) {
        const int r = idx[(size_t)i * k + j];
        const double w = 1.0 / (fmax(0.0, dist[(size_t)i * k + j]) + 1e-6);
        num[0] += w * ref[2 * r];
        num[1] += w * ref[2 * r + 1];
        den += w;
      }
      /* the reference divides per dimension with its own running
       * denominator; both dimensions accumulate the same weights */
      y0 = num[0] / den;
      y1 = num[1] / den;
    } else {
      double med[2];
      for (int d = 0; d < 2; ++d) {
        for (int j = 0; j < k; ++j) vals[j] = ref[2 * idx[(size_t)i * k + j] + d];
        qsort(vals, (size_t)k, sizeof(double), cmp_double);
        const int mid = k / 2;
        med[d] = (k & 1) ? vals[mid] : 0.5 * (vals[mid] + vals[mid - 1]);
      }
      y0 = med[0];
      y1 = med[1];
    }
    double u0 = 0, u1 = 0, g0n = 1, g1n = 1;
    for (int it = 0; it < early_iter + n_iter; ++it) {
      const int early = it < early_iter;
      const double ex = early ? early_exag : exag, mom = early ? mom0 : mom1;
      const int iter_seed = seed + 65537 * (it + 1);
      uint32_t st = or_mix_uint32((uint32_t)iter_seed ^ ((uint32_t)(i + 1) * 0x9e3779b9u));
      double g0 = 0.0, g1 = 0.0, sum_q = DBL_MIN;
      if (exact) {
        for (int r = 0; r < n_ref; ++r) {
          const double dx = y0 - ref[2 * r], dy = y1 - ref[2 * r + 1];
          qv[r] = 1.0 / (1.0 + dx * dx + dy * dy);
          sum_q += qv[r];
        }
        const double inv = 1.0 / sum_q;
        for (int r = 0; r < n_ref; ++r) {
          const double c = -(qv[r] * qv[r]) * inv;
          g0 += c * (y0 - ref[2 * r]);
          g1 += c * (y1 - ref[2 * r + 1]);
        }
      } else {
        for (int m = 0; m < n_negatives; ++m) {
          const int r = or_uniform_index(&st, n_ref);
          sampled[m] = r;
          const double dx = y0 - ref[2 * r], dy = y1 - ref[2 * r + 1];
          qv[m] = 1.0 / (1.0 + dx * dx + dy * dy);
          sum_q += qv[m];
        }
        const double inv = 1.0 / sum_q;
        for (int m = 0; m < n_negatives; ++m) {
          const int r = sampled[m];
          const double c = -(qv[m] * qv[m]) * inv;
          g0 += c * (y0 - ref[2 * r]);
          g1 += c * (y1 - ref[2 * r + 1]);
        }
      }
      for (int j = 0; j < k; ++j) {
        const int r = idx[(size_t)i * k + j];
        const double dx = y0 - ref[2 * r], dy = y1 - ref[2 * r + 1];
        const double q = 1.0 / (1.0 + dx * dx + dy * dy);
        const double c = ex * p[(size_t)i * k + j] * q;
        g0 += c * dx;
        g1 += c * dy;
      }
      const double gn2 = g0 * g0 + g1 * g1;
      if (gn2 > max_grad * max_grad) {
        const double s = max_grad / (sqrt(gn2) + 1e-12);
        g0 *= s;
        g1 *= s;
      }
      double* gain[2] = {&g0n, &g1n};
      double* upd[2] = {&u0, &u1};
      const double grad[2] = {g0, g1};
      for (int d = 0; d < 2; ++d) {
        const double su = *upd[d] == 0.0 ? 0.0 : (*upd[d] < 0 ? -1.0 : 1.0);
        const double sg = grad[d] == 0.0 ? 0.0 : (grad[d] < 0 ? -1.0 : 1.0);
        if (su != sg) *gain[d] += 0.2;
        else *gain[d] = *gain[d] * 0.8 + 0.01;
        if (*gain[d] < 0.01) *gain[d] = 0.01;
        *upd[d] = mom * *upd[d] - lr * *gain[d] * grad[d];
      }
      const double sn2 = u0 * u0 + u1 * u1;
      if (sn2 > max_step * max_step) {
        const double s = max_step / (sqrt(sn2) + 1e-12);
        u0 *= s;
        u1 *= s;
      }
      y0 += u0;
      y1 += u1;
    }
    out[2 * i] = y0;
    out[2 * i + 1] = y1;
  }
  free(p);
  free(qv);
  free(sampled);
  free(vals);
  return 0;
}
