/*
 * oracle.c -- FP64 CPU restatement of the fastEmbedR embedding-optimizer
 * path (t-SNE affinities, FFT-grid repulsion, sparse attraction, optimizer
 * update; UMAP epoch schedule; fixed-reference t-SNE transform).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * test suite, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py.  The product library never links, loads or calls it.
 *
 * Every function cites the reference function it restates (paths relative
 * to the fastEmbedR source tree).  Arithmetic is double precision
 * throughout, so the float32 HIP kernels can be compared against it with
 * tolerances.  Arrays are row-major; neighbour indices are 0-based.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_OK 0
#define OR_ERR_RANGE 1
#define OR_ERR_DIST 2
#define OR_ERR_EMPTY 3
#define OR_ERR_NORM 4
#define OR_ERR_ARG 5
#define OR_ERR_ALLOC 6

/* ------------------------------------------------------------------------ */
/* Perplexity calibration: src/tsne_neighbors.cpp:452-530                     */
/* (compute_row_probabilities_flat, FP64 bisection on beta with a min-d^2     */
/* shift, tolerance 1e-5, 200 iterations, tied-row fallback).                 */
/* ------------------------------------------------------------------------ */
void or_row_probabilities(const double* dist, int k, double perplexity,
                          double* row_p) {
  double min_d2 = DBL_MAX, max_d2 = 0.0;
  for (int j = 0; j < k; ++j) {
    const double d2 = dist[j] * dist[j];
    if (d2 < min_d2) min_d2 = d2;
    if (d2 > max_d2) max_d2 = d2;
    row_p[j] = 0.0;
  }
  const double spread = max_d2 - min_d2;
  if (spread <= DBL_EPSILON * fmax(1.0, max_d2)) {
    for (int j = 0; j < k; ++j) row_p[j] = 1.0 / (double)k;
    return;
  }
  int found = 0;
  double beta = 1.0, min_beta = -DBL_MAX, max_beta = DBL_MAX;
  double sum_p = DBL_MIN;
  const double target = log(perplexity);
  for (int iter = 0; !found && iter < 200; ++iter) {
    sum_p = DBL_MIN;
    for (int j = 0; j < k; ++j) {
      const double d2 = dist[j] * dist[j] - min_d2;
      const double p = exp(-beta * d2);
      row_p[j] = p;
      sum_p += p;
    }
    double entropy = 0.0;
    for (int j = 0; j < k; ++j) {
      entropy += beta * ((dist[j] * dist[j] - min_d2) * row_p[j]);
    }
    entropy = entropy / sum_p + log(sum_p);
    const double diff = entropy - target;
    if (fabs(diff) < 1e-5) {
      found = 1;
    } else if (diff > 0.0) {
      min_beta = beta;
      beta = (max_beta == DBL_MAX || max_beta == -DBL_MAX) ? beta * 2.0
                                                           : (beta + max_beta) / 2.0;
    } else {
      max_beta = beta;
      beta = (min_beta == -DBL_MAX || min_beta == DBL_MAX) ? beta / 2.0
                                                           : (beta + min_beta) / 2.0;
    }
    if (!isfinite(beta)) break;
  }
  if (!isfinite(sum_p) || sum_p <= DBL_MIN) {
    int tied = 0;
    for (int j = 0; j < k; ++j) {
      const double d2 = dist[j] * dist[j];
      if (fabs(d2 - min_d2) <= DBL_EPSILON * fmax(1.0, min_d2)) ++tied;
    }
    const double mass = 1.0 / (double)(tied > 1 ? tied : 1);
    for (int j = 0; j < k; ++j) {
      const double d2 = dist[j] * dist[j];
      row_p[j] = fabs(d2 - min_d2) <= DBL_EPSILON * fmax(1.0, min_d2) ? mass : 0.0;
    }
    return;
  }
  const double inv = 1.0 / sum_p;
  for (int j = 0; j < k; ++j) row_p[j] *= inv;
}

/* ------------------------------------------------------------------------ */
/* Symmetrisation into CSR: src/tsne_neighbors.cpp:616-724                   */
/* (build_tsne_probabilities_float).  Directed rows p_{j|i} are keyed by the  */
/* unordered pair, duplicates are summed, P_ij = sum * 0.5 / total_mass, and  */
/* both directions are stored.  Within a row the columns come out ascending,  */
/* exactly as the reference's key-sorted fill produces them.                  */
/* Capacity of col/val must be >= 2*n*k.  Returns nnz or a negative error.    */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint64_t key;
  double value;
} or_edge;

static int or_edge_cmp(const void* a, const void* b) {
  const uint64_t ka = ((const or_edge*)a)->key, kb = ((const or_edge*)b)->key;
  return ka < kb ? -1 : (ka > kb ? 1 : 0);
}

long or_symmetrize(int n, int k, const int* idx, const double* p_dir,
                   int* row_ptr, int* col, double* val) {
  const size_t total = (size_t)n * (size_t)k;
  or_edge* edges = (or_edge*)malloc(sizeof(or_edge) * (total ? total : 1));
  if (!edges) return -OR_ERR_ALLOC;
  size_t m = 0;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < k; ++j) {
      const int nb = idx[(size_t)i * k + j];
      if (nb == i) continue;
      const uint32_t a = (uint32_t)(i < nb ? i : nb), b = (uint32_t)(i < nb ? nb : i);
      edges[m].key = ((uint64_t)a << 32) | b;
      edges[m].value = p_dir[(size_t)i * k + j];
      ++m;
    }
  }
  if (m == 0) {
    free(edges);
    return -OR_ERR_EMPTY;
  }
  qsort(edges, m, sizeof(or_edge), or_edge_cmp);
  for (int i = 0; i <= n; ++i) row_ptr[i] = 0;
  size_t w = 0;
  double total_mass = 0.0;
  for (size_t r = 0; r < m;) {
    const uint64_t key = edges[r].key;
    double sum = 0.0;
    while (r < m && edges[r].key == key) sum += edges[r++].value;
    edges[w].key = key;
    edges[w].value = sum;
    ++w;
    total_mass += sum;
    row_ptr[(key >> 32) + 1]++;
    row_ptr[(key & 0xffffffffu) + 1]++;
  }
  if (!isfinite(total_mass) || total_mass <= 0.0) {
    free(edges);
    return -OR_ERR_NORM;
  }
  for (int i = 0; i < n; ++i) row_ptr[i + 1] += row_ptr[i];
  int* fill = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  if (!fill) {
    free(edges);
    return -OR_ERR_ALLOC;
  }
  memcpy(fill, row_ptr, sizeof(int) * (size_t)(n + 1));
  const double norm = 0.5 / total_mass;
  for (size_t e = 0; e < w; ++e) {
    const int a = (int)(edges[e].key >> 32), b = (int)(edges[e].key & 0xffffffffu);
    const double v = edges[e].value * norm;
    int pos = fill[a]++;
    col[pos] = b;
    val[pos] = v;
    pos = fill[b]++;
    col[pos] = a;
    val[pos] = v;
  }
  free(fill);
  free(edges);
  return (long)row_ptr[n];
}

/* Validation and full affinity build: tsne_neighbors.cpp:623-640 then the two
 * functions above.  dist is row-major n x k Euclidean (not squared). */
long or_build_p(int n, int k, const int* idx, const double* dist, double perplexity,
                int* row_ptr, int* col, double* val, double* p_dir_out) {
  for (size_t e = 0; e < (size_t)n * k; ++e) {
    if (idx[e] < 0 || idx[e] >= n) return -OR_ERR_RANGE;
    if (!isfinite(dist[e]) || dist[e] < 0.0) return -OR_ERR_DIST;
  }
  double* p_dir = p_dir_out ? p_dir_out : (double*)calloc((size_t)n * k, sizeof(double));
  if (!p_dir) return -OR_ERR_ALLOC;
  for (int i = 0; i < n; ++i) {
    or_row_probabilities(dist + (size_t)i * k, k, perplexity, p_dir + (size_t)i * k);
  }
  const long nnz = or_symmetrize(n, k, idx, p_dir, row_ptr, col, val);
  if (!p_dir_out) free(p_dir);
  return nnz;
}

/* ------------------------------------------------------------------------ */
/* Forces.                                                                    */
/* ------------------------------------------------------------------------ */

/* Sparse attraction: tsne_neighbors.cpp:1061-1088. out += exag*P_ij*q_ij*(yi-yj) */
void or_attractive(int n, const int* row_ptr, const int* col, const double* val,
                   const double* y, double exag, double* out) {
  for (int i = 0; i < n; ++i) {
    double gx = 0.0, gy = 0.0;
    for (int p = row_ptr[i]; p < row_ptr[i + 1]; ++p) {
      const int j = col[p];
      const double dx = y[2 * i] - y[2 * j], dy = y[2 * i + 1] - y[2 * j + 1];
      const double q = 1.0 / (1.0 + dx * dx + dy * dy);
      const double c = exag * val[p] * q;
      gx += c * dx;
      gy += c * dy;
    }
    out[2 * i] += gx;
    out[2 * i + 1] += gy;
  }
}

/* Exact O(n^2) repulsion: tsne_neighbors.cpp:1090-1150 (repulsive part).
 * out_i = -sum_j q_ij^2 (yi - yj) / sum_q ; returns sum_q = sum_{i!=j} q_ij. */
double or_exact_repulsive(int n, const double* y, double* out) {
  double sum_q = 0.0;
  for (int i = 0; i < 2 * n; ++i) out[i] = 0.0;
  for (int i = 0; i < n; ++i) {
    for (int j = i + 1; j < n; ++j) {
      const double dx = y[2 * i] - y[2 * j], dy = y[2 * i + 1] - y[2 * j + 1];
      const double q = 1.0 / (1.0 + dx * dx + dy * dy);
      sum_q += 2.0 * q;
      const double c = -(q * q);
      out[2 * i] += c * dx;
      out[2 * i + 1] += c * dy;
      out[2 * j] -= c * dx;
      out[2 * j + 1] -= c * dy;
    }
  }
  const double inv = 1.0 / fmax(sum_q, (double)FLT_MIN);
  for (int i = 0; i < 2 * n; ++i) out[i] *= inv;
  return sum_q;
}

/* sum_{i != j} 1/(1+|yi-yj|^2): tsne_neighbors.cpp:1034-1059 */
double or_sum_q(int n, const double* y) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    for (int j = i + 1; j < n; ++j) {
      const double dx = y[2 * i] - y[2 * j], dy = y[2 * i + 1] - y[2 * j + 1];
      s += 2.0 / (1.0 + dx * dx + dy * dy);
    }
  }
  return fmax(s, (double)FLT_MIN);
}

/* KL(P||Q): tsne_neighbors.cpp:1338-1365 (with its 1e-9 guards). */
double or_kl(int n, const int* row_ptr, const int* col, const double* val,
             const double* y) {
  const double sum_q = or_sum_q(n, y);
  double kl = 0.0;
  for (int i = 0; i < n; ++i) {
    for (int p = row_ptr[i]; p < row_ptr[i + 1]; ++p) {
      const int j = col[p];
      const double dx = y[2 * i] - y[2 * j], dy = y[2 * i + 1] - y[2 * j + 1];
      const double q = (1.0 / (1.0 + dx * dx + dy * dy)) / sum_q;
      kl += val[p] * log((val[p] + 1e-9) / (q + 1e-9));
    }
  }
  return kl;
}

/* ------------------------------------------------------------------------ */
/* FP64 complex FFT (radix-2, exact per-stage twiddles).  Used to restate     */
/* compute_fft_grid_convolution_workspace_t (tsne_neighbors.cpp:956-1018).   */
/* ------------------------------------------------------------------------ */
static void or_fft1(double* re, double* im, int n, int inverse) {
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len >> 1;
    for (int j = 0; j < half; ++j) {
      const double ang = (inverse ? 2.0 : -2.0) * M_PI * (double)j / (double)len;
      const double wr = cos(ang), wi = sin(ang);
      for (int i = 0; i < n; i += len) {
        const double ur = re[i + j], ui = im[i + j];
        const double xr = re[i + j + half], xi = im[i + j + half];
        const double vr = xr * wr - xi * wi, vi = xr * wi + xi * wr;
        re[i + j] = ur + vr; im[i + j] = ui + vi;
        re[i + j + half] = ur - vr; im[i + j + half] = ui - vi;
      }
    }
  }
  if (inverse) {
    for (int i = 0; i < n; ++i) { re[i] /= n; im[i] /= n; }
  }
}

/* 2D transform of an F x F row-major complex array (rows then columns). */
void or_fft2(double* re, double* im, int f, int inverse) {
  double* cr = (double*)malloc(sizeof(double) * f);
  double* ci = (double*)malloc(sizeof(double) * f);
  for (int r = 0; r < f; ++r) or_fft1(re + (size_t)r * f, im + (size_t)r * f, f, inverse);
  for (int c = 0; c < f; ++c) {
    for (int r = 0; r < f; ++r) { cr[r] = re[(size_t)r * f + c]; ci[r] = im[(size_t)r * f + c]; }
    or_fft1(cr, ci, f, inverse);
    for (int r = 0; r < f; ++r) { re[(size_t)r * f + c] = cr[r]; im[(size_t)r * f + c] = ci[r]; }
  }
  free(cr);
  free(ci);
}

/* Grid geometry: tsne_neighbors.cpp:1168-1184 (CUDA twin at
 * embedding_cuda_kernels.cpp:2567-2583).  geom = {lower_x, lower_y, spacing}. */
void or_grid_geometry(int n, const double* y, int g, double* geom) {
  double min_x = y[0], max_x = y[0], min_y = y[1], max_y = y[1];
  for (int i = 1; i < n; ++i) {
    min_x = fmin(min_x, y[2 * i]); max_x = fmax(max_x, y[2 * i]);
    min_y = fmin(min_y, y[2 * i + 1]); max_y = fmax(max_y, y[2 * i + 1]);
  }
  const double cx = 0.5 * (min_x + max_x), cy = 0.5 * (min_y + max_y);
  double span = fmax(max_x - min_x, max_y - min_y);
  if (!isfinite(span) || span <= 0.0) span = 1.0;
  const double half = 0.55 * span + 1.0e-3;
  geom[0] = cx - half;
  geom[1] = cy - half;
  geom[2] = (2.0 * half) / (double)(g - 1);
}

static double or_bilinear(const double* grid, int g, double gx, double gy) {
  const double cx = fmax(0.0, fmin((double)(g - 1), gx));
  const double cy = fmax(0.0, fmin((double)(g - 1), gy));
  int x0 = (int)floor(cx), y0 = (int)floor(cy);
  if (x0 > g - 2) x0 = g - 2;
  if (y0 > g - 2) y0 = g - 2;
  if (x0 < 0) x0 = 0;
  if (y0 < 0) y0 = 0;
  const double tx = cx - x0, ty = cy - y0;
  return (1 - tx) * (1 - ty) * grid[(size_t)y0 * g + x0] +
         tx * (1 - ty) * grid[(size_t)y0 * g + x0 + 1] +
         (1 - tx) * ty * grid[(size_t)(y0 + 1) * g + x0] +
         tx * ty * grid[(size_t)(y0 + 1) * g + x0 + 1];
}

/* FFT-grid repulsion: tsne_neighbors.cpp:1152-1260 (compute_gradient_fft_grid_f)
 * with the convolution of :956-1018, in FP64.
 * rep (n x 2) receives -(y*q2 - xq2, y*q2 - yq2)/Z.  *z_out receives
 * Z = max(sum_i bilinear(q)(y_i) - n, FLT_MIN).  grids (optional, 4*g*g)
 * receives q, q2, xq2, yq2 on the g x g nodes; splat (optional, 3*g*g)
 * receives mass, mass*x, mass*y. */
int or_fft_repulsive(int n, const double* y, int g, double* rep, double* z_out,
                     double* grids, double* splat) {
  if (g < 4 || (g & (g - 1))) return OR_ERR_ARG;
  const int f = 2 * g;
  const size_t gg = (size_t)g * g, ff = (size_t)f * f;
  double geom[3];
  or_grid_geometry(n, y, g, geom);
  const double lower_x = geom[0], lower_y = geom[1], h = geom[2], inv_h = 1.0 / h;
  double* mass = (double*)calloc(3 * gg, sizeof(double));
  double* mx = mass + gg;
  double* my = mass + 2 * gg;
  double* gpos = (double*)malloc(sizeof(double) * 2 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    const double xc = y[2 * i], yc = y[2 * i + 1];
    const double cx = fmax(0.0, fmin((double)(g - 1), (xc - lower_x) * inv_h));
    const double cy = fmax(0.0, fmin((double)(g - 1), (yc - lower_y) * inv_h));
    int x0 = (int)floor(cx), y0 = (int)floor(cy);
    if (x0 > g - 2) x0 = g - 2;
    if (y0 > g - 2) y0 = g - 2;
    const double tx = cx - x0, ty = cy - y0;
    gpos[2 * i] = cx;
    gpos[2 * i + 1] = cy;
    const double w[4] = {(1 - tx) * (1 - ty), tx * (1 - ty), (1 - tx) * ty, tx * ty};
    const size_t p[4] = {(size_t)y0 * g + x0, (size_t)y0 * g + x0 + 1,
                         (size_t)(y0 + 1) * g + x0, (size_t)(y0 + 1) * g + x0 + 1};
    for (int c = 0; c < 4; ++c) {
      mass[p[c]] += w[c];
      mx[p[c]] += w[c] * xc;
      my[p[c]] += w[c] * yc;
    }
  }
  if (splat) memcpy(splat, mass, sizeof(double) * 3 * gg);

  /* kernels at wrapped offsets |dx|,|dy| <= g-1 on the f x f torus */
  double* kq_re = (double*)calloc(ff, sizeof(double));
  double* kq_im = (double*)calloc(ff, sizeof(double));
  double* k2_re = (double*)calloc(ff, sizeof(double));
  double* k2_im = (double*)calloc(ff, sizeof(double));
  for (int dy = -(g - 1); dy <= g - 1; ++dy) {
    const int yy = dy < 0 ? dy + f : dy;
    for (int dx = -(g - 1); dx <= g - 1; ++dx) {
      const int xx = dx < 0 ? dx + f : dx;
      const double ox = dx * h, oy = dy * h;
      const double q = 1.0 / (1.0 + ox * ox + oy * oy);
      kq_re[(size_t)yy * f + xx] = q;
      k2_re[(size_t)yy * f + xx] = q * q;
    }
  }
  or_fft2(kq_re, kq_im, f, 0);
  or_fft2(k2_re, k2_im, f, 0);

  double* in_re[3];
  double* in_im[3];
  for (int c = 0; c < 3; ++c) {
    in_re[c] = (double*)calloc(ff, sizeof(double));
    in_im[c] = (double*)calloc(ff, sizeof(double));
    const double* src = mass + c * gg;
    for (int r = 0; r < g; ++r)
      for (int x = 0; x < g; ++x) in_re[c][(size_t)r * f + x] = src[(size_t)r * g + x];
    or_fft2(in_re[c], in_im[c], f, 0);
  }
  double* out = (double*)malloc(sizeof(double) * 4 * gg);
  double* wr = (double*)malloc(sizeof(double) * ff);
  double* wi = (double*)malloc(sizeof(double) * ff);
  /* q = M*Kq, q2 = M*Kq2, xq2 = Mx*Kq2, yq2 = My*Kq2 */
  const int src_ch[4] = {0, 0, 1, 2};
  for (int o = 0; o < 4; ++o) {
    const double* ar = in_re[src_ch[o]];
    const double* ai = in_im[src_ch[o]];
    const double* br = o == 0 ? kq_re : k2_re;
    const double* bi = o == 0 ? kq_im : k2_im;
    for (size_t t = 0; t < ff; ++t) {
      wr[t] = ar[t] * br[t] - ai[t] * bi[t];
      wi[t] = ar[t] * bi[t] + ai[t] * br[t];
    }
    or_fft2(wr, wi, f, 1);
    for (int r = 0; r < g; ++r)
      for (int x = 0; x < g; ++x) out[o * gg + (size_t)r * g + x] = wr[(size_t)r * f + x];
  }
  if (grids) memcpy(grids, out, sizeof(double) * 4 * gg);
  double sq = 0.0;
  for (int i = 0; i < n; ++i) sq += or_bilinear(out, g, gpos[2 * i], gpos[2 * i + 1]);
  const double z = fmax(sq - (double)n, (double)FLT_MIN);
  if (z_out) *z_out = z;
  for (int i = 0; i < n; ++i) {
    const double q2 = or_bilinear(out + gg, g, gpos[2 * i], gpos[2 * i + 1]);
    const double xq2 = or_bilinear(out + 2 * gg, g, gpos[2 * i], gpos[2 * i + 1]);
    const double yq2 = or_bilinear(out + 3 * gg, g, gpos[2 * i], gpos[2 * i + 1]);
    rep[2 * i] = -(y[2 * i] * q2 - xq2) / z;
    rep[2 * i + 1] = -(y[2 * i + 1] * q2 - yq2) / z;
  }
  free(out); free(wr); free(wi);
  for (int c = 0; c < 3; ++c) { free(in_re[c]); free(in_im[c]); }
  free(kq_re); free(kq_im); free(k2_re); free(k2_im);
  free(gpos);
  free(mass);
  return OR_OK;
}

/* Linear (zero-padded) convolution of a g x g grid with the wrapped Cauchy
 * kernel table of or_fft_repulsive, evaluated directly in O(g^4).  Used only
 * to pin the FFT convolution above on tiny grids. */
void or_direct_convolution(int g, double h, const double* src, int power, double* out) {
  for (int y0 = 0; y0 < g; ++y0) {
    for (int x0 = 0; x0 < g; ++x0) {
      double s = 0.0;
      for (int y1 = 0; y1 < g; ++y1) {
        for (int x1 = 0; x1 < g; ++x1) {
          const double ox = (x0 - x1) * h, oy = (y0 - y1) * h;
          const double q = 1.0 / (1.0 + ox * ox + oy * oy);
          s += (power == 1 ? q : q * q) * src[(size_t)y1 * g + x1];
        }
      }
      out[(size_t)y0 * g + x0] = s;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Optimizer update: tsne_neighbors.cpp:1293-1336 (apply_open_tsne_update_f)  */
/* and zero_mean_f :1280-1291, in FP64.                                       */
/* ------------------------------------------------------------------------ */
static double or_sign(double x) { return x == 0.0 ? 0.0 : (x < 0.0 ? -1.0 : 1.0); }

void or_update(int n, double* y, double* upd, double* gains, const double* grad,
               double lr, double momentum, double min_gain, double max_step) {
  const int clip = isfinite(max_step) && max_step > 0.0;
  for (int i = 0; i < n; ++i) {
    double norm2 = 0.0;
    for (int d = 0; d < 2; ++d) {
      const int t = 2 * i + d;
      if (or_sign(upd[t]) != or_sign(grad[t])) gains[t] += 0.2;
      else gains[t] = gains[t] * 0.8 + min_gain;
      if (gains[t] < min_gain) gains[t] = min_gain;
      upd[t] = momentum * upd[t] - lr * gains[t] * grad[t];
      norm2 += upd[t] * upd[t];
    }
    double scale = 1.0;
    if (clip && norm2 > max_step * max_step) scale = max_step / sqrt(fmax(norm2, (double)FLT_MIN));
    for (int d = 0; d < 2; ++d) {
      upd[2 * i + d] *= scale;
      y[2 * i + d] += upd[2 * i + d];
    }
  }
}

void or_zero_mean(int n, double* y) {
  double mx = 0.0, my = 0.0;
  for (int i = 0; i < n; ++i) { mx += y[2 * i]; my += y[2 * i + 1]; }
  mx /= n;
  my /= n;
  for (int i = 0; i < n; ++i) { y[2 * i] -= mx; y[2 * i + 1] -= my; }
}

/* Grid-size rules: CPU tsne_neighbors.cpp:403-413, CUDA
 * embedding_cuda_kernels.cpp:5610-5623 (environment overrides omitted). */
int or_grid_size_cpu(int n) { return n >= 50000 ? 256 : 128; }
int or_grid_size_gpu(int n) { return n >= 100000 ? 1024 : (n < 20000 ? 256 : 512); }

/* ------------------------------------------------------------------------ */
/* Whole optimisation in FP64: the phase loop of knn_tsne_opentsne_float_cpp  */
/* (tsne_neighbors.cpp:1976-2070): per iteration FFT-grid gradient (or exact  */
/* if g == 0), update, zero mean.  y (n x 2) holds the init on entry and the  */
/* layout on exit; the init is centred first (:1953).                         */
/* ------------------------------------------------------------------------ */
int or_tsne_optimize(int n, const int* row_ptr, const int* col, const double* val,
                     double* y, int g, int early_iter, int n_iter, double early_exag,
                     double exag, double lr_early, double lr_normal, double mom0,
                     double mom1, double min_gain, double max_step) {
  double* grad = (double*)malloc(sizeof(double) * 2 * (size_t)n);
  double* upd = (double*)calloc(2 * (size_t)n, sizeof(double));
  double* gains = (double*)malloc(sizeof(double) * 2 * (size_t)n);
  if (!grad || !upd || !gains) return OR_ERR_ALLOC;
  for (size_t t = 0; t < 2 * (size_t)n; ++t) gains[t] = 1.0;
  or_zero_mean(n, y);
  for (int it = 0; it < early_iter + n_iter; ++it) {
    const int early = it < early_iter;
    const double ex = early ? early_exag : exag;
    if (g > 0) {
      double z;
      or_fft_repulsive(n, y, g, grad, &z, NULL, NULL);
    } else {
      or_exact_repulsive(n, y, grad);
    }
    or_attractive(n, row_ptr, col, val, y, ex, grad);
    or_update(n, y, upd, gains, grad, early ? lr_early : lr_normal, early ? mom0 : mom1,
              min_gain, max_step);
    or_zero_mean(n, y);
  }
  free(grad);
  free(upd);
  free(gains);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Exact rank metrics: src/embedding_utils.cpp:1582-1752                      */
/* (trustworthiness, continuity, kNN preservation, mean neighbour rank error) */
/* high: n x ph, low: n x pl (row-major).  out[4] for the single k.           */
/* ------------------------------------------------------------------------ */
static const double* or_rank_row;
static int or_rank_cmp(const void* a, const void* b) {
  const int l = *(const int*)a, r = *(const int*)b;
  const double dl = or_rank_row[l], dr = or_rank_row[r];
  if (dl < dr || (dl == dr && l < r)) return -1;
  return 1;
}

int or_structure_metrics(int n, int ph, const double* high, int pl, const double* low,
                         int k, double* out) {
  if (n < 3 || k < 1 || k >= n || 2 * n - 3 * k - 1 <= 0) return OR_ERR_ARG;
  double* hd = (double*)malloc(sizeof(double) * (size_t)n * n);
  double* ld = (double*)malloc(sizeof(double) * (size_t)n * n);
  int* ho = (int*)malloc(sizeof(int) * n);
  int* lo = (int*)malloc(sizeof(int) * n);
  int* hr = (int*)malloc(sizeof(int) * n);
  int* lr = (int*)malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      double a = 0.0, b = 0.0;
      for (int c = 0; c < ph; ++c) { const double d = high[(size_t)i * ph + c] - high[(size_t)j * ph + c]; a += d * d; }
      for (int c = 0; c < pl; ++c) { const double d = low[(size_t)i * pl + c] - low[(size_t)j * pl + c]; b += d * d; }
      hd[(size_t)i * n + j] = a;
      ld[(size_t)i * n + j] = b;
    }
  }
  double shared = 0, trust = 0, cont = 0, rerr = 0;
  for (int i = 0; i < n; ++i) {
    int pos = 0;
    for (int j = 0; j < n; ++j) if (j != i) { ho[pos] = j; lo[pos] = j; ++pos; }
    or_rank_row = hd + (size_t)i * n;
    qsort(ho, n - 1, sizeof(int), or_rank_cmp);
    or_rank_row = ld + (size_t)i * n;
    qsort(lo, n - 1, sizeof(int), or_rank_cmp);
    for (int r = 0; r < n - 1; ++r) { hr[ho[r]] = r + 1; lr[lo[r]] = r + 1; }
    for (int r = 0; r < k; ++r) {
      const int rih = hr[lo[r]], ril = lr[ho[r]];
      if (rih <= k) shared += 1.0;
      if (rih > k) trust += rih - k;
      if (ril > k) cont += ril - k;
      rerr += fabs((double)(rih - (r + 1)));
    }
  }
  const double norm = 2.0 / ((double)n * k * (2.0 * n - 3.0 * k - 1.0));
  out[0] = fmax(0.0, fmin(1.0, 1.0 - norm * trust));
  out[1] = fmax(0.0, fmin(1.0, 1.0 - norm * cont));
  out[2] = shared / ((double)n * k);
  out[3] = rerr / ((double)n * k);
  free(hd); free(ld); free(ho); free(lo); free(hr); free(lr);
  return OR_OK;
}
