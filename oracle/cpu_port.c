/*
 * cpu_port.c -- float32, multi-threaded restatement of the reference CPU
 * t-SNE optimiser (knn_tsne_opentsne_float_cpp, src/tsne_neighbors.cpp:
 * 1871-2121) used only as the timed CPU baseline and as a float32 trajectory
 * check.
 *
 * TEST INFRASTRUCTURE ONLY (cpu_baseline leg of bench.py and the test suite).
 * The product library never links, loads or calls it.
 *
 * It keeps the reference's stages, precisions and threading so its wall time
 * stands for the reference CPU path on the same host:
 *   - float32 perplexity search with the min-d^2 shift (:532-614), rows in
 *     parallel;
 *   - packed (pair key, value) edges, one global sort, FP64 duplicate sums,
 *     CSR with both directions (:616-724);
 *   - serial bounds and serial bilinear splat (:1168-1226);
 *   - radix-2 float FFT with bit reversal and the w *= root twiddle
 *     recurrence (:726-834), rows then columns in parallel; 5 forward and
 *     4 inverse F x F transforms per iteration (:956-1018);
 *   - Z and gather in parallel (:1228-1258), CSR row-parallel attraction
 *     (:1061-1088), gains/momentum/clip update (:1297-1336), FP64 mean
 *     centring (:1280-1291);
 *   - a persistent worker pool with contiguous chunking (:113-284).
 * Built with -O2 and no -ffast-math / -march=native (docs/implementation.md:
 * 47-53).
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef float _Complex cf;

/* ---------------------------- worker pool ------------------------------- */
typedef void (*range_fn)(int begin, int end, int tid, void* ctx);

typedef struct {
  int n_threads;
  pthread_t* threads;
  pthread_mutex_t mu;
  pthread_cond_t cv_start;
  pthread_cond_t cv_done;
  unsigned long generation;
  int pending;
  int stop;
  range_fn fn;
  void* ctx;
  int n;
  int chunk;
} pool_t;

typedef struct {
  pool_t* pool;
  int tid;
} worker_arg;

static void* pool_worker(void* p) {
  worker_arg* a = (worker_arg*)p;
  pool_t* pool = a->pool;
  const int tid = a->tid;
  unsigned long seen = 0;
  for (;;) {
    pthread_mutex_lock(&pool->mu);
    while (!pool->stop && pool->generation == seen) pthread_cond_wait(&pool->cv_start, &pool->mu);
    if (pool->stop) {
      pthread_mutex_unlock(&pool->mu);
      break;
    }
    seen = pool->generation;
    range_fn fn = pool->fn;
    void* ctx = pool->ctx;
    const int n = pool->n, chunk = pool->chunk;
    pthread_mutex_unlock(&pool->mu);
    const int begin = tid * chunk;
    const int end = begin + chunk < n ? begin + chunk : n;
    if (begin < end) fn(begin, end, tid, ctx);
    pthread_mutex_lock(&pool->mu);
    if (--pool->pending == 0) pthread_cond_signal(&pool->cv_done);
    pthread_mutex_unlock(&pool->mu);
  }
  free(a);
  return NULL;
}

static pool_t* pool_create(int n_threads) {
  pool_t* pool = (pool_t*)calloc(1, sizeof(pool_t));
  pool->n_threads = n_threads < 1 ? 1 : n_threads;
  pthread_mutex_init(&pool->mu, NULL);
  pthread_cond_init(&pool->cv_start, NULL);
  pthread_cond_init(&pool->cv_done, NULL);
  pool->threads = (pthread_t*)calloc((size_t)pool->n_threads, sizeof(pthread_t));
  for (int t = 1; t < pool->n_threads; ++t) {
    worker_arg* a = (worker_arg*)malloc(sizeof(worker_arg));
    a->pool = pool;
    a->tid = t;
    pthread_create(&pool->threads[t], NULL, pool_worker, a);
  }
  return pool;
}

static void pool_destroy(pool_t* pool) {
  pthread_mutex_lock(&pool->mu);
  pool->stop = 1;
  pthread_cond_broadcast(&pool->cv_start);
  pthread_mutex_unlock(&pool->mu);
  for (int t = 1; t < pool->n_threads; ++t) pthread_join(pool->threads[t], NULL);
  free(pool->threads);
  pthread_mutex_destroy(&pool->mu);
  pthread_cond_destroy(&pool->cv_start);
  pthread_cond_destroy(&pool->cv_done);
  free(pool);
}

/* contiguous chunks, thread 0 runs the first chunk on the caller */
static void parallel_for(pool_t* pool, int n, range_fn fn, void* ctx) {
  if (n <= 0) return;
  const int threads = pool->n_threads < n ? pool->n_threads : n;
  const int chunk = (n + threads - 1) / threads;
  if (threads <= 1) {
    fn(0, n, 0, ctx);
    return;
  }
  pthread_mutex_lock(&pool->mu);
  pool->fn = fn;
  pool->ctx = ctx;
  pool->n = n;
  pool->chunk = chunk;
  pool->pending = pool->n_threads - 1;
  pool->generation++;
  pthread_cond_broadcast(&pool->cv_start);
  pthread_mutex_unlock(&pool->mu);
  fn(0, chunk < n ? chunk : n, 0, ctx);
  pthread_mutex_lock(&pool->mu);
  while (pool->pending > 0) pthread_cond_wait(&pool->cv_done, &pool->mu);
  pthread_mutex_unlock(&pool->mu);
}

static double now_sec(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* --------------------------- affinities ---------------------------------- */
/* compute_row_probabilities_float: tsne_neighbors.cpp:532-614 */
static void row_probabilities_f(const float* d, int k, double perplexity, float* row_p) {
  float min_d2 = FLT_MAX, max_d2 = 0.0f;
  for (int j = 0; j < k; ++j) {
    const float d2 = d[j] * d[j];
    if (d2 < min_d2) min_d2 = d2;
    if (d2 > max_d2) max_d2 = d2;
    row_p[j] = 0.0f;
  }
  const float spread = max_d2 - min_d2;
  if (spread <= FLT_EPSILON * fmaxf(1.0f, max_d2)) {
    for (int j = 0; j < k; ++j) row_p[j] = 1.0f / (float)k;
    return;
  }
  int found = 0;
  float beta = 1.0f, min_beta = -FLT_MAX, max_beta = FLT_MAX, sum_p = FLT_MIN;
  const float target = (float)log(perplexity);
  for (int iter = 0; !found && iter < 200; ++iter) {
    sum_p = FLT_MIN;
    for (int j = 0; j < k; ++j) {
      const float d2 = d[j] * d[j] - min_d2;
      const float p = expf(-beta * d2);
      row_p[j] = p;
      sum_p += p;
    }
    float entropy = 0.0f;
    for (int j = 0; j < k; ++j) entropy += beta * ((d[j] * d[j] - min_d2) * row_p[j]);
    entropy = entropy / sum_p + logf(sum_p);
    const float diff = entropy - target;
    if (fabsf(diff) < 1e-5f) {
      found = 1;
    } else if (diff > 0.0f) {
      min_beta = beta;
      beta = (max_beta == FLT_MAX || max_beta == -FLT_MAX) ? beta * 2.0f : (beta + max_beta) * 0.5f;
    } else {
      max_beta = beta;
      beta = (min_beta == -FLT_MAX || min_beta == FLT_MAX) ? beta * 0.5f : (beta + min_beta) * 0.5f;
    }
    if (!isfinite(beta)) break;
  }
  if (!isfinite(sum_p) || sum_p <= FLT_MIN) {
    int tied = 0;
    for (int j = 0; j < k; ++j) {
      const float d2 = d[j] * d[j];
      if (fabsf(d2 - min_d2) <= FLT_EPSILON * fmaxf(1.0f, min_d2)) ++tied;
    }
    const float mass = 1.0f / (float)(tied > 1 ? tied : 1);
    for (int j = 0; j < k; ++j) {
      const float d2 = d[j] * d[j];
      row_p[j] = fabsf(d2 - min_d2) <= FLT_EPSILON * fmaxf(1.0f, min_d2) ? mass : 0.0f;
    }
    return;
  }
  const float inv = 1.0f / sum_p;
  for (int j = 0; j < k; ++j) row_p[j] *= inv;
}

typedef struct {
  uint64_t key;
  float value;
} packed_edge;

static int packed_cmp(const void* a, const void* b) {
  const uint64_t ka = ((const packed_edge*)a)->key, kb = ((const packed_edge*)b)->key;
  return ka < kb ? -1 : (ka > kb ? 1 : 0);
}

typedef struct {
  int n, k;
  const int* idx;
  const float* dist;
  double perplexity;
  packed_edge** local;
  size_t* local_n;
} aff_ctx;

static void aff_rows(int begin, int end, int tid, void* c) {
  aff_ctx* a = (aff_ctx*)c;
  const int k = a->k;
  float* row_p = (float*)malloc(sizeof(float) * (size_t)k);
  packed_edge* e = (packed_edge*)malloc(sizeof(packed_edge) * (size_t)(end - begin) * k + 1);
  size_t m = 0;
  for (int i = begin; i < end; ++i) {
    row_probabilities_f(a->dist + (size_t)i * k, k, a->perplexity, row_p);
    for (int j = 0; j < k; ++j) {
      const int nb = a->idx[(size_t)i * k + j];
      if (nb == i) continue;
      const uint32_t lo = (uint32_t)(i < nb ? i : nb), hi = (uint32_t)(i < nb ? nb : i);
      e[m].key = ((uint64_t)lo << 32) | hi;
      e[m].value = row_p[j];
      ++m;
    }
  }
  a->local[tid] = e;
  a->local_n[tid] = m;
  free(row_p);
}

typedef struct {
  int n;
  int* row_ptr;
  int* col;
  float* val;
} csr_f;

static int build_probabilities_f(pool_t* pool, int n, int k, const int* idx, const float* dist,
                                 double perplexity, csr_f* p) {
  for (size_t e = 0; e < (size_t)n * k; ++e) {
    if (idx[e] < 0 || idx[e] >= n) return 1;
    if (!isfinite(dist[e]) || dist[e] < 0.0f) return 2;
  }
  aff_ctx a = {n, k, idx, dist, perplexity, NULL, NULL};
  a.local = (packed_edge**)calloc((size_t)pool->n_threads, sizeof(packed_edge*));
  a.local_n = (size_t*)calloc((size_t)pool->n_threads, sizeof(size_t));
  parallel_for(pool, n, aff_rows, &a);
  size_t total = 0;
  for (int t = 0; t < pool->n_threads; ++t) total += a.local_n[t];
  packed_edge* edges = (packed_edge*)malloc(sizeof(packed_edge) * (total ? total : 1));
  size_t off = 0;
  for (int t = 0; t < pool->n_threads; ++t) {
    if (a.local[t]) {
      memcpy(edges + off, a.local[t], sizeof(packed_edge) * a.local_n[t]);
      off += a.local_n[t];
      free(a.local[t]);
    }
  }
  free(a.local);
  free(a.local_n);
  if (total == 0) {
    free(edges);
    return 3;
  }
  qsort(edges, total, sizeof(packed_edge), packed_cmp);
  p->n = n;
  p->row_ptr = (int*)calloc((size_t)n + 1, sizeof(int));
  size_t w = 0;
  double total_mass = 0.0;
  for (size_t r = 0; r < total;) {
    const uint64_t key = edges[r].key;
    double sum = 0.0;
    while (r < total && edges[r].key == key) sum += (double)edges[r++].value;
    edges[w].key = key;
    edges[w].value = (float)sum;
    ++w;
    total_mass += sum;
    p->row_ptr[(key >> 32) + 1]++;
    p->row_ptr[(key & 0xffffffffu) + 1]++;
  }
  if (!isfinite(total_mass) || total_mass <= 0.0) {
    free(edges);
    return 4;
  }
  for (int i = 0; i < n; ++i) p->row_ptr[i + 1] += p->row_ptr[i];
  const size_t nnz = (size_t)p->row_ptr[n];
  p->col = (int*)malloc(sizeof(int) * nnz);
  p->val = (float*)malloc(sizeof(float) * nnz);
  int* fill = (int*)malloc(sizeof(int) * ((size_t)n + 1));
  memcpy(fill, p->row_ptr, sizeof(int) * ((size_t)n + 1));
  const float norm = (float)(0.5 / total_mass);
  for (size_t e = 0; e < w; ++e) {
    const int lo = (int)(edges[e].key >> 32), hi = (int)(edges[e].key & 0xffffffffu);
    const float v = edges[e].value * norm;
    int pos = fill[lo]++;
    p->col[pos] = hi;
    p->val[pos] = v;
    pos = fill[hi]++;
    p->col[pos] = lo;
    p->val[pos] = v;
  }
  free(fill);
  free(edges);
  return 0;
}

/* ------------------------------ FFT ------------------------------------- */
/* FftPlanT / fft_1d_t / fft_2d_t: tsne_neighbors.cpp:726-834 */
typedef struct {
  int size;
  int bits;
  int* bit_reverse;
  cf* fwd_roots;
  cf* inv_roots;
  cf* scratch; /* size * n_threads */
} fft_plan;

static void plan_init(fft_plan* pl, int size, int n_threads) {
  pl->size = size;
  pl->bits = 0;
  for (int v = size; v > 1; v >>= 1) ++pl->bits;
  pl->bit_reverse = (int*)malloc(sizeof(int) * (size_t)size);
  for (int i = 0; i < size; ++i) {
    unsigned v = (unsigned)i, r = 0;
    for (int b = 0; b < pl->bits; ++b) {
      r = (r << 1) | (v & 1u);
      v >>= 1;
    }
    pl->bit_reverse[i] = (int)r;
  }
  pl->fwd_roots = (cf*)malloc(sizeof(cf) * (size_t)(pl->bits + 1));
  pl->inv_roots = (cf*)malloc(sizeof(cf) * (size_t)(pl->bits + 1));
  int s = 0;
  for (int len = 2; len <= size; len <<= 1, ++s) {
    const float angle = (float)6.283185307179586476925286766559 / (float)len;
    pl->fwd_roots[s] = cosf(angle) + sinf(angle) * I;
    pl->inv_roots[s] = cosf(-angle) + sinf(-angle) * I;
  }
  pl->scratch = (cf*)malloc(sizeof(cf) * (size_t)size * (size_t)n_threads);
}

static void plan_free(fft_plan* pl) {
  free(pl->bit_reverse);
  free(pl->fwd_roots);
  free(pl->inv_roots);
  free(pl->scratch);
}

static void fft_1d(cf* a, int n, int inverse, const fft_plan* pl) {
  for (int i = 1; i < n; ++i) {
    const int j = pl->bit_reverse[i];
    if (i < j) {
      cf t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
  const cf* roots = inverse ? pl->inv_roots : pl->fwd_roots;
  int stage = 0;
  for (int len = 2; len <= n; len <<= 1, ++stage) {
    const cf root = roots[stage];
    const int half = len >> 1;
    for (int i = 0; i < n; i += len) {
      cf w = 1.0f;
      for (int j = 0; j < half; ++j) {
        const cf u = a[i + j];
        const cf v = a[i + j + half] * w;
        a[i + j] = u + v;
        a[i + j + half] = u - v;
        w *= root;
      }
    }
  }
  if (inverse) {
    const float scale = 1.0f / (float)n;
    for (int i = 0; i < n; ++i) a[i] *= scale;
  }
}

typedef struct {
  cf* values;
  int size;
  int inverse;
  const fft_plan* pl;
} fft2_ctx;

static void fft2_rows(int b, int e, int tid, void* c) {
  fft2_ctx* x = (fft2_ctx*)c;
  (void)tid;
  for (int r = b; r < e; ++r) fft_1d(x->values + (size_t)r * x->size, x->size, x->inverse, x->pl);
}

static void fft2_cols(int b, int e, int tid, void* c) {
  fft2_ctx* x = (fft2_ctx*)c;
  const int s = x->size;
  cf* col = x->pl->scratch + (size_t)tid * s;
  for (int cc = b; cc < e; ++cc) {
    for (int r = 0; r < s; ++r) col[r] = x->values[(size_t)r * s + cc];
    fft_1d(col, s, x->inverse, x->pl);
    for (int r = 0; r < s; ++r) x->values[(size_t)r * s + cc] = col[r];
  }
}

static void fft_2d(pool_t* pool, cf* values, int size, int inverse, const fft_plan* pl) {
  fft2_ctx x = {values, size, inverse, pl};
  parallel_for(pool, size, fft2_rows, &x);
  parallel_for(pool, size, fft2_cols, &x);
}

/* ------------------------- FFT-grid gradient ----------------------------- */
typedef struct {
  int g, f, n;
  float *mass, *mass_x, *mass_y, *gx, *gy;
  float *q, *q2, *xq2, *yq2;
  cf *mass_fft, *mass_x_fft, *mass_y_fft, *kq, *kq2, *work;
  double* partial;
  fft_plan plan;
} grid_ws;

typedef struct {
  grid_ws* ws;
  const float* src;
  cf* dst;
  const cf* a;
  const cf* b;
  float* out;
  float spacing;
  const csr_f* p;
  const float* y;
  float* grad;
  float exag;
  float inv_sum_q;
} conv_ctx;

static void to_fft_rows(int b, int e, int tid, void* c) {
  conv_ctx* x = (conv_ctx*)c;
  (void)tid;
  const int g = x->ws->g, f = x->ws->f;
  for (int yc = b; yc < e; ++yc) {
    cf* row = x->dst + (size_t)yc * f;
    for (int i = 0; i < f; ++i) row[i] = 0.0f;
    if (yc < g)
      for (int xc = 0; xc < g; ++xc) row[xc] = x->src[(size_t)yc * g + xc];
  }
}

static void kernel_rows(int b, int e, int tid, void* c) {
  conv_ctx* x = (conv_ctx*)c;
  (void)tid;
  const int g = x->ws->g, f = x->ws->f;
  for (int row = b; row < e; ++row) {
    const int dy = row - (g - 1);
    const int yy = dy < 0 ? dy + f : dy;
    const float yo = (float)dy * x->spacing;
    for (int dx = -(g - 1); dx <= g - 1; ++dx) {
      const int xx = dx < 0 ? dx + f : dx;
      const float xo = (float)dx * x->spacing;
      const float d2 = xo * xo + yo * yo;
      const float q = 1.0f / (1.0f + d2);
      x->ws->kq[(size_t)yy * f + xx] = q;
      x->ws->kq2[(size_t)yy * f + xx] = q * q;
    }
  }
}

static void zero_kernels(int b, int e, int tid, void* c) {
  conv_ctx* x = (conv_ctx*)c;
  (void)tid;
  for (int i = b; i < e; ++i) {
    x->ws->kq[i] = 0.0f;
    x->ws->kq2[i] = 0.0f;
  }
}

static void multiply(int b, int e, int tid, void* c) {
  conv_ctx* x = (conv_ctx*)c;
  (void)tid;
  for (int i = b; i < e; ++i) x->ws->work[i] = x->a[i] * x->b[i];
}

static void from_fft_rows(int b, int e, int tid, void* c) {
  conv_ctx* x = (conv_ctx*)c;
  (void)tid;
  const int g = x->ws->g, f = x->ws->f;
  for (int yc = b; yc < e; ++yc)
    for (int xc = 0; xc < g; ++xc) x->out[(size_t)yc * g + xc] = crealf(x->ws->work[(size_t)yc * f + xc]);
}

static float bilinear_f(const float* grid, int g, float gx, float gy) {
  const float cx = fmaxf(0.0f, fminf((float)(g - 1), gx));
  const float cy = fmaxf(0.0f, fminf((float)(g - 1), gy));
  int x0 = (int)floorf(cx), y0 = (int)floorf(cy);
  if (x0 > g - 2) x0 = g - 2;
  if (y0 > g - 2) y0 = g - 2;
  if (x0 < 0) x0 = 0;
  if (y0 < 0) y0 = 0;
  const float tx = cx - (float)x0, ty = cy - (float)y0;
  return (1.0f - tx) * (1.0f - ty) * grid[(size_t)y0 * g + x0] + tx * (1.0f - ty) * grid[(size_t)y0 * g + x0 + 1] +
         (1.0f - tx) * ty * grid[(size_t)(y0 + 1) * g + x0] + tx * ty * grid[(size_t)(y0 + 1) * g + x0 + 1];
}

static void sum_q_rows(int b, int e, int tid, void* c) {
  conv_ctx* x = (conv_ctx*)c;
  double s = 0.0;
  for (int i = b; i < e; ++i) s += (double)bilinear_f(x->ws->q, x->ws->g, x->ws->gx[i], x->ws->gy[i]);
  x->ws->partial[tid] = s;
}

static void gather_rows(int b, int e, int tid, void* c) {
  conv_ctx* x = (conv_ctx*)c;
  (void)tid;
  grid_ws* ws = x->ws;
  for (int i = b; i < e; ++i) {
    const float px = ws->gx[i], py = ws->gy[i];
    const float q2 = bilinear_f(ws->q2, ws->g, px, py);
    const float xq2 = bilinear_f(ws->xq2, ws->g, px, py);
    const float yq2 = bilinear_f(ws->yq2, ws->g, px, py);
    x->grad[2 * i] = -(x->y[2 * i] * q2 - xq2) * x->inv_sum_q;
    x->grad[2 * i + 1] = -(x->y[2 * i + 1] * q2 - yq2) * x->inv_sum_q;
  }
}

static void attract_rows(int b, int e, int tid, void* c) {
  conv_ctx* x = (conv_ctx*)c;
  (void)tid;
  const csr_f* p = x->p;
  for (int i = b; i < e; ++i) {
    for (int pos = p->row_ptr[i]; pos < p->row_ptr[i + 1]; ++pos) {
      const int j = p->col[pos];
      const float dx = x->y[2 * i] - x->y[2 * j];
      const float dy = x->y[2 * i + 1] - x->y[2 * j + 1];
      const float d2 = dx * dx + dy * dy;
      const float q = 1.0f / (1.0f + d2);
      const float coeff = x->exag * p->val[pos] * q;
      x->grad[2 * i] += coeff * dx;
      x->grad[2 * i + 1] += coeff * dy;
    }
  }
}

static void grid_ws_init(grid_ws* ws, int g, int n, int n_threads) {
  ws->g = g;
  ws->f = 2 * g;
  ws->n = n;
  const size_t gg = (size_t)g * g, ff = (size_t)ws->f * ws->f;
  ws->mass = (float*)malloc(sizeof(float) * gg);
  ws->mass_x = (float*)malloc(sizeof(float) * gg);
  ws->mass_y = (float*)malloc(sizeof(float) * gg);
  ws->q = (float*)malloc(sizeof(float) * gg);
  ws->q2 = (float*)malloc(sizeof(float) * gg);
  ws->xq2 = (float*)malloc(sizeof(float) * gg);
  ws->yq2 = (float*)malloc(sizeof(float) * gg);
  ws->gx = (float*)malloc(sizeof(float) * (size_t)n);
  ws->gy = (float*)malloc(sizeof(float) * (size_t)n);
  ws->mass_fft = (cf*)malloc(sizeof(cf) * ff);
  ws->mass_x_fft = (cf*)malloc(sizeof(cf) * ff);
  ws->mass_y_fft = (cf*)malloc(sizeof(cf) * ff);
  ws->kq = (cf*)malloc(sizeof(cf) * ff);
  ws->kq2 = (cf*)malloc(sizeof(cf) * ff);
  ws->work = (cf*)malloc(sizeof(cf) * ff);
  ws->partial = (double*)malloc(sizeof(double) * (size_t)n_threads);
  plan_init(&ws->plan, ws->f, n_threads);
}

static void grid_ws_free(grid_ws* ws) {
  free(ws->mass); free(ws->mass_x); free(ws->mass_y);
  free(ws->q); free(ws->q2); free(ws->xq2); free(ws->yq2);
  free(ws->gx); free(ws->gy);
  free(ws->mass_fft); free(ws->mass_x_fft); free(ws->mass_y_fft);
  free(ws->kq); free(ws->kq2); free(ws->work); free(ws->partial);
  plan_free(&ws->plan);
}

/* compute_gradient_fft_grid_f: tsne_neighbors.cpp:1152-1260 */
static void gradient_fft_f(pool_t* pool, const csr_f* p, const float* y, int n, float exag,
                           grid_ws* ws, float* grad) {
  const int g = ws->g, f = ws->f;
  for (size_t t = 0; t < 2 * (size_t)n; ++t) grad[t] = 0.0f;
  float min_x = y[0], max_x = y[0], min_y = y[1], max_y = y[1];
  for (int i = 1; i < n; ++i) {
    min_x = fminf(min_x, y[2 * i]); max_x = fmaxf(max_x, y[2 * i]);
    min_y = fminf(min_y, y[2 * i + 1]); max_y = fmaxf(max_y, y[2 * i + 1]);
  }
  const float cx = 0.5f * (min_x + max_x), cy = 0.5f * (min_y + max_y);
  float span = fmaxf(max_x - min_x, max_y - min_y);
  if (!isfinite(span) || span <= 0.0f) span = 1.0f;
  const float half = 0.55f * span + 1.0e-3f;
  const float lower_x = cx - half, lower_y = cy - half;
  const float spacing = (2.0f * half) / (float)(g - 1);
  const float inv_spacing = 1.0f / spacing;
  const size_t gg = (size_t)g * g;
  memset(ws->mass, 0, sizeof(float) * gg);
  memset(ws->mass_x, 0, sizeof(float) * gg);
  memset(ws->mass_y, 0, sizeof(float) * gg);
  for (int i = 0; i < n; ++i) {
    const float xc = y[2 * i], yc = y[2 * i + 1];
    const float rx = (xc - lower_x) * inv_spacing, ry = (yc - lower_y) * inv_spacing;
    const float clx = fmaxf(0.0f, fminf((float)(g - 1), rx));
    const float cly = fmaxf(0.0f, fminf((float)(g - 1), ry));
    int x0 = (int)floorf(clx), y0 = (int)floorf(cly);
    if (x0 > g - 2) x0 = g - 2;
    if (y0 > g - 2) y0 = g - 2;
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    const float tx = clx - (float)x0, ty = cly - (float)y0;
    ws->gx[i] = clx;
    ws->gy[i] = cly;
    const float w00 = (1.0f - tx) * (1.0f - ty), w10 = tx * (1.0f - ty);
    const float w01 = (1.0f - tx) * ty, w11 = tx * ty;
    const size_t p00 = (size_t)y0 * g + x0, p10 = p00 + 1, p01 = p00 + g, p11 = p01 + 1;
    ws->mass[p00] += w00; ws->mass[p10] += w10; ws->mass[p01] += w01; ws->mass[p11] += w11;
    ws->mass_x[p00] += w00 * xc; ws->mass_x[p10] += w10 * xc;
    ws->mass_x[p01] += w01 * xc; ws->mass_x[p11] += w11 * xc;
    ws->mass_y[p00] += w00 * yc; ws->mass_y[p10] += w10 * yc;
    ws->mass_y[p01] += w01 * yc; ws->mass_y[p11] += w11 * yc;
  }
  conv_ctx c;
  memset(&c, 0, sizeof(c));
  c.ws = ws;
  c.spacing = spacing;
  c.src = ws->mass; c.dst = ws->mass_fft; parallel_for(pool, f, to_fft_rows, &c);
  c.src = ws->mass_x; c.dst = ws->mass_x_fft; parallel_for(pool, f, to_fft_rows, &c);
  c.src = ws->mass_y; c.dst = ws->mass_y_fft; parallel_for(pool, f, to_fft_rows, &c);
  parallel_for(pool, f * f, zero_kernels, &c);
  parallel_for(pool, 2 * g - 1, kernel_rows, &c);
  fft_2d(pool, ws->mass_fft, f, 0, &ws->plan);
  fft_2d(pool, ws->mass_x_fft, f, 0, &ws->plan);
  fft_2d(pool, ws->mass_y_fft, f, 0, &ws->plan);
  fft_2d(pool, ws->kq, f, 0, &ws->plan);
  fft_2d(pool, ws->kq2, f, 0, &ws->plan);
  const cf* as[4] = {ws->mass_fft, ws->mass_fft, ws->mass_x_fft, ws->mass_y_fft};
  const cf* bs[4] = {ws->kq, ws->kq2, ws->kq2, ws->kq2};
  float* outs[4] = {ws->q, ws->q2, ws->xq2, ws->yq2};
  for (int o = 0; o < 4; ++o) {
    c.a = as[o];
    c.b = bs[o];
    parallel_for(pool, f * f, multiply, &c);
    fft_2d(pool, ws->work, f, 1, &ws->plan);
    c.out = outs[o];
    parallel_for(pool, g, from_fft_rows, &c);
  }
  for (int t = 0; t < pool->n_threads; ++t) ws->partial[t] = 0.0;
  parallel_for(pool, n, sum_q_rows, &c);
  double s = 0.0;
  for (int t = 0; t < pool->n_threads; ++t) s += ws->partial[t];
  c.inv_sum_q = (float)(1.0 / fmax(s - (double)n, (double)FLT_MIN));
  c.y = y;
  c.grad = grad;
  parallel_for(pool, n, gather_rows, &c);
  c.p = p;
  c.exag = exag;
  parallel_for(pool, n, attract_rows, &c);
}

/* exact repulsion (compute_gradient_pair_symmetric_f, :1090-1150), used for
 * the n <= 3000 "auto" route; single-threaded here (timing is not taken on it) */
static void gradient_exact_f(const csr_f* p, const float* y, int n, float exag, float* grad) {
  double sum_q = 0.0;
  for (size_t t = 0; t < 2 * (size_t)n; ++t) grad[t] = 0.0f;
  for (int i = 0; i < n - 1; ++i) {
    for (int j = i + 1; j < n; ++j) {
      const float dx = y[2 * i] - y[2 * j], dy = y[2 * i + 1] - y[2 * j + 1];
      const float q = 1.0f / (1.0f + dx * dx + dy * dy);
      sum_q += 2.0 * (double)q;
      const float c = -(q * q);
      grad[2 * i] += c * dx; grad[2 * i + 1] += c * dy;
      grad[2 * j] -= c * dx; grad[2 * j + 1] -= c * dy;
    }
  }
  const float inv = (float)(1.0 / fmax(sum_q, (double)FLT_MIN));
  for (size_t t = 0; t < 2 * (size_t)n; ++t) grad[t] *= inv;
  for (int i = 0; i < n; ++i) {
    for (int pos = p->row_ptr[i]; pos < p->row_ptr[i + 1]; ++pos) {
      const int j = p->col[pos];
      const float dx = y[2 * i] - y[2 * j], dy = y[2 * i + 1] - y[2 * j + 1];
      const float q = 1.0f / (1.0f + dx * dx + dy * dy);
      const float coeff = exag * p->val[pos] * q;
      grad[2 * i] += coeff * dx;
      grad[2 * i + 1] += coeff * dy;
    }
  }
}

typedef struct {
  float *y, *upd, *gains;
  const float* grad;
  float lr, mom, min_gain, max_step;
  int clip;
} upd_ctx;

static float sign_f(float x) { return x == 0.0f ? 0.0f : (x < 0.0f ? -1.0f : 1.0f); }

/* apply_open_tsne_update_f: tsne_neighbors.cpp:1297-1336 */
static void update_rows(int b, int e, int tid, void* c) {
  upd_ctx* u = (upd_ctx*)c;
  (void)tid;
  const float max2 = u->max_step * u->max_step;
  for (int i = b; i < e; ++i) {
    float norm2 = 0.0f;
    for (int d = 0; d < 2; ++d) {
      const int t = 2 * i + d;
      if (sign_f(u->upd[t]) != sign_f(u->grad[t])) u->gains[t] += 0.2f;
      else u->gains[t] = u->gains[t] * 0.8f + u->min_gain;
      if (u->gains[t] < u->min_gain) u->gains[t] = u->min_gain;
      u->upd[t] = u->mom * u->upd[t] - u->lr * u->gains[t] * u->grad[t];
      norm2 += u->upd[t] * u->upd[t];
    }
    float scale = 1.0f;
    if (u->clip && norm2 > max2) scale = u->max_step / sqrtf(fmaxf(norm2, FLT_MIN));
    for (int d = 0; d < 2; ++d) {
      u->upd[2 * i + d] *= scale;
      u->y[2 * i + d] += u->upd[2 * i + d];
    }
  }
}

static void zero_mean_f(float* y, int n) {
  double mx = 0.0, my = 0.0;
  for (int i = 0; i < n; ++i) { mx += y[2 * i]; my += y[2 * i + 1]; }
  mx /= (double)n;
  my /= (double)n;
  for (int i = 0; i < n; ++i) { y[2 * i] -= (float)mx; y[2 * i + 1] -= (float)my; }
}

/*
 * Optimisation only, from a given symmetric P (0-based CSR): `iters`
 * iterations of one phase (gradient, update, zero mean) on y (n x 2, in
 * place).  Used to time a bounded sample of the reference CPU iteration on a
 * large workload without repeating its affinity stage.  *seconds receives
 * the wall time of the loop.
 */
int port_tsne_iterate(int n, const int* row_ptr, const int* col, const float* val, float* y,
                      int iters, double exag, double lr, double mom, double min_gain,
                      double max_step, int grid, int n_threads, double* seconds) {
  if (n < 2 || n_threads < 1 || iters < 0) return 5;
  pool_t* pool = pool_create(n_threads);
  csr_f p = {n, (int*)row_ptr, (int*)col, (float*)val};
  float* grad = (float*)calloc(2 * (size_t)n, sizeof(float));
  float* upd = (float*)calloc(2 * (size_t)n, sizeof(float));
  float* gains = (float*)malloc(sizeof(float) * 2 * (size_t)n);
  for (size_t t = 0; t < 2 * (size_t)n; ++t) gains[t] = 1.0f;
  const int g = grid == 0 ? (n >= 50000 ? 256 : 128) : grid;
  grid_ws ws;
  memset(&ws, 0, sizeof(ws));
  grid_ws_init(&ws, g, n, pool->n_threads);
  upd_ctx u = {y, upd, gains, grad, (float)lr, (float)mom, (float)min_gain, (float)max_step, 0};
  u.clip = isfinite(max_step) && max_step > 0.0;
  const double t0 = now_sec();
  for (int it = 0; it < iters; ++it) {
    gradient_fft_f(pool, &p, y, n, (float)exag, &ws, grad);
    parallel_for(pool, n, update_rows, &u);
    zero_mean_f(y, n);
  }
  if (seconds) *seconds = now_sec() - t0;
  grid_ws_free(&ws);
  free(grad);
  free(upd);
  free(gains);
  pool_destroy(pool);
  return 0;
}

/*
 * Full reference-algorithm run.  idx/dist are row-major n x k (0-based).
 * y_init (n x 2, row-major) is required (seeded explicit init, as in the
 * measurement plan).  grid: 0 -> reference CPU rule (128 / 256), < 0 ->
 * exact repulsion.  lr <= 0 -> n / phase exaggeration.  timings[0] =
 * affinity seconds, timings[1] = optimisation seconds.  Returns 0 on success.
 */
int port_tsne_run(int n, int k, const int* idx, const float* dist, const float* y_init,
                  double perplexity, int early_iter, int n_iter, double early_exag,
                  double exag, double lr, double mom0, double mom1, double min_gain,
                  double max_step, int grid, int n_threads, float* out, double* timings) {
  if (n < 2 || k < 1 || n_threads < 1) return 5;
  pool_t* pool = pool_create(n_threads);
  const double t0 = now_sec();
  csr_f p;
  memset(&p, 0, sizeof(p));
  const int rc = build_probabilities_f(pool, n, k, idx, dist, perplexity, &p);
  const double t1 = now_sec();
  if (rc != 0) {
    pool_destroy(pool);
    return rc;
  }
  float* y = out;
  memcpy(y, y_init, sizeof(float) * 2 * (size_t)n);
  zero_mean_f(y, n);
  float* grad = (float*)calloc(2 * (size_t)n, sizeof(float));
  float* upd = (float*)calloc(2 * (size_t)n, sizeof(float));
  float* gains = (float*)malloc(sizeof(float) * 2 * (size_t)n);
  for (size_t t = 0; t < 2 * (size_t)n; ++t) gains[t] = 1.0f;
  const int g = grid == 0 ? (n >= 50000 ? 256 : 128) : grid;
  grid_ws ws;
  memset(&ws, 0, sizeof(ws));
  if (g > 0) grid_ws_init(&ws, g, n, pool->n_threads);
  upd_ctx u = {y, upd, gains, grad, 0, 0, (float)min_gain, (float)max_step, 0};
  u.clip = isfinite(max_step) && max_step > 0.0;
  for (int it = 0; it < early_iter + n_iter; ++it) {
    const int early = it < early_iter;
    const double ex = early ? early_exag : exag;
    u.lr = (float)(lr > 0.0 ? lr : (double)n / fmax(ex, DBL_MIN));
    u.mom = (float)(early ? mom0 : mom1);
    if (g > 0) gradient_fft_f(pool, &p, y, n, (float)ex, &ws, grad);
    else gradient_exact_f(&p, y, n, (float)ex, grad);
    parallel_for(pool, n, update_rows, &u);
    zero_mean_f(y, n);
  }
  const double t2 = now_sec();
  if (timings) {
    timings[0] = t1 - t0;
    timings[1] = t2 - t1;
  }
  if (g > 0) grid_ws_free(&ws);
  free(grad);
  free(upd);
  free(gains);
  free(p.row_ptr);
  free(p.col);
  free(p.val);
  pool_destroy(pool);
  return 0;
}

/* compute_row_probabilities_float (tsne_neighbors.cpp:532-614) over the rows
 * of a row-major n x k distance matrix, exposed for its pin against the
 * reference's own function */
void port_row_probabilities_f(const float* dist, int n, int k, double perplexity, float* out) {
  for (int i = 0; i < n; ++i)
    row_probabilities_f(dist + (size_t)i * k, k, perplexity, out + (size_t)i * k);
}
