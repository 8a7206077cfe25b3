/*
 * oracle_umap_cpu.c -- restatement of the reference CPU UMAP optimizer,
 * optimize_layout_csr (src/fast_knn_umap.cpp:3088-3558), for 2D layouts
 * (TEST INFRASTRUCTURE ONLY: the checker and the UMAP CPU baseline; the
 * product never links it).
 *
 * The reference picks one of three forms by thread count and size:
 *   threads = effective_cpu_threads(n_threads, active edges)   :86-89, 3224
 *           = max(1, min(n_threads, 4, active edges));
 *   threads > 1 and n >= 10000: hogwild, float32, per-thread TauPrng
 *                                 negatives, relaxed atomic adds  :3225-3351
 *   threads > 1 and n < 10000:  per-thread float deltas over sync
 *                                 batches, FP64 layout             :3353-3456
 *   otherwise:                   serial, FP64                      :3458-3557
 * Shared by all three: the active edges (weight >= max_w / n_epochs, self and
 * out-of-range entries dropped) with epochs_per_sample = max_w / max(w, 1e-6)
 * in FP64 rounded to float (:3107-3182), the accumulated float schedule
 * (epoch_of_next_sample += period; n_neg = floor((e - next_neg) / period_neg),
 * next_neg += n_neg * period_neg), alpha = lr (1 - e / E) (hogwild: a float
 * alpha0 stepped down by alpha0 / E each epoch), clip +-4.
 * Helpers restated: mix_seed :513-520, deterministic_vertex :532-542,
 * TauPrng / make_tau_prng :659-701, umap_pow :156-186, umap_powf_fast
 * :188-215, clip4f :146-148, add_delta_2d :2663-2679.
 *
 * The hogwild form with one thread is deterministic; with more it is not,
 * in the reference too (threads race on shared coordinates).
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static uint32_t mix_seed(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

/* deterministic_vertex(n, seed, epoch, edge, sample): the edge index form */
static int det_vertex(int n, int seed, int epoch, size_t edge, int sample) {
  uint32_t x = (uint32_t)seed;
  x ^= (uint32_t)((uint32_t)epoch * 0x9e3779b9u);
  x ^= (uint32_t)((edge + 1u) * 0x85ebca6bu);
  x ^= (uint32_t)((uint32_t)(sample + 1) * 0xc2b2ae35u);
  return (int)(mix_seed(x) % (uint32_t)n);
}

typedef struct {
  uint32_t state;
} tau_t;

static tau_t tau_make(uint64_t s0, uint64_t s1, uint64_t s2) {
  tau_t t;
  t.state = mix_seed((uint32_t)s0 ^ (uint32_t)(s1 << 7) ^ (uint32_t)(s2 << 17));
  if (t.state == 0u) t.state = 0x9e3779b9u;
  return t;
}

static uint32_t tau_next(tau_t* t) {
  uint32_t x = t->state;
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  t->state = x == 0u ? 0x9e3779b9u : x;
  return t->state;
}

static tau_t tau_for(int seed, int epoch, size_t window_end, int thread_id) {
  const uint32_t base = mix_seed((uint32_t)seed ^ (uint32_t)((uint32_t)(epoch + 1) * 0x9e3779b9u) ^
                                 (uint32_t)((uint32_t)(thread_id + 1) * 0x85ebca6bu) ^
                                 (uint32_t)((window_end + 1u) * 0xc2b2ae35u));
  return tau_make(mix_seed(base ^ 0xa341316cu), mix_seed(base ^ 0xc8013ea4u),
                  mix_seed(base ^ 0xad90777du));
}

/* exponent-interpolated x^b, double and float forms */
static double pow_interp_d(double x, double b) {
  if (x <= 0.0) return 0.0;
  uint64_t xb;
  memcpy(&xb, &x, 8);
  const double bias = 1072632447.0;
  const int whole = (int)b;
  const double frac = b - (double)whole;
  const double high = (double)(uint32_t)(xb >> 32);
  const uint64_t ih = (uint64_t)(frac * (high - bias) + bias);
  const uint64_t bits = (ih & 0xffffffffull) << 32;
  double fp;
  memcpy(&fp, &bits, 8);
  double ip = 1.0, base = x;
  for (int e = whole; e > 0; e >>= 1) {
    if (e & 1) ip *= base;
    base *= base;
  }
  return ip * fp;
}

static float pow_interp_f(float x, float b) {
  if (x <= 0.0f) return 0.0f;
  uint32_t xb;
  memcpy(&xb, &x, 4);
  const float bias = 1064866805.0f;
  const int whole = (int)b;
  const float frac = b - (float)whole;
  const uint32_t bits = (uint32_t)(frac * ((float)xb - bias) + bias);
  float fp;
  memcpy(&fp, &bits, 4);
  float ip = 1.0f, base = x;
  for (int e = whole; e > 0; e >>= 1) {
    if (e & 1) ip *= base;
    base *= base;
  }
  return ip * fp;
}

static float clip4f(float x) { return x < -4.0f ? -4.0f : (x > 4.0f ? 4.0f : x); }
static double clipd(double x) { return x < -4.0 ? -4.0 : (x > 4.0 ? 4.0 : x); }

/* a relaxed atomic float add (a compare-and-swap loop, as atomic_add_float) */
static void atomic_addf(float* target, float v) {
  float old;
  __atomic_load(target, &old, __ATOMIC_RELAXED);
  float want = old + v;
  while (!__atomic_compare_exchange(target, &old, &want, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
    want = old + v;
}

static float atomic_loadf(float* p) {
  float v;
  __atomic_load(p, &v, __ATOMIC_RELAXED);
  return v;
}

/*
 * The active edges of a CSR graph in the reference's order (:3107-3174).
 * Returns their count (-1: no usable edge, -2: none sampled by n_epochs).
 * rows / tails / eps hold at least nnz entries.
 */
long long or_umap_cpu_active(int n, const int* offsets, const int* neighbors,
                             const float* weights, int index_offset, int n_epochs, int* rows,
                             int* tails, float* eps) {
  double max_w = 0.0;
  for (int r = 0; r < n; ++r)
    for (int pos = offsets[r]; pos < offsets[r + 1]; ++pos) {
      const int nb = neighbors[pos] - index_offset;
      if (nb >= 0 && nb < n && nb != r && (double)weights[pos] > max_w) max_w = weights[pos];
    }
  if (max_w <= 0.0) return -1;
  const double min_w = n_epochs > 0 ? max_w / (double)n_epochs : 0.0;
  long long m = 0;
  for (int r = 0; r < n; ++r)
    for (int pos = offsets[r]; pos < offsets[r + 1]; ++pos) {
      const int nb = neighbors[pos] - index_offset;
      if (nb < 0 || nb >= n || nb == r || !((double)weights[pos] >= min_w)) continue;
      rows[m] = r;
      tails[m] = nb;
      const double w = (double)weights[pos] > 1e-6 ? (double)weights[pos] : 1e-6;
      eps[m] = (float)(max_w / w);
      ++m;
    }
  return m > 0 ? m : -2;
}

/* ------------------------------ hogwild ---------------------------------- */
typedef struct {
  const int* rows;
  const int* tails;
  const float* eps;
  const float* eps_neg;
  float* next;
  float* next_neg;
  float* x;
  float* y;
  int n, n_epochs, nsr, seed, threads;
  float af, bf, attr_c, rep_c, alpha0, alpha_step;
  size_t* begin;
  pthread_barrier_t* barrier;
} hog_ctx;

typedef struct {
  hog_ctx* c;
  int t;
} hog_arg;

static void* hog_worker(void* p) {
  hog_arg* a = (hog_arg*)p;
  hog_ctx* c = a->c;
  const int t = a->t;
  const size_t begin = c->begin[t], end = c->begin[t + 1];
  const float eps = FLT_EPSILON;
  float alpha = c->alpha0;
  for (int epoch = 0; epoch < c->n_epochs; ++epoch) {
    tau_t prng = tau_for(c->seed, epoch, end, t);
    for (size_t i = begin; i < end; ++i) {
      if (c->next[i] > (float)epoch) continue;
      const int j = c->rows[i], k = c->tails[i];
      const float dx = atomic_loadf(&c->x[j]) - atomic_loadf(&c->x[k]);
      const float dy = atomic_loadf(&c->y[j]) - atomic_loadf(&c->y[k]);
      const float s = dx * dx + dy * dy;
      const float d2 = eps < s ? s : eps;
      const float dp = pow_interp_f(d2, c->bf);
      const float coef = c->attr_c * dp / (d2 * (c->af * dp + 1.0f));
      const float gx = clip4f(coef * dx) * alpha, gy = clip4f(coef * dy) * alpha;
      atomic_addf(&c->x[j], gx);
      atomic_addf(&c->y[j], gy);
      atomic_addf(&c->x[k], -gx);
      atomic_addf(&c->y[k], -gy);
      c->next[i] += c->eps[i];
      int nneg = 0;
      if (c->nsr > 0 && (float)epoch >= c->next_neg[i]) {
        nneg = (int)floorf(((float)epoch - c->next_neg[i]) / c->eps_neg[i]);
        if (nneg < 0) nneg = 0;
      }
      for (int q = 0; q < nneg; ++q) {
        const int v = (int)(tau_next(&prng) % (uint32_t)c->n);
        if (v == j) continue;
        const float ndx = atomic_loadf(&c->x[j]) - atomic_loadf(&c->x[v]);
        const float ndy = atomic_loadf(&c->y[j]) - atomic_loadf(&c->y[v]);
        const float ns = ndx * ndx + ndy * ndy;
        const float nd2 = eps < ns ? ns : eps;
        const float np = pow_interp_f(nd2, c->bf);
        const float rep = c->rep_c / ((0.001f + nd2) * (c->af * np + 1.0f));
        atomic_addf(&c->x[j], clip4f(rep * ndx) * alpha);
        atomic_addf(&c->y[j], clip4f(rep * ndy) * alpha);
      }
      if (nneg > 0) c->next_neg[i] += (float)nneg * c->eps_neg[i];
    }
    if ((epoch + 1) % 50 == 0 || epoch == c->n_epochs - 1) pthread_barrier_wait(c->barrier);
    alpha -= c->alpha_step;
  }
  return NULL;
}

/* shared schedule state of the three forms */
static int schedule_init(long long m, const float* eps, int nsr, float** next, float** eps_neg,
                         float** next_neg) {
  *next = (float*)malloc(sizeof(float) * (size_t)m);
  *eps_neg = (float*)malloc(sizeof(float) * (size_t)m);
  *next_neg = (float*)malloc(sizeof(float) * (size_t)m);
  if (!*next || !*eps_neg || !*next_neg) return 6;
  for (long long i = 0; i < m; ++i) {
    (*next)[i] = eps[i];
    if (nsr > 0) {
      (*eps_neg)[i] = eps[i] / (float)nsr;
      (*next_neg)[i] = (*eps_neg)[i];
    } else {
      (*eps_neg)[i] = INFINITY;
      (*next_neg)[i] = INFINITY;
    }
  }
  return 0;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int run_hogwild(long long m, const int* rows, const int* tails, const float* eps, int n,
                       float* layout, int n_epochs, int nsr, double a, double b, double gamma,
                       double lr, int seed, int threads) {
  float *next, *eps_neg, *next_neg;
  if (schedule_init(m, eps, nsr, &next, &eps_neg, &next_neg)) return 6;
  float* x = (float*)malloc(sizeof(float) * (size_t)n);
  float* y = (float*)malloc(sizeof(float) * (size_t)n);
  size_t* begin = (size_t*)malloc(sizeof(size_t) * (size_t)(threads + 1));
  for (int i = 0; i < n; ++i) {
    x[i] = layout[2 * (size_t)i];
    y[i] = layout[2 * (size_t)i + 1];
  }
  begin[0] = 0;
  begin[threads] = (size_t)m;
  for (int t = 1; t < threads; ++t) {
    size_t bd = (size_t)m * (size_t)t / (size_t)threads;
    while (bd < (size_t)m && rows[bd] == rows[bd - 1]) ++bd;
    begin[t] = bd;
  }
  pthread_barrier_t barrier;
  pthread_barrier_init(&barrier, NULL, (unsigned)threads);
  hog_ctx c = {rows, tails, eps, eps_neg, next, next_neg, x, y, n, n_epochs, nsr, seed, threads,
               (float)a, (float)b, (float)(-2.0 * a * b), (float)(2.0 * gamma * b), (float)lr,
               n_epochs > 0 ? (float)lr / (float)n_epochs : 0.0f, begin, &barrier};
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  hog_arg* args = (hog_arg*)malloc(sizeof(hog_arg) * (size_t)threads);
  for (int t = 0; t < threads; ++t) args[t] = (hog_arg){&c, t};
  for (int t = 1; t < threads; ++t) pthread_create(&th[t], NULL, hog_worker, &args[t]);
  hog_worker(&args[0]);
  for (int t = 1; t < threads; ++t) pthread_join(th[t], NULL);
  pthread_barrier_destroy(&barrier);
  for (int i = 0; i < n; ++i) {
    layout[2 * (size_t)i] = x[i];
    layout[2 * (size_t)i + 1] = y[i];
  }
  free(th); free(args); free(begin); free(x); free(y); free(next); free(eps_neg); free(next_neg);
  return 0;
}

/* --------------------------- FP64 edge updates --------------------------- */
static double attract_coef_d(double d2, double a, double b) {
  if (!(d2 > 0.0)) return 0.0;
  const double p = pow_interp_d(d2, b);
  return -2.0 * a * b * (p / d2) / (a * p + 1.0);
}

static double repel_coef_d(double d2, double a, double b, double gamma) {
  if (!(d2 > 0.0)) return 0.0;
  return 2.0 * gamma * b / ((0.001 + d2) * (a * pow_interp_d(d2, b) + 1.0));
}

static int neg_count(int epoch, float next_neg, float eps_neg, int nsr) {
  int nneg = 0;
  if (nsr > 0 && (float)epoch >= next_neg) {
    nneg = (int)floorf(((float)epoch - next_neg) / eps_neg);
    if (nneg < 0) nneg = 0;
  }
  return nneg;
}

/* serial form (:3458-3557), 2D branch */
static int run_serial(long long m, const int* rows, const int* tails, const float* eps, int n,
                      float* layout, int n_epochs, int nsr, double a, double b, double gamma,
                      double lr, int seed) {
  float *next, *eps_neg, *next_neg;
  if (schedule_init(m, eps, nsr, &next, &eps_neg, &next_neg)) return 6;
  double* ex = (double*)malloc(sizeof(double) * (size_t)n);
  double* ey = (double*)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    ex[i] = layout[2 * (size_t)i];
    ey[i] = layout[2 * (size_t)i + 1];
  }
  for (int epoch = 0; epoch < n_epochs; ++epoch) {
    const double alpha = lr * (1.0 - (double)epoch / n_epochs);
    for (long long i = 0; i < m; ++i) {
      if (next[i] > (float)epoch) continue;
      const int j = rows[i], k = tails[i];
      const double dx = ex[j] - ex[k], dy = ey[j] - ey[k];
      const double coef = attract_coef_d(dx * dx + dy * dy, a, b);
      const double gx = clipd(coef * dx), gy = clipd(coef * dy);
      ex[j] += gx * alpha;
      ey[j] += gy * alpha;
      ex[k] -= gx * alpha;
      ey[k] -= gy * alpha;
      next[i] += eps[i];
      const int nneg = neg_count(epoch, next_neg[i], eps_neg[i], nsr);
      for (int q = 0; q < nneg; ++q) {
        const int v = det_vertex(n, seed, epoch, (size_t)i, q);
        if (v == j) continue;
        const double ndx = ex[j] - ex[v], ndy = ey[j] - ey[v];
        const double rep = repel_coef_d(ndx * ndx + ndy * ndy, a, b, gamma);
        ex[j] += clipd(rep * ndx) * alpha;
        ey[j] += clipd(rep * ndy) * alpha;
      }
      if (nneg > 0) next_neg[i] += (float)nneg * eps_neg[i];
    }
  }
  for (int i = 0; i < n; ++i) {
    layout[2 * (size_t)i] = (float)ex[i];
    layout[2 * (size_t)i + 1] = (float)ey[i];
  }
  free(ex); free(ey); free(next); free(eps_neg); free(next_neg);
  return 0;
}

/* add_delta_2d (:2663-2679) */
static void add_delta(float* delta, int n, int i, double dx, double dy, double scale, int other_too,
                      int other) {
  const double gx = clipd(scale * dx), gy = clipd(scale * dy);
  delta[i] += (float)gx;
  delta[(size_t)n + i] += (float)gy;
  if (other_too) {
    delta[other] -= (float)gx;
    delta[(size_t)n + other] -= (float)gy;
  }
}

/*
 * The delta form (:3353-3456).  Every thread reads the layout as it stood at
 * the start of the sync batch and writes only its own delta and its own edges'
 * schedule, and thread 0 adds the deltas in thread order after a barrier, so
 * running the threads one after another gives the reference's result.
 */
static int run_delta(long long m, const int* rows, const int* tails, const float* eps, int n,
                     float* layout, int n_epochs, int nsr, double a, double b, double gamma,
                     double lr, int seed, int threads) {
  float *next, *eps_neg, *next_neg;
  if (schedule_init(m, eps, nsr, &next, &eps_neg, &next_neg)) return 6;
  const int batches = n >= 3000 ? (2 * threads < 2 ? 2 : (2 * threads > 8 ? 8 : 2 * threads))
                                : (threads < 4 ? threads : 4);
  float* deltas = (float*)calloc((size_t)threads * 2 * (size_t)n, sizeof(float));
  double* ex = (double*)malloc(sizeof(double) * (size_t)n);
  double* ey = (double*)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    ex[i] = layout[2 * (size_t)i];
    ey[i] = layout[2 * (size_t)i + 1];
  }
  const size_t um = (size_t)m;
  for (int epoch = 0; epoch < n_epochs; ++epoch) {
    const double alpha = lr * (1.0 - (double)epoch / n_epochs);
    for (int batch = 0; batch < batches; ++batch) {
      const size_t b0 = um * (size_t)batch / (size_t)batches;
      const size_t b1 = um * (size_t)(batch + 1) / (size_t)batches;
      for (int t = 0; t < threads; ++t) {
        float* delta = deltas + (size_t)t * 2 * (size_t)n;
        memset(delta, 0, sizeof(float) * 2 * (size_t)n);
        const size_t lo = b0 + (b1 - b0) * (size_t)t / (size_t)threads;
        const size_t hi = b0 + (b1 - b0) * (size_t)(t + 1) / (size_t)threads;
        for (size_t i = lo; i < hi; ++i) {
          if (next[i] > (float)epoch) continue;
          const int j = rows[i], k = tails[i];
          const double dx = ex[j] - ex[k], dy = ey[j] - ey[k];
          add_delta(delta, n, j, dx, dy, attract_coef_d(dx * dx + dy * dy, a, b), 1, k);
          next[i] += eps[i];
          const int nneg = neg_count(epoch, next_neg[i], eps_neg[i], nsr);
          for (int q = 0; q < nneg; ++q) {
            const int v = det_vertex(n, seed, epoch, i, q);
            if (v == j) continue;
            const double ndx = ex[j] - ex[v], ndy = ey[j] - ey[v];
            add_delta(delta, n, j, ndx, ndy, repel_coef_d(ndx * ndx + ndy * ndy, a, b, gamma), 0, v);
          }
          if (nneg > 0) next_neg[i] += (float)nneg * eps_neg[i];
        }
      }
      for (int t = 0; t < threads; ++t) {
        const float* delta = deltas + (size_t)t * 2 * (size_t)n;
        for (int i = 0; i < n; ++i) {
          ex[i] += alpha * delta[i];
          ey[i] += alpha * delta[(size_t)n + i];
        }
      }
    }
  }
  for (int i = 0; i < n; ++i) {
    layout[2 * (size_t)i] = (float)ex[i];
    layout[2 * (size_t)i + 1] = (float)ey[i];
  }
  free(deltas); free(ex); free(ey); free(next); free(eps_neg); free(next_neg);
  return 0;
}

/*
 * The whole optimizer from the active edges: `form` -1 picks as the
 * reference does from n_threads and n, 0 serial, 1 delta, 2 hogwild (with
 * `threads` threads after the reference's cap).  layout: row-major n x 2
 * float, in place (the reference's FP64 forms round back to float here).
 * *seconds: wall time.  Returns the form used, or < 0 on bad arguments.
 */
int or_umap_cpu_optimize(long long m, const int* rows, const int* tails, const float* eps, int n,
                         float* layout, int n_epochs, int nsr, double a, double b, double gamma,
                         double lr, int seed, int n_threads, int form, double* seconds) {
  if (m < 1 || n < 2 || n_epochs < 0 || n_threads < 1) return -1;
  int threads = n_threads < 4 ? n_threads : 4;
  if (m < threads) threads = (int)m;
  if (threads < 1) threads = 1;
  if (form < 0) form = threads > 1 ? (n >= 10000 ? 2 : 1) : 0;
  const double t0 = now_s();
  int rc = 0;
  if (n_epochs > 0) {
    if (form == 2) rc = run_hogwild(m, rows, tails, eps, n, layout, n_epochs, nsr, a, b, gamma, lr, seed, threads);
    else if (form == 1) rc = run_delta(m, rows, tails, eps, n, layout, n_epochs, nsr, a, b, gamma, lr, seed, threads);
    else rc = run_serial(m, rows, tails, eps, n, layout, n_epochs, nsr, a, b, gamma, lr, seed);
  }
  if (seconds) *seconds = now_s() - t0;
  return rc ? -rc : form;
}

/* helpers exposed for their pins against the reference's own code */
int or_umap_cpu_det_vertex(int n, int seed, int epoch, long long edge, int sample) {
  return det_vertex(n, seed, epoch, (size_t)edge, sample);
}

void or_umap_cpu_tau_vertices(int seed, int epoch, long long window_end, int thread_id, int n,
                              int count, int* out) {
  tau_t t = tau_for(seed, epoch, (size_t)window_end, thread_id);
  for (int q = 0; q < count; ++q) out[q] = (int)(tau_next(&t) % (uint32_t)n);
}

double or_umap_cpu_pow(double x, double b) { return pow_interp_d(x, b); }
