/* Exact Euclidean k-nearest neighbours, FP64 brute force (test oracle).
 *
 * The reference's exact route (native_cuda_knn_impl,
 * src/native_knn_cuda_impl.cpp:520-700) asks an external brute-force search
 * (FAISS bfKnn or cuVS brute_force, not part of the reference) for the k + 1
 * nearest rows of every row and then finalize_cuvs_knn_kernel
 * (src/embedding_cuda_kernels.cpp:1428-1480) drops the row itself, keeps the
 * first k others and turns squared distances into distances.  The result
 * this restates is the mathematical one that search approximates: for every
 * query row q, the k rows j != q with the smallest squared distance
 * sum_c (x[q][c] - x[j][c])^2, formed in double from the float32 inputs in
 * column order, ties by the smaller row index, distances as
 * (float) sqrt(d2).
 *
 * data: column-major n x p float32 (element (i, c) at c * n + i).
 * queries: nq row numbers (NULL: every row, nq = n).
 * out_idx, out_dist: row-major nq x k, 0-based indices.
 */
#include <math.h>
#include <stdlib.h>

typedef struct {
  double d;
  int j;
} or_nb;

/* a before b: smaller distance, then smaller index */
static int or_nb_less(or_nb a, or_nb b) { return a.d < b.d || (a.d == b.d && a.j < b.j); }

/* max-heap on (d, j): the worst kept neighbour at the root */
static void or_sift_down(or_nb* h, int m, int i) {
  for (;;) {
    int l = 2 * i + 1, r = l + 1, w = i;
    if (l < m && or_nb_less(h[w], h[l])) w = l;
    if (r < m && or_nb_less(h[w], h[r])) w = r;
    if (w == i) return;
    or_nb t = h[i];
    h[i] = h[w];
    h[w] = t;
    i = w;
  }
}

static void or_sift_up(or_nb* h, int i) {
  while (i > 0) {
    int p = (i - 1) / 2;
    if (!or_nb_less(h[p], h[i])) return;
    or_nb t = h[i];
    h[i] = h[p];
    h[p] = t;
    i = p;
  }
}

static int or_nb_cmp(const void* a, const void* b) {
  const or_nb* x = (const or_nb*)a;
  const or_nb* y = (const or_nb*)b;
  if (or_nb_less(*x, *y)) return -1;
  if (or_nb_less(*y, *x)) return 1;
  return 0;
}

int or_exact_knn(const float* data, int n, int p, int k, const int* queries, int nq,
                 int* out_idx, float* out_dist) {
  if (n < 2 || p < 1 || k < 1 || k >= n) return 1;
  double* d2 = (double*)malloc(sizeof(double) * (size_t)n);
  or_nb* heap = (or_nb*)malloc(sizeof(or_nb) * (size_t)k);
  if (!d2 || !heap) {
    free(d2);
    free(heap);
    return 1;
  }
  for (int t = 0; t < nq; ++t) {
    const int q = queries ? queries[t] : t;
    for (int j = 0; j < n; ++j) d2[j] = 0.0;
    /* column by column: every d2[j] still sums its terms in column order */
    for (int c = 0; c < p; ++c) {
      const float* col = data + (size_t)c * n;
      const double xq = (double)col[q];
      for (int j = 0; j < n; ++j) {
        const double diff = (double)col[j] - xq;
        d2[j] = d2[j] + diff * diff;
      }
    }
    int m = 0;
    for (int j = 0; j < n; ++j) {
      if (j == q) continue;
      or_nb e = {d2[j], j};
      if (m < k) {
        heap[m] = e;
        or_sift_up(heap, m);
        ++m;
      } else if (or_nb_less(e, heap[0])) {
        heap[0] = e;
        or_sift_down(heap, m, 0);
      }
    }
    qsort(heap, (size_t)m, sizeof(or_nb), or_nb_cmp);
    for (int r = 0; r < k; ++r) {
      out_idx[(size_t)t * k + r] = heap[r].j;
      out_dist[(size_t)t * k + r] = (float)sqrt(heap[r].d);
    }
  }
  free(d2);
  free(heap);
  return 0;
}

/* KNN structure score (knn_structure_score_cpp, src/embedding_utils.cpp:
 * 1386-1580; the GPU form structure_score_rows_kernel,
 * src/embedding_cuda_kernels.cpp:1827-1944).  layout: column-major n x dim
 * double; indices: column-major index_rows x limit, 1-based; keep: keep_n
 * 1-based rows; compact: indices row kk belongs to keep[kk] (else to the
 * row itself); labels: n factor codes (NULL or n_levels = 0: none).  For
 * every kept row: its first preserve_k high-dimensional neighbours (the
 * targets) and its preserve_k nearest rows in the layout by (d2, index);
 * preservation, local trustworthiness and continuity as the reference forms
 * them, and the majority label of the layout neighbours (ties by the
 * smaller label; rows whose own label is outside 1..n_levels are not
 * counted, as the GPU form).  totals: the six sums of the GPU form
 * (preservation, trust, continuity, label hits, label rows, scored rows),
 * added in row order. */
static double or_layout_d2(const double* y, int n, int dim, int i, int j) {
  double d2 = 0.0;
  for (int c = 0; c < dim; ++c) {
    const double diff = y[(size_t)c * n + i] - y[(size_t)c * n + j];
    d2 = d2 + diff * diff;
  }
  return d2;
}

int or_knn_structure_score(const double* layout, int n, int dim, const int* indices,
                           int index_rows, int limit, const int* keep, int keep_n,
                           int compact, int preserve_k, const int* labels, int n_levels,
                           double* totals) {
  if (n < 2 || preserve_k < 1 || preserve_k > limit) return 1;
  for (int m = 0; m < 6; ++m) totals[m] = 0.0;
  or_nb* low = (or_nb*)malloc(sizeof(or_nb) * (size_t)preserve_k);
  or_nb* tg = (or_nb*)malloc(sizeof(or_nb) * (size_t)preserve_k);
  int* cnt = (int*)malloc(sizeof(int) * (size_t)(n_levels + 1));
  if (!low || !tg || !cnt) {
    free(low);
    free(tg);
    free(cnt);
    return 1;
  }
  for (int kk = 0; kk < keep_n; ++kk) {
    const int q = keep[kk] - 1;
    if (q < 0 || q >= n) continue;
    const int row = compact ? kk : q;
    if (row < 0 || row >= index_rows) continue;
    int T = 0;
    for (int r = 0; r < preserve_k; ++r) {
      const int nb = indices[(size_t)r * index_rows + row] - 1;
      if (nb < 0 || nb >= n) continue;
      or_nb e = {or_layout_d2(layout, n, dim, q, nb), nb};
      tg[T++] = e;
    }
    qsort(tg, (size_t)T, sizeof(or_nb), or_nb_cmp);
    /* the preserve_k nearest rows of the layout */
    int m = 0;
    for (int c = 0; c < n; ++c) {
      if (c == q) continue;
      or_nb e = {or_layout_d2(layout, n, dim, q, c), c};
      if (m < preserve_k) {
        low[m] = e;
        or_sift_up(low, m);
        ++m;
      } else if (or_nb_less(e, low[0])) {
        low[0] = e;
        or_sift_down(low, m, 0);
      }
    }
    if (m < preserve_k) continue;
    qsort(low, (size_t)m, sizeof(or_nb), or_nb_cmp);
    int shared = 0;
    double trust_penalty = 0.0;
    for (int r = 0; r < preserve_k; ++r) {
      int rank = limit + 1;
      for (int s = 0; s < limit; ++s)
        if (indices[(size_t)s * index_rows + row] - 1 == low[r].j) {
          rank = s + 1;
          break;
        }
      if (rank <= preserve_k) ++shared;
      if (rank > preserve_k) trust_penalty += (double)(rank - preserve_k);
    }
    /* low rank of each target: 1 + the rows that come before it */
    double cont_penalty = 0.0;
    for (int t = 0; t < T; ++t) {
      int before = 0;
      for (int c = 0; c < n; ++c) {
        if (c == q) continue;
        or_nb e = {or_layout_d2(layout, n, dim, q, c), c};
        if (or_nb_less(e, tg[t])) ++before;
      }
      const int low_rank = 1 + before;
      if (low_rank > preserve_k) cont_penalty += (double)(low_rank - preserve_k);
    }
    double hit = 0.0, hit_n = 0.0;
    if (labels && n_levels > 0) {
      const int truth = labels[q];
      if (truth >= 1 && truth <= n_levels) {
        for (int l = 0; l <= n_levels; ++l) cnt[l] = 0;
        for (int r = 0; r < preserve_k; ++r) {
          const int l = labels[low[r].j];
          if (l >= 1 && l <= n_levels) ++cnt[l];
        }
        int best = 0, best_count = 0;
        for (int l = 1; l <= n_levels; ++l)
          if (cnt[l] > best_count) {
            best_count = cnt[l];
            best = l;
          }
        if (best_count > 0) {
          hit = best == truth ? 1.0 : 0.0;
          hit_n = 1.0;
        }
      }
    }
    const double k = (double)preserve_k;
    const double trust_denom = k * (double)(limit + 1 - preserve_k > 1 ? limit + 1 - preserve_k : 1);
    const double cont_denom = k * (double)(n - preserve_k > 1 ? n - preserve_k : 1);
    double trust = 1.0 - trust_penalty / trust_denom;
    double cont = 1.0 - cont_penalty / cont_denom;
    trust = trust < 0.0 ? 0.0 : (trust > 1.0 ? 1.0 : trust);
    cont = cont < 0.0 ? 0.0 : (cont > 1.0 ? 1.0 : cont);
    totals[0] += (double)shared / k;
    totals[1] += trust;
    totals[2] += cont;
    totals[3] += hit;
    totals[4] += hit_n;
    totals[5] += 1.0;
  }
  free(low);
  free(tg);
  free(cnt);
  return 0;
}

/* Exact k-NN of query rows among data rows (no row is dropped): the result
 * the reference's exact query route approximates (native_cuda_query_knn_impl,
 * src/native_knn_cuda_impl.cpp:878-1010, with finalize_cuvs_query_knn_kernel,
 * src/embedding_cuda_kernels.cpp:1482-1522).  data: column-major n x p;
 * query: column-major m x p; out: row-major m x k, 0-based. */
int or_exact_query_knn(const float* data, int n, const float* query, int m, int p, int k,
                       int* out_idx, float* out_dist) {
  if (n < 1 || m < 1 || p < 1 || k < 1 || k > n) return 1;
  double* d2 = (double*)malloc(sizeof(double) * (size_t)n);
  or_nb* heap = (or_nb*)malloc(sizeof(or_nb) * (size_t)k);
  if (!d2 || !heap) {
    free(d2);
    free(heap);
    return 1;
  }
  for (int t = 0; t < m; ++t) {
    for (int j = 0; j < n; ++j) d2[j] = 0.0;
    for (int c = 0; c < p; ++c) {
      const float* col = data + (size_t)c * n;
      const double xq = (double)query[(size_t)c * m + t];
      for (int j = 0; j < n; ++j) {
        const double diff = (double)col[j] - xq;
        d2[j] = d2[j] + diff * diff;
      }
    }
    int h = 0;
    for (int j = 0; j < n; ++j) {
      or_nb e = {d2[j], j};
      if (h < k) {
        heap[h] = e;
        or_sift_up(heap, h);
        ++h;
      } else if (or_nb_less(e, heap[0])) {
        heap[0] = e;
        or_sift_down(heap, h, 0);
      }
    }
    qsort(heap, (size_t)h, sizeof(or_nb), or_nb_cmp);
    for (int r = 0; r < k; ++r) {
      out_idx[(size_t)t * k + r] = heap[r].j;
      out_dist[(size_t)t * k + r] = (float)sqrt(heap[r].d);
    }
  }
  free(d2);
  free(heap);
  return 0;
}
