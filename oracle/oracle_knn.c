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
