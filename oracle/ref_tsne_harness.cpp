// extern "C" access to the reference's t-SNE CPU functions compiled by
// build_ref.sh from src/tsne_neighbors.cpp (the line ranges it lists sit in
// namespace ref_tsne above this file in the same translation unit).
//
// TEST INFRASTRUCTURE ONLY: used by the tests to pin the oracle and the CPU
// port, and by bench.py's cpu_baseline leg; the product never loads it.
//
// The only logic here is the sequencing the reference's driver does around
// those functions: knn_tsne_opentsne_float_cpp (src/tsne_neighbors.cpp:
// 1871-2121) creates a ParallelExecutor of `threads` workers and scopes it
// (:1912-1913), and run_phase (:2002-2012) does, per iteration,
// compute_gradient_f, apply_open_tsne_update_f and zero_mean_f.
#include <cstdint>

namespace ref_tsne {
namespace {

SparseProbabilitiesF make_csr(int n, const int* row_ptr, const int* col, const float* val) {
  SparseProbabilitiesF p;
  const int nnz = row_ptr[n];
  p.row_ptr.assign(row_ptr, row_ptr + n + 1);
  p.col.assign(col, col + nnz);
  p.val.assign(val, val + nnz);
  return p;
}

double seconds_since(const std::chrono::steady_clock::time_point& t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace
}  // namespace ref_tsne

using namespace ref_tsne;

extern "C" {

// compute_row_probabilities_float for every row of a row-major n x k matrix
// of (non-squared) neighbour distances.
void ref_row_probabilities_float(const float* dist, int n, int k, double perplexity, float* out) {
  const std::vector<float> d(dist, dist + static_cast<std::size_t>(n) * k);
  for (int i = 0; i < n; ++i)
    compute_row_probabilities_float(d, i, k, perplexity, out + static_cast<std::size_t>(i) * k);
}

uint32_t ref_mix_uint32(uint32_t v) { return mix_uint32(v); }

// `count` values of xorshift32 from `state`, and `count` draws of
// uniform_index(state, n) from the same start state.
void ref_xorshift_stream(uint32_t state, int count, int n, uint32_t* raw, int* index) {
  uint32_t a = state, b = state;
  for (int t = 0; t < count; ++t) {
    raw[t] = xorshift32(a);
    index[t] = uniform_index(b, n);
  }
}

int ref_fft_grid_size(int n) { return tsne_fft_grid_size(n); }

// The optimizer gradient of a fixed row-major n x 2 layout: FFT-grid
// repulsion (grid > 0: that grid; grid == 0: the reference's CPU rule) or the
// exact pair form (grid < 0), plus the sparse attraction.
int ref_tsne_gradient(int n, const int* row_ptr, const int* col, const float* val,
                      const float* layout, float exaggeration, int grid, int threads,
                      float* grad_out) {
  if (n < 2 || threads < 1) return 1;
  const SparseProbabilitiesF p = make_csr(n, row_ptr, col, val);
  const std::vector<float> y(layout, layout + 2 * static_cast<std::size_t>(n));
  std::vector<float> grad(y.size(), 0.0f);
  ParallelExecutor executor(threads);
  ParallelExecutorScope scope(&executor);
  if (grid < 0) {
    compute_gradient_pair_symmetric_f(p, y, n, 2, exaggeration, threads, grad);
  } else {
    FftGridWorkspaceT<float> ws;
    compute_gradient_fft_grid_f(p, y, n, 2, exaggeration, threads, &ws, grad, grid);
  }
  std::copy(grad.begin(), grad.end(), grad_out);
  return 0;
}

double ref_tsne_sum_q(const float* layout, int n, int threads) {
  const std::vector<float> y(layout, layout + 2 * static_cast<std::size_t>(n));
  ParallelExecutor executor(threads);
  ParallelExecutorScope scope(&executor);
  return static_cast<double>(compute_sum_q_f(y, n, 2, threads));
}

double ref_tsne_kl(int n, const int* row_ptr, const int* col, const float* val,
                   const float* layout, int threads) {
  const SparseProbabilitiesF p = make_csr(n, row_ptr, col, val);
  const std::vector<float> y(layout, layout + 2 * static_cast<std::size_t>(n));
  ParallelExecutor executor(threads);
  ParallelExecutorScope scope(&executor);
  return evaluate_kl_f(p, y, n, 2, threads);
}

// `iters` iterations of one phase on the row-major n x 2 layout (in place),
// with the optimizer state `update` and `gains` (n x 2, in place): the
// reference's run_phase body.  grid as for ref_tsne_gradient.  *seconds gets
// the wall time of the iterations.
int ref_tsne_iterate(int n, const int* row_ptr, const int* col, const float* val, float* layout,
                     float* update_io, float* gains_io, int iters, float exaggeration,
                     float learning_rate, float momentum, float min_gain, float max_step_norm,
                     int grid, int threads, double* seconds) {
  if (n < 2 || threads < 1 || iters < 0) return 1;
  const SparseProbabilitiesF p = make_csr(n, row_ptr, col, val);
  const std::size_t m = 2 * static_cast<std::size_t>(n);
  std::vector<float> y(layout, layout + m), update(update_io, update_io + m),
      gains(gains_io, gains_io + m), grad(m, 0.0f);
  ParallelExecutor executor(threads);
  ParallelExecutorScope scope(&executor);
  FftGridWorkspaceT<float> ws;
  const auto t0 = std::chrono::steady_clock::now();
  for (int it = 0; it < iters; ++it) {
    if (grid < 0)
      compute_gradient_pair_symmetric_f(p, y, n, 2, exaggeration, threads, grad);
    else
      compute_gradient_fft_grid_f(p, y, n, 2, exaggeration, threads, &ws, grad, grid);
    apply_open_tsne_update_f(y, update, gains, grad, n, 2, learning_rate, momentum, min_gain,
                             max_step_norm, threads);
    zero_mean_f(y, n, 2);
  }
  if (seconds) *seconds = seconds_since(t0);
  std::copy(y.begin(), y.end(), layout);
  std::copy(update.begin(), update.end(), update_io);
  std::copy(gains.begin(), gains.end(), gains_io);
  return 0;
}

}  // extern "C"
