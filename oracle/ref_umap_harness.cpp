// extern "C" access to the reference's UMAP CPU functions compiled by
// build_ref.sh from src/fast_knn_umap.cpp (the line ranges it lists sit in
// namespace ref_umap above this file), and the reference CPU hogwild epochs
// of optimize_layout_csr (:3088-3558) around the reference's own per-epoch
// edge loop (:3273-3325), which build_ref.sh inserts at the marker below.
//
// TEST INFRASTRUCTURE ONLY: used by the tests to pin the oracle, and by
// bench.py's CPU baseline leg; the product never loads it.
#include <cstdint>

using namespace ref_umap;

extern "C" {

float ref_umap_powf_fast(float x, float b) { return umap_powf_fast(x, b); }
double ref_umap_pow(double x, double b) { return umap_pow(x, b); }
float ref_clip4f(float x) { return clip4f(x); }
uint32_t ref_mix_seed(uint32_t x) { return mix_seed(x); }

int ref_deterministic_vertex(int n, int seed, int epoch, long long edge, int sample) {
  return deterministic_vertex(n, seed, epoch, static_cast<std::size_t>(edge), sample);
}

int ref_deterministic_vertex_cuda_style(int n, int seed, int epoch, int head, long long edge,
                                        int sample) {
  return deterministic_vertex_cuda_style(n, seed, epoch, head, static_cast<std::size_t>(edge),
                                         sample);
}

int ref_positive_samples(float period, int epoch) {
  return positive_samples_this_epoch_umap_schedule(period, epoch);
}

int ref_negative_samples(float period, int negative_sample_rate, int epoch) {
  return negative_samples_this_epoch_umap_schedule(period, negative_sample_rate, epoch);
}

// `count` negative vertices from the TauPrng of (seed, epoch, window_end,
// thread_id).
void ref_tau_vertices(int seed, int epoch, long long window_end, int thread_id, int n, int count,
                      int* out) {
  TauPrng prng = make_tau_prng(seed, epoch, static_cast<std::size_t>(window_end), thread_id);
  for (int t = 0; t < count; ++t) out[t] = prng.vertex(n);
}

void ref_find_ab(double spread, double min_dist, double* a, double* b) {
  const auto ab = find_ab_params(spread, min_dist);
  *a = ab.first;
  *b = ab.second;
}

int ref_effective_cpu_threads(int n_threads, int n_items) {
  return effective_cpu_threads(n_threads, n_items);
}

// The hogwild epochs of optimize_layout_csr (taken there for 2D, more than one
// thread and n >= 10000; any n and thread count here).  The active edges
// (rows, tails, epochs_per_sample) are given in the reference's order, CSR
// row by row; xy is row-major n x 2 float (in place).  The setup restates
// :3166-3182 (first samples at one period, negative period = period / rate)
// and :3199-3264 (float constants, atomic coordinates, worker ranges cut at
// row boundaries, alpha stepped down each epoch, a barrier every 50 epochs).
int ref_umap_hogwild(const int* active_rows_in, const int* active_tails_in,
                     const float* epochs_per_sample_in, long long n_active, int n, float* xy,
                     int n_epochs, int negative_sample_rate, double a, double b, double gamma,
                     double learning_rate, int seed, int threads, double* seconds) {
  if (n < 2 || n_active < 1 || threads < 1 || n_epochs < 0) return 1;
  const std::size_t na = static_cast<std::size_t>(n_active);
  const std::vector<int> active_rows(active_rows_in, active_rows_in + na);
  const std::vector<int> active_tails(active_tails_in, active_tails_in + na);
  const std::vector<float> epochs_per_sample(epochs_per_sample_in, epochs_per_sample_in + na);
  std::vector<float> epoch_of_next_sample(na), epochs_per_negative_sample(na),
      epoch_of_next_negative_sample(na);
  for (std::size_t i = 0; i < na; ++i) {
    epoch_of_next_sample[i] = epochs_per_sample[i];
    if (negative_sample_rate > 0) {
      epochs_per_negative_sample[i] = epochs_per_sample[i] / negative_sample_rate;
      epoch_of_next_negative_sample[i] = epochs_per_negative_sample[i];
    } else {
      epochs_per_negative_sample[i] = std::numeric_limits<float>::infinity();
      epoch_of_next_negative_sample[i] = std::numeric_limits<float>::infinity();
    }
  }
  const int* active_rows_ptr = active_rows.data();
  const int* active_tails_ptr = active_tails.data();
  const float* epochs_per_sample_ptr = epochs_per_sample.data();
  const float* epochs_per_negative_sample_ptr = epochs_per_negative_sample.data();
  const float af = static_cast<float>(a);
  const float bf = static_cast<float>(b);
  const float attraction_const = static_cast<float>(-2.0 * a * b);
  const float repulsion_const = static_cast<float>(2.0 * gamma * b);
  const float alpha0 = static_cast<float>(learning_rate);
  const float alpha_step = n_epochs > 0 ? alpha0 / static_cast<float>(n_epochs) : 0.0f;
  const float eps = std::numeric_limits<float>::epsilon();
  std::vector<std::atomic<float>> emb_xf(static_cast<std::size_t>(n));
  std::vector<std::atomic<float>> emb_yf(static_cast<std::size_t>(n));
  for (int i = 0; i < n; ++i) {
    emb_xf[static_cast<std::size_t>(i)].store(xy[2 * static_cast<std::size_t>(i)],
                                              std::memory_order_relaxed);
    emb_yf[static_cast<std::size_t>(i)].store(xy[2 * static_cast<std::size_t>(i) + 1],
                                              std::memory_order_relaxed);
  }
  std::atomic<float>* x = emb_xf.data();
  std::atomic<float>* y = emb_yf.data();
  ReusableBarrier barrier(threads);
  std::vector<std::size_t> worker_begin(static_cast<std::size_t>(threads) + 1u, 0u);
  worker_begin[static_cast<std::size_t>(threads)] = na;
  for (int t = 1; t < threads; ++t) {
    std::size_t boundary = na * static_cast<std::size_t>(t) / threads;
    while (boundary < na && active_rows_ptr[boundary] == active_rows_ptr[boundary - 1u])
      ++boundary;
    worker_begin[static_cast<std::size_t>(t)] = boundary;
  }
  auto run_worker = [&](const int t) {
    const std::size_t begin = worker_begin[static_cast<std::size_t>(t)];
    const std::size_t end = worker_begin[static_cast<std::size_t>(t + 1)];
    float alpha = alpha0;
    for (int epoch = 0; epoch < n_epochs; ++epoch) {
// REF_HOGWILD_EDGE_LOOP
      if ((epoch + 1) % 50 == 0 || epoch == n_epochs - 1) barrier.wait();
      alpha -= alpha_step;
    }
  };
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> workers;
  for (int t = 1; t < threads; ++t) workers.emplace_back(run_worker, t);
  run_worker(0);
  for (auto& w : workers) w.join();
  if (seconds)
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int i = 0; i < n; ++i) {
    xy[2 * static_cast<std::size_t>(i)] = x[i].load(std::memory_order_relaxed);
    xy[2 * static_cast<std::size_t>(i) + 1] = y[i].load(std::memory_order_relaxed);
  }
  return 0;
}

}  // extern "C"
