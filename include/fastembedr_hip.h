/*
 * fastembedr_hip.h -- C-ABI of the MI355X (gfx950) embedding-optimizer
 * backend.  Plain pointers and sizes only; every function returns 0 on
 * success and 1 on failure (message from fastembedr_hip_last_error()), and
 * nothing throws across this boundary.
 *
 * Layout conventions are the reference CUDA ABI's
 * (src/embedding_cuda_impl.cpp:789-797, 933-945, 2435-2523):
 *   indices / distances  R column-major n x k: element (i, j) at j*n + i,
 *                        int32 ids with `index_offset` (0 or 1), Euclidean
 *                        (not squared), self excluded;
 *   init / out           row-major interleaved n x 2 float32.
 * Host-pointer entry points copy their inputs to the device and free every
 * device allocation before returning; plan handles own resident state until
 * destroyed.  Calls are synchronous and run on the current device.
 */
#ifndef FASTEMBEDR_HIP_H
#define FASTEMBEDR_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- availability / errors / policy ---------------------------------- */

/* Replaces fastembedr_cuda_available (src/embedding_cuda_kernels.cpp:3833-3841). */
int fastembedr_hip_available(void);

/* Replaces fastembedr_cuda_embedding_last_error (embedding_cuda_kernels.cpp:3829-3831). */
const char* fastembedr_hip_last_error(void);

/* Replaces fastembedr_cuda_opentsne_fft_grid_size (embedding_cuda_kernels.cpp:5625-5627):
 * 256 for n < 20000, 1024 for n >= 100000, else 512; FASTEMBEDR_TSNE_FFT_GRID
 * in {16, 32, ..., 1024} overrides (resolve_cuda_tsne_fft_grid_size :5610-5623). */
int fastembedr_hip_opentsne_fft_grid_size(int n);

/* ---- one-shot t-SNE (reference: embedding_cuda_kernels.cpp:6178-6264) ---
 * Same argument order as the CUDA functions plus `grid_size` (0 = the rule
 * above, -1 = exact pair repulsion, the reference CPU route for small n).
 * n_components must be 2; k <= 256.  learning_rate_auto != 0 uses
 * n / exaggeration of each phase (:6077-6082).  Without init the layout is
 * N(0, 1e-4^2) from std::mt19937(seed) (tsne_neighbors.cpp:1944-1950).      */

/* Replaces fastembedr_cuda_opentsne_fft_from_knn (double distances, :6178-6205). */
int fastembedr_hip_opentsne_fft_from_knn(const int* indices, const double* distances,
                                         const float* init, int has_init, int n, int k,
                                         int n_components, float perplexity,
                                         int early_exaggeration_iter, int n_iter,
                                         float early_exaggeration, float exaggeration,
                                         float learning_rate, int learning_rate_auto,
                                         float initial_momentum, float final_momentum,
                                         float min_gain, float max_step_norm, unsigned int seed,
                                         int index_offset, int grid_size, float* out);

/* Replaces fastembedr_cuda_opentsne_fft_from_knn_float (:6207-6234). */
int fastembedr_hip_opentsne_fft_from_knn_float(const int* indices, const float* distances,
                                               const float* init, int has_init, int n, int k,
                                               int n_components, float perplexity,
                                               int early_exaggeration_iter, int n_iter,
                                               float early_exaggeration, float exaggeration,
                                               float learning_rate, int learning_rate_auto,
                                               float initial_momentum, float final_momentum,
                                               float min_gain, float max_step_norm,
                                               unsigned int seed, int index_offset,
                                               int grid_size, float* out);

/* Replaces fastembedr_cuda_opentsne_fft_from_device_knn_float (:6236-6264):
 * indices/distances are device pointers (borrowed); init/out are host. */
int fastembedr_hip_opentsne_fft_from_device_knn_float(
    const int* indices, const float* distances, const float* init, int has_init, int n, int k,
    int n_components, float perplexity, int early_exaggeration_iter, int n_iter,
    float early_exaggeration, float exaggeration, float learning_rate, int learning_rate_auto,
    float initial_momentum, float final_momentum, float min_gain, float max_step_norm,
    unsigned int seed, int index_offset, int grid_size, float* out);

/* ---- force diagnostic ------------------------------------------------
 * HIP counterpart of opentsne_force_diagnostic_cpp (tsne_neighbors.cpp:
 * 1698-1817): from a fixed row-major n x 2 layout, the attractive force
 * (exaggeration included), the FFT-grid repulsive force and their sum (the
 * optimizer gradient, = objective gradient / 4), the grid estimate of
 * Z = sum_{i != j} q_ij, and optionally the symmetric P as 0-based CSR
 * (row_ptr n+1, col/val capacity 2*n*k; *nnz receives the count).  Host
 * pointers; any output may be NULL. */
int fastembedr_hip_tsne_gradient_diagnostic(const int* indices, const float* distances, int n,
                                            int k, int index_offset, float perplexity,
                                            float exaggeration, int grid_size,
                                            const float* layout, float* attractive,
                                            float* repulsive, float* gradient, double* sum_q,
                                            int* row_ptr, int* col, float* val,
                                            long long* nnz);

/* Replaces opentsne_force_diagnostic_cpp (tsne_neighbors.cpp:1698-1817), the
 * whole bundle: from a fixed row-major n x 2 layout (used as given, not
 * centred), the attractive force (exaggeration included), the exact pair
 * repulsion (compute_gradient_pair_symmetric_f, :1090-1150) and the FFT-grid
 * repulsion at `grid_size` (16..1024, a power of two), the two optimizer
 * gradients (attraction + repulsion; the objective gradient is 4 times
 * these), the exact Z = sum_{i != j} q_ij rounded to float (compute_sum_q_f,
 * :1034-1059), the KL divergence of P against q / Z (evaluate_kl_f,
 * :1338-1365), and optionally the symmetric P as 0-based CSR (row_ptr n+1,
 * col/val capacity 2*n*k).  P is the FP64-searched device P.  Host pointers;
 * any output may be NULL.  The exact terms cost O(n^2). */
int fastembedr_hip_opentsne_force_diagnostic(
    const int* indices, const float* distances, int n, int k, int index_offset, float perplexity,
    float exaggeration, int grid_size, const float* layout, float* attractive,
    float* repulsive_exact, float* gradient_exact, float* repulsive_fft, float* gradient_fft,
    double* sum_q, double* kl, int* row_ptr, int* col, float* val, long long* nnz);

/* ---- resident plans ----------------------------------------------------
 * No counterpart in the reference (which frees everything per call,
 * embedding_cuda_kernels.cpp:88-127): a plan keeps P, the layout, the
 * optimizer state, the FFT work arrays and captured iteration graphs on the
 * device so a caller can run phases, read the layout, time kernels, and
 * serve repeated requests without re-uploading.  Handles are opaque. */
typedef struct fastembedr_hip_tsne_plan fastembedr_hip_tsne_plan;

/* indices/distances host (device_inputs = 0) or device (1) pointers in the
 * column-major layout above; init is host row-major n x 2 or NULL.
 * grid_size: 0 = the grid rule, a power of two 16..1024, or -1 = exact pair
 * repulsion every iteration (the reference CPU route for n <= 3000,
 * compute_gradient_pair_symmetric_f, tsne_neighbors.cpp:1090-1150; O(n^2)). */
int fastembedr_hip_tsne_plan_create(const int* indices, const float* distances,
                                    int device_inputs, int n, int k, int index_offset,
                                    float perplexity, int grid_size, const float* init,
                                    unsigned int seed, fastembedr_hip_tsne_plan** plan);

/* As _plan_create, with the launch choices the autotuner picks per size
 * bucket (fastembedr_amd/tune.py): attraction 0 = built in (every pair once
 * where most pairs fall in staged tiles, else segment-blocked), 1 = row
 * groups, 2 = segment-blocked with LDS windows, 3 = every pair once;
 * pair_workgroups = workgroups of the pair form (0 = two per compute unit,
 * at most one per chunk).  The attraction form fixes the order of the
 * attraction sums (forms agree to about 1e-5, not bit for bit); the
 * workgroup count changes no bit. */
int fastembedr_hip_tsne_plan_create_shaped(const int* indices, const float* distances,
                                           int device_inputs, int n, int k, int index_offset,
                                           float perplexity, int grid_size, int attraction,
                                           int pair_workgroups, const float* init,
                                           unsigned int seed, fastembedr_hip_tsne_plan** plan);

/* Enqueue n_iter optimizer iterations with one phase's constants (chunks of
 * up to 50 iterations, each of 4 or more replayed from a captured hipGraph
 * unless FASTEMBEDR_HIP_GRAPH_CAPTURE=0).
 * Asynchronous: call _plan_sync or _plan_layout to wait. */
int fastembedr_hip_tsne_plan_run(fastembedr_hip_tsne_plan* plan, int n_iter, float exaggeration,
                                 float momentum, float learning_rate, float min_gain,
                                 float max_step_norm);

/* Captures, without running anything, the iteration graphs that
 * _plan_run(plan, n_iter, <same constants>) will replay when it is called
 * after `after` further iterations (runs are cut into chunks of at most 50
 * iterations; every chunk of 4 or more replays a graph keyed by its length,
 * constants and starting buffer parity).  Lets a caller move graph
 * instantiation out of a latency- or time-critical region. */
int fastembedr_hip_tsne_plan_prepare(fastembedr_hip_tsne_plan* plan, int n_iter,
                                     float exaggeration, float momentum, float learning_rate,
                                     float min_gain, float max_step_norm, long long after);

int fastembedr_hip_tsne_plan_sync(fastembedr_hip_tsne_plan* plan);

/* Centred layout, host row-major n x 2. */
int fastembedr_hip_tsne_plan_layout(fastembedr_hip_tsne_plan* plan, float* out);

/* info[0..7] = n, k, grid size G, nnz of P, lanes per CSR row, device bytes,
 * iterations run, CSR setup microseconds. */
int fastembedr_hip_tsne_plan_info(fastembedr_hip_tsne_plan* plan, long long* info);

/* The attraction kernel the plan chose: info[0] = 0 row groups, 1 segment-
 * blocked (layout reads), 2 segment-blocked (LDS windows), 3 every pair once
 * (staged tiles); info[1] = entries streamed per iteration (padded segments,
 * or staged pairs), info[2] = rest-stream entries (pair form), info[3] =
 * workgroups (groups, or chunk ranges), info[4] = chunks, info[5] = partial
 * slots (pair form; 0 otherwise).  info holds 6 values. */
int fastembedr_hip_tsne_plan_attraction_info(fastembedr_hip_tsne_plan* plan, long long* info);

/* Copies the plan's symmetric P to host: row_ptr (n+1), col (nnz, 0-based,
 * ascending per row), val (nnz) -- the affinity_row_ptr0 / affinity_col1 /
 * affinity_weight fields of opentsne_force_diagnostic_cpp
 * (tsne_neighbors.cpp:1793-1808), with 0-based columns. */
int fastembedr_hip_tsne_plan_affinities(fastembedr_hip_tsne_plan* plan, int* row_ptr, int* col,
                                        float* val);

/* KL divergence of the plan's P against the current layout with the exact Z
 * (evaluate_kl_f, tsne_neighbors.cpp:1338-1365, the cost the reference records
 * every 50 iterations with record_costs; opentsne_kl_diagnostic_cpp
 * :1656-1692), and that Z rounded to float.  Waits for queued iterations;
 * O(n^2) for Z. */
int fastembedr_hip_tsne_plan_kl(fastembedr_hip_tsne_plan* plan, double* kl, double* sum_q);

/* Runs n_iter eager iterations with hipEvents between the six kernels on
 * the plan's stream; kernel_ms[0..5] = mean ms of splat, conv_rows_fwd,
 * conv_cols, conv_rows_inv, attract, update; kernel_ms[6] = mean ms per
 * iteration. */
int fastembedr_hip_tsne_plan_profile(fastembedr_hip_tsne_plan* plan, int n_iter,
                                     float exaggeration, float momentum, float learning_rate,
                                     float min_gain, float max_step_norm, double* kernel_ms);

void fastembedr_hip_tsne_plan_destroy(fastembedr_hip_tsne_plan* plan);

/* ---- batched jobs ------------------------------------------------------
 * No reference counterpart (the reference runs one job per call).  Many
 * t-SNE jobs of uneven size advance together through job-table kernels:
 * every iteration of every resident job costs 1 + 3 * (#distinct grid sizes)
 * + (#distinct lane widths) launches, and each job gives the same bits as
 * its single-job plan.  Jobs are admitted while the device-memory budget
 * allows, run their own schedules, and are removed by the caller; _add and
 * _remove may be called between _step calls (continuous batching). */
typedef struct fastembedr_hip_tsne_job {
  const int* indices;     /* column-major n x k */
  const float* distances; /* column-major n x k */
  int n, k, index_offset, grid_size;
  float perplexity;
  const float* init; /* row-major n x 2 or NULL (mt19937(seed) N(0, 1e-4^2)) */
  unsigned int seed;
  int early_exaggeration_iter, n_iter;
  float early_exaggeration, exaggeration, learning_rate_early, learning_rate_normal;
  float initial_momentum, final_momentum, min_gain, max_step_norm;
  int attraction;      /* as fastembedr_hip_tsne_plan_create_shaped (0 = built in) */
  int pair_workgroups; /* as fastembedr_hip_tsne_plan_create_shaped (0 = built in) */
} fastembedr_hip_tsne_job;

typedef struct fastembedr_hip_tsne_batch fastembedr_hip_tsne_batch;

/* hbm_budget_bytes <= 0: no budget. */
int fastembedr_hip_tsne_batch_create(int max_jobs, long long hbm_budget_bytes,
                                     fastembedr_hip_tsne_batch** batch);

/* Device bytes a job of this shape holds while resident, and at its peak
 * during setup (for admission decisions). */
long long fastembedr_hip_tsne_job_bytes(int n, int k, int grid_size, int setup_peak);

/* Builds the job's P and state and makes it resident.  Returns 0 and the
 * slot, 2 when the budget or the slot table does not admit it now (retry
 * after removing jobs), 1 on error.  indices/distances are host pointers
 * (device_inputs = 0) or device pointers (1). */
int fastembedr_hip_tsne_batch_add(fastembedr_hip_tsne_batch* batch,
                                  const fastembedr_hip_tsne_job* job, int device_inputs,
                                  int* slot);

/* Enqueue n_iter iterations of every resident, unfinished job (async). */
int fastembedr_hip_tsne_batch_step(fastembedr_hip_tsne_batch* batch, int n_iter);

/* Waits for the batch and reports per slot (max_jobs entries) the iterations
 * done (-1 = empty slot) and whether the job still runs. */
int fastembedr_hip_tsne_batch_status(fastembedr_hip_tsne_batch* batch, int* iterations,
                                     int* active);

/* Centred layout of a slot, host row-major n x 2. */
int fastembedr_hip_tsne_batch_layout(fastembedr_hip_tsne_batch* batch, int slot, float* out);

int fastembedr_hip_tsne_batch_remove(fastembedr_hip_tsne_batch* batch, int slot);

void fastembedr_hip_tsne_batch_destroy(fastembedr_hip_tsne_batch* batch);

/* ---- serving session -------------------------------------------------------
 * No reference counterpart (the reference runs one job per call and frees
 * everything before it returns, embedding_cuda_kernels.cpp:88-127).  A
 * session keeps a t-SNE job batch, a fixed reference layout and the
 * transforms' workspace pools resident on the current device and serves
 * requests continuously: embed jobs join the batch when the HBM budget and
 * a slot allow and advance in 25-iteration ticks with every other resident
 * job; transforms run on their own streams beside the ticks, with device
 * workspace and pinned staging taken from pools carved out of the budget.
 * One engine thread inside the session makes every device call; submit and
 * poll may be called from any thread.  Inputs are copied at submit. */
typedef struct fastembedr_hip_session fastembedr_hip_session;

/* reference_layout: row-major n_reference x 2 (NULL / 0: no transforms).
 * hbm_budget_bytes <= 0: no budget.  transform_pool_bytes <= 0: one eighth
 * of the budget (1 GiB without one); the batch gets the rest of the budget.
 * The pool is allocated on the device and as pinned host memory. */
int fastembedr_hip_session_create(const float* reference_layout, int n_reference, int max_jobs,
                                  long long hbm_budget_bytes, long long transform_pool_bytes,
                                  fastembedr_hip_session** session);

/* Queues a t-SNE job (host inputs, fastembedr_hip_tsne_batch_add's rules).
 * Returns 1 with the message when the job is invalid; *request receives its
 * id. */
int fastembedr_hip_session_submit_embed(fastembedr_hip_session* session,
                                        const fastembedr_hip_tsne_job* job, long long* request);

/* Queues a UMAP job (host inputs, fastembedr_hip_umap_batch_add's rules;
 * the job struct is declared with the UMAP batch below).  Its graph is built
 * when it joins the session's UMAP batch, which advances every resident job
 * by 25 epochs per tick beside the t-SNE ticks; the two batches share the
 * budget left after the transform pool.  Returns 1 with the message when the
 * job is invalid. */
struct fastembedr_hip_umap_job;
int fastembedr_hip_session_submit_umap(fastembedr_hip_session* session,
                                       const struct fastembedr_hip_umap_job* job,
                                       long long* request);

/* Queues a transform against the session's reference: the arguments of
 * fastembedr_hip_transform_reference_run.  Returns 1 with the message when
 * the request is invalid or larger than the transform pool. */
int fastembedr_hip_session_submit_transform(
    fastembedr_hip_session* session, const int* indices, const float* distances, int n_query,
    int k, int index_offset, const float* init, int initialization, double perplexity, int n_iter,
    int early_exaggeration_iter, double learning_rate, double early_exaggeration,
    double exaggeration, double initial_momentum, double final_momentum, double max_grad_norm,
    double max_step_norm, int n_negatives, int exact_repulsion_threshold, int seed,
    long long* request);

/* *state: 0 queued, 1 running, 2 done, 3 failed.  wait_ms > 0 waits up to
 * that long for the request to end.  Done: the layout (row-major, the job's
 * n or the query count x 2) is copied to out (may be NULL) and the request
 * is released; failed: the request's message becomes the last error and the
 * request is released.  latency_ms: submit to end (-1 while it runs).
 * Returns 1 for an unknown or already released id. */
int fastembedr_hip_session_poll(fastembedr_hip_session* session, long long request, int wait_ms,
                                int* state, float* out, double* latency_ms);

/* stats[0..11]: submitted, completed, failed, queued embeds, queued
 * transforms, running embeds, running transforms, transform pool bytes in
 * use, engine ticks (t-SNE and UMAP), transform pool bytes, queued UMAP
 * jobs, running UMAP jobs. */
int fastembedr_hip_session_stats(fastembedr_hip_session* session, long long* stats);

/* Queued requests fail, work on the device is waited for, everything is
 * freed. */
void fastembedr_hip_session_destroy(fastembedr_hip_session* session);

/* ---- UMAP --------------------------------------------------------------- */

/* Replaces fastembedr_cuda_umap_optimize_coo (embedding_cuda_kernels.cpp:
 * 5483-5608): COO heads/tails (0-based), weights, epochs_per_sample
 * (= max_w / max(w, 1e-6)), init/out row-major n x 2, n_epochs epochs with
 * negative_sample_rate negatives, a/b curve, repulsion gamma, hash seed.
 * optimizer_mode 0 = in-place float atomics (the reference's hogwild form),
 * 1 = epoch-start reads with a 64-bit fixed-point delta (deterministic). */
int fastembedr_hip_umap_optimize_coo(const int* heads, const int* tails, const float* weights,
                                     const float* epochs_per_sample, const float* init, int n,
                                     int n_edges, int n_epochs, int negative_sample_rate,
                                     float learning_rate, float a, float b,
                                     float repulsion_strength, unsigned int seed,
                                     int optimizer_mode, float* out);

/* Same with heads/tails/weights/epochs_per_sample as device pointers
 * (borrowed); init/out host.  No reference counterpart (its UMAP entry takes
 * host buffers only); for callers whose fuzzy graph already lives on the GPU. */
int fastembedr_hip_umap_optimize_coo_device(const int* heads, const int* tails,
                                            const float* weights, const float* epochs_per_sample,
                                            const float* init, int n, int n_edges, int n_epochs,
                                            int negative_sample_rate, float learning_rate, float a,
                                            float b, float repulsion_strength, unsigned int seed,
                                            int optimizer_mode, float* out);

/* The epochs of fastembedr_hip_umap_optimize_coo_device with no device
 * allocation, for callers that cannot allocate (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No counterpart in
 * the reference.  Every pointer is device memory; `layout` (row-major n x 2)
 * holds the start layout and receives the result in place.  The work is
 * enqueued on `stream` (a hipStream_t; NULL = the null stream) and the call
 * returns without waiting.  optimizer_mode 1 keeps its fixed-point delta in
 * a file-scope device array of 2 x 2,097,152 64-bit words (32 MiB per GPU,
 * held until the process ends), so it takes n <= 2,097,152 and rejects larger
 * inputs; mode 0 needs no scratch.  The array is shared by every call on a
 * GPU and cleared at the start of each call: calls that overlap on one GPU
 * (concurrent streams or threads) race and must not be issued.  The result
 * equals fastembedr_hip_umap_optimize_coo_device's bit for bit in mode 1. */
int fastembedr_hip_umap_optimize_coo_device_noalloc(
    const int* heads, const int* tails, const float* weights, const float* epochs_per_sample,
    float* layout, int n, int n_edges, int n_epochs, int negative_sample_rate,
    float learning_rate, float a, float b, float repulsion_strength, unsigned int seed,
    int optimizer_mode, void* stream);

/* The sampling schedule of one epoch as the epoch kernel computes it
 * (positive_samples_this_epoch_umap_schedule, negative_samples_..._schedule,
 * deterministic_vertex; embedding_cuda_kernels.cpp:262-274, 3044-3084):
 * per edge n_pos, n_neg and the first max_neg negative ids (-1 = skipped as
 * head or tail).  No reference counterpart; exposes the schedule for exact
 * parity checks.  Host pointers. */
int fastembedr_hip_umap_schedule(const int* heads, const int* tails,
                                 const float* epochs_per_sample, int n_edges, int n, int epoch,
                                 int negative_sample_rate, unsigned int seed, int max_neg,
                                 int* n_pos, int* n_neg, int* neg_ids);

/* The UMAP fuzzy graph, built and kept on the device.  Replaces
 * umap_graph_csr_float_cpp (src/fast_knn_umap.cpp:3995-4048) at the default
 * edge budget (edge_budget >= n_cols; fast_knn_umap.R:611-640 passes
 * n_neighbors), i.e. build_graph_csr_native_direct_values (:1562-1780):
 * smooth_knn_dist sigmas and rhos, memberships exp(-(d - rho)/sigma), the
 * union f + r - f*r above 1e-6 and epochs_per_sample = max_w / max(w, 1e-6).
 * indices / distances: column-major n x k_total (float32 distances, as the
 * reference converts them), columns col_start .. col_start + n_cols - 1;
 * index_offset -1 detects it as the reference does (1 when every used index
 * lies in 1..n); device_inputs: the two arrays are device pointers. */
typedef struct fastembedr_hip_umap_graph fastembedr_hip_umap_graph;

int fastembedr_hip_umap_graph_create(const int* indices, const float* distances, int n,
                                     int k_total, int col_start, int n_cols, int edge_budget,
                                     int index_offset, int device_inputs,
                                     fastembedr_hip_umap_graph** graph);

/* nnz and max_weight of the graph. */
int fastembedr_hip_umap_graph_info(fastembedr_hip_umap_graph* graph, long long* nnz,
                                   float* max_weight);

/* The CSR to host buffers (offsets n+1, the rest nnz; any may be NULL), as
 * the reference's list(offsets, neighbors, weights, epochs_per_sample),
 * 0-based neighbours in ascending order within a row. */
int fastembedr_hip_umap_graph_export(fastembedr_hip_umap_graph* graph, int* offsets,
                                     int* neighbors, float* weights, float* epochs_per_sample);

/* Borrowed device pointers to the graph as COO (heads = row of each entry),
 * valid until destroy: the inputs of fastembedr_hip_umap_optimize_coo_device.
 * No reference counterpart (the reference returns the graph to the host). */
int fastembedr_hip_umap_graph_device(fastembedr_hip_umap_graph* graph, const int** heads,
                                     const int** tails, const float** weights,
                                     const float** epochs_per_sample);

/* UMAP epochs on the resident graph: the arguments and result of
 * fastembedr_hip_umap_optimize_coo on the graph's COO (its heads, neighbours,
 * weights and epochs_per_sample).  The graph is symmetric with equal weights
 * both ways, which the deterministic mode (optimizer_mode 1) uses to apply
 * each pair's attraction through the heads alone (the same bits as the COO
 * entry point).  No reference counterpart (the reference runs the epochs
 * from a host graph, umap_cuda_optimize_csr_cpp). */
int fastembedr_hip_umap_graph_optimize(fastembedr_hip_umap_graph* graph, const float* init,
                                       int n_epochs, int negative_sample_rate,
                                       float learning_rate, float a, float b,
                                       float repulsion_strength, unsigned int seed,
                                       int optimizer_mode, float* out);

void fastembedr_hip_umap_graph_destroy(fastembedr_hip_umap_graph* graph);

/* ---- batched UMAP jobs ------------------------------------------------------
 * No reference counterpart (the reference runs one job per call).  Many UMAP
 * jobs of uneven size advance together: each job's fuzzy graph
 * (fastembedr_hip_umap_graph_create at the default edge budget) and layout
 * stay resident, and one epoch of every resident job costs two launches per
 * optimizer mode present plus one, replayed in captured 25-epoch chunks.  A
 * job in optimizer_mode 1 gives the bits fastembedr_hip_umap_graph_optimize
 * gives on its graph alone; mode 0 is the reference's hogwild form, which is
 * not reproducible alone or batched. */
typedef struct fastembedr_hip_umap_job {
  const int* indices;     /* column-major n x k_total */
  const float* distances; /* column-major n x k_total, float32 */
  int n, k_total;
  int n_neighbors;        /* the graph uses columns 0 .. n_neighbors - 1 */
  int index_offset;       /* 0, 1, or -1 to detect it as the reference does */
  const float* init;      /* host row-major n x 2 */
  int n_epochs, negative_sample_rate;
  float learning_rate, a, b, repulsion_strength;
  unsigned int seed;
  int optimizer_mode;     /* 0 in-place float atomics, 1 fixed-point delta */
} fastembedr_hip_umap_job;

typedef struct fastembedr_hip_umap_batch fastembedr_hip_umap_batch;

/* hbm_budget_bytes <= 0: no budget. */
int fastembedr_hip_umap_batch_create(int max_jobs, long long hbm_budget_bytes,
                                     fastembedr_hip_umap_batch** batch);

/* Device bytes a UMAP job of this shape holds while resident (the graph at
 * its worst case of 2 n n_neighbors entries), and at its peak during the
 * graph build. */
long long fastembedr_hip_umap_job_bytes(int n, int n_neighbors, int optimizer_mode,
                                        int setup_peak);

/* Builds the job's fuzzy graph and uploads its start layout.  Returns 0 and
 * the slot, 2 when the budget or the slot table does not admit it now, 1 on
 * error.  indices/distances are host (device_inputs = 0) or device (1)
 * pointers. */
int fastembedr_hip_umap_batch_add(fastembedr_hip_umap_batch* batch,
                                  const fastembedr_hip_umap_job* job, int device_inputs,
                                  int* slot);

/* Enqueue n_epochs epochs of every resident, unfinished job (async). */
int fastembedr_hip_umap_batch_step(fastembedr_hip_umap_batch* batch, int n_epochs);

/* Waits for the batch and reports per slot (max_jobs entries) the epochs
 * done (-1 = empty slot) and whether the job still runs. */
int fastembedr_hip_umap_batch_status(fastembedr_hip_umap_batch* batch, int* epochs, int* active);

/* Layout of a slot, host row-major n x 2. */
int fastembedr_hip_umap_batch_layout(fastembedr_hip_umap_batch* batch, int slot, float* out);

int fastembedr_hip_umap_batch_remove(fastembedr_hip_umap_batch* batch, int slot);

void fastembedr_hip_umap_batch_destroy(fastembedr_hip_umap_batch* batch);

/* ---- fixed-reference t-SNE transform ------------------------------------
 * The reference binds transform_tsne_cuda_cpp (R/transform_tsne.R:275) but
 * defines it nowhere; the argument list follows transform_tsne_metal_cpp
 * (src/embedding_metal_bridge.cpp:323-342) and the semantics
 * transform_tsne_cpp (src/tsne_neighbors.cpp:2124-2373), two dimensions.
 * reference_layout: row-major n_reference x 2; indices/distances:
 * column-major n_query x k into the reference; init: row-major n_query x 2
 * or NULL; initialization 0 = median, 1 = weighted, 2 = random;
 * max_grad_norm / max_step_norm <= 0 or non-finite mean no clip. */
int fastembedr_hip_tsne_transform(const float* reference_layout, int n_reference,
                                  const int* indices, const double* distances, int n_query, int k,
                                  int index_offset, const float* init, int initialization,
                                  double perplexity, int n_iter, int early_exaggeration_iter,
                                  double learning_rate, double early_exaggeration,
                                  double exaggeration, double initial_momentum,
                                  double final_momentum, double max_grad_norm,
                                  double max_step_norm, int n_negatives,
                                  int exact_repulsion_threshold, int seed, float* out);

/* A reference layout kept resident for repeated transform requests (no
 * reference counterpart: serving without re-uploading the reference). */
typedef struct fastembedr_hip_transform_reference fastembedr_hip_transform_reference;

int fastembedr_hip_transform_reference_create(const float* reference_layout, int n_reference,
                                              fastembedr_hip_transform_reference** ref);

/* As fastembedr_hip_tsne_transform against the resident reference, with
 * float32 distances. */
int fastembedr_hip_transform_reference_run(fastembedr_hip_transform_reference* ref,
                                           const int* indices, const float* distances,
                                           int n_query, int k, int index_offset, const float* init,
                                           int initialization, double perplexity, int n_iter,
                                           int early_exaggeration_iter, double learning_rate,
                                           double early_exaggeration, double exaggeration,
                                           double initial_momentum, double final_momentum,
                                           double max_grad_norm, double max_step_norm,
                                           int n_negatives, int exact_repulsion_threshold,
                                           int seed, float* out);

void fastembedr_hip_transform_reference_destroy(fastembedr_hip_transform_reference* ref);

/* ---- embedding quality ---------------------------------------------------
 * Replaces exact_structure_metrics_cpp (src/embedding_utils.cpp:1582-1752):
 * high (n x high_cols) and low (n x low_cols) are column-major FP64 (R
 * matrices), 3 <= n <= 5000, every requested k with 1 <= k < n and
 * 2n - 3k - 1 > 0.  out is column-major nk x 4: trustworthiness,
 * continuity, knn_preservation, mean_neighbor_rank_error (the reference's
 * column order); equal to the reference's values bit for bit. */
int fastembedr_hip_exact_structure_metrics(const double* high, int n, int high_cols,
                                           const double* low, int low_cols,
                                           const int* requested_k, int nk, double* out);

/* Replaces fastembedr_cuda_knn_structure_score
 * (src/embedding_cuda_kernels.cpp:4570-4700, the GPU form of
 * knn_structure_score_cpp, src/embedding_utils.cpp:1386-1580), same
 * arguments: layout column-major n x 2 FP64; indices column-major
 * index_rows x high_rank_limit, 1-based (index_rows = n, or keep_n with
 * compact_indices); keep: keep_n 1-based rows; labels: n factor codes or
 * NULL (n_label_levels 0); 1 <= preserve_k <= min(high_rank_limit, 256).
 * totals[6]: sums over the scored rows of preservation, local
 * trustworthiness, local continuity, label hits, labelled rows, and the
 * number of scored rows, added in row order. */
int fastembedr_hip_knn_structure_score(const double* layout, const int* indices, const int* keep,
                                       const int* labels, int n, int index_rows,
                                       int high_rank_limit, int preserve_k, int keep_n,
                                       int compact_indices, int n_label_levels, double* totals);

/* ---- exact k-nearest neighbours -----------------------------------------
 * Replaces the exact route of native_cuda_knn_impl
 * (src/native_knn_cuda_impl.cpp:520-700: FAISS bfKnn or cuVS brute force
 * for the k + 1 nearest rows, then finalize_cuvs_knn_kernel,
 * src/embedding_cuda_kernels.cpp:1428-1480, dropping the row itself) for
 * the Euclidean metric.  data: column-major n x p float32 (an R matrix);
 * 1 <= k <= min(n - 1, 256).  indices/distances: column-major n x k (the
 * GPU-resident list contract of embedding_cuda_impl.cpp:2435-2523:
 * "column_major_query_by_k"), indices 0- or 1-based by index_base (the
 * reference's list uses 1), the row itself excluded, ties by the smaller
 * index, distances Euclidean float32.  The neighbours are exact: squared
 * distances in FP64 from the float32 inputs, not the float32 expanded form
 * of the reference's external search.  device_inputs / device_outputs: the
 * pointers are device memory.  scanned_rows (may be NULL): how many rows the
 * error bound did not settle and were found by a full FP64 scan. */
int fastembedr_hip_exact_knn(const float* data, int n, int p, int k, int device_inputs,
                             int index_base, int* indices, float* distances, int device_outputs,
                             int* scanned_rows);

/* Replaces the exact route of native_cuda_query_knn_impl
 * (src/native_knn_cuda_impl.cpp:878-1010, FAISS bfKnn / cuVS brute force of
 * the query rows against the data rows, then finalize_cuvs_query_knn_kernel,
 * src/embedding_cuda_kernels.cpp:1482-1522): the k nearest data rows of
 * every query row (no row dropped), 1 <= k <= min(n, 256).  data:
 * column-major n x p, query: column-major n_query x p, float32; outputs
 * column-major n_query x k, indices into data with index_base; exact as
 * fastembedr_hip_exact_knn.  The query rows of the t-SNE transform. */
int fastembedr_hip_exact_query_knn(const float* data, int n, const float* query, int n_query,
                                   int p, int k, int device_inputs, int index_base, int* indices,
                                   float* distances, int device_outputs, int* scanned_rows);

#ifdef __cplusplus
}
#endif

#endif /* FASTEMBEDR_HIP_H */
