// Internal interface between the serving session (serve.hip) and the engines
// it drives: the fixed-reference transform (transform.hip) and the t-SNE job
// batch (tsne.hip).  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <memory>

#include "../../include/fastembedr_hip.h"

namespace fe {

struct TransformReference;
struct TransformTask;

// The resident reference behind a C-ABI handle.
const TransformReference* transform_reference_of(const fastembedr_hip_transform_reference* ref);

void transform_task_destroy(TransformTask* t);
struct TransformTaskDeleter {
  void operator()(TransformTask* t) const { transform_task_destroy(t); }
};
using TransformTaskPtr = std::unique_ptr<TransformTask, TransformTaskDeleter>;

// Checks a transform request against the reference (the messages of
// fastembedr_hip_transform_reference_run, including the KNN range and
// distance checks, done on the host) and builds its host-side inputs.  The
// KNN rows themselves are not copied.
int transform_task_create(const TransformReference& R, const int* indices, const float* distances,
                          int n_query, int k, int index_offset, const float* init,
                          int initialization, double perplexity, int n_iter,
                          int early_exaggeration_iter, double learning_rate,
                          double early_exaggeration, double exaggeration,
                          double initial_momentum, double final_momentum, double max_grad_norm,
                          double max_step_norm, int n_negatives, int exact_repulsion_threshold,
                          int seed, TransformTaskPtr* out);

// Device workspace and pinned host staging a request needs.
std::size_t transform_task_device_bytes(const TransformTask& t);
std::size_t transform_task_pinned_bytes(const TransformTask& t);
int transform_task_queries(const TransformTask& t);

// Copies the request's inputs into `pinned` (host, page-locked) and enqueues
// on `s` the uploads, the P rows, the start, the optimisation and the
// download of the layout and the error flags into `pinned`.  Returns without
// waiting; `dev` and `pinned` stay owned by the caller until the stream has
// passed the work.
int transform_task_enqueue(const TransformReference& R, const TransformTask& t, const int* indices,
                           const float* distances, void* dev, void* pinned, hipStream_t s);

// After the enqueued work has finished: the result (row-major n_query x 2)
// inside `pinned`, or 1 with the error the device checks raised.
int transform_task_result(const TransformTask& t, const void* pinned, const float** layout);

// The stream a batch enqueues its iterations on (fastembedr_hip_tsne_batch_step).
hipStream_t tsne_batch_stream(const fastembedr_hip_tsne_batch* batch);
// A batch's budget (<= 0: none) and the device bytes its jobs hold: the
// session moves budget between its t-SNE and UMAP batches as jobs come and go.
void tsne_batch_set_budget(fastembedr_hip_tsne_batch* batch, long long budget);
long long tsne_batch_used(const fastembedr_hip_tsne_batch* batch);

hipStream_t umap_batch_stream(const fastembedr_hip_umap_batch* batch);
void umap_batch_set_budget(fastembedr_hip_umap_batch* batch, long long budget);
long long umap_batch_used(const fastembedr_hip_umap_batch* batch);
// Checks a UMAP job as fastembedr_hip_umap_batch_add does before it builds
// anything.
int umap_job_validate(const fastembedr_hip_umap_job& j);

// Checks a t-SNE job as fastembedr_hip_tsne_batch_add does before it builds
// anything (shape, schedule, learning rates).
int tsne_job_validate(const fastembedr_hip_tsne_job& j);

}  // namespace fe
