// Continuous-batching session on one GPU: the serving form of the backend.
//
// No reference counterpart: the reference runs one job per call and frees
// every device buffer before it returns (embedding_cuda_kernels.cpp:88-127).
// A session keeps everything resident and serves two kinds of request:
//
//  * embed: a whole t-SNE job (KNN in, layout out).  It joins the session's
//    job batch (job-table kernels, captured 25-iteration ticks) when the HBM
//    budget and a slot allow; every tick advances all resident jobs, so a job
//    that arrives while others run starts at the next tick and shares their
//    kernels.  A finished job is read back and removed, which frees its
//    budget for waiting ones.
//  * umap: a whole UMAP job (KNN in, layout out): its fuzzy graph is built
//    when it joins the session's UMAP batch, which advances every resident
//    job by 25 epochs per tick on its own stream, beside the t-SNE ticks.
//  * transform: placement of query points into the fixed reference layout.
//    Its device workspace and pinned host staging come from two pools carved
//    out of the budget at creation; it is enqueued on one of the session's
//    transform streams without waiting, so it overlaps the tick on the device.
//
// One engine thread owns every device call after creation.  Callers submit
// from any thread (inputs are copied at submit, so the caller may free them)
// and poll by request id.  The engine loop:
//   1. retires transforms whose completion event has passed;
//   2. for each batch (t-SNE, UMAP) whose tick in flight has passed: reads
//      its status, retires finished jobs, admits waiting jobs in arrival
//      order while the budget allows, and enqueues the next tick (the two
//      batches share the budget left after the transform pool);
//   3. admits waiting transforms in arrival order while a stream and pool
//      space are free.
// A job the budget cannot take with nothing resident, or a transform larger
// than the pools, fails instead of waiting for ever.  Results do not depend
// on what else runs: batched jobs give the bits of their single-job plan and
// a transform is one deterministic launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fastembedr_hip.h"
#include "fe_common.hpp"
#include "serve_iface.hpp"

namespace fe {
namespace {

using Clock = std::chrono::steady_clock;

constexpr int kTransformStreams = 2;
constexpr int kTick = 25;  // iterations per engine tick (one captured batch graph)

// First-fit ranges of a fixed pool (offsets in bytes, 256-byte granules);
// freed ranges merge with their neighbours.
class RangePool {
 public:
  void init(std::size_t cap) {
    free_.clear();
    cap_ = cap;
    if (cap) free_[0] = cap;
  }
  std::size_t capacity() const { return cap_; }
  std::size_t used() const { return used_.load(); }
  long long take(std::size_t bytes) {
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const std::size_t off = it->first, rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      used_ += bytes;
      return static_cast<long long>(off);
    }
    return -1;
  }
  void give(long long off, std::size_t bytes) {
    if (off < 0) return;
    bytes = align_up(bytes);
    used_ -= bytes;
    auto it = free_.emplace(static_cast<std::size_t>(off), bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<std::size_t, std::size_t> free_;
  std::size_t cap_ = 0;
  std::atomic<std::size_t> used_{0};  // read by stats() outside the engine thread
};

enum RequestState { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3 };

enum RequestKind { kEmbed = 0, kTransform = 1, kUmap = 2 };

struct Request {
  long long id = 0;
  int kind = kEmbed;
  int state = kQueued;
  std::string error;
  Clock::time_point submitted, finished;
  int rows = 0;
  std::vector<float> result;
  // embed / UMAP job: the job with pointers into the session's copies
  fastembedr_hip_tsne_job job{};
  fastembedr_hip_umap_job ujob{};
  std::vector<int> idx;
  std::vector<float> dist, init;
  // transform
  TransformTaskPtr task;
  std::size_t dev_bytes = 0, pin_bytes = 0;
  long long dev_off = -1, pin_off = -1;
  int stream = -1;
};

using RequestPtr = std::shared_ptr<Request>;

}  // namespace

class Session {
 public:
  ~Session() { shutdown(); }

  int create(const float* reference_layout, int n_reference, int max_jobs, long long budget,
             long long pool_bytes) {
    if (max_jobs < 1) return fail("max_jobs must be positive");
    FE_CHECK(hipGetDevice(&device_));
    if (pool_bytes <= 0) pool_bytes = budget > 0 ? budget / 8 : (1ll << 30);
    if (budget > 0 && pool_bytes >= budget)
      return fail("the transform pool must be smaller than the HBM budget");
    if (reference_layout && n_reference > 0) {
      if (fastembedr_hip_transform_reference_create(reference_layout, n_reference, &ref_)) return 1;
      pool_bytes_ = static_cast<std::size_t>(pool_bytes);
      FE_CHECK(hipMalloc(&dev_pool_, pool_bytes_));
      FE_CHECK(hipHostMalloc(&pin_pool_, pool_bytes_, hipHostMallocDefault));
      dev_ranges_.init(pool_bytes_);
      pin_ranges_.init(pool_bytes_);
      for (int i = 0; i < kTransformStreams; ++i) {
        FE_CHECK(hipStreamCreateWithFlags(&tf_stream_[i], hipStreamNonBlocking));
        FE_CHECK(hipEventCreateWithFlags(&tf_event_[i], hipEventDisableTiming));
      }
    }
    embed_budget_ = budget > 0 ? budget - static_cast<long long>(pool_bytes_) : 0;
    if (fastembedr_hip_tsne_batch_create(max_jobs, embed_budget_, &batch_)) return 1;
    if (fastembedr_hip_umap_batch_create(max_jobs, embed_budget_, &ubatch_)) return 1;
    FE_CHECK(hipEventCreateWithFlags(&utick_event_, hipEventDisableTiming));
    max_jobs_ = max_jobs;
    FE_CHECK(hipEventCreateWithFlags(&tick_event_, hipEventDisableTiming));
    engine_ = std::thread([this] { engine(); });
    return 0;
  }

  int submit_embed(const fastembedr_hip_tsne_job& j, long long* id) {
    if (tsne_job_validate(j)) return 1;
    auto r = std::make_shared<Request>();
    r->kind = kEmbed;
    r->rows = j.n;
    const std::size_t nk = static_cast<std::size_t>(j.n) * j.k;
    r->idx.assign(j.indices, j.indices + nk);
    r->dist.assign(j.distances, j.distances + nk);
    if (j.init) r->init.assign(j.init, j.init + 2 * static_cast<std::size_t>(j.n));
    r->job = j;
    r->job.indices = r->idx.data();
    r->job.distances = r->dist.data();
    r->job.init = j.init ? r->init.data() : nullptr;
    return enqueue(std::move(r), id);
  }

  int submit_transform(const int* indices, const float* distances, int n_query, int k,
                       int index_offset, const float* init, int initialization,
                       double perplexity, int n_iter, int early_iter, double lr,
                       double early_exag, double exag, double mom0, double mom1, double max_grad,
                       double max_step, int n_negatives, int threshold, int seed, long long* id) {
    if (!ref_) return fail("the session has no reference layout for transforms");
    auto r = std::make_shared<Request>();
    r->kind = kTransform;
    TransformTaskPtr t;
    if (transform_task_create(*transform_reference_of(ref_), indices, distances, n_query, k,
                              index_offset, init, initialization, perplexity, n_iter, early_iter,
                              lr, early_exag, exag, mom0, mom1, max_grad, max_step, n_negatives,
                              threshold, seed, &t))
      return 1;
    r->dev_bytes = transform_task_device_bytes(*t);
    r->pin_bytes = transform_task_pinned_bytes(*t);
    if (align_up(r->dev_bytes) > pool_bytes_ || align_up(r->pin_bytes) > pool_bytes_)
      return fail("the transform needs " + std::to_string(r->dev_bytes) +
                  " bytes of device workspace, more than the session's transform pool (" +
                  std::to_string(pool_bytes_) + " bytes)");
    r->task = std::move(t);
    r->rows = n_query;
    const std::size_t total = static_cast<std::size_t>(n_query) * k;
    r->idx.assign(indices, indices + total);
    r->dist.assign(distances, distances + total);
    return enqueue(std::move(r), id);
  }

  int submit_umap(const fastembedr_hip_umap_job& j, long long* id) {
    if (umap_job_validate(j)) return 1;
    auto r = std::make_shared<Request>();
    r->kind = kUmap;
    r->rows = j.n;
    const std::size_t nk = static_cast<std::size_t>(j.n) * j.k_total;
    r->idx.assign(j.indices, j.indices + nk);
    r->dist.assign(j.distances, j.distances + nk);
    r->init.assign(j.init, j.init + 2 * static_cast<std::size_t>(j.n));
    r->ujob = j;
    r->ujob.indices = r->idx.data();
    r->ujob.distances = r->dist.data();
    r->ujob.init = r->init.data();
    return enqueue(std::move(r), id);
  }

  // state of a request; copies the result and forgets the request once it is
  // done or failed
  int poll(long long id, int wait_ms, int* state, float* out, double* latency_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = requests_.find(id);
    if (it == requests_.end()) return fail("unknown or already collected request id");
    RequestPtr r = it->second;
    if (wait_ms > 0)
      done_cv_.wait_for(lk, std::chrono::milliseconds(wait_ms),
                        [&] { return r->state == kDone || r->state == kFailed; });
    if (state) *state = r->state;
    if (r->state == kDone || r->state == kFailed) {
      if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->finished - r->submitted).count();
      if (r->state == kDone && out) std::memcpy(out, r->result.data(), r->result.size() * sizeof(float));
      if (r->state == kFailed) set_error(r->error);
      requests_.erase(it);
    } else if (latency_ms) {
      *latency_ms = -1.0;
    }
    return 0;
  }

  int stats(long long* s) {
    std::lock_guard<std::mutex> lk(mu_);
    s[0] = submitted_;
    s[1] = completed_;
    s[2] = failed_;
    s[3] = queued_embed_;
    s[4] = queued_tf_;
    s[5] = running_embed_;
    s[6] = running_tf_;
    s[7] = static_cast<long long>(dev_ranges_.used());
    s[8] = ticks_;
    s[9] = static_cast<long long>(pool_bytes_);
    s[10] = queued_umap_;
    s[11] = running_umap_;
    return 0;
  }

  void shutdown() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    work_cv_.notify_all();
    if (engine_.joinable()) engine_.join();
    if (batch_) fastembedr_hip_tsne_batch_destroy(batch_);
    batch_ = nullptr;
    if (ubatch_) fastembedr_hip_umap_batch_destroy(ubatch_);
    ubatch_ = nullptr;
    if (utick_event_) (void)hipEventDestroy(utick_event_);
    utick_event_ = nullptr;
    for (int i = 0; i < kTransformStreams; ++i) {
      if (tf_stream_[i]) (void)hipStreamSynchronize(tf_stream_[i]);
      if (tf_event_[i]) (void)hipEventDestroy(tf_event_[i]);
      if (tf_stream_[i]) (void)hipStreamDestroy(tf_stream_[i]);
      tf_event_[i] = nullptr;
      tf_stream_[i] = nullptr;
    }
    if (tick_event_) (void)hipEventDestroy(tick_event_);
    tick_event_ = nullptr;
    if (dev_pool_) (void)hipFree(dev_pool_);
    if (pin_pool_) (void)hipHostFree(pin_pool_);
    dev_pool_ = nullptr;
    pin_pool_ = nullptr;
    if (ref_) fastembedr_hip_transform_reference_destroy(ref_);
    ref_ = nullptr;
  }

 private:
  int enqueue(RequestPtr r, long long* id) {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) return fail("the session is shutting down");
    if (broken_) return fail("the serving engine stopped after an error");
    r->id = next_id_++;
    r->submitted = Clock::now();
    requests_[r->id] = r;
    incoming_.push_back(r);
    ++submitted_;
    ++*queued_of(r->kind);
    if (id) *id = r->id;
    work_cv_.notify_all();
    return 0;
  }

  long long* queued_of(int kind) {
    return kind == kEmbed ? &queued_embed_ : kind == kUmap ? &queued_umap_ : &queued_tf_;
  }
  long long* running_of(int kind) {
    return kind == kEmbed ? &running_embed_ : kind == kUmap ? &running_umap_ : &running_tf_;
  }

  // the budget the batches share, less what the other batch holds
  long long budget_left(long long other_used) const {
    return std::max(1ll, embed_budget_ - other_used);
  }

  // engine-side state changes (under the lock: poll reads them)
  void finish(const RequestPtr& r, bool ok, const std::string& error) {
    std::lock_guard<std::mutex> lk(mu_);
    if (r->state == kRunning) --*running_of(r->kind);
    else if (r->state == kQueued) --*queued_of(r->kind);
    r->state = ok ? kDone : kFailed;
    r->error = error;
    r->finished = Clock::now();
    if (ok) ++completed_;
    else ++failed_;
    r->idx.clear();
    r->idx.shrink_to_fit();
    r->dist.clear();
    r->dist.shrink_to_fit();
    r->init.clear();
    r->task.reset();
    done_cv_.notify_all();
  }

  void started(const RequestPtr& r) {
    std::lock_guard<std::mutex> lk(mu_);
    r->state = kRunning;
    --*queued_of(r->kind);
    ++*running_of(r->kind);
  }

  bool retire_transforms() {
    bool progressed = false;
    for (int i = 0; i < kTransformStreams; ++i) {
      RequestPtr& r = tf_running_[i];
      if (!r) continue;
      const hipError_t q = hipEventQuery(tf_event_[i]);
      if (q == hipErrorNotReady) continue;
      progressed = true;
      const void* pinned = static_cast<unsigned char*>(pin_pool_) + r->pin_off;
      const float* layout = nullptr;
      std::string error;
      bool ok = false;
      if (q != hipSuccess) {
        error = std::string("transform: ") + hipGetErrorString(q);
      } else if (transform_task_result(*r->task, pinned, &layout)) {
        error = last_error();
      } else {
        r->result.assign(layout, layout + 2 * static_cast<std::size_t>(r->rows));
        ok = true;
      }
      dev_ranges_.give(r->dev_off, r->dev_bytes);
      pin_ranges_.give(r->pin_off, r->pin_bytes);
      finish(r, ok, error);
      r.reset();
    }
    return progressed;
  }

  bool admit_transforms() {
    bool progressed = false;
    while (!tf_waiting_.empty()) {
      int stream = -1;
      for (int i = 0; i < kTransformStreams; ++i)
        if (!tf_running_[i]) {
          stream = i;
          break;
        }
      if (stream < 0) break;
      RequestPtr r = tf_waiting_.front();
      const long long d = dev_ranges_.take(r->dev_bytes);
      const long long p = d >= 0 ? pin_ranges_.take(r->pin_bytes) : -1;
      if (d < 0 || p < 0) {
        dev_ranges_.give(d, r->dev_bytes);
        break;  // retried when a running transform returns its ranges
      }
      tf_waiting_.pop_front();
      r->dev_off = d;
      r->pin_off = p;
      r->stream = stream;
      started(r);
      void* dev = static_cast<unsigned char*>(dev_pool_) + d;
      void* pinned = static_cast<unsigned char*>(pin_pool_) + p;
      if (transform_task_enqueue(*transform_reference_of(ref_), *r->task, r->idx.data(),
                                 r->dist.data(), dev, pinned, tf_stream_[stream]) ||
          check(hipEventRecord(tf_event_[stream], tf_stream_[stream]), "hipEventRecord")) {
        const std::string error = last_error();
        (void)hipStreamSynchronize(tf_stream_[stream]);
        dev_ranges_.give(d, r->dev_bytes);
        pin_ranges_.give(p, r->pin_bytes);
        finish(r, false, error);
      } else {
        tf_running_[stream] = r;
      }
      progressed = true;
    }
    return progressed;
  }

  // between ticks: retire finished jobs, admit waiting ones, enqueue a tick
  bool advance_batch() {
    bool progressed = false;
    if (!embed_running_.empty()) {
      std::vector<int> iters(max_jobs_), active(max_jobs_);
      if (fastembedr_hip_tsne_batch_status(batch_, iters.data(), active.data())) {
        fail_running_jobs(last_error());
        return true;
      }
      for (auto it = embed_running_.begin(); it != embed_running_.end();) {
        if (active[it->first]) {
          ++it;
          continue;
        }
        RequestPtr r = it->second;
        r->result.resize(2 * static_cast<std::size_t>(r->rows));
        const bool ok = fastembedr_hip_tsne_batch_layout(batch_, it->first, r->result.data()) == 0;
        const std::string error = ok ? std::string() : last_error();
        (void)fastembedr_hip_tsne_batch_remove(batch_, it->first);
        finish(r, ok, error);
        it = embed_running_.erase(it);
        progressed = true;
      }
    }
    while (!embed_waiting_.empty()) {
      RequestPtr r = embed_waiting_.front();
      int slot = -1;
      if (embed_budget_ > 0) tsne_batch_set_budget(batch_, budget_left(umap_batch_used(ubatch_)));
      const int rc = fastembedr_hip_tsne_batch_add(batch_, &r->job, 0, &slot);
      // retried when a job of either batch leaves
      if (rc == 2 && (!embed_running_.empty() || !umap_running_.empty())) break;
      embed_waiting_.pop_front();
      progressed = true;
      if (rc == 0) {
        started(r);
        embed_running_[slot] = r;
        continue;
      }
      finish(r, false, rc == 2 ? "not admissible: " + last_error() : last_error());
    }
    if (!embed_running_.empty()) {
      if (fastembedr_hip_tsne_batch_step(batch_, kTick) ||
          check(hipEventRecord(tick_event_, tsne_batch_stream(batch_)), "hipEventRecord")) {
        fail_running_jobs(last_error());
        return true;
      }
      tick_in_flight_ = true;
      std::lock_guard<std::mutex> lk(mu_);
      ++ticks_;
    }
    return progressed;
  }

  // the same for the UMAP batch
  bool advance_umap() {
    bool progressed = false;
    if (!umap_running_.empty()) {
      std::vector<int> epochs(max_jobs_), active(max_jobs_);
      if (fastembedr_hip_umap_batch_status(ubatch_, epochs.data(), active.data())) {
        fail_running_umap(last_error());
        return true;
      }
      for (auto it = umap_running_.begin(); it != umap_running_.end();) {
        if (active[it->first]) {
          ++it;
          continue;
        }
        RequestPtr r = it->second;
        r->result.resize(2 * static_cast<std::size_t>(r->rows));
        const bool ok = fastembedr_hip_umap_batch_layout(ubatch_, it->first, r->result.data()) == 0;
        const std::string error = ok ? std::string() : last_error();
        (void)fastembedr_hip_umap_batch_remove(ubatch_, it->first);
        finish(r, ok, error);
        it = umap_running_.erase(it);
        progressed = true;
      }
    }
    while (!umap_waiting_.empty()) {
      RequestPtr r = umap_waiting_.front();
      int slot = -1;
      if (embed_budget_ > 0) umap_batch_set_budget(ubatch_, budget_left(tsne_batch_used(batch_)));
      const int rc = fastembedr_hip_umap_batch_add(ubatch_, &r->ujob, 0, &slot);
      if (rc == 2 && (!embed_running_.empty() || !umap_running_.empty())) break;
      umap_waiting_.pop_front();
      progressed = true;
      if (rc == 0) {
        started(r);
        umap_running_[slot] = r;
        continue;
      }
      finish(r, false, rc == 2 ? "not admissible: " + last_error() : last_error());
    }
    if (!umap_running_.empty()) {
      if (fastembedr_hip_umap_batch_step(ubatch_, kTick) ||
          check(hipEventRecord(utick_event_, umap_batch_stream(ubatch_)), "hipEventRecord")) {
        fail_running_umap(last_error());
        return true;
      }
      utick_in_flight_ = true;
      std::lock_guard<std::mutex> lk(mu_);
      ++ticks_;
    }
    return progressed;
  }

  void fail_running_umap(const std::string& error) {
    for (auto& e : umap_running_) {
      (void)fastembedr_hip_umap_batch_remove(ubatch_, e.first);
      finish(e.second, false, error);
    }
    umap_running_.clear();
  }

  void fail_running_jobs(const std::string& error) {
    for (auto& e : embed_running_) {
      (void)fastembedr_hip_tsne_batch_remove(batch_, e.first);
      finish(e.second, false, error);
    }
    embed_running_.clear();
  }

  // The engine loop; an exception (host memory) fails every request and
  // stops the engine, after which submits are refused.
  void engine() {
    (void)hipSetDevice(device_);
    try {
      engine_loop();
    } catch (...) {
      sync_all_streams();
      std::lock_guard<std::mutex> lk(mu_);
      broken_ = true;
      for (auto& e : requests_) {
        RequestPtr r = e.second;
        if (r->state == kQueued || r->state == kRunning) {
          r->state = kFailed;
          r->error = "the serving engine stopped: out of host memory";
          r->finished = Clock::now();
          ++failed_;
        }
      }
      queued_embed_ = queued_tf_ = queued_umap_ = 0;
      running_embed_ = running_tf_ = running_umap_ = 0;
      done_cv_.notify_all();
    }
  }

  // waits for all device work of the session (the engine's error path)
  void sync_all_streams() {
    for (int i = 0; i < kTransformStreams; ++i)
      if (tf_stream_[i]) (void)hipStreamSynchronize(tf_stream_[i]);
    if (batch_) (void)hipStreamSynchronize(tsne_batch_stream(batch_));
    if (ubatch_) (void)hipStreamSynchronize(umap_batch_stream(ubatch_));
  }

  void engine_loop() {
    for (;;) {
      bool stopping = false;
      {
        std::unique_lock<std::mutex> lk(mu_);
        const bool busy = tick_in_flight_ || utick_in_flight_ ||
                          std::any_of(tf_running_, tf_running_ + kTransformStreams,
                                      [](const RequestPtr& r) { return !!r; });
        const bool waiting = !embed_waiting_.empty() || !tf_waiting_.empty() ||
                             !umap_waiting_.empty();
        if (stop_ && !busy) break;
        if (!busy && !waiting && incoming_.empty())
          work_cv_.wait(lk, [&] { return stop_ || !incoming_.empty(); });
        while (!incoming_.empty()) {
          RequestPtr r = incoming_.front();
          incoming_.pop_front();
          (r->kind == kEmbed ? embed_waiting_ : r->kind == kUmap ? umap_waiting_ : tf_waiting_)
              .push_back(r);
        }
        stopping = stop_;
      }
      if (stopping) {
        // queued requests fail; work on the device is waited for
        for (auto& r : embed_waiting_) finish(r, false, "the session was destroyed");
        for (auto& r : umap_waiting_) finish(r, false, "the session was destroyed");
        for (auto& r : tf_waiting_) finish(r, false, "the session was destroyed");
        embed_waiting_.clear();
        umap_waiting_.clear();
        tf_waiting_.clear();
      }
      bool progressed = retire_transforms();
      if (tick_in_flight_) {
        const hipError_t q = hipEventQuery(tick_event_);
        if (q != hipErrorNotReady) {
          tick_in_flight_ = false;
          progressed = true;
          if (q != hipSuccess) fail_running_jobs(std::string("batch tick: ") + hipGetErrorString(q));
        }
      }
      if (!tick_in_flight_) {
        if (stopping) {
          if (!embed_running_.empty()) fail_running_jobs("the session was destroyed");
        } else {
          progressed |= advance_batch();
        }
      }
      if (utick_in_flight_) {
        const hipError_t q = hipEventQuery(utick_event_);
        if (q != hipErrorNotReady) {
          utick_in_flight_ = false;
          progressed = true;
          if (q != hipSuccess) fail_running_umap(std::string("UMAP tick: ") + hipGetErrorString(q));
        }
      }
      if (!utick_in_flight_) {
        if (stopping) {
          if (!umap_running_.empty()) fail_running_umap("the session was destroyed");
        } else {
          progressed |= advance_umap();
        }
      }
      if (!stopping) progressed |= admit_transforms();
      if (!progressed) {
        // work on the device: look again shortly, or at once for a new request
        std::unique_lock<std::mutex> lk(mu_);
        work_cv_.wait_for(lk, std::chrono::microseconds(100),
                          [&] { return !stopping && (stop_ || !incoming_.empty()); });
      }
    }
  }

  int device_ = 0;
  int max_jobs_ = 0;
  fastembedr_hip_tsne_batch* batch_ = nullptr;
  fastembedr_hip_transform_reference* ref_ = nullptr;
  std::size_t pool_bytes_ = 0;
  void* dev_pool_ = nullptr;
  void* pin_pool_ = nullptr;
  RangePool dev_ranges_, pin_ranges_;
  hipStream_t tf_stream_[kTransformStreams] = {};
  hipEvent_t tf_event_[kTransformStreams] = {};
  RequestPtr tf_running_[kTransformStreams];
  hipEvent_t tick_event_ = nullptr;
  bool tick_in_flight_ = false;
  fastembedr_hip_umap_batch* ubatch_ = nullptr;
  hipEvent_t utick_event_ = nullptr;
  bool utick_in_flight_ = false;
  long long embed_budget_ = 0;  // both batches together (<= 0: none)
  std::deque<RequestPtr> embed_waiting_, tf_waiting_, umap_waiting_;
  std::map<int, RequestPtr> embed_running_, umap_running_;
  std::thread engine_;

  std::mutex mu_;
  std::condition_variable work_cv_, done_cv_;
  bool stop_ = false;
  bool broken_ = false;
  std::deque<RequestPtr> incoming_;
  std::unordered_map<long long, RequestPtr> requests_;
  long long next_id_ = 1;
  long long submitted_ = 0, completed_ = 0, failed_ = 0, ticks_ = 0;
  long long queued_embed_ = 0, queued_tf_ = 0, running_embed_ = 0, running_tf_ = 0;
  long long queued_umap_ = 0, running_umap_ = 0;
};

}  // namespace fe

struct fastembedr_hip_session {
  fe::Session impl;
};

extern "C" {

int fastembedr_hip_session_create(const float* reference_layout, int n_reference, int max_jobs,
                                  long long hbm_budget_bytes, long long transform_pool_bytes,
                                  fastembedr_hip_session** session) {
  fe::clear_error();
  if (!session) return fe::fail("null session handle");
  *session = nullptr;
  auto* s = new (std::nothrow) fastembedr_hip_session;
  if (!s) return fe::fail("out of host memory");
  if (fe::guarded("session create", [&] {
        return s->impl.create(reference_layout, n_reference, max_jobs, hbm_budget_bytes,
                              transform_pool_bytes);
      })) {
    const std::string error = fe::last_error();
    delete s;
    return fe::fail(error);
  }
  *session = s;
  return 0;
}

int fastembedr_hip_session_submit_embed(fastembedr_hip_session* session,
                                        const fastembedr_hip_tsne_job* job, long long* request) {
  fe::clear_error();
  if (!session || !job) return fe::fail("null session or job");
  return fe::guarded("session submit", [&] { return session->impl.submit_embed(*job, request); });
}

int fastembedr_hip_session_submit_transform(
    fastembedr_hip_session* session, const int* indices, const float* distances, int n_query,
    int k, int index_offset, const float* init, int initialization, double perplexity, int n_iter,
    int early_exaggeration_iter, double learning_rate, double early_exaggeration,
    double exaggeration, double initial_momentum, double final_momentum, double max_grad_norm,
    double max_step_norm, int n_negatives, int exact_repulsion_threshold, int seed,
    long long* request) {
  fe::clear_error();
  if (!session) return fe::fail("null session");
  return fe::guarded("session submit", [&] {
    return session->impl.submit_transform(
        indices, distances, n_query, k, index_offset, init, initialization, perplexity, n_iter,
        early_exaggeration_iter, learning_rate, early_exaggeration, exaggeration,
        initial_momentum, final_momentum, max_grad_norm, max_step_norm, n_negatives,
        exact_repulsion_threshold, seed, request);
  });
}

int fastembedr_hip_session_submit_umap(fastembedr_hip_session* session,
                                       const fastembedr_hip_umap_job* job, long long* request) {
  fe::clear_error();
  if (!session || !job) return fe::fail("null session or job");
  return fe::guarded("session submit", [&] { return session->impl.submit_umap(*job, request); });
}

int fastembedr_hip_session_poll(fastembedr_hip_session* session, long long request, int wait_ms,
                                int* state, float* out, double* latency_ms) {
  fe::clear_error();
  if (!session) return fe::fail("null session");
  return fe::guarded("session poll", [&] {
    return session->impl.poll(request, wait_ms, state, out, latency_ms);
  });
}

int fastembedr_hip_session_stats(fastembedr_hip_session* session, long long* stats) {
  fe::clear_error();
  if (!session || !stats) return fe::fail("null session or stats");
  return session->impl.stats(stats);
}

void fastembedr_hip_session_destroy(fastembedr_hip_session* session) { delete session; }

}  // extern "C"
