// Shared host/device plumbing for the fastEmbedR HIP backend (gfx950).
//
// Error convention (mirrors embedding_cuda_kernels.cpp:65-82): every C-ABI
// entry returns 0 on success and 1 on failure, with the message kept in a
// thread-local string that fastembedr_hip_last_error() returns.  Nothing
// throws across the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <utility>
#include <vector>

namespace fe {

void set_error(const std::string& message);
const std::string& last_error();
void clear_error();

inline int fail(const std::string& message) {
  set_error(message);
  return 1;
}

inline int check(hipError_t code, const char* where) {
  if (code == hipSuccess) return 0;
  set_error(std::string(where) + ": " + hipGetErrorString(code));
  return 1;
}

// Runs f() and turns an exception into the error convention above: nothing
// throws across the C-ABI.
template <class F>
int guarded(const char* what, F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return fail(std::string(what) + ": out of host memory");
  } catch (const std::exception& e) {
    return fail(std::string(what) + ": " + e.what());
  } catch (...) {
    return fail(std::string(what) + ": unexpected error");
  }
}

#define FE_CHECK(expr)                        \
  do {                                        \
    if (::fe::check((expr), #expr)) return 1; \
  } while (0)

#define FE_LAUNCH_CHECK(name)                                   \
  do {                                                          \
    if (::fe::check(hipGetLastError(), name " launch")) return 1; \
  } while (0)

inline std::size_t align_up(std::size_t bytes, std::size_t a = 256) {
  return (bytes + a - 1) / a * a;
}

// One device allocation carved into aligned sub-buffers (the reference's
// bump workspace, embedding_cuda_kernels.cpp:88-127).  Owned; freed in the
// destructor.
class Arena {
 public:
  Arena() = default;
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;
  ~Arena() { release(); }

  void release() {
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
    cap_ = off_ = 0;
  }

  // Reserve `bytes` (already the sum of align_up'd requests).
  int init(std::size_t bytes, const char* label) {
    release();
    cap_ = align_up(bytes);
    if (cap_ == 0) return 0;
    std::size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
      // preflight: require bytes <= free - free/20 (embedding_cuda_kernels.cpp:234-251)
      if (cap_ > free_b - free_b / 20) {
        return fail(std::string(label) + ": HIP memory request exceeds available memory; required " +
                    std::to_string(cap_) + " bytes, free " + std::to_string(free_b) + " bytes");
      }
    }
    return check(hipMalloc(&base_, cap_), label);
  }

  template <typename T>
  T* take(std::size_t count) {
    const std::size_t bytes = align_up(count * sizeof(T));
    if (count == 0) return nullptr;
    if (!base_ || off_ + bytes > cap_) return nullptr;
    T* p = reinterpret_cast<T*>(static_cast<unsigned char*>(base_) + off_);
    off_ += bytes;
    return p;
  }

  std::size_t capacity() const { return cap_; }

 private:
  void* base_ = nullptr;
  std::size_t cap_ = 0;
  std::size_t off_ = 0;
};

// Byte-size planner for an Arena: call add<T>(count) for each buffer first.
struct ArenaPlan {
  std::size_t bytes = 0;
  template <typename T>
  void add(std::size_t count) {
    if (count) bytes += align_up(count * sizeof(T));
  }
};

class Stream {
 public:
  Stream() = default;
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  ~Stream() {
    if (s_) (void)hipStreamDestroy(s_);
  }
  int init() { return check(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "hipStreamCreate"); }
  hipStream_t get() const { return s_; }

 private:
  hipStream_t s_ = nullptr;
};

// A ring of timing-free events for fork/join points between two streams.  An
// enqueue that forks takes the next pair; a wait refers to the record made
// just before it, so reusing the ring after a full turn is safe on a stream
// and inside one capture of up to size() / 2 forks.
class EventRing {
 public:
  EventRing() = default;
  EventRing(const EventRing&) = delete;
  EventRing& operator=(const EventRing&) = delete;
  ~EventRing() {
    for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
  }
  int init(int count) {
    for (int i = 0; i < count; ++i) {
      hipEvent_t e = nullptr;
      if (check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate")) return 1;
      ev_.push_back(e);
    }
    return 0;
  }
  bool ready() const { return !ev_.empty(); }
  hipEvent_t next() {
    hipEvent_t e = ev_[at_];
    at_ = (at_ + 1) % ev_.size();
    return e;
  }
  std::size_t size() const { return ev_.size(); }

 private:
  std::vector<hipEvent_t> ev_;
  std::size_t at_ = 0;
};

class GraphExec {
 public:
  GraphExec() = default;
  GraphExec(const GraphExec&) = delete;
  GraphExec& operator=(const GraphExec&) = delete;
  ~GraphExec() { reset(); }
  void reset() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    exec_ = nullptr;
  }
  int instantiate(hipGraph_t g) {
    reset();
    const hipError_t rc = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    return check(rc, "hipGraphInstantiate");
  }
  hipGraphExec_t get() const { return exec_; }
  explicit operator bool() const { return exec_ != nullptr; }

 private:
  hipGraphExec_t exec_ = nullptr;
};

}  // namespace fe
