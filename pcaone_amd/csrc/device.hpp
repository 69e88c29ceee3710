// HIP runtime helpers: status checks and an owning device buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <utility>

#include "common.hpp"

namespace pcaone {

#define PCAONE_HIP(call)                                                                          \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      ::pcaone::fail(::pcaone::kErrDevice, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)

// Checks the launch of the kernel just enqueued.
inline void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) fail(kErrDevice, std::string("launch ") + what + ": " + hipGetErrorString(e));
}

template <class T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t n) { resize(n); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  // A move carries the allocation with its capacity; the source is left empty.
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_), cap_(o.cap_) { o.p_ = nullptr, o.n_ = o.cap_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_, n_ = o.n_, cap_ = o.cap_;
      o.p_ = nullptr, o.n_ = o.cap_ = 0;
    }
    return *this;
  }
  // Reallocates only when growing; contents are not preserved.
  void resize(size_t n) {
    if (n <= cap_) {
      n_ = n;
      return;
    }
    release();
    if (n == 0) return;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess)
      fail(kErrDevice, "hipMalloc(" + std::to_string(n * sizeof(T)) + " bytes): " + hipGetErrorString(e));
    p_ = static_cast<T*>(p);
    n_ = cap_ = n;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = cap_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  T* p_ = nullptr;
  size_t n_ = 0, cap_ = 0;
};

// Page-locked host memory (hipHostMalloc), for staging host <-> device copies.
class PinnedBuf {
 public:
  PinnedBuf() = default;
  ~PinnedBuf() { release(); }
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  // Reallocates only when growing; contents are not preserved.
  void resize(size_t bytes) {
    if (bytes <= cap_) return;
    release();
    void* p = nullptr;
    hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
    if (e != hipSuccess)
      fail(kErrDevice, "hipHostMalloc(" + std::to_string(bytes) + " bytes): " + hipGetErrorString(e));
    p_ = static_cast<unsigned char*>(p);
    cap_ = bytes;
  }
  void release() {
    if (p_) (void)hipHostFree(p_);
    p_ = nullptr;
    cap_ = 0;
  }
  unsigned char* get() const { return p_; }

 private:
  unsigned char* p_ = nullptr;
  size_t cap_ = 0;
};

// Owning HIP event (timing disabled: used for ordering only).
class Event {
 public:
  Event() { PCAONE_HIP(hipEventCreateWithFlags(&ev_, hipEventDisableTiming)); }
  ~Event() { (void)hipEventDestroy(ev_); }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
  hipEvent_t get() const { return ev_; }

 private:
  hipEvent_t ev_ = nullptr;
};

}  // namespace pcaone
