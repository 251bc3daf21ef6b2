// Error type, HIP status checks and grow-only device buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>

namespace lm {

// Every failure surfaces as lm::Error; the C ABI turns it into a status code
// plus a thread-local message (Python re-raises it as RuntimeError).
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define LM_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t lm_e_ = (call);                                                          \
    if (lm_e_ != hipSuccess)                                                            \
      throw ::lm::Error(std::string("[libmobility_amd] HIP error: ") +                 \
                        hipGetErrorString(lm_e_) + " (" #call ")");                     \
  } while (0)

// Device allocation that only grows.  The solver keeps one per role, which
// plays the part of the reference's cached pool allocator
// (include/memory/allocator.h:127-292): steady-state calls allocate nothing.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }

  void* reserve(size_t bytes) {
    if (bytes > cap_) {
      release();
      size_t want = bytes < 256 ? 256 : bytes;
      LM_HIP(hipMalloc(&p_, want));
      cap_ = want;
    }
    return p_;
  }
  // Grow while keeping the first `keep` bytes.
  void* grow_keep(size_t bytes, size_t keep, hipStream_t s) {
    if (bytes <= cap_) return p_;
    void* np = nullptr;
    LM_HIP(hipMalloc(&np, bytes));
    if (p_ && keep) LM_HIP(hipMemcpyAsync(np, p_, keep < cap_ ? keep : cap_, hipMemcpyDeviceToDevice, s));
    LM_HIP(hipStreamSynchronize(s));
    release();
    p_ = np;
    cap_ = bytes;
    return p_;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    cap_ = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p_);
  }
  size_t capacity() const { return cap_; }

 private:
  void* p_ = nullptr;
  size_t cap_ = 0;
};

// Pinned host allocation that only grows (small device<->host exchanges of
// per-system flags and errors during a solve).
class HostBuf {
 public:
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() { release(); }
  // flags: hipHostMalloc flags of a new allocation (the first reserve that
  // allocates decides them)
  void* reserve(size_t bytes, unsigned flags = hipHostMallocDefault) {
    if (bytes > cap_) {
      release();
      size_t want = bytes < 256 ? 256 : bytes;
      LM_HIP(hipHostMalloc(&p_, want, flags));
      cap_ = want;
    }
    return p_;
  }
  // Address of the allocation in the current device's address space.
  template <class T>
  T* device_as() const {
    void* d = nullptr;
    LM_HIP(hipHostGetDevicePointer(&d, p_, 0));
    return static_cast<T*>(d);
  }
  void release() {
    if (p_) (void)hipHostFree(p_);
    p_ = nullptr;
    cap_ = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p_);
  }

 private:
  void* p_ = nullptr;
  size_t cap_ = 0;
};

}  // namespace lm
