// Shared plumbing of libnapmo_gpu.so: error state, HIP checks, device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace napmo {

// Thread-local last error; the C ABI never lets an exception escape.
void set_error(const std::string &msg);
void clear_error();

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// On failure the runtime's per-thread error state is cleared (hipGetLastError) so that
// a failed call does not resurface at the next unrelated launch check.
#define NAPMO_HIP(call)                                                                     \
  do {                                                                                      \
    hipError_t _e = (call);                                                                 \
    if (_e != hipSuccess) {                                                                 \
      (void)hipGetLastError();                                                              \
      throw ::napmo::Error(std::string(#call) + ": " + hipGetErrorString(_e) + " at " +     \
                           __FILE__ + ":" + std::to_string(__LINE__));                      \
    }                                                                                       \
  } while (0)

#define NAPMO_REQUIRE(cond, msg)                                                            \
  do {                                                                                      \
    if (!(cond)) throw ::napmo::Error(msg);                                                 \
  } while (0)

// Runs f, converting exceptions into the thread-local error; returns 0 / -1.
template <class F> int guarded(F &&f) {
  try {
    clear_error();
    f();
    return 0;
  } catch (const std::exception &e) {
    set_error(e.what());
    return -1;
  } catch (...) {
    set_error("unknown error");
    return -1;
  }
}

// Owning device allocation (hipMalloc / hipFree).
template <class T> class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { resize(n); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  DeviceBuffer(DeviceBuffer &&o) noexcept : ptr_(o.ptr_), n_(o.n_) { o.ptr_ = nullptr, o.n_ = 0; }
  DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
    if (this != &o) {
      release();
      ptr_ = o.ptr_, n_ = o.n_;
      o.ptr_ = nullptr, o.n_ = 0;
    }
    return *this;
  }
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) NAPMO_HIP(hipMalloc(&ptr_, n * sizeof(T)));
    n_ = n;
  }
  void release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr, n_ = 0;
  }
  T *get() const { return ptr_; }
  size_t size() const { return n_; }
  void upload(const T *h, size_t n, hipStream_t s) {
    resize(n);
    if (n) NAPMO_HIP(hipMemcpyAsync(ptr_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T> &h, hipStream_t s) { upload(h.data(), h.size(), s); }
  void download(T *h, size_t n, hipStream_t s) const {
    if (n) NAPMO_HIP(hipMemcpyAsync(h, ptr_, n * sizeof(T), hipMemcpyDeviceToHost, s));
  }

 private:
  T *ptr_ = nullptr;
  size_t n_ = 0;
};

// Ensures a device exists; throws a clear error otherwise (no CPU fallback).
void require_device();
// Largest dynamic LDS allocation one workgroup may request on the current device.
size_t max_lds_per_block();

inline long div_up(long a, long b) { return (a + b - 1) / b; }
inline long round_up(long a, long b) { return div_up(a, b) * b; }

}  // namespace napmo
