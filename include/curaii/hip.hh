// RAII and error wrappers over the HIP runtime, in the shape of the reference's curaii
// layer (src/curaii/include/curaii/cuda.hh:29-97: CudaError + CUDA_CHECK, unique
// device/host pointers with make_unique_*, a stream owner).  Here: HipError + HIP_CHECK,
// and HFX_CHECK turning a libhfx status into the exception the reference's nodes throw:
// std::invalid_argument for invalid settings, HipError for a HIP runtime failure (the
// reference's CudaError), HfxError for an unsupported configuration or an exhausted device.
#pragma once

#include <hip/hip_runtime_api.h>

#include <memory>
#include <stdexcept>
#include <string>

#include "hfx.h"

namespace curaii {

class HipError : public std::runtime_error {
 public:
  HipError(hipError_t e, const char *expr, const char *file, int line)
      : std::runtime_error(std::string(expr) + " failed at " + file + ":" + std::to_string(line) + ": " +
                           hipGetErrorString(e)),
        code(e) {}
  // a HIP failure reported through libhfx's C ABI, which carries its text but not the code
  explicit HipError(const std::string &msg) : std::runtime_error(msg), code(hipErrorUnknown) {}
  hipError_t code;
};

#define HIP_CHECK(expr)                                                     \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) throw ::curaii::HipError(e_, #expr, __FILE__, __LINE__); \
  } while (0)

class HfxError : public std::runtime_error {
 public:
  HfxError(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
  int code;
};

inline void hfx_check(int status, const char *expr) {
  if (status == HFX_OK) return;
  const std::string msg = std::string(expr) + ": " + hfx_last_error();
  if (status == HFX_E_INVALID) throw std::invalid_argument(msg);
  if (status == HFX_E_HIP) throw HipError(msg);
  throw HfxError(status, msg);
}

#define HFX_CHECK(expr) ::curaii::hfx_check((expr), #expr)

struct DeviceDeleter {
  void operator()(void *p) const noexcept { (void)hipFree(p); }
};
template <typename T> using unique_device_ptr = std::unique_ptr<T, DeviceDeleter>;

template <typename T> unique_device_ptr<T> make_unique_device_ptr(size_t count) {
  void *p = nullptr;
  HIP_CHECK(hipMalloc(&p, count * sizeof(T) > 0 ? count * sizeof(T) : 16));
  return unique_device_ptr<T>(static_cast<T *>(p));
}

struct HostDeleter {
  void operator()(void *p) const noexcept { (void)hipHostFree(p); }
};
template <typename T> using unique_host_ptr = std::unique_ptr<T, HostDeleter>;

// Pinned host memory (allocated pinned, never registered after the fact).
template <typename T> unique_host_ptr<T> make_unique_host_ptr(size_t count) {
  void *p = nullptr;
  HIP_CHECK(hipHostMalloc(&p, count * sizeof(T) > 0 ? count * sizeof(T) : 16, hipHostMallocDefault));
  return unique_host_ptr<T>(static_cast<T *>(p));
}

// Owns a default-priority stream.
class HipStream {
 public:
  HipStream() { HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking)); }
  ~HipStream() {
    if (s_) (void)hipStreamDestroy(s_);
  }
  HipStream(const HipStream &) = delete;
  HipStream &operator=(const HipStream &) = delete;
  hipStream_t get() const { return s_; }

 private:
  hipStream_t s_ = nullptr;
};

}  // namespace curaii
