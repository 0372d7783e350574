// Owning device buffer for host drivers: the HIP counterpart of the
// reference's wl::device_ptr (src/cuda_utils.cuh:112-265).  Non-copyable,
// size-checked copies that throw std::out_of_range, HIP failures as
// std::runtime_error.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

namespace wlamd {

inline void hip_throw(hipError_t err, const char* what)
{
  if (err != hipSuccess) {
    throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(err) + " in " + what);
  }
}

template<typename T>
class device_buffer {
 public:
  explicit device_buffer(size_t n = 0) : n_(n)
  {
    if (n_ > 0) {
      hip_throw(hipMalloc(&p_, n_ * sizeof(T)), "hipMalloc");
      hip_throw(hipMemset(p_, 0, n_ * sizeof(T)), "hipMemset");
    }
  }
  explicit device_buffer(const std::vector<T>& host) : device_buffer(host.size()) { to_device(host); }
  ~device_buffer()
  {
    if (p_ != nullptr) {
      (void)hipFree(p_);
    }
  }
  device_buffer(const device_buffer&) = delete;
  device_buffer& operator=(const device_buffer&) = delete;

  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }

  void to_device(const std::vector<T>& host, size_t offset = 0)
  {
    if (offset + host.size() > n_) {
      throw std::out_of_range("device_buffer: host-to-device copy out of range");
    }
    if (!host.empty()) {
      hip_throw(hipMemcpy(p_ + offset, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice), "H2D");
    }
  }
  void to_host(std::vector<T>& host, size_t offset = 0, size_t count = static_cast<size_t>(-1)) const
  {
    if (count == static_cast<size_t>(-1)) {
      count = host.size();
    }
    if (offset + count > n_ || count > host.size()) {
      throw std::out_of_range("device_buffer: device-to-host copy out of range");
    }
    if (count > 0) {
      hip_throw(hipMemcpy(host.data(), p_ + offset, count * sizeof(T), hipMemcpyDeviceToHost), "D2H");
    }
  }
  std::vector<T> to_host() const
  {
    std::vector<T> h(n_);
    to_host(h);
    return h;
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

}  // namespace wlamd
