// Self-test of the owning device buffer that the host drivers use
// (wl_device.h), case for case the checks the reference runs on its
// device_ptr (test/test_device_ptr.cu): size of an empty buffer, a buffer
// initialised from a host range, a kernel writing into it, and every
// out-of-range copy refused with std::out_of_range.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <vector>

#include "wl_amd.h"
#include "wl_common.h"
#include "wl_device.h"

namespace wlamd {
namespace {

__global__ void k_iota(int* p, int n, int base)
{
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) {
    p[i] = base + i;
  }
}

template<typename F>
bool throws_out_of_range(F&& f)
{
  try {
    f();
  } catch (const std::out_of_range&) {
    return true;
  }
  return false;
}

}  // namespace
}  // namespace wlamd

using namespace wlamd;

extern "C" {

int wl_selftest_device_buffer(int* failed_mask)
{
  return guarded([&] {
    require(failed_mask != nullptr, "selftest_device_buffer: failed_mask is required");
    int m = 0;
    {  // empty buffer of a given size, zero-filled
      device_buffer<int> d(1000);
      const std::vector<int> h = d.to_host();
      bool zero = h.size() == 1000;
      for (int v : h) {
        zero = zero && v == 0;
      }
      m |= (d.size() == 1000 && zero) ? 0 : 1;
    }
    {  // initialised from a host range
      const std::vector<int> src(1000, 5);
      device_buffer<int> d(src);
      m |= (d.size() == 1000 && d.to_host() == src) ? 0 : 2;
    }
    {  // a kernel writes the buffer
      device_buffer<int> d(1000);
      k_iota<<<grid_for(1000, 256), 256>>>(d.data(), 1000, 7);
      WL_LAUNCH_CHECK();
      WL_HIP(hipDeviceSynchronize());
      std::vector<int> want(1000);
      for (int i = 0; i < 1000; ++i) {
        want[i] = 7 + i;
      }
      m |= d.to_host() == want ? 0 : 4;
    }
    {  // host-to-device copy larger than the buffer
      device_buffer<int> d(10);
      const std::vector<int> h(11, 1);
      m |= throws_out_of_range([&] { d.to_device(h); }) ? 0 : 8;
    }
    {  // device-to-host copy into a host range longer than the buffer
      device_buffer<int> d(10);
      std::vector<int> h(11);
      m |= throws_out_of_range([&] { d.to_host(h); }) ? 0 : 16;
    }
    {  // offset + count beyond the buffer
      device_buffer<int> d(10);
      std::vector<int> h(5);
      m |= throws_out_of_range([&] { d.to_host(h, 6, 5); }) ? 0 : 32;
    }
    {  // count larger than the host range
      device_buffer<int> d(100);
      std::vector<int> h(5);
      m |= throws_out_of_range([&] { d.to_host(h, 0, 15); }) ? 0 : 64;
    }
    {  // offset past the end of the buffer
      device_buffer<int> d(10);
      std::vector<int> h(1);
      m |= throws_out_of_range([&] { d.to_host(h, 10, 1); }) ? 0 : 128;
    }
    {  // an in-range copy at an offset reads the right elements
      device_buffer<int> d(std::vector<int>{0, 1, 2, 3, 4, 5, 6, 7});
      std::vector<int> h(3);
      d.to_host(h, 4, 3);
      m |= (h == std::vector<int>{4, 5, 6}) ? 0 : 256;
    }
    *failed_mask = m;
  });
}

}  // extern "C"
