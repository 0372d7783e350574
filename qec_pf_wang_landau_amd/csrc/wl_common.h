// Internal helpers shared by the HIP translation units of libwl_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <initializer_list>
#include <stdexcept>
#include <string>

#include "wl_math.h"

namespace wlamd {

// Records the message of the last failed C-ABI call (thread-local).
void set_last_error(const std::string& msg);

// HIP API failures become std::runtime_error("HIP error <name> in <file> line
// <n>"), the convention of the reference's handle_cuda (src/cuda_utils.cuh:22-28).
inline void hip_check(hipError_t err, const char* file, int line)
{
  if (err != hipSuccess) {
    throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(err) + " in file " + file + " line " +
                             std::to_string(line));
  }
}

#define WL_HIP(call) ::wlamd::hip_check((call), __FILE__, __LINE__)
#define WL_LAUNCH_CHECK() ::wlamd::hip_check(hipGetLastError(), __FILE__, __LINE__)

// Runs `fn` and converts exceptions to a status code for the C ABI: the
// boundary never throws (0 = ok, 1 = invalid argument, 2 = runtime error).
template<typename F>
int guarded(F&& fn)
{
  try {
    fn();
    return 0;
  } catch (const std::invalid_argument& e) {
    set_last_error(e.what());
    return 1;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return 2;
  }
}

inline void require(bool ok, const char* what)
{
  if (!ok) {
    throw std::invalid_argument(what);
  }
}

// Throws std::invalid_argument naming the entry point when a buffer the call
// reads or writes is null.  Entry points check after their empty-call early
// return, so a call with nothing to do needs no buffers.
inline void require_buffers(const char* what, std::initializer_list<const void*> ptrs)
{
  for (const void* p : ptrs) {
    if (p == nullptr) {
      throw std::invalid_argument(std::string(what) + ": a required buffer is null");
    }
  }
}

inline unsigned grid_for(int64_t n, int block)
{
  return static_cast<unsigned>((n + block - 1) / block);
}

inline int round_up64(int n)
{
  return ((n + 63) / 64) * 64;
}

// Energy change of flipping site (i, j) of an int8 +-1 lattice with the
// reference's coupling layout: J[i*Y+j] is the bond (i,j)-(i,j+1), J[N+i*Y+j]
// the bond (i,j)-(i+1,j) (src/wl.cu:52-156).  Returns the new energy.
__device__ __forceinline__ int flip_energy_i8(const int8_t* s, const int8_t* J, int X, int Y, int boundary, int E, int i,
                                              int j)
{
  const int N = X * Y;
  const int ipp = (i + 1 < X) ? i + 1 : 0;
  const int inn = (i - 1 >= 0) ? i - 1 : X - 1;
  const int jpp = (j + 1 < Y) ? j + 1 : 0;
  const int jnn = (j - 1 >= 0) ? j - 1 : Y - 1;
  int up = s[inn * Y + j] * J[N + inn * Y + j];
  int left = s[i * Y + jnn] * J[i * Y + jnn];
  int down = s[ipp * Y + j] * J[N + i * Y + j];
  int right = s[i * Y + jpp] * J[i * Y + j];
  if (boundary == 1) {
    up = (i > 0) ? up : 0;
    down = (i < X - 1) ? down : 0;
    left = (j > 0) ? left : 0;
    right = (j < Y - 1) ? right : 0;
  }
  return E + 2 * s[i * Y + j] * (up + left + down + right);
}

}  // namespace wlamd
