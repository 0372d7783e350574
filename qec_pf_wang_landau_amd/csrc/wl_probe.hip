// LDS read-bandwidth probe: the measured LDS peak of the device, against
// which the batched kernels' LDS traffic (34 B per attempt) is compared.
//
// Each workgroup of 256 threads fills a 32 KiB LDS buffer, then streams
// `reps` x 8 reads of 16 bytes per thread (ds_read_b128); lane l of a wave
// reads the l-th 16-byte chunk of a 1 KiB row, so a wave's read touches every
// bank once.  The row moves every repetition, so no read can be hoisted out
// of the loop, and the XOR of everything read is written to the sink.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "wl_amd.h"
#include "wl_common.h"

namespace wlamd {
namespace {

constexpr int kProbeThreads = 256;
constexpr int kProbeChunks = 2048;  // 16-byte chunks: 32 KiB of LDS per workgroup
constexpr int kProbeUnroll = 8;

__global__ void __launch_bounds__(kProbeThreads) k_lds_read(int reps, uint32_t* sink)
{
  __shared__ uint4 buf[kProbeChunks];
  const int tid = threadIdx.x;
  for (int i = tid; i < kProbeChunks; i += kProbeThreads) {
    const uint32_t v = static_cast<uint32_t>(i) * 2654435761u;
    buf[i] = make_uint4(v, v ^ 1u, v ^ 2u, v ^ 3u);
  }
  __syncthreads();
  uint4 acc[kProbeUnroll];
#pragma unroll
  for (int u = 0; u < kProbeUnroll; ++u) {
    acc[u] = make_uint4(0u, 0u, 0u, 0u);
  }
  int base = tid;
  for (int r = 0; r < reps; ++r) {
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) {
      const uint4 v = buf[(base + u * kProbeThreads) & (kProbeChunks - 1)];
      acc[u].x ^= v.x;
      acc[u].y ^= v.y;
      acc[u].z ^= v.z;
      acc[u].w ^= v.w;
    }
    base += 64;
  }
  uint32_t x = 0u;
#pragma unroll
  for (int u = 0; u < kProbeUnroll; ++u) {
    x ^= acc[u].x ^ acc[u].y ^ acc[u].z ^ acc[u].w;
  }
  sink[static_cast<int64_t>(blockIdx.x) * kProbeThreads + tid] = x;
}

}  // namespace
}  // namespace wlamd

using namespace wlamd;

extern "C" {

int wl_probe_lds_read(int blocks, int reps, uint32_t* d_sink, void* stream)
{
  return guarded([&] {
    require(blocks >= 1 && blocks <= (1 << 20), "probe_lds_read: 1 <= blocks <= 2^20");
    require(reps >= 1, "probe_lds_read: reps >= 1");
    require(d_sink != nullptr, "probe_lds_read: sink buffer of blocks x 256 words is required");
    k_lds_read<<<blocks, kProbeThreads, 0, reinterpret_cast<hipStream_t>(stream)>>>(reps, d_sink);
    WL_LAUNCH_CHECK();
  });
}

double wl_probe_lds_read_bytes(int blocks, int reps)
{
  return static_cast<double>(blocks) * kProbeThreads * static_cast<double>(reps) * kProbeUnroll * 16.0;
}

}  // extern "C"
