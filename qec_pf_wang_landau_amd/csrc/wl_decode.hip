// Device-side end of the batched form: stitch every finished task's
// per-interval ln g into one normalised ln g(E) (src/result_handling.cpp:
// 368-431, see wl_stitch.h) and take log Z at the task's inverse temperature,
// so only a few bytes per task come back to the host for the logical-class
// decision.  One lane per task: the work per task is O(I^2 len) scalar steps,
// tiny next to the Wang-Landau run that produced it.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "wl_amd.h"
#include "wl_common.h"
#include "wl_stitch.h"

namespace wlamd {
namespace {

__global__ void k_stitch_tasks(const wl_task* tasks, int num_tasks, const double* logG, const int* start,
                               const int* end, const int* cond_sample, const double* beta, unsigned char* work,
                               const int64_t* work_off, int* out_E, double* out_lng, const int64_t* out_off,
                               int* out_n, double* log_z)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= num_tasks) {
    return;
  }
  const wl_task t = tasks[q];
  if (cond_sample[q] != -1) {
    out_n[q] = 0;
    log_z[q] = 0.0;
    return;
  }
  const int* st = start + t.interval_off;
  const int* en = end + t.interval_off;
  int total = 0;
  for (int i = 0; i < t.I; ++i) {
    total += en[i] - st[i] + 1;
  }
  unsigned char* w = work + work_off[q];
  stitch_work wk;
  wk.vals = reinterpret_cast<double*>(w);
  wk.nxt = reinterpret_cast<int*>(w + 8 * static_cast<int64_t>(total));
  wk.key_iv = wk.nxt + total;
  wk.key_e = wk.key_iv + t.I;
  wk.pres = reinterpret_cast<uint8_t*>(wk.key_e + t.I);
  const int cap = en[t.I - 1] - st[0] + 1;
  int* oE = out_E + out_off[q];
  double* oG = out_lng + out_off[q];
  const int n = stitch_dense(logG + t.hist_off, st, en, t.I, t.W, t.X, t.Y, wk, oE, oG, cap);
  out_n[q] = n;
  log_z[q] = (n > 0) ? log_partition(oE, oG, n, beta[q]) : 0.0;
}

}  // namespace
}  // namespace wlamd

using namespace wlamd;

extern "C" {

size_t wl_stitch_work_bytes(int num_intervals, int total_len)
{
  // vals f64, nxt int, key_iv/key_e int per interval, pres u8 — rounded to 8
  const size_t b = 8 * static_cast<size_t>(total_len) + 4 * static_cast<size_t>(total_len) +
                   8 * static_cast<size_t>(num_intervals) + static_cast<size_t>(total_len);
  return (b + 7) & ~static_cast<size_t>(7);
}

int wl_stitch_tasks(const wl_task* d_tasks, int num_tasks, const double* d_logG, const int* d_start, const int* d_end,
                    const int* d_cond_sample, const double* d_beta, void* d_work, const int64_t* d_work_off,
                    int* d_out_E, double* d_out_lng, const int64_t* d_out_off, int* d_out_n, double* d_log_z,
                    void* stream)
{
  return guarded([&] {
    require(num_tasks >= 0, "stitch_tasks: bad task count");
    if (num_tasks == 0) {
      return;
    }
    require_buffers("stitch_tasks", {d_tasks, d_logG, d_start, d_end, d_cond_sample, d_beta, d_work, d_work_off,
                                     d_out_E, d_out_lng, d_out_off, d_out_n, d_log_z});
    k_stitch_tasks<<<grid_for(num_tasks, 64), 64, 0, reinterpret_cast<hipStream_t>(stream)>>>(
      d_tasks, num_tasks, d_logG, d_start, d_end, d_cond_sample, d_beta, static_cast<unsigned char*>(d_work),
      d_work_off, d_out_E, d_out_lng, d_out_off, d_out_n, d_log_z);
    WL_LAUNCH_CHECK();
  });
}

}  // extern "C"
