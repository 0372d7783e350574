// Exact density of states g(E) of small random-bond Ising lattices
// (N = X Y <= 40) by Gray-code enumeration on the GPU: the exact answer that
// Wang-Landau's ln g(E) approximates, and an exact decoder for lattices of up
// to 6x6 (ln Z per error class from the counts).
//
// Spin 0 is held at +1 (E(-s) = E(s) without a field, so every count is
// doubled at the end) and the other N-1 spins run through the binary
// reflected Gray code: state g(i) = i ^ (i >> 1), and g(i) differs from
// g(i-1) in bit ctz(i).  Each thread walks one chunk of 2^c consecutive
// codes; the chunk start i0 = t 2^c has c zero low bits, so for 0 < m < 2^c
// the bit flipped at step m is ctz(m): the same for every lane.
//
// The whole state is one u64.  Per site k, LDS holds the neighbour bits
// nb[k], the coupling bits jm[k] (bit n set when the bond k-n is
// antiferromagnetic) and the degree.  Flipping k changes the energy by
// 2 deg - 4 u, where u = number of unsatisfied bonds at k:
// x = popcount((state ^ jm[k]) & nb[k]), u = x if spin k is +1, deg - x
// otherwise.  (Periodic lattices need X, Y >= 3 so that the four neighbours
// of a site are distinct bits.)
//
// Energies follow the reference (src/wl.cu:261-362): E = -sum over right
// bonds J[i Y + j] s_ij s_i,j+1 and down bonds J[N + i Y + j] s_ij s_i+1,j,
// wrapping for periodic boundaries.
//
// Histograms: one wave per workgroup, one private counter column per lane
// in LDS (hist[E][lane]: each lane increments only its own counters, so the
// LDS adds never contend and need no return), reduced into the sample's u64
// counts at the end.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "wl_amd.h"
#include "wl_common.h"

namespace wlamd {
namespace {

constexpr int kMaxSites = 40;
constexpr int kLanes = 64;

__global__ void __launch_bounds__(kLanes) k_exact_dos(const int8_t* J, int X, int Y, int boundary, int chunk_log2,
                                                     unsigned long long* counts)
{
  __shared__ unsigned long long nb[kMaxSites], jm[kMaxSites];
  __shared__ int deg[kMaxSites];
  extern __shared__ uint32_t hist[];  // [4N + 1][kLanes]
  const int N = X * Y;
  const int bins = 4 * N + 1;
  const int s = blockIdx.y;
  const int lane = threadIdx.x;
  const int8_t* Js = J + static_cast<int64_t>(s) * 2 * N;

  for (int k = lane; k < N; k += kLanes) {
    const int i = k / Y, j = k - (k / Y) * Y;
    unsigned long long m = 0, a = 0;
    int d = 0;
    auto add = [&](int ni, int nj, int8_t bond) {
      const int n = ni * Y + nj;
      m |= 1ull << n;
      a |= static_cast<unsigned long long>(bond < 0) << n;
      ++d;
    };
    if (j + 1 < Y) {
      add(i, j + 1, Js[i * Y + j]);
    } else if (boundary == 0) {
      add(i, 0, Js[i * Y + j]);
    }
    if (j > 0) {
      add(i, j - 1, Js[i * Y + j - 1]);
    } else if (boundary == 0) {
      add(i, Y - 1, Js[i * Y + Y - 1]);
    }
    if (i + 1 < X) {
      add(i + 1, j, Js[N + i * Y + j]);
    } else if (boundary == 0) {
      add(0, j, Js[N + i * Y + j]);
    }
    if (i > 0) {
      add(i - 1, j, Js[N + (i - 1) * Y + j]);
    } else if (boundary == 0) {
      add(X - 1, j, Js[N + (X - 1) * Y + j]);
    }
    nb[k] = m;
    jm[k] = a;
    deg[k] = d;
  }
  for (int x = lane; x < bins * kLanes; x += kLanes) {
    hist[x] = 0u;
  }
  __syncthreads();

  const unsigned long long t = static_cast<unsigned long long>(blockIdx.x) * kLanes + lane;
  const unsigned long long i0 = t << chunk_log2;
  unsigned long long state = (i0 ^ (i0 >> 1)) << 1;  // bits 1..N-1; spin 0 stays +1
  // energy of the chunk's first state: every bond counted once from both ends
  int unsat2 = 0, bonds2 = 0;
  for (int k = 0; k < N; ++k) {
    const int x = __popcll((state ^ jm[k]) & nb[k]);
    unsat2 += ((state >> k) & 1ull) ? deg[k] - x : x;
    bonds2 += deg[k];
  }
  int e = (2 * unsat2 - bonds2) / 2;  // E = -(bonds - 2 unsatisfied)
  atomicAdd(&hist[(e + 2 * N) * kLanes + lane], 1u);  // own column: no contention, no return value
  const unsigned long long steps = 1ull << chunk_log2;
  for (unsigned long long m = 1; m < steps; ++m) {
    const int k = __ffsll(static_cast<long long>(m));  // ctz(m) + 1: the state bit that flips
    const int x = __popcll((state ^ jm[k]) & nb[k]);
    const int u = ((state >> k) & 1ull) ? deg[k] - x : x;
    e += 2 * deg[k] - 4 * u;
    state ^= 1ull << k;
    atomicAdd(&hist[(e + 2 * N) * kLanes + lane], 1u);
  }
  __syncthreads();
  for (int b = lane; b < bins; b += kLanes) {
    unsigned long long c = 0;
    for (int l = 0; l < kLanes; ++l) {
      c += hist[b * kLanes + l];
    }
    if (c != 0) {
      atomicAdd(&counts[static_cast<int64_t>(s) * bins + b], 2ull * c);
    }
  }
}

__global__ void k_zero_u64(unsigned long long* p, int64_t n)
{
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    p[i] = 0ull;
  }
}

}  // namespace
}  // namespace wlamd

using namespace wlamd;

extern "C" {

int wl_exact_dos(const int8_t* d_J, int X, int Y, int boundary, int num_samples, uint64_t* d_counts, void* stream)
{
  return guarded([&] {
    const int N = X * Y;
    require(X >= 2 && Y >= 2 && N <= kMaxSites, "exact_dos: lattice must have 2 <= X, Y and X Y <= 40");
    require(boundary == 0 || boundary == 1, "exact_dos: boundary must be 0 (periodic) or 1 (open)");
    require(boundary == 1 || (X >= 3 && Y >= 3), "exact_dos: periodic lattices need X, Y >= 3");
    require(num_samples >= 0 && num_samples <= 65535, "exact_dos: 0..65535 samples per call");
    require(N - 1 >= 6, "exact_dos: at least 7 sites (64 Gray codes, one per lane)");
    if (num_samples == 0) {
      return;
    }
    require_buffers("exact_dos", {d_J, d_counts});
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int bins = 4 * N + 1;
    const int64_t total = static_cast<int64_t>(num_samples) * bins;
    k_zero_u64<<<grid_for(total, 256), 256, 0, st>>>(reinterpret_cast<unsigned long long*>(d_counts), total);
    WL_LAUNCH_CHECK();
    // 2^(N-1) codes per sample over up to 2^16 threads of 64-lane workgroups
    const int codes_log2 = N - 1;
    const int threads_log2 = codes_log2 < 16 ? codes_log2 : 16;
    const int chunk_log2 = codes_log2 - threads_log2;
    const int blocks = (1 << threads_log2) / kLanes;
    const size_t lds = static_cast<size_t>(bins) * kLanes * sizeof(uint32_t);
    dim3 grid(static_cast<unsigned>(blocks), static_cast<unsigned>(num_samples));
    k_exact_dos<<<grid, kLanes, lds, st>>>(d_J, X, Y, boundary, chunk_log2,
                                           reinterpret_cast<unsigned long long*>(d_counts));
    WL_LAUNCH_CHECK();
  });
}

}  // extern "C"
