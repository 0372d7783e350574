// Host-side domain logic around the kernels: the interval partition
// (src/interval.cpp), spectrum extraction and seed-lattice selection from a
// prerun (src/file_io.cpp:558-659), stitching (src/result_handling.cpp) and
// log Z for the logical-class decision.  C entry points in include/wl_host.h.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "wl_host.h"
#include "wl_math.h"
#include "wl_stitch.h"

namespace wlamd {
void set_last_error(const std::string& msg);
}

namespace {

template<typename F>
int guarded_host(F&& fn)
{
  try {
    fn();
    return 0;
  } catch (const std::invalid_argument& e) {
    wlamd::set_last_error(e.what());
    return 1;
  } catch (const std::exception& e) {
    wlamd::set_last_error(e.what());
    return 2;
  }
}

// A required argument that is null is an invalid argument (status 1).
void require_args(const char* what, std::initializer_list<const void*> ptrs)
{
  for (const void* p : ptrs) {
    if (p == nullptr) {
      throw std::invalid_argument(std::string(what) + ": a required argument is null");
    }
  }
}

}  // namespace

extern "C" {

// src/interval.cpp:5-41.  The float arithmetic of the reference is kept:
// len = trunc(R / (1 + o (I - 1))) and step = trunc(o len) in float.
int wl_interval_results(int e_min, int e_max, unsigned int num_intervals, int num_walkers, float overlap, int* h_start,
                        int* h_end, int64_t* len_histogram_over_all_walkers, int* len_interval)
{
  return guarded_host([&] {
    require_args("interval_results", {h_start, h_end, len_histogram_over_all_walkers, len_interval});
    if (overlap > 1 || overlap < 0) {
      throw std::invalid_argument("Overlap is out of bounds.");
    }
    if (num_intervals == 0) {
      throw std::invalid_argument("The interval arguments create empty intervals.");
    }
    const int e_range = e_max - e_min + 1;
    const float denom = 1.0F + overlap * static_cast<float>(num_intervals - 1);
    const int len = static_cast<int>(static_cast<float>(e_range) / denom);
    const int step = static_cast<int>(overlap * static_cast<float>(len));
    int64_t total = 0;
    int s = e_min;
    for (unsigned int i = 0; i < num_intervals; ++i) {
      h_start[i] = s;
      h_end[i] = (i + 1 < num_intervals) ? s + len - 1 : e_max;
      total += static_cast<int64_t>(num_walkers) * (h_end[i] - h_start[i] + 1);
      if (h_start[i] >= h_end[i]) {
        throw std::invalid_argument("The interval arguments create empty intervals.");
      }
      s += step;
    }
    *len_histogram_over_all_walkers = total;
    *len_interval = len;
  });
}

// src/file_io.cpp:558-616 for one sample: E_min/E_max are the first and last
// visited energies of the prerun histogram over [-2N, 2N]; the mask is the
// histogram's 0/1 flags between them.
int wl_energy_spectrum(const uint8_t* hist, int len_hist, int X, int Y, int* E_min, int* E_max, int8_t* mask,
                       int* len_mask)
{
  return guarded_host([&] {
    if (len_hist != 4 * X * Y + 1) {
      throw std::invalid_argument("histogram length does not match 4XY+1");
    }
    require_args("energy_spectrum", {hist, E_min, E_max, mask, len_mask});
    int first = -1, last = -1;
    for (int e = 0; e < len_hist; ++e) {
      if (hist[e] == 1) {
        if (first < 0) {
          first = e;
        }
        last = e;
      }
    }
    if (first < 0) {
      throw std::runtime_error("Error: Histogram does not contain 1.");
    }
    *E_min = -2 * X * Y + first;
    *E_max = -2 * X * Y + last;
    *len_mask = last - first + 1;
    for (int e = first; e <= last; ++e) {
      mask[e - first] = static_cast<int8_t>(hist[e]);
    }
  });
}

// Spacing of the energies present in a spectrum mask (gcd of their
// differences; 0 for a single energy or no mask).
int wl_energy_stride(const int8_t* mask, int len)
{
  if (mask == nullptr) {
    return 0;
  }
  int g = 0, prev = -1;
  for (int e = 0; e < len; ++e) {
    if (mask[e] == 1) {
      if (prev >= 0) {
        g = std::gcd(g, e - prev);
      }
      prev = e;
    }
  }
  return g;
}

// src/file_io.cpp:618-659 for one sample: for every interval, the lowest-energy
// found lattice with energy in [start, end], replicated for its W walkers.
int wl_select_seed_lattices(const int8_t* found, const int* found_energy, int num_found, int N, const int* start,
                            const int* end, int I, int W, int8_t* out)
{
  return guarded_host([&] {
    if (num_found > 0) {
      require_args("select_seed_lattices", {found, found_energy});
    }
    if (I > 0) {
      require_args("select_seed_lattices", {start, end, out});
    }
    std::map<int, int> by_energy;  // energy -> index of the found lattice
    for (int q = 0; q < num_found; ++q) {
      by_energy.emplace(found_energy[q], q);
    }
    for (int k = 0; k < I; ++k) {
      int pick = -1;
      for (const auto& [e, q] : by_energy) {
        if (e >= start[k] && e <= end[k]) {
          pick = q;
          break;
        }
      }
      if (pick < 0) {
        throw std::runtime_error("Error: Did not find lattice configuration with suiting energy for interval [" +
                                 std::to_string(start[k]) + ", " + std::to_string(end[k]) + "]");
      }
      for (int w = 0; w < W; ++w) {
        std::memcpy(out + (static_cast<int64_t>(k) * W + w) * N, found + static_cast<int64_t>(pick) * N,
                    static_cast<size_t>(N));
      }
    }
  });
}

int wl_stitch_log_g(const double* logG_sample, const int* start, const int* end, int I, int W, int X, int Y,
                    int* out_E, double* out_lng, int max_out, int* n_out)
{
  return guarded_host([&] {
    if (I < 1 || W < 1) {
      throw std::invalid_argument("stitch: bad shape");
    }
    require_args("stitch_log_g", {logG_sample, start, end, out_E, out_lng, n_out});
    int total = 0;
    for (int i = 0; i < I; ++i) {
      total += end[i] - start[i] + 1;
    }
    std::vector<double> vals(total);
    std::vector<uint8_t> pres(total);
    std::vector<int> nxt(total), kiv(I), ke(I);
    wlamd::stitch_work wk{vals.data(), pres.data(), nxt.data(), kiv.data(), ke.data()};
    const int n = wlamd::stitch_dense(logG_sample, start, end, I, W, X, Y, wk, out_E, out_lng, max_out);
    if (n == -2) {
      throw std::runtime_error("stitch: no overlapping intervals to stitch");
    }
    if (n == -3) {
      throw std::runtime_error("stitch: stitching energy not found in one of the intervals");
    }
    if (n == -4) {
      throw std::invalid_argument("stitch: output buffer too small");
    }
    *n_out = n;
  });
}

double wl_log_partition(const int* E, const double* lng, int n, double beta)
{
  if (n > 0 && (E == nullptr || lng == nullptr)) {
    return std::nan("");
  }
  return wlamd::log_partition(E, lng, n, beta);
}

double wl_nishimori_beta(double p)
{
  return 0.5 * wlamd::wl_log((1.0 - p) / p);
}

// Class 0 ('i') is the coset of the sampled error itself, so a sample fails
// unless class 0 is the most likely class.  Among m tied maxima the decoder
// picks uniformly, which fails with probability 1 - 1/m when class 0 is one
// of them.
int wl_decide_classes(const double* log_z, int num_samples, int num_classes, int* decision, double* fail_weight)
{
  return guarded_host([&] {
    if (num_samples < 0 || num_classes < 1) {
      throw std::invalid_argument("decide_classes: need num_samples >= 0 and num_classes >= 1");
    }
    if (num_samples > 0 && (log_z == nullptr || decision == nullptr || fail_weight == nullptr)) {
      throw std::invalid_argument("decide_classes: log_z, decision and fail_weight are required");
    }
    for (int s = 0; s < num_samples; ++s) {
      const double* z = log_z + static_cast<int64_t>(s) * num_classes;
      bool finite = true;
      int best = 0;
      for (int c = 0; c < num_classes; ++c) {
        finite = finite && std::isfinite(z[c]);
        if (z[c] > z[best]) {
          best = c;
        }
      }
      if (!finite) {
        decision[s] = -1;
        fail_weight[s] = std::nan("");
        continue;
      }
      int ties = 0;
      for (int c = 0; c < num_classes; ++c) {
        ties += z[c] == z[best] ? 1 : 0;
      }
      decision[s] = best;
      fail_weight[s] = z[0] == z[best] ? 1.0 - 1.0 / static_cast<double>(ties) : 1.0;
    }
  });
}

// Wilson score interval of a binomial proportion k/n at z standard errors:
// (k + z^2/2 -+ z sqrt(k (n-k)/n + z^2/4)) / (n + z^2).
int wl_wilson_interval(double failures, int n, double z, double* lo, double* hi)
{
  return guarded_host([&] {
    if (n < 1 || failures < 0.0 || failures > static_cast<double>(n) || !(z > 0.0)) {
      throw std::invalid_argument("wilson_interval: need n >= 1, 0 <= failures <= n and z > 0");
    }
    if (lo == nullptr || hi == nullptr) {
      throw std::invalid_argument("wilson_interval: lo and hi are required");
    }
    const double nn = static_cast<double>(n);
    const double z2 = z * z;
    const double centre = failures + 0.5 * z2;
    const double half = z * std::sqrt(failures * (nn - failures) / nn + 0.25 * z2);
    *lo = std::max(0.0, (centre - half) / (nn + z2));
    *hi = std::min(1.0, (centre + half) / (nn + z2));
  });
}

}  // extern "C"
