// prerun: disorder samples, logical-error chain, spectrum discovery and seed
// lattices per energy interval, written to init/.../prerun_results.h5.
// Same flags and outputs as the reference's src/prerun.cu:15-167.

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "wl_amd.h"
#include "wl_device.h"
#include "wl_host.h"
#include "wl_io.h"

using wlamd::device_buffer;

namespace {

void ok(int rc, const char* what)
{
  if (rc != 0) {
    throw std::runtime_error(std::string(what) + ": " + wl_last_error());
  }
}

}  // namespace

int main(int argc, char** argv)
{
  wlio::prerun_options o;
  int code = 0;
  if (!wlio::parse_prerun_options(argc, argv, o, code)) {
    return code;
  }
  try {
    const int N = o.X * o.Y;
    const int e_min = -2 * N;
    const int e_max = 2 * N;
    std::vector<int> iv_start(o.num_intervals), iv_end(o.num_intervals);
    int64_t len_all = 0;
    int len_interval = 0;
    ok(wl_interval_results(e_min, e_max, o.num_intervals, 1, 1.0F, iv_start.data(), iv_end.data(), &len_all,
                           &len_interval),
       "interval partition");
    wlio::log("Intervals for the run");
    for (unsigned k = 0; k < o.num_intervals; ++k) {
      wlio::log(std::to_string(iv_start[k]) + " " + std::to_string(iv_end[k]));
    }
    const int S = o.num_interactions;
    const int Wp = o.walker_per_interactions;
    const size_t total_walker = static_cast<size_t>(S) * Wp;
    const size_t total_intervals = static_cast<size_t>(S) * o.num_intervals;
    const size_t len_hist = static_cast<size_t>(e_max - e_min + 1);

    device_buffer<uint64_t> d_hist(S * len_hist);
    device_buffer<uint64_t> d_seed_offset(total_walker);
    device_buffer<int8_t> d_lattice(total_walker * N);
    device_buffer<float> d_probs(total_walker);
    device_buffer<int8_t> d_J(static_cast<size_t>(S) * 2 * N);
    device_buffer<int> d_energy(total_walker);
    device_buffer<int8_t> d_found(total_intervals * N);
    device_buffer<int> d_flag(total_intervals);
    device_buffer<int> d_interval_energy(total_intervals);
    device_buffer<int> d_off_walker(total_walker);
    device_buffer<int> d_off_interval(total_intervals);

    ok(wl_init_interactions(d_J.data(), N, S, o.seed, o.prob_interactions, nullptr), "init_interactions");
    switch (o.logical_error_type) {
      case 'h':
        ok(wl_apply_x_horizontal_error(d_J.data(), o.X, o.Y, S, nullptr), "horizontal chain");
        wlio::log("Applied horizontal error cycle.");
        break;
      case 'v':
        ok(wl_apply_x_vertical_error(d_J.data(), o.X, o.Y, S, nullptr), "vertical chain");
        wlio::log("Applied vertical error cycle.");
        break;
      case 'c':
        ok(wl_apply_x_horizontal_error(d_J.data(), o.X, o.Y, S, nullptr), "horizontal chain");
        ok(wl_apply_x_vertical_error(d_J.data(), o.X, o.Y, S, nullptr), "vertical chain");
        wlio::log("Applied both horizontal and vertical error cycle.");
        break;
      default:
        wlio::log("Trivial error class.");
        break;
    }
    ok(wl_init_lattice(d_lattice.data(), d_probs.data(), o.X, o.Y, total_walker, o.seed, Wp, nullptr),
       "init_lattice");
    ok(wl_init_offsets_lattice(d_off_walker.data(), o.X, o.Y, total_walker, nullptr), "offsets");
    ok(wl_init_offsets_lattice(d_off_interval.data(), o.X, o.Y, total_intervals, nullptr), "offsets");
    ok(wl_calc_energy(o.boundary, d_energy.data(), d_lattice.data(), d_J.data(), d_off_walker.data(), o.X, o.Y,
                      total_walker, Wp, nullptr),
       "calc_energy");
    ok(wl_wang_landau_pre_run(d_energy.data(), d_flag.data(), d_lattice.data(), d_found.data(), d_hist.data(),
                              d_seed_offset.data(), d_J.data(), d_off_walker.data(), e_min, e_max, o.num_iterations,
                              o.X, o.Y, o.seed, o.num_intervals, len_interval, total_walker, Wp, o.boundary, nullptr),
       "wang_landau_pre_run");
    ok(wl_calc_energy(o.boundary, d_interval_energy.data(), d_found.data(), d_J.data(), d_off_interval.data(), o.X,
                      o.Y, total_intervals, static_cast<int>(o.num_intervals), nullptr),
       "calc_energy (found lattices)");
    wlamd::hip_throw(hipDeviceSynchronize(), "prerun");

    if (!o.skip_output) {
      const std::string path = wlio::build_prerun_path(o);
      wlio::create_directories(path);
      wlio::write_prerun_results(d_hist.to_host(), d_J.to_host(), d_found.to_host(), d_interval_energy.to_host(), o,
                                 len_hist, path);
    }
    wlio::log("Finished prerun for Lattice " + std::to_string(o.X) + "x" + std::to_string(o.Y) + ", probability " +
              std::to_string(o.prob_interactions) + ", error type " + std::string(1, o.logical_error_type) + " and " +
              std::to_string(S) + " interactions");
  } catch (const std::exception& e) {
    wlio::log(e.what(), wlio::log_level::error);
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
