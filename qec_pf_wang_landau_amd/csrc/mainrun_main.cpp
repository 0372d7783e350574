// mainrun: reads init/.../prerun_results.h5, runs the replica-exchange
// Wang-Landau loop to ln f <= beta for every disorder sample, stitches the
// per-interval ln g and appends results/.../mainrun_results.h5.  Same flags,
// paths and output layout as the reference's src/mainrun.cu:30-428; the loop
// runs either as the plain launch sequence or (--batched) in the batched
// persistent kernel, which gives identical results.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <numeric>
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
  if (rc != 0 && rc != 3) {
    throw std::runtime_error(std::string(what) + ": " + wl_last_error());
  }
}

}  // namespace

int main(int argc, char** argv)
{
  wlio::mainrun_options o;
  int code = 0;
  if (!wlio::parse_mainrun_options(argc, argv, o, code)) {
    return code;
  }
  const auto t_start = std::chrono::steady_clock::now();
  try {
    const int S = o.num_interactions, I = o.num_intervals, W = o.walker_per_interval;
    const int X = o.X, Y = o.Y, N = X * Y;
    const int total_walker = S * I * W;
    const size_t len_prerun_hist = static_cast<size_t>(4 * N + 1);
    const std::string init_path = wlio::build_mainrun_path(o);

    // Spectrum per sample (src/file_io.cpp:558-616).
    std::map<int, std::vector<uint8_t>> hist;
    wlio::read_prerun_histograms(hist, init_path, len_prerun_hist);
    std::vector<int8_t> spectrum;
    std::vector<int> off_spec, e_min(S), e_max(S), strides(S);
    for (int s = 0; s < S; ++s) {
      auto it = hist.find(s);
      if (it == hist.end()) {
        throw std::runtime_error("Error: Histogram for interaction " + std::to_string(s) + " was not found.");
      }
      std::vector<int8_t> mask(len_prerun_hist);
      int len_mask = 0;
      if (wl_energy_spectrum(it->second.data(), static_cast<int>(len_prerun_hist), X, Y, &e_min[s], &e_max[s],
                             mask.data(), &len_mask) != 0) {
        throw std::runtime_error(wl_last_error());
      }
      off_spec.push_back(static_cast<int>(spectrum.size()));
      spectrum.insert(spectrum.end(), mask.begin(), mask.begin() + len_mask);
      strides[s] = wl_energy_stride(mask.data(), len_mask);
    }
    // Intervals per sample (src/mainrun.cu:79-100).
    std::vector<int> start, end, len_hist;
    for (int s = 0; s < S; ++s) {
      std::vector<int> st(I), en(I);
      int64_t tot = 0;
      int len = 0;
      if (wl_interval_results(e_min[s], e_max[s], I, W, o.overlap_decimal, st.data(), en.data(), &tot, &len) != 0) {
        throw std::invalid_argument(wl_last_error());
      }
      start.insert(start.end(), st.begin(), st.end());
      end.insert(end.end(), en.begin(), en.end());
      len_hist.push_back(static_cast<int>(tot));
      wlio::log("Last interval: " + std::to_string(st.back()) + " - " + std::to_string(en.back()) +
                ", other interval len: " + std::to_string(len) + ", num intervals: " + std::to_string(I));
    }
    // Couplings and seed lattices (src/file_io.cpp:598-659).
    std::map<int, std::vector<int8_t>> inter;
    wlio::read_prerun_interactions(inter, init_path, X, Y);
    std::vector<int8_t> J;
    for (int s = 0; s < S; ++s) {
      auto it = inter.find(s);
      if (it == inter.end()) {
        throw std::runtime_error("Error: Interaction for disorder offset " + std::to_string(s) + " was not found.");
      }
      J.insert(J.end(), it->second.begin(), it->second.end());
    }
    std::map<int, std::map<int, std::vector<int8_t>>> lats;
    wlio::read_prerun_lattices(lats, init_path, X, Y);
    std::vector<int8_t> lattice(static_cast<size_t>(total_walker) * N);
    for (int s = 0; s < S; ++s) {
      auto it = lats.find(s);
      if (it == lats.end()) {
        throw std::runtime_error("Error: No Lattices found for disorder offset " + std::to_string(s) + ".");
      }
      std::vector<int8_t> found;
      std::vector<int> found_e;
      for (const auto& [e, l] : it->second) {
        found_e.push_back(e);
        found.insert(found.end(), l.begin(), l.end());
      }
      if (wl_select_seed_lattices(found.data(), found_e.data(), static_cast<int>(found_e.size()), N,
                                  start.data() + s * I, end.data() + s * I, I, W,
                                  lattice.data() + static_cast<size_t>(s) * I * W * N) != 0) {
        throw std::runtime_error(wl_last_error());
      }
    }
    const int64_t total_len = std::accumulate(len_hist.begin(), len_hist.end(), int64_t{0});

    device_buffer<int> d_spec_off(off_spec), d_len_hist(len_hist), d_start(start), d_end(end);
    device_buffer<int8_t> d_spec(spectrum), d_J(J), d_lattice(lattice);
    device_buffer<uint64_t> d_H(total_len), d_off_iter(total_walker);
    device_buffer<double> d_logG(total_len);
    device_buffer<double> d_factor(std::vector<double>(total_walker, std::exp(1.0)));
    device_buffer<int> d_off_hist(total_walker), d_off_lattice(total_walker), d_idx(total_walker),
      d_energy(total_walker);
    device_buffer<int8_t> d_cond(static_cast<size_t>(S) * I), d_flag(total_walker);
    device_buffer<int> d_cond_sample(S);

    ok(wl_init_offsets_lattice(d_off_lattice.data(), X, Y, total_walker, nullptr), "offsets");
    ok(wl_init_offsets_histogram(d_off_hist.data(), d_start.data(), d_end.data(), d_len_hist.data(), I, total_walker,
                                 W, nullptr),
       "histogram offsets");
    ok(wl_init_indices(d_idx.data(), total_walker, W, nullptr), "indices");
    ok(wl_calc_energy(o.boundary, d_energy.data(), d_lattice.data(), d_J.data(), d_off_lattice.data(), X, Y,
                      total_walker, I * W, nullptr),
       "calc_energy");
    ok(wl_check_energy_ranges(d_flag.data(), d_energy.data(), d_start.data(), d_end.data(), total_walker, W, nullptr),
       "check_energy_ranges");
    const auto flags = d_flag.to_host();
    if (std::any_of(flags.begin(), flags.end(), [](int8_t f) { return f != 0; })) {
      throw std::runtime_error("Calculated energies do not match intervals");
    }
    const double before = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    wlio::log("Execution time before Wang Landau has started: " + std::to_string(before) + " seconds.");

    auto elapsed = [&] {
      return std::chrono::duration_cast<std::chrono::seconds>(std::chrono::steady_clock::now() - t_start).count();
    };
    const int64_t chunk = 20;  // rounds between time-limit checks
    bool new_energy = false;
    if (!o.batched) {
      int64_t counter = 1;
      for (;;) {
        int64_t rounds = 0;
        const int rc =
          wl_mainrun_loop(d_lattice.data(), d_J.data(), d_H.data(), d_logG.data(), d_factor.data(), d_energy.data(),
                          d_off_iter.data(), d_off_lattice.data(), d_idx.data(), d_cond.data(), d_cond_sample.data(),
                          d_start.data(), d_end.data(), d_off_hist.data(), d_spec.data(), d_spec_off.data(),
                          d_len_hist.data(), S, I, W, X, Y, o.boundary, o.num_iterations, o.alpha, o.beta,
                          o.replica_exchange_offset, o.seed, o.exchange_compat ? 1 : 0, chunk, &counter, &rounds,
                          nullptr);
        ok(rc, "mainrun loop");
        if (rc == 3) {
          new_energy = true;
          break;
        }
        if (rounds < chunk) {
          break;
        }
        if (elapsed() >= o.time_limit) {
          wlio::log("Time limit reached. Terminating execution.", wlio::log_level::warn);
          break;
        }
      }
    } else {
      std::vector<wl_task> tasks(S);
      const auto off_hist = d_off_hist.to_host();
      for (int s = 0; s < S; ++s) {
        wl_task& t = tasks[s];
        t = wl_task{};
        t.X = X;
        t.Y = Y;
        t.boundary = o.boundary;
        t.I = I;
        t.W = W;
        t.rng_key = o.seed + s;
        t.num_iterations = o.num_iterations;
        t.exchange_every = o.replica_exchange_offset;
        t.energy_stride = (strides[s] == 2 || strides[s] == 4) ? strides[s] : 1;
        t.parity_base = s * I;
        t.flags = o.exchange_compat ? 1 : 0;
        t.interval_off = s * I;
        t.spectrum_off = off_spec[s];
        t.walker_off = static_cast<int64_t>(s) * I * W;
        t.hist_off = off_hist[static_cast<size_t>(s) * I * W];
        t.lattice_off = static_cast<int64_t>(s) * I * W * N;
        t.couplings_off = static_cast<int64_t>(s) * 2 * N;
        t.alpha = o.alpha;
        t.beta = o.beta;
      }
      device_buffer<wl_task> d_tasks(tasks);
      device_buffer<int64_t> d_counter(std::vector<int64_t>(S, 1)), d_rounds(S);
      device_buffer<int> d_status(S), d_queue(1);
      wl_fused_launch_plan plan{};
      if (wl_fused_prepare(tasks.data(), S, start.data(), end.data(), &plan) != 0) {
        throw std::invalid_argument(wl_last_error());
      }
      for (;;) {
        ok(wl_fused_launch(&plan, d_tasks.data(), d_lattice.data(), d_J.data(), d_H.data(), d_logG.data(),
                           d_factor.data(), d_energy.data(), d_off_iter.data(), d_off_lattice.data(), d_idx.data(),
                           d_cond.data(), d_cond_sample.data(), d_start.data(), d_end.data(), d_spec.data(),
                           d_counter.data(), d_rounds.data(), d_status.data(), d_queue.data(), nullptr, 50 * chunk,
                           nullptr),
           "batched kernel");
        wlamd::hip_throw(hipDeviceSynchronize(), "batched kernel");
        const auto status = d_status.to_host();
        if (std::any_of(status.begin(), status.end(), [](int v) { return v >= 2; })) {
          new_energy = std::any_of(status.begin(), status.end(), [](int v) { return v == 2; });
          if (!new_energy) {
            throw std::runtime_error("batched kernel rejected a task (status >= 3)");
          }
          break;
        }
        if (std::all_of(status.begin(), status.end(), [](int v) { return v == 1; })) {
          break;
        }
        if (elapsed() >= o.time_limit) {
          wlio::log("Time limit reached. Terminating execution.", wlio::log_level::warn);
          break;
        }
      }
    }
    if (new_energy) {
      wlio::log("Error: Found new energy", wlio::log_level::error);
      return EXIT_FAILURE;
    }

    // Stitch and write finished samples (src/mainrun.cu:385-421).
    const auto cond_sample = d_cond_sample.to_host();
    const auto off_hist = d_off_hist.to_host();
    const std::string result_path = wlio::build_mainrun_result_path(o);
    wlio::create_directories(result_path);
    const std::string stamp = wlio::timestamp_group_name();
    for (int s = 0; s < S; ++s) {
      if (cond_sample[s] != -1) {
        wlio::log("Sample " + std::to_string(s) + " did not converge before the time limit; not written.",
                  wlio::log_level::warn);
        continue;
      }
      std::vector<double> g(static_cast<size_t>(len_hist[s]));
      d_logG.to_host(g, static_cast<size_t>(off_hist[static_cast<size_t>(s) * I * W]), g.size());
      const int cap = end[s * I + I - 1] - start[s * I] + 1;
      std::vector<int> E(cap);
      std::vector<double> lng(cap);
      int n = 0;
      if (wl_stitch_log_g(g.data(), start.data() + s * I, end.data() + s * I, I, W, X, Y, E.data(), lng.data(), cap,
                          &n) != 0) {
        throw std::runtime_error(wl_last_error());
      }
      E.resize(n);
      lng.resize(n);
      wlio::write_results(E, lng, s, o.seed, result_path, stamp, wl_version());
    }
  } catch (const std::exception& e) {
    wlio::log(e.what(), wlio::log_level::error);
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
