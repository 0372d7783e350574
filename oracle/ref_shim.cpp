// ORACLE — test infrastructure only.
//
// C entry point over the reference's own interval partition
// (src/interval.cpp:5-41), compiled from /root/reference/src by
// oracle/Makefile into oracle/_ref/.  Used to pin the oracle's and the
// engine's interval partition against the reference implementation itself.
#include <cstdint>
#include <stdexcept>

#include "interval.hpp"

extern "C" __attribute__((visibility("default"))) int wlref_interval_results(int e_min,
                                                                             int e_max,
                                                                             unsigned int num_intervals,
                                                                             int num_walkers,
                                                                             float overlap,
                                                                             int* h_start,
                                                                             int* h_end,
                                                                             int64_t* len_histogram_over_all_walkers,
                                                                             int* len_interval)
{
  try {
    wl::interval_results r{e_min, e_max, num_intervals, num_walkers, overlap};
    for (unsigned int i = 0; i < num_intervals; ++i) {
      h_start[i] = r.h_start[i];
      h_end[i] = r.h_end[i];
    }
    *len_histogram_over_all_walkers = r.len_histogram_over_all_walkers;
    *len_interval = r.len_interval;
    return 0;
  } catch (const std::invalid_argument&) {
    return -1;
  }
}
