// decode_results: logical-class decisions and the logical error rate of one
// (lattice, p) point from the mainrun results of its four error classes:
//
//   decode_results --prob P [--z Z] FILE_i FILE_v FILE_h FILE_c
//
// Each FILE is a mainrun_results.h5 (/{disorder}/{seed}/{timestamp}/log_g
// {Key int E, Value double ln g}, src/utils.cpp:137-236) written by mainrun
// for error class i, v, h or c of the same disorder samples (same prerun
// seed, so disorder id d names the same couplings up to the class chain).
// Every disorder id present in all four files is decoded; when a file holds
// several runs of one id, the last one in group order is used.  ln Z_c is
// taken at the Nishimori temperature of p from the stitched, normalised ln g;
// the decision is the most likely class and the sample fails unless it is
// class i (ties count the probability of a uniform pick missing i).
// Output, one line per decoded sample and a summary:
//   <disorder> <ln Z_i> <ln Z_v> <ln Z_h> <ln Z_c> <decision>
//   # samples <n> failures <f> rate <r> wilson <lo> <hi> z <Z> missing <m>
// "missing" counts disorder ids found in some but not all four files.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "wl_amd.h"
#include "wl_host.h"
#include "wl_io.h"

namespace {

const char* const kClasses = "ivhc";

int usage(const char* argv0)
{
  std::fprintf(stderr, "usage: %s --prob P [--z Z] FILE_i FILE_v FILE_h FILE_c\n", argv0);
  return EXIT_FAILURE;
}

}  // namespace

int main(int argc, char** argv)
{
  double prob = -1.0;
  double z = 1.96;
  std::vector<std::string> files;
  for (int a = 1; a < argc; ++a) {
    const std::string arg = argv[a];
    if ((arg == "--prob" || arg == "-p") && a + 1 < argc) {
      prob = std::strtod(argv[++a], nullptr);
    } else if (arg == "--z" && a + 1 < argc) {
      z = std::strtod(argv[++a], nullptr);
    } else if (!arg.empty() && arg[0] == '-') {
      return usage(argv[0]);
    } else {
      files.push_back(arg);
    }
  }
  if (files.size() != 4 || !(prob > 0.0 && prob < 0.5)) {
    if (files.size() == 4) {
      std::fprintf(stderr, "--prob must lie in (0, 0.5)\n");
    }
    return usage(argv[0]);
  }
  try {
    const double beta = wl_nishimori_beta(prob);
    // disorder id -> ln Z per class (NaN until that class's file provides it)
    std::map<int, std::vector<double>> log_z;
    for (int c = 0; c < 4; ++c) {
      std::map<int, const wlio::result_record*> last;
      const std::vector<wlio::result_record> recs = wlio::read_results(files[c]);
      for (const wlio::result_record& r : recs) {
        last[r.disorder_id] = &r;
      }
      for (const auto& [id, r] : last) {
        if (r->E.empty()) {
          throw std::runtime_error("empty log_g for disorder " + std::to_string(id) + " in " + files[c]);
        }
        auto it = log_z.try_emplace(id, std::vector<double>(4, std::nan(""))).first;
        it->second[c] = wl_log_partition(r->E.data(), r->lng.data(), static_cast<int>(r->E.size()), beta);
      }
    }
    std::vector<int> ids;
    std::vector<double> table;
    int missing = 0;
    for (const auto& [id, zs] : log_z) {
      bool all = true;
      for (double v : zs) {
        all = all && v == v;
      }
      if (!all) {
        ++missing;
        continue;
      }
      ids.push_back(id);
      table.insert(table.end(), zs.begin(), zs.end());
    }
    const int n = static_cast<int>(ids.size());
    std::vector<int> decision(n);
    std::vector<double> fail(n);
    if (wl_decide_classes(table.data(), n, 4, decision.data(), fail.data()) != 0) {
      throw std::runtime_error(wl_last_error());
    }
    double failures = 0.0;
    int decided = 0;
    for (int s = 0; s < n; ++s) {
      const double* row = &table[static_cast<size_t>(s) * 4];
      std::printf("%d %.17g %.17g %.17g %.17g %c\n", ids[s], row[0], row[1], row[2], row[3],
                  decision[s] >= 0 ? kClasses[decision[s]] : '-');
      if (decision[s] >= 0) {
        failures += fail[s];
        ++decided;
      }
    }
    double lo = 0.0, hi = 1.0;
    if (decided > 0 && wl_wilson_interval(failures, decided, z, &lo, &hi) != 0) {
      throw std::runtime_error(wl_last_error());
    }
    std::printf("# samples %d failures %.17g rate %.17g wilson %.17g %.17g z %g missing %d\n", decided, failures,
                decided > 0 ? failures / decided : 0.0, lo, hi, z, missing);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
