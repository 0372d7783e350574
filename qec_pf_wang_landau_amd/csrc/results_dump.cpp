// results_dump: prints every ln g record of a mainrun_results.h5 file as
// text, for checking the command-line runs without an HDF5 reader:
//   # <disorder> <seed> <timestamp> <entries> <CodeVersion>
//   <E> <ln g with 17 significant digits>
// Layout read: /{disorder}/{seed}/{timestamp}/log_g {Key int, Value double}
// (src/utils.cpp:137-236).

#include <cstdio>
#include <cstdlib>
#include <exception>

#include "wl_io.h"

int main(int argc, char** argv)
{
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <mainrun_results.h5>\n", argv[0]);
    return EXIT_FAILURE;
  }
  try {
    for (const wlio::result_record& r : wlio::read_results(argv[1])) {
      std::printf("# %d %d %s %zu %s\n", r.disorder_id, r.seed, r.timestamp.c_str(), r.E.size(), r.version.c_str());
      for (size_t i = 0; i < r.E.size(); ++i) {
        std::printf("%d %.17g\n", r.E[i], r.lng[i]);
      }
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
