// prerun_dump: prints a prerun_results.h5 (src/file_io.cpp:23-192 layout) as
// text, for checks of the prerun without an HDF5 reader:
//   prerun_dump <directory holding prerun_results.h5> X Y [histogram length]
// prints, per disorder sample s,
//   # histogram <s> <len>        then the energy flags over [-2XY, 2XY] as 0/1
//                                (len = 4XY+1 unless given)
//   # interaction <s> <2XY>      then the couplings, space separated
//   # lattice <s> <E>            then the seed lattice stored for energy E
// in the order of the file's groups.

#include <cstdio>
#include <cstdlib>
#include <exception>
#include <map>
#include <string>
#include <vector>

#include "wl_io.h"

int main(int argc, char** argv)
{
  if (argc != 4 && argc != 5) {
    std::fprintf(stderr, "usage: %s <prerun result directory> X Y [histogram length]\n", argv[0]);
    return EXIT_FAILURE;
  }
  const std::string dir = argv[1];
  const int X = std::atoi(argv[2]);
  const int Y = std::atoi(argv[3]);
  if (X < 1 || Y < 1) {
    std::fprintf(stderr, "X and Y must be positive\n");
    return EXIT_FAILURE;
  }
  try {
    const size_t len = argc == 5 ? static_cast<size_t>(std::atoll(argv[4])) : static_cast<size_t>(4 * X * Y + 1);
    std::map<int, std::vector<uint8_t>> hist;
    std::map<int, std::vector<int8_t>> inter;
    std::map<int, std::map<int, std::vector<int8_t>>> lat;
    wlio::read_prerun_histograms(hist, dir, len);
    wlio::read_prerun_interactions(inter, dir, X, Y);
    wlio::read_prerun_lattices(lat, dir, X, Y);
    for (const auto& [s, h] : hist) {
      std::printf("# histogram %d %zu\n", s, h.size());
      for (uint8_t f : h) {
        std::putchar(f ? '1' : '0');
      }
      std::putchar('\n');
    }
    for (const auto& [s, j] : inter) {
      std::printf("# interaction %d %zu\n", s, j.size());
      for (size_t i = 0; i < j.size(); ++i) {
        std::printf(i ? " %d" : "%d", static_cast<int>(j[i]));
      }
      std::putchar('\n');
    }
    for (const auto& [s, byE] : lat) {
      for (const auto& [E, l] : byE) {
        std::printf("# lattice %d %d\n", s, E);
        for (size_t i = 0; i < l.size(); ++i) {
          std::printf(i ? " %d" : "%d", static_cast<int>(l[i]));
        }
        std::putchar('\n');
      }
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
