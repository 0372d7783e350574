// Checks of the prerun/mainrun file layer (qec_pf_wang_landau_amd/csrc/wl_io.cpp)
// against the vectors of the reference's own file tests (test/test_file_io.cpp:
// histogram, interaction and lattice write/read; test/test_result_handling.cpp:
// the ln g result dataset and its CodeVersion attribute) and its path scheme
// (test/integration_mainrun.py:91-95).  Built and run by tests/test_file_io.py.
//
//   io_check <scratch directory>
// prints one line per failed check and exits non-zero if any failed.

#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <map>
#include <string>
#include <vector>

#include "wl_io.h"

namespace {

int failures = 0;

#define CHECK(cond)                                                          \
  do {                                                                       \
    if (!(cond)) {                                                           \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
      ++failures;                                                            \
    }                                                                        \
  } while (0)

template<typename F>
bool throws(F&& f)
{
  try {
    f();
  } catch (const std::exception&) {
    return true;
  }
  return false;
}

}  // namespace

int main(int argc, char** argv)
{
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
    return 2;
  }
  const std::string root = argv[1];
  CHECK(wlio::have_hdf5());

  // --- prerun_results.h5: the reference's histogram / interaction / lattice vectors
  wlio::prerun_options o;
  o.X = 4;
  o.Y = 4;
  o.num_interactions = 2;
  o.num_intervals = 4;
  const size_t len_histogram = 5;
  const std::vector<uint64_t> hist = {99, 0, 0, 0, 98, 0, 0, 97, 0, 96};
  const std::vector<int8_t> inter = {
    1, -1, 1, 1, 1, 1, -1, 1, 1, 1, -1, 1, 1, -1, -1, 1,  1,  1, 1, 1, 1, 1, 1, 1, 1, -1, 1, 1, 1, 1, -1, 1,
    1, 1,  1, 1, 1, 1, 1,  1, 1, 1, 1,  1, 1, 1,  1,  -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -1, 1, 1, 1, 1, 1,  -1};
  const std::vector<int8_t> a = {1, -1, 1, 1, 1, 1, -1, 1, 1, 1, -1, 1, 1, -1, -1, 1};
  const std::vector<int8_t> b = {1, 1, 1, -1, 1, 1, 1, -1, -1, 1, 1, 1, 1, 1, 1, 1};
  const std::vector<int8_t> c = {1, 1, 1, -1, -1, 1, 1, 1, 1, 1, -1, 1, 1, -1, 1, 1};
  const std::vector<int8_t> z(16, 0);
  std::vector<int8_t> lat;
  for (const auto* v : {&a, &b, &c, &z, &a, &b, &z, &c}) {
    lat.insert(lat.end(), v->begin(), v->end());
  }
  const std::vector<int> energies = {0, 10, 20, 30, 40, 50, 60, 70};
  const std::string pre = root + "/prerun";
  wlio::create_directories(pre);
  wlio::write_prerun_results(hist, inter, lat, energies, o, len_histogram, pre);
  CHECK(throws([&] { wlio::write_prerun_results(hist, inter, lat, energies, o, len_histogram, pre); }));

  std::map<int, std::vector<uint8_t>> h;
  wlio::read_prerun_histograms(h, pre, len_histogram);
  CHECK(h.size() == 2);
  CHECK((h[0] == std::vector<uint8_t>{1, 0, 0, 0, 1}));
  CHECK((h[1] == std::vector<uint8_t>{0, 0, 1, 0, 1}));

  std::map<int, std::vector<int8_t>> in;
  wlio::read_prerun_interactions(in, pre, 4, 4);
  CHECK(in.size() == 2);
  CHECK((in[0] == std::vector<int8_t>(inter.begin(), inter.begin() + 32)));
  CHECK((in[1] == std::vector<int8_t>(inter.begin() + 32, inter.end())));

  std::map<int, std::map<int, std::vector<int8_t>>> l;
  wlio::read_prerun_lattices(l, pre, 4, 4);
  CHECK(l.size() == 2);
  CHECK(l[0].size() == 3 && l[1].size() == 3);  // all-zero lattices are not written
  CHECK(l[0][0] == a && l[0][10] == b && l[0][20] == c);
  CHECK(l[1][40] == a && l[1][50] == b && l[1][70] == c);
  CHECK(l[0].count(30) == 0 && l[1].count(60) == 0);

  // --- mainrun_results.h5: /{disorder}/{seed}/{timestamp}/log_g with CodeVersion
  const std::string res = root + "/results";
  wlio::create_directories(res);
  const std::vector<int> E = {-32, -24, -20, 24, 32};
  const std::vector<double> g1 = {0.6931471805599453, 3.4657359027997265, 4.1588830833596715, 3.4657359027997265,
                                  0.6931471805599453};
  const std::vector<double> g2 = {0.5, 1.25, 2.0, 1.25, 0.5};
  wlio::write_results(E, g1, 3, 17, res, "run-a", "v1.2.3");
  wlio::write_results(E, g2, 3, 17, res, "run-b", "v1.2.3");
  wlio::write_results(E, g1, 5, 18, res, "run-a", "v1.2.3");
  const std::vector<wlio::result_record> recs = wlio::read_results(res + "/mainrun_results.h5");
  CHECK(recs.size() == 3);
  int seen = 0;
  for (const auto& r : recs) {
    CHECK(r.E == E);
    CHECK(r.version == "v1.2.3");
    if (r.disorder_id == 3 && r.timestamp == "run-a") {
      CHECK(r.seed == 17 && r.lng == g1);
      ++seen;
    } else if (r.disorder_id == 3 && r.timestamp == "run-b") {
      CHECK(r.seed == 17 && r.lng == g2);
      ++seen;
    } else if (r.disorder_id == 5) {
      CHECK(r.seed == 18 && r.lng == g1);
      ++seen;
    }
  }
  CHECK(seen == 3);

  // --- the reference's result-writing vector (test/test_result_handling.cpp:43-105):
  // two interval maps flattened in order, disorder 100, seed 42
  const std::string res2 = root + "/results2";
  wlio::create_directories(res2);
  wlio::write_results({1, 2, 3, 4}, {0.5, 1.5, 2.5, 3.5}, 100, 42, res2, "timestamp_01:01:2024",
                      "release_0001-10-h76if76a-dirty");
  const std::vector<wlio::result_record> r2 = wlio::read_results(res2 + "/mainrun_results.h5");
  CHECK(r2.size() == 1);
  if (r2.size() == 1) {
    CHECK(r2[0].disorder_id == 100 && r2[0].seed == 42 && r2[0].timestamp == "timestamp_01:01:2024");
    CHECK((r2[0].E == std::vector<int>{1, 2, 3, 4}));
    CHECK((r2[0].lng == std::vector<double>{0.5, 1.5, 2.5, 3.5}));
    CHECK(r2[0].version == "release_0001-10-h76if76a-dirty");
  }

  // --- paths (test/integration_mainrun.py:91-95)
  wlio::prerun_options po;
  po.task_id = 0;
  po.seed = 42;
  po.boundary = 0;
  po.prob_interactions = 0.1f;
  po.X = 4;
  po.Y = 6;
  po.logical_error_type = 'v';
  CHECK(wlio::build_prerun_path(po) == "init/task_id_0/seed_42/periodic/prob_0.100000/X_4_Y_6/error_class_v");
  wlio::mainrun_options mo;
  mo.task_id = 3;
  mo.seed = 7;
  mo.boundary = 1;
  mo.prob_interactions = 0.04f;
  mo.X = 6;
  mo.Y = 6;
  mo.logical_error_type = 'c';
  CHECK(wlio::build_mainrun_path(mo) == "init/task_id_3/seed_7/open/prob_0.040000/X_6_Y_6/error_class_c");
  CHECK(wlio::build_mainrun_result_path(mo) == "results/task_id_3/seed_7/open/prob_0.040000/X_6_Y_6/error_class_c");

  if (failures == 0) {
    std::printf("io_check: all checks passed\n");
  }
  return failures == 0 ? 0 : 1;
}
