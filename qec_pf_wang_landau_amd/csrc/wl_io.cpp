// See wl_io.h.  HDF5 is optional at build time (WL_WITH_HDF5); without it the
// executables run but cannot read or write result files.

#include "wl_io.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <functional>
#include <iomanip>
#include <iostream>
#include <random>
#include <sstream>
#include <stdexcept>

#ifdef WL_WITH_HDF5
#include <hdf5.h>
#endif

namespace wlio {

void log(std::string_view msg, log_level level)
{
  const auto now = std::chrono::system_clock::now();
  const std::time_t t = std::chrono::system_clock::to_time_t(now);
  const char c = level == log_level::info ? 'I' : (level == log_level::warn ? 'W' : 'E');
  std::tm tm{};
  localtime_r(&t, &tm);
  std::clog << c << ' ' << std::put_time(&tm, "%H:%M:%S - %d/%m/%Y") << " -- " << msg << '\n';
}

std::string boundary_to_str(int boundary)
{
  if (boundary == 0) {
    return "periodic";
  }
  if (boundary == 1) {
    return "open";
  }
  throw std::runtime_error("Unsupported boundary condition");
}

// ------------------------------------------------------------ options -----

namespace {

struct opt {
  std::string long_name;
  char short_name;  // 0 if none
  bool takes_value;
  bool required;
  std::string help;
  std::function<void(const std::string&)> set;
  bool seen = false;
};

int to_int(const std::string& s)
{
  size_t pos = 0;
  const int v = std::stoi(s, &pos);
  if (pos != s.size()) {
    throw std::invalid_argument("not an integer: " + s);
  }
  return v;
}

double to_double(const std::string& s)
{
  size_t pos = 0;
  const double v = std::stod(s, &pos);
  if (pos != s.size()) {
    throw std::invalid_argument("not a number: " + s);
  }
  return v;
}

void print_help(const char* prog, const std::string& title, const std::vector<opt>& opts)
{
  std::cout << prog << "\n\n" << title << ":\n";
  for (const auto& o : opts) {
    std::cout << "  ";
    if (o.short_name) {
      std::cout << '-' << o.short_name << " [ --" << o.long_name << " ]";
    } else {
      std::cout << "--" << o.long_name;
    }
    if (o.takes_value) {
      std::cout << " arg";
    }
    std::cout << "    " << o.help << '\n';
  }
}

// Boost.Program_options-like parsing: -x 4, -x4, --X 4, --X=4.
// Returns "help"/"version" when those flags are present.
std::string parse(int argc, char** argv, std::vector<opt>& opts)
{
  std::string special;
  for (int a = 1; a < argc; ++a) {
    std::string arg = argv[a];
    opt* hit = nullptr;
    std::string value;
    bool have_value = false;
    if (arg.rfind("--", 0) == 0) {
      std::string name = arg.substr(2);
      const auto eq = name.find('=');
      if (eq != std::string::npos) {
        value = name.substr(eq + 1);
        name = name.substr(0, eq);
        have_value = true;
      }
      for (auto& o : opts) {
        if (o.long_name == name) {
          hit = &o;
        }
      }
    } else if (arg.size() >= 2 && arg[0] == '-') {
      for (auto& o : opts) {
        if (o.short_name == arg[1]) {
          hit = &o;
        }
      }
      if (arg.size() > 2) {
        value = arg.substr(2);
        have_value = true;
      }
    }
    if (hit == nullptr) {
      throw std::invalid_argument("unrecognised option '" + arg + "'");
    }
    if (hit->takes_value && !have_value) {
      if (a + 1 >= argc) {
        throw std::invalid_argument("the required argument for option '--" + hit->long_name + "' is missing");
      }
      value = argv[++a];
    }
    hit->seen = true;
    if (hit->long_name == "help" || hit->long_name == "version") {
      special = hit->long_name;
    }
    if (hit->set) {
      hit->set(value);
    }
  }
  return special;
}

bool finish(int argc, char** argv, const std::string& title, std::vector<opt>& opts, int& exit_code,
            const char* version_label)
{
  std::string special;
  try {
    special = parse(argc, argv, opts);
  } catch (const std::exception& e) {
    std::cerr << e.what() << '\n';
    print_help(argv[0], title, opts);
    exit_code = EXIT_FAILURE;
    return false;
  }
  if (special == "help") {
    print_help(argv[0], title, opts);
    exit_code = EXIT_SUCCESS;
    return false;
  }
  if (special == "version") {
    std::cout << version_label << "qec_pf_wang_landau_amd 0.1 (gfx950)\n";
    exit_code = EXIT_SUCCESS;
    return false;
  }
  for (const auto& o : opts) {
    if (o.required && !o.seen) {
      std::cerr << "the option '--" << o.long_name << "' is required but missing\n";
      print_help(argv[0], title, opts);
      exit_code = EXIT_SUCCESS;  // the reference returns success after printing help
      return false;
    }
  }
  return true;
}

int seed_from_os()
{
  std::random_device rd;
  return static_cast<int>(rd());
}

int map_boundary(int b)
{
  if (b != 0 && b != 1) {
    throw std::runtime_error("Unsupported boundary condition " + std::to_string(b));
  }
  return b;
}

}  // namespace

bool parse_prerun_options(int argc, char** argv, prerun_options& o, int& exit_code)
{
  bool seed_set = false;
  int boundary = 0;
  std::vector<opt> opts = {
    {"help", 'h', false, false, "Help", nullptr},
    {"X", 'x', true, true, "Ising lattice, horizontal dimension", [&](auto v) { o.X = to_int(v); }},
    {"Y", 'y', true, true, "Ising lattice, vertical dimension", [&](auto v) { o.Y = to_int(v); }},
    {"prob", 'p', true, false, "Bit-flip error probablities",
     [&](auto v) { o.prob_interactions = static_cast<float>(to_double(v)); }},
    {"nit", 'n', true, true, "Number of WL iterations", [&](auto v) { o.num_iterations = to_int(v); }},
    {"nw", 'w', true, true, "Number of walkers per disorder sample",
     [&](auto v) { o.walker_per_interactions = to_int(v); }},
    {"seed", 's', true, false, "Random seed", [&](auto v) { o.seed = to_int(v); seed_set = true; }},
    {"num-intervals", 'i', true, true, "Number of intervals in the energy spectrum",
     [&](auto v) { o.num_intervals = static_cast<unsigned>(to_int(v)); }},
    {"logical-error", 'e', true, false, "type of logical error coset representative",
     [&](auto v) { o.logical_error_type = v.empty() ? 'i' : v[0]; }},
    {"disorder-samples", 'r', true, true, "Number of disorder samples",
     [&](auto v) { o.num_interactions = to_int(v); }},
    {"boundary", 0, true, false, "Boundary condition", [&](auto v) { boundary = to_int(v); }},
    {"task_id", 'd', true, true, "Task ID", [&](auto v) { o.task_id = to_int(v); }},
    {"skip-output", 0, false, false, "Skip writing results to the filesystem",
     [&](auto) { o.skip_output = true; }},
    {"version", 'v', false, false, "Show version", nullptr},
  };
  if (!finish(argc, argv, "WL pre-run usage", opts, exit_code, "WL prerun version: ")) {
    return false;
  }
  if (!seed_set) {
    o.seed = seed_from_os();
  }
  o.boundary = map_boundary(boundary);
  return true;
}

bool parse_mainrun_options(int argc, char** argv, mainrun_options& o, int& exit_code)
{
  bool seed_set = false;
  int boundary = 0;
  std::vector<opt> opts = {
    {"help", 0, false, false, "Help", nullptr},
    {"X", 'x', true, true, "X", [&](auto v) { o.X = to_int(v); }},
    {"Y", 'y', true, true, "Y", [&](auto v) { o.Y = to_int(v); }},
    {"num_iterations", 'n', true, true, "Number of iterations", [&](auto v) { o.num_iterations = to_int(v); }},
    {"prob_interactions", 'p', true, true, "Interaction probabilities",
     [&](auto v) { o.prob_interactions = static_cast<float>(to_double(v)); }},
    {"alpha", 'a', true, true, "Alpha", [&](auto v) { o.alpha = to_double(v); }},
    {"beta", 'b', true, true, "Beta", [&](auto v) { o.beta = to_double(v); }},
    {"num_intervals", 'i', true, true, "Number of intervals", [&](auto v) { o.num_intervals = to_int(v); }},
    {"walker_per_interval", 'w', true, true, "Number of walkers per interval",
     [&](auto v) { o.walker_per_interval = to_int(v); }},
    {"overlap_decimal", 'o', true, true, "Overlap of neighbouring intervals",
     [&](auto v) { o.overlap_decimal = static_cast<float>(to_double(v)); }},
    {"seed", 's', true, false, "Random seed (Run)", [&](auto v) { o.seed = to_int(v); seed_set = true; }},
    {"logical_error", 'e', true, false, "Type of logical error",
     [&](auto v) { o.logical_error_type = v.empty() ? 'I' : v[0]; }},
    {"repetitions_interactions", 'r', true, true, "Repetitions interactions",
     [&](auto v) { o.num_interactions = to_int(v); }},
    {"replica_exchange_offsets", 'c', true, true, "Rounds between replica exchanges",
     [&](auto v) { o.replica_exchange_offset = to_int(v); }},
    {"boundary", 0, true, false, "Boundary condition", [&](auto v) { boundary = to_int(v); }},
    {"task_id", 'd', true, true, "Task ID", [&](auto v) { o.task_id = to_int(v); }},
    {"time_limit", 'f', true, true, "Time Limit", [&](auto v) { o.time_limit = to_int(v); }},
    {"version", 'v', false, false, "Show version", nullptr},
    {"exchange-compat", 0, false, false,
     "Reference exchange schedule: exchange only if -c 1, pair intervals by global index parity",
     [&](auto) { o.exchange_compat = true; }},
    {"batched", 0, false, false, "Run every sample in the batched persistent kernel",
     [&](auto) { o.batched = true; }},
  };
  if (!finish(argc, argv, "WL main usage", opts, exit_code, "WL mainrun version: ")) {
    return false;
  }
  if (!seed_set) {
    o.seed = seed_from_os();
  }
  o.boundary = map_boundary(boundary);
  return true;
}

// -------------------------------------------------------------- paths -----

namespace {
template<typename O>
std::string tail_path(const O& o)
{
  std::stringstream ss;
  ss << "/seed_" << o.seed << "/" << boundary_to_str(o.boundary) << "/prob_" << std::fixed << std::setprecision(6)
     << o.prob_interactions << "/X_" << o.X << "_Y_" << o.Y << "/error_class_" << o.logical_error_type;
  return ss.str();
}
}  // namespace

std::string build_prerun_path(const prerun_options& o)
{
  return "init/task_id_" + std::to_string(o.task_id) + tail_path(o);
}

std::string build_mainrun_path(const mainrun_options& o)
{
  return "init/task_id_" + std::to_string(o.task_id) + tail_path(o);
}

std::string build_mainrun_result_path(const mainrun_options& o)
{
  return "results/task_id_" + std::to_string(o.task_id) + tail_path(o);
}

void create_directories(const std::string& path)
{
  if (!std::filesystem::exists(path)) {
    std::filesystem::create_directories(path);
  }
}

std::string timestamp_group_name()
{
  const std::time_t t = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
  char buf[64];
  ctime_r(&t, buf);
  std::string s(buf);
  if (!s.empty() && s.back() == '\n') {
    s.pop_back();
  }
  std::replace(s.begin(), s.end(), ' ', '_');
  return s;
}

// ---------------------------------------------------------------- HDF5 ----

#ifdef WL_WITH_HDF5

bool have_hdf5()
{
  return true;
}

namespace {

hid_t check(hid_t id, const std::string& what)
{
  if (id < 0) {
    throw std::runtime_error(what);
  }
  return id;
}

// Closes an HDF5 handle with the given closer at scope exit.
struct h5 {
  hid_t id;
  herr_t (*close)(hid_t);
  h5(hid_t i, herr_t (*c)(hid_t), const std::string& what) : id(check(i, what)), close(c) {}
  ~h5()
  {
    if (id >= 0) {
      close(id);
    }
  }
  h5(const h5&) = delete;
  h5& operator=(const h5&) = delete;
  operator hid_t() const { return id; }
};

hid_t open_or_create_group(hid_t parent, const std::string& name)
{
  if (H5Lexists(parent, name.c_str(), H5P_DEFAULT) > 0) {
    return check(H5Gopen2(parent, name.c_str(), H5P_DEFAULT), "Error: Could not open existing group " + name);
  }
  return check(H5Gcreate2(parent, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
               "Error: Could not create group " + name);
}

void write_1d(hid_t group, const char* name, hid_t file_type, hid_t mem_type, const void* data, size_t n)
{
  hsize_t dims[1] = {static_cast<hsize_t>(n)};
  h5 space(H5Screate_simple(1, dims, nullptr), H5Sclose, "Failed to create h5 dataspace");
  h5 ds(H5Dcreate2(group, name, file_type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
        "Failed to create h5 dataset.");
  check(H5Dwrite(ds, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "Failed to write h5 dataset.");
}

std::vector<std::string> child_names(hid_t group)
{
  std::vector<std::string> names;
  hsize_t n = 0;
  H5Gget_num_objs(group, &n);
  for (hsize_t i = 0; i < n; ++i) {
    char buf[256];
    if (H5Gget_objname_by_idx(group, i, buf, sizeof buf) < 0) {
      throw std::runtime_error("Error: Could not retrieve a group name.");
    }
    names.emplace_back(buf);
  }
  return names;
}

template<typename T>
std::vector<T> read_1d(hid_t group, const char* name, hid_t mem_type, size_t expect, const std::string& what)
{
  h5 ds(H5Dopen2(group, name, H5P_DEFAULT), H5Dclose, "Error: Could not open HDF5 " + what + " dataset");
  h5 space(H5Dget_space(ds), H5Sclose, "Could not get " + what + " dataspace");
  if (H5Sget_simple_extent_ndims(space) != 1) {
    throw std::runtime_error("number dimension of dataspace not 1.");
  }
  hsize_t dims[1];
  H5Sget_simple_extent_dims(space, dims, nullptr);
  if (dims[0] != expect) {
    throw std::runtime_error(what + " length " + std::to_string(expect) + " does not coincide with length written "
                             "to file " + std::to_string(dims[0]));
  }
  std::vector<T> out(expect);
  check(H5Dread(ds, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()), "Could not read " + what);
  return out;
}

hid_t open_read(const std::string& path)
{
  return check(H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), "Error: Could not open HDF5 file:" + path);
}

}  // namespace

void write_prerun_results(const std::vector<uint64_t>& histograms, const std::vector<int8_t>& interactions,
                          const std::vector<int8_t>& lattices, const std::vector<int>& energies,
                          const prerun_options& o, size_t len_histogram, const std::string& dir)
{
  const std::string path = dir + "/prerun_results.h5";
  if (std::filesystem::exists(path)) {
    throw std::runtime_error("Prerun result file " + path + " already exists, aborting.");
  }
  h5 file(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, "Failed to create h5 file.");
  const int S = o.num_interactions;
  const size_t N = static_cast<size_t>(o.X) * o.Y;
  {
    // Histogram: 0/1 flags written from u8 into a u64 dataset (src/file_io.cpp:23-78).
    h5 g(H5Gcreate2(file, "Histogram", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose,
         "Failed to create Histogram group");
    std::vector<uint8_t> bin(len_histogram);
    for (int s = 0; s < S; ++s) {
      for (size_t e = 0; e < len_histogram; ++e) {
        bin[e] = histograms[s * len_histogram + e] != 0 ? 1 : 0;
      }
      h5 sg(H5Gcreate2(g, std::to_string(s).c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose,
            "Failed to create h5 group");
      write_1d(sg, "Histogram", H5T_NATIVE_ULLONG, H5T_NATIVE_UCHAR, bin.data(), len_histogram);
    }
  }
  {
    h5 g(H5Gcreate2(file, "Interaction", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose,
         "Failed to create Interaction group");
    for (int s = 0; s < S; ++s) {
      h5 sg(H5Gcreate2(g, std::to_string(s).c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose,
            "Failed to create h5 group");
      write_1d(sg, "Interaction", H5T_NATIVE_SCHAR, H5T_NATIVE_SCHAR, interactions.data() + s * 2 * N, 2 * N);
    }
  }
  {
    // Lattice/{sample}/{energy}/Lattice for every prerun interval that found one.
    h5 g(H5Gcreate2(file, "Lattice", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "Failed to create Lattice group");
    const size_t I = o.num_intervals;
    for (int s = 0; s < S; ++s) {
      h5 sg(H5Gcreate2(g, std::to_string(s).c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose,
            "Failed to create disorder group for prerun lattice result");
      for (size_t k = 0; k < I; ++k) {
        const int8_t* lat = lattices.data() + (s * I + k) * N;
        if (lat[0] == 0) {
          continue;
        }
        h5 eg(H5Gcreate2(sg, std::to_string(energies[s * I + k]).c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
              H5Gclose, "Failed to create h5 interval group for lattice write");
        write_1d(eg, "Lattice", H5T_NATIVE_SCHAR, H5T_NATIVE_SCHAR, lat, N);
      }
    }
  }
}

void read_prerun_histograms(std::map<int, std::vector<uint8_t>>& out, const std::string& dir, size_t len_histogram)
{
  h5 file(open_read(dir + "/prerun_results.h5"), H5Fclose, "open");
  h5 root(H5Gopen2(file, "/Histogram", H5P_DEFAULT), H5Gclose, "Error: Could not open root group");
  for (const auto& name : child_names(root)) {
    h5 g(H5Gopen2(root, name.c_str(), H5P_DEFAULT), H5Gclose, "Error: Could not open HDF5 group:" + name);
    out[std::stoi(name)] = read_1d<uint8_t>(g, "Histogram", H5T_NATIVE_UCHAR, len_histogram, "Histogram");
  }
}

void read_prerun_interactions(std::map<int, std::vector<int8_t>>& out, const std::string& dir, int X, int Y)
{
  h5 file(open_read(dir + "/prerun_results.h5"), H5Fclose, "open");
  h5 root(H5Gopen2(file, "/Interaction", H5P_DEFAULT), H5Gclose, "Error: Could not open root group");
  for (const auto& name : child_names(root)) {
    h5 g(H5Gopen2(root, name.c_str(), H5P_DEFAULT), H5Gclose, "Error: Could not open HDF5 group:" + name);
    out[std::stoi(name)] =
      read_1d<int8_t>(g, "Interaction", H5T_NATIVE_SCHAR, static_cast<size_t>(2 * X * Y), "Interaction");
  }
}

void read_prerun_lattices(std::map<int, std::map<int, std::vector<int8_t>>>& out, const std::string& dir, int X,
                          int Y)
{
  h5 file(open_read(dir + "/prerun_results.h5"), H5Fclose, "open");
  h5 root(H5Gopen2(file, "/Lattice", H5P_DEFAULT), H5Gclose, "Error: Could not open root group");
  for (const auto& sname : child_names(root)) {
    h5 sg(H5Gopen2(root, sname.c_str(), H5P_DEFAULT), H5Gclose, "Failed to open disorder group");
    for (const auto& ename : child_names(sg)) {
      h5 eg(H5Gopen2(sg, ename.c_str(), H5P_DEFAULT), H5Gclose, "Could not open energy subgroup");
      out[std::stoi(sname)][std::stoi(ename)] =
        read_1d<int8_t>(eg, "Lattice", H5T_NATIVE_SCHAR, static_cast<size_t>(X * Y), "lattice");
    }
  }
}

void write_results(const std::vector<int>& E, const std::vector<double>& lng, int disorder_id, int seed,
                   const std::string& dir, const std::string& timestamp, const std::string& version)
{
  struct entry {
    int key;
    double value;
  };
  std::vector<entry> rows(E.size());
  for (size_t i = 0; i < E.size(); ++i) {
    rows[i] = {E[i], lng[i]};
  }
  const std::string path = dir + "/mainrun_results.h5";
  hid_t fid = std::filesystem::exists(path) ? H5Fopen(path.c_str(), H5F_ACC_RDWR, H5P_DEFAULT)
                                            : H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  h5 file(fid, H5Fclose, "Error: Could not create or open result file.");
  h5 dg(open_or_create_group(file, "/" + std::to_string(disorder_id)), H5Gclose, "group");
  h5 sg(open_or_create_group(dg, std::to_string(seed)), H5Gclose, "group");
  h5 tg(open_or_create_group(sg, timestamp), H5Gclose, "group");
  h5 ctype(H5Tcreate(H5T_COMPOUND, sizeof(entry)), H5Tclose, "compound type");
  H5Tinsert(ctype, "Key", HOFFSET(entry, key), H5T_NATIVE_INT);
  H5Tinsert(ctype, "Value", HOFFSET(entry, value), H5T_NATIVE_DOUBLE);
  hsize_t dims[1] = {rows.size()};
  h5 space(H5Screate_simple(1, dims, nullptr), H5Sclose, "dataspace");
  h5 ds(H5Dcreate2(tg, "log_g", ctype, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
        "Error: Could not create log_g dataset in run group.");
  check(H5Dwrite(ds, ctype, H5S_ALL, H5S_ALL, H5P_DEFAULT, rows.data()), "Error: Could not write data to dataset.");
  h5 aspace(H5Screate(H5S_SCALAR), H5Sclose, "Error: Could not create attribute dataspace.");
  h5 stype(H5Tcopy(H5T_C_S1), H5Tclose, "string type");
  H5Tset_size(stype, 256);
  H5Tset_strpad(stype, H5T_STR_NULLTERM);
  char buf[256] = {0};
  std::strncpy(buf, version.c_str(), sizeof buf - 1);
  h5 attr(H5Acreate2(ds, "CodeVersion", stype, aspace, H5P_DEFAULT, H5P_DEFAULT), H5Aclose,
          "Error: Could not create attribute for dataset.");
  check(H5Awrite(attr, stype, buf), "Error: Could not write attribute to dataset.");
}

std::vector<result_record> read_results(const std::string& path)
{
  struct entry {
    int key;
    double value;
  };
  std::vector<result_record> out;
  h5 file(open_read(path), H5Fclose, "Error: Could not open HDF5 file:" + path);
  for (const std::string& dname : child_names(file)) {
    h5 dg(H5Gopen2(file, dname.c_str(), H5P_DEFAULT), H5Gclose, "disorder group");
    for (const std::string& sname : child_names(dg)) {
      h5 sg(H5Gopen2(dg, sname.c_str(), H5P_DEFAULT), H5Gclose, "seed group");
      for (const std::string& tname : child_names(sg)) {
        h5 tg(H5Gopen2(sg, tname.c_str(), H5P_DEFAULT), H5Gclose, "run group");
        h5 ds(H5Dopen2(tg, "log_g", H5P_DEFAULT), H5Dclose, "Error: Could not open log_g dataset");
        h5 space(H5Dget_space(ds), H5Sclose, "log_g dataspace");
        hsize_t dims[1] = {0};
        H5Sget_simple_extent_dims(space, dims, nullptr);
        h5 ctype(H5Tcreate(H5T_COMPOUND, sizeof(entry)), H5Tclose, "compound type");
        H5Tinsert(ctype, "Key", HOFFSET(entry, key), H5T_NATIVE_INT);
        H5Tinsert(ctype, "Value", HOFFSET(entry, value), H5T_NATIVE_DOUBLE);
        std::vector<entry> rows(dims[0]);
        check(H5Dread(ds, ctype, H5S_ALL, H5S_ALL, H5P_DEFAULT, rows.data()), "Could not read log_g");
        result_record r;
        r.disorder_id = std::stoi(dname);
        r.seed = std::stoi(sname);
        r.timestamp = tname;
        for (const entry& e : rows) {
          r.E.push_back(e.key);
          r.lng.push_back(e.value);
        }
        if (H5Aexists(ds, "CodeVersion") > 0) {
          h5 attr(H5Aopen(ds, "CodeVersion", H5P_DEFAULT), H5Aclose, "CodeVersion attribute");
          h5 stype(H5Tcopy(H5T_C_S1), H5Tclose, "string type");
          H5Tset_size(stype, 256);
          char buf[257] = {0};
          check(H5Aread(attr, stype, buf), "Could not read CodeVersion");
          r.version = buf;
        }
        out.push_back(std::move(r));
      }
    }
  }
  return out;
}

#else

bool have_hdf5()
{
  return false;
}

std::vector<result_record> read_results(const std::string&)
{
  throw std::runtime_error("built without HDF5: result files are unavailable");
}

namespace {
[[noreturn]] void no_hdf5()
{
  throw std::runtime_error("built without HDF5: result files are unavailable");
}
}  // namespace

void write_prerun_results(const std::vector<uint64_t>&, const std::vector<int8_t>&, const std::vector<int8_t>&,
                          const std::vector<int>&, const prerun_options&, size_t, const std::string&)
{
  no_hdf5();
}
void read_prerun_histograms(std::map<int, std::vector<uint8_t>>&, const std::string&, size_t)
{
  no_hdf5();
}
void read_prerun_interactions(std::map<int, std::vector<int8_t>>&, const std::string&, int, int)
{
  no_hdf5();
}
void read_prerun_lattices(std::map<int, std::map<int, std::vector<int8_t>>>&, const std::string&, int, int)
{
  no_hdf5();
}
void write_results(const std::vector<int>&, const std::vector<double>&, int, int, const std::string&,
                   const std::string&, const std::string&)
{
  no_hdf5();
}

#endif

}  // namespace wlio
