// Command-line options, paths, logging and HDF5 files of the prerun/mainrun
// executables.  Same flags (src/program_options.cpp:37-157), same directory
// scheme (src/utils.cpp:88-128) and same HDF5 layout (src/file_io.cpp:23-192,
// src/utils.cpp:137-236) as the reference.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <string_view>
#include <vector>

namespace wlio {

enum class log_level { info, warn, error };

// "I HH:MM:SS - dd/mm/YYYY -- msg" on std::clog (src/utils.cpp:45-68).
void log(std::string_view msg, log_level level = log_level::info);

std::string boundary_to_str(int boundary);

struct prerun_options {
  int X = 0, Y = 0;
  float prob_interactions = 0;
  int num_iterations = 0;
  int walker_per_interactions = 0;
  int seed = 0;
  int task_id = 0;
  unsigned int num_intervals = 0;
  char logical_error_type = 'i';
  int num_interactions = 0;
  int boundary = 0;
  bool skip_output = false;
};

struct mainrun_options {
  int X = 0, Y = 0;
  int num_iterations = 0;
  float prob_interactions = 0;
  double alpha = 0, beta = 0;
  int num_intervals = 0;
  int walker_per_interval = 0;
  float overlap_decimal = 0;
  int seed = 0;
  char logical_error_type = 'I';
  int num_interactions = 0;
  int boundary = 0;
  int replica_exchange_offset = 0;
  int task_id = 0;
  int time_limit = 0;
  // additions
  bool exchange_compat = false;  // reference exchange schedule and parity
  bool batched = false;          // run the batched persistent kernel
};

// Parse; returns false when the program should exit (help, version, error).
// exit_code receives the code to exit with in that case.
bool parse_prerun_options(int argc, char** argv, prerun_options& o, int& exit_code);
bool parse_mainrun_options(int argc, char** argv, mainrun_options& o, int& exit_code);

std::string build_prerun_path(const prerun_options& o);
std::string build_mainrun_path(const mainrun_options& o);
std::string build_mainrun_result_path(const mainrun_options& o);
void create_directories(const std::string& path);
std::string timestamp_group_name();

bool have_hdf5();

// prerun_results.h5 (src/file_io.cpp:194-217); throws if the file exists.
void write_prerun_results(const std::vector<uint64_t>& histograms, const std::vector<int8_t>& interactions,
                          const std::vector<int8_t>& lattices, const std::vector<int>& energies,
                          const prerun_options& o, size_t len_histogram, const std::string& dir);

void read_prerun_histograms(std::map<int, std::vector<uint8_t>>& out, const std::string& dir, size_t len_histogram);
void read_prerun_interactions(std::map<int, std::vector<int8_t>>& out, const std::string& dir, int X, int Y);
void read_prerun_lattices(std::map<int, std::map<int, std::vector<int8_t>>>& out, const std::string& dir, int X,
                          int Y);

// Appends /{disorder_id}/{seed}/{timestamp}/log_g {Key int, Value double}
// with a 256-byte CodeVersion attribute (src/utils.cpp:137-236).
void write_results(const std::vector<int>& E, const std::vector<double>& lng, int disorder_id, int seed,
                   const std::string& dir, const std::string& timestamp, const std::string& version);

// One /{disorder}/{seed}/{timestamp}/log_g dataset of a results file.
struct result_record {
  int disorder_id = 0;
  int seed = 0;
  std::string timestamp;
  std::vector<int> E;
  std::vector<double> lng;
  std::string version;
};

// Every log_g dataset of mainrun_results.h5, in group order.
std::vector<result_record> read_results(const std::string& file);

}  // namespace wlio
