// ORACLE — test infrastructure only.
//
// A single-threaded CPU restatement of the reference's replica-exchange
// Wang-Landau path (src/wl.cu, src/mainrun.cu, src/prerun.cu,
// src/interval.cpp, src/file_io.cpp, src/result_handling.cpp), written as
// plain loops over the same flat arrays the reference kernels use.  It is
// the checker for the HIP engine: only tests, the smoke check and the CPU
// baseline of the benchmark may load it.  Nothing in the product links it.
//
// Random numbers, exp/log/sqrt: the engine defines its streams as
// Philox4x32-10 (Salmon et al., SC'11) keyed (seed + sample, tag) with
// counter (block, slot).  WL attempt a of a walker uses words 2(a&1) (site)
// and 2(a&1)+1 (acceptance) of block a/2 of its walk stream; its shuffle and
// exchange draws use word d%4 of block d/4 of its exchange stream; its
// offset_iter packs attempts (low 40 bits) and exchange draws (high 24).
// A word x gives the index floor(x n / 2^32) or the uniform (x + 1/2) 2^-32.
// exp is Cody-Waite reduction + degree-13 Taylor, log is the atanh series,
// sqrt is a Newton reciprocal-square-root iteration.  They are restated here
// from those published methods so that the oracle and the device produce the
// same bits.
//
// Deviations from the reference that the engine defines (each is also a
// deviation of the engine, so both sides agree):
//  D1  exchange every c rounds (reference: counter never incremented,
//      src/mainrun.cu:241,346) unless compat mode is on;
//  D2  converged-interval branch tests E_new in [start,end] with && (src/wl.cu:1161);
//  D3  averaging sums walkers in index order, no atomics (src/wl.cu:779);
//  D4  initial-lattice probability is a dedicated draw (src/wl.cu:450-454);
//  D5  prerun walkers of one sample move in lock step: decisions read the
//      histogram as it was at the start of the step (src/wl.cu:534-556);
//  D6  energies outside [E_min, E_max] count as "new energy" (src/wl.cu:1091);
//  D7  stitching skips shared keys without a successor (result_handling.cpp:125);
//  D8  stitching checks the iterator against the map it searched (result_handling.cpp:209);
//  D10 exchange parity is the interval index within the sample unless
//      compat mode asks for the global block index (src/wl.cu:830);
//  D12 ln f is log(f) of the stored factor, computed with the shared log;
//  D13 exchange acceptance is min(1, exp(a + b)) (reference exp(a)*exp(b), src/wl.cu:872);
//  D14 every WL attempt consumes two words (the reference draws the second
//      only for in-range proposals, src/wl.cu:1114), shuffle/exchange draws
//      come from a separate stream, and indices are floor(x n / 2^32)
//      (reference trunc(u (n - 1 + 0.999999)), src/wl.cu:66).
//
// Parity status per function is listed in the header of tests/test_oracle.py.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <tuple>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#define WLO_API extern "C" __attribute__((visibility("default")))

namespace {

// ---------------------------------------------------------------- RNG ------

void philox_round(uint32_t c[4], const uint32_t k[2])
{
  const uint64_t a = static_cast<uint64_t>(0xD2511F53u) * c[0];
  const uint64_t b = static_cast<uint64_t>(0xCD9E8D57u) * c[2];
  const uint32_t out0 = static_cast<uint32_t>(b >> 32) ^ c[1] ^ k[0];
  const uint32_t out1 = static_cast<uint32_t>(b);
  const uint32_t out2 = static_cast<uint32_t>(a >> 32) ^ c[3] ^ k[1];
  const uint32_t out3 = static_cast<uint32_t>(a);
  c[0] = out0;
  c[1] = out1;
  c[2] = out2;
  c[3] = out3;
}

void philox10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  philox_round(c, k);
  for (int r = 1; r < 10; ++r) {
    k[0] += 0x9E3779B9u;
    k[1] += 0xBB67AE85u;
    philox_round(c, k);
  }
  std::memcpy(out, c, sizeof c);
}

void block_at(int key0, uint32_t tag, uint32_t slot, uint64_t blk, uint32_t out[4])
{
  const uint32_t ctr[4] = {static_cast<uint32_t>(blk & 0xffffffffu), static_cast<uint32_t>(blk >> 32), slot, 0u};
  const uint32_t key[2] = {static_cast<uint32_t>(key0), tag};
  philox10(ctr, key, out);
}

uint32_t word_at(int key0, uint32_t tag, uint32_t slot, uint64_t draw)
{
  uint32_t out[4];
  block_at(key0, tag, slot, draw / 4, out);
  return out[draw % 4];
}

double unit(uint32_t x)
{
  return (static_cast<double>(x) + 0.5) / 4294967296.0;
}

double uniform_at(int key0, uint32_t tag, uint32_t slot, uint64_t draw)
{
  return unit(word_at(key0, tag, slot, draw));
}

constexpr uint32_t TAG_WALK = 0, TAG_DISORDER = 1, TAG_LATTICE = 2, TAG_PRERUN = 3, TAG_EXCHANGE = 4;
constexpr uint64_t ATTEMPT_MASK = (1ull << 40) - 1;
constexpr uint64_t EXCHANGE_ONE = 1ull << 40;

// floor(x n / 2^32): the index a word selects among n.
int pick(uint32_t x, int n)
{
  return static_cast<int>((static_cast<uint64_t>(x) * static_cast<uint64_t>(n)) >> 32);
}

// Site and acceptance words of WL attempt a of a walker.
void attempt_words(int key0, uint32_t slot, uint64_t a, uint32_t& xs, uint32_t& xa)
{
  uint32_t out[4];
  block_at(key0, TAG_WALK, slot, a / 2, out);
  xs = out[2 * (a % 2)];
  xa = out[2 * (a % 2) + 1];
}

// Exchange draw of a walker: advances the high part of its counter.
uint32_t exchange_draw(int key0, uint32_t slot, uint64_t& counter)
{
  const uint32_t x = word_at(key0, TAG_EXCHANGE, slot, counter >> 40);
  counter += EXCHANGE_ONE;
  return x;
}

double from_bits(uint64_t b)
{
  double d;
  std::memcpy(&d, &b, 8);
  return d;
}
uint64_t to_bits(double d)
{
  uint64_t b;
  std::memcpy(&b, &d, 8);
  return b;
}

double exp_det(double x)
{
  if (!(x > -745.2)) {
    return 0.0;
  }
  if (x > 709.78) {
    return std::numeric_limits<double>::infinity();
  }
  const double t = x * 0x1.71547652b82fep+0;
  const double kd = (t + 0x1.8p52) - 0x1.8p52;
  const int k = static_cast<int>(kd);
  const double r = std::fma(-kd, 0x1.a39ef35793c76p-33, std::fma(-kd, 0x1.62e42fee00000p-1, x));
  static const double c[14] = {1.0,
                               1.0,
                               0x1.0000000000000p-1,
                               0x1.5555555555555p-3,
                               0x1.5555555555555p-5,
                               0x1.1111111111111p-7,
                               0x1.6c16c16c16c17p-10,
                               0x1.a01a01a01a01ap-13,
                               0x1.a01a01a01a01ap-16,
                               0x1.71de3a556c734p-19,
                               0x1.27e4fb7789f5cp-22,
                               0x1.ae64567f544e4p-26,
                               0x1.1eed8eff8d898p-29,
                               0x1.6124613a86d09p-33};
  double p = c[13];
  for (int i = 12; i >= 0; --i) {
    p = std::fma(p, r, c[i]);
  }
  const int k1 = k / 2;
  const int k2 = k - k1;
  return (p * from_bits(static_cast<uint64_t>(k1 + 1023) << 52)) * from_bits(static_cast<uint64_t>(k2 + 1023) << 52);
}

double log_det(double x)
{
  if (!(x > 0.0)) {
    return x == 0.0 ? -std::numeric_limits<double>::infinity() : std::numeric_limits<double>::quiet_NaN();
  }
  const uint64_t b = to_bits(x);
  int e = static_cast<int>((b >> 52) & 0x7ff) - 1023;
  double m = from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  if (m > 0x1.6a09e667f3bcdp+0) {
    m *= 0.5;
    e += 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  static const double inv_odd[10] = {0x1.5555555555555p-2,
                                     0x1.999999999999ap-3,
                                     0x1.2492492492492p-3,
                                     0x1.c71c71c71c71cp-4,
                                     0x1.745d1745d1746p-4,
                                     0x1.3b13b13b13b14p-4,
                                     0x1.1111111111111p-4,
                                     0x1.e1e1e1e1e1e1ep-5,
                                     0x1.af286bca1af28p-5,
                                     0x1.8618618618618p-5};
  double q = inv_odd[9];
  for (int i = 8; i >= 0; --i) {
    q = std::fma(q, z, inv_odd[i]);
  }
  const double ts = 2.0 * s;
  const double lm = std::fma(ts * z, q, ts);
  const double ed = static_cast<double>(e);
  return std::fma(ed, 0x1.62e42fee00000p-1, std::fma(ed, 0x1.a39ef35793c76p-33, lm));
}

double sqrt_det(double x)
{
  if (!(x > 0.0)) {
    return x == 0.0 ? 0.0 : std::numeric_limits<double>::quiet_NaN();
  }
  double y = from_bits(0x5fe6eb50c7b537a9ull - (to_bits(x) >> 1));
  const double hx = 0.5 * x;
  for (int i = 0; i < 5; ++i) {
    y = y * std::fma(-hx * y, y, 1.5);
  }
  const double s = x * y;
  return std::fma(std::fma(-s, s, x), 0.5 * y, s);
}

double accept_p(double d)
{
  return d >= 0.0 ? 1.0 : exp_det(d);
}

// ------------------------------------------------ RBIM energy and delta ----

// Energy of one lattice: periodic src/wl.cu:261-300, open src/wl.cu:302-340.
int lattice_energy(const int8_t* s, const int8_t* J, int X, int Y, int boundary)
{
  const int N = X * Y;
  int e = 0;
  for (int i = 0; i < X; ++i) {
    for (int j = 0; j < Y; ++j) {
      if (boundary == 0) {
        const int inn = (i - 1 >= 0) ? i - 1 : X - 1;
        const int jnn = (j - 1 >= 0) ? j - 1 : Y - 1;
        e += s[i * Y + j] * (s[inn * Y + j] * J[N + inn * Y + j] + s[i * Y + jnn] * J[i * Y + jnn]);
      } else {
        const int up = (i > 0) ? s[(i - 1) * Y + j] * J[N + (i - 1) * Y + j] : 0;
        const int left = (j > 0) ? s[i * Y + j - 1] * J[i * Y + j - 1] : 0;
        e += s[i * Y + j] * (up + left);
      }
    }
  }
  return -e;
}

// New energy after flipping site (i, j): src/wl.cu:52-156.
int flip_energy(const int8_t* s, const int8_t* J, int X, int Y, int boundary, int E, int i, int j)
{
  const int N = X * Y;
  const int ipp = (i + 1 < X) ? i + 1 : 0;
  const int inn = (i - 1 >= 0) ? i - 1 : X - 1;
  const int jpp = (j + 1 < Y) ? j + 1 : 0;
  const int jnn = (j - 1 >= 0) ? j - 1 : Y - 1;
  int cu = 1, cd = 1, cl = 1, cr = 1;
  if (boundary == 1) {
    cu = (i > 0);
    cd = (i < X - 1);
    cl = (j > 0);
    cr = (j < Y - 1);
  }
  const int up = cu * s[inn * Y + j] * J[N + inn * Y + j];
  const int left = cl * s[i * Y + jnn] * J[i * Y + jnn];
  const int down = cd * s[ipp * Y + j] * J[N + i * Y + j];
  const int right = cr * s[i * Y + jpp] * J[i * Y + j];
  const int diff = -2 * s[i * Y + j] * (up + left + down + right);
  return E - diff;
}

}  // namespace

// ============================================================ C ABI ========

WLO_API void wlo_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out)
{
  philox10(ctr, key, out);
}
WLO_API double wlo_uniform(int key0, uint32_t tag, uint32_t slot, uint64_t draw)
{
  return uniform_at(key0, tag, slot, draw);
}
WLO_API int wlo_pick(uint32_t x, int n)
{
  return pick(x, n);
}
WLO_API uint32_t wlo_word(int key0, uint32_t tag, uint32_t slot, uint64_t draw)
{
  return word_at(key0, tag, slot, draw);
}
WLO_API double wlo_exp(double x)
{
  return exp_det(x);
}
WLO_API double wlo_log(double x)
{
  return log_det(x);
}
WLO_API double wlo_sqrt(double x)
{
  return sqrt_det(x);
}

// calc_energy (src/wl.cu:342-362); walkers of sample s use couplings s.
WLO_API void wlo_calc_energy(int boundary,
                             int* E,
                             const int8_t* lattice,
                             const int8_t* J,
                             const int* off_lattice,
                             int X,
                             int Y,
                             int64_t num_lattices,
                             int walker_per_interactions)
{
  for (int64_t t = 0; t < num_lattices; ++t) {
    const int64_t s = t / walker_per_interactions;
    E[t] = lattice_energy(lattice + off_lattice[t], J + s * 2 * X * Y, X, Y, boundary);
  }
}

// init_interactions (src/wl.cu:364-386): J = -1 iff uniform < p, one draw per bond.
WLO_API void wlo_init_interactions(int8_t* J, int64_t lattice_size, int num_interactions, int seed, float prob)
{
  for (int s = 0; s < num_interactions; ++s) {
    for (int64_t b = 0; b < 2 * lattice_size; ++b) {
      const double u = uniform_at(seed + s, TAG_DISORDER, 0, static_cast<uint64_t>(b));
      J[s * 2 * lattice_size + b] = (u < static_cast<double>(prob)) ? -1 : 1;
    }
  }
}

// apply_x_horizontal_error (src/wl.cu:388-407): flip the down bonds of row 0.
WLO_API void wlo_apply_x_horizontal_error(int8_t* J, int X, int Y, int num_interactions)
{
  for (int s = 0; s < num_interactions; ++s) {
    for (int col = 0; col < Y; ++col) {
      J[s * 2 * X * Y + X * Y + col] *= -1;
    }
  }
}

// apply_x_vertical_error (src/wl.cu:409-428): flip the right bonds of column 0.
WLO_API void wlo_apply_x_vertical_error(int8_t* J, int X, int Y, int num_interactions)
{
  for (int s = 0; s < num_interactions; ++s) {
    for (int row = 0; row < X; ++row) {
      J[s * 2 * X * Y + row * Y] *= -1;
    }
  }
}

// init_lattice (src/wl.cu:430-457) with D4: lattice l of sample s draws its
// probability q as draw 0 of slot l%W and spin i as draw 1+i.
WLO_API void wlo_init_lattice(int8_t* lattice,
                              float* probs,
                              int X,
                              int Y,
                              int64_t num_lattices,
                              int seed,
                              int walker_per_interactions)
{
  const int N = X * Y;
  for (int64_t l = 0; l < num_lattices; ++l) {
    const int s = static_cast<int>(l / walker_per_interactions);
    const uint32_t slot = static_cast<uint32_t>(l % walker_per_interactions);
    const double q = uniform_at(seed + s, TAG_LATTICE, slot, 0);
    if (probs != nullptr) {
      probs[l] = static_cast<float>(q);
    }
    for (int i = 0; i < N; ++i) {
      lattice[l * N + i] = (uniform_at(seed + s, TAG_LATTICE, slot, 1 + static_cast<uint64_t>(i)) < q) ? -1 : 1;
    }
  }
}

// init_offsets_lattice (src/wl.cu:459-471)
WLO_API void wlo_init_offsets_lattice(int* off, int X, int Y, int64_t num_lattices)
{
  for (int64_t t = 0; t < num_lattices; ++t) {
    off[t] = static_cast<int>(t * X * Y);
  }
}

// wang_landau_pre_run (src/wl.cu:473-560) with store_lattice (227-257), D5:
// for each step all walkers of a sample decide against the histogram as it
// stood at the start of the step, then the counts are added; among walkers
// that accepted into a still-empty prerun interval, the lowest index stores.
WLO_API void wlo_wang_landau_pre_run(int* E,
                                     int* flag_found_interval,
                                     int8_t* lattice,
                                     int8_t* found_interval_lattice,
                                     uint64_t* histogram,
                                     uint64_t* seed_offset,
                                     const int8_t* J,
                                     const int* off_lattice,
                                     int E_min,
                                     int E_max,
                                     int num_wl_iterations,
                                     int X,
                                     int Y,
                                     int seed,
                                     int num_intervals,
                                     int len_interval,
                                     int64_t num_total_walker,
                                     int walker_per_interactions,
                                     int boundary)
{
  const int N = X * Y;
  const int len_hist = E_max - E_min + 1;
  const int S = static_cast<int>(num_total_walker / walker_per_interactions);
  const int Wp = walker_per_interactions;
  std::vector<int> newE(Wp), site(Wp);
  std::vector<char> acc(Wp);
  for (int s = 0; s < S; ++s) {
    uint64_t* H = histogram + static_cast<int64_t>(s) * len_hist;
    const int8_t* Js = J + static_cast<int64_t>(s) * 2 * N;
    for (int it = 0; it < num_wl_iterations; ++it) {
      for (int w = 0; w < Wp; ++w) {
        const int64_t t = static_cast<int64_t>(s) * Wp + w;
        const int8_t* lat = lattice + off_lattice[t];
        const uint32_t xs = word_at(seed + s, TAG_PRERUN, static_cast<uint32_t>(w), seed_offset[t]);
        seed_offset[t] += 1;
        const int sidx = pick(xs, N);
        site[w] = sidx;
        newE[w] = flip_energy(lat, Js, X, Y, boundary, E[t], sidx / Y, sidx % Y);
        const double prob = exp_det(static_cast<double>(H[E[t] - E_min]) - static_cast<double>(H[newE[w] - E_min]));
        const double u2 = uniform_at(seed + s, TAG_PRERUN, static_cast<uint32_t>(w), seed_offset[t]);
        seed_offset[t] += 1;
        acc[w] = u2 < prob;
      }
      for (int w = 0; w < Wp; ++w) {
        const int64_t t = static_cast<int64_t>(s) * Wp + w;
        if (acc[w]) {
          lattice[off_lattice[t] + site[w]] *= -1;
          E[t] = newE[w];
          H[newE[w] - E_min] += 1;
        } else {
          H[E[t] - E_min] += 1;
        }
      }
      for (int w = 0; w < Wp; ++w) {
        if (!acc[w]) {
          continue;
        }
        const int64_t t = static_cast<int64_t>(s) * Wp + w;
        int ii = (E[t] - E_min) / len_interval;
        if (ii >= num_intervals) {
          ii = num_intervals - 1;
        }
        int& flag = flag_found_interval[s * num_intervals + ii];
        if (flag == 0) {
          flag = 1;
          std::memcpy(found_interval_lattice + (static_cast<int64_t>(s) * num_intervals + ii) * N,
                      lattice + off_lattice[t],
                      static_cast<size_t>(N));
        }
      }
    }
  }
}

// check_energy_ranges (src/wl.cu:562-583)
WLO_API void wlo_check_energy_ranges(int8_t* flags,
                                     const int* E,
                                     const int* start,
                                     const int* end,
                                     int64_t num_total_walker,
                                     int walker_per_interval)
{
  for (int64_t t = 0; t < num_total_walker; ++t) {
    const int64_t k = t / walker_per_interval;
    if (E[t] > end[k] || E[t] < start[k]) {
      flags[t] = 1;
    }
  }
}

// init_indices (src/wl.cu:585-594)
WLO_API void wlo_init_indices(int* idx, int64_t num_total_walker, int walker_per_interval)
{
  for (int64_t t = 0; t < num_total_walker; ++t) {
    idx[t] = static_cast<int>(t % walker_per_interval);
  }
}

// reset_d_cond (src/wl.cu:596-613)
WLO_API void wlo_reset_d_cond(int8_t* cond, const double* factor, int total_intervals, double beta, int walker_per_interval)
{
  const double ebeta = exp_det(beta);
  for (int k = 0; k < total_intervals; ++k) {
    if (cond[k] == 1 && factor[static_cast<int64_t>(k) * walker_per_interval] > ebeta) {
      cond[k] = 0;
    }
  }
}

// init_offsets_histogram (src/wl.cu:615-652): base_s + k W len_0 + w len_k.
WLO_API void wlo_init_offsets_histogram(int* off,
                                        const int* start,
                                        const int* end,
                                        const int* len_histograms,
                                        int num_intervals,
                                        int total_walker,
                                        int walker_per_interval)
{
  const int W = walker_per_interval;
  const int S = total_walker / (num_intervals * W);
  int base = 0;
  for (int s = 0; s < S; ++s) {
    const int len0 = end[s * num_intervals] - start[s * num_intervals] + 1;
    for (int k = 0; k < num_intervals; ++k) {
      const int lenk = end[s * num_intervals + k] - start[s * num_intervals + k] + 1;
      for (int w = 0; w < W; ++w) {
        off[(s * num_intervals + k) * W + w] = base + k * W * len0 + w * lenk;
      }
    }
    base += len_histograms[s];
  }
}

// check_interactions_finished (src/wl.cu:662-708)
WLO_API void wlo_check_interactions_finished(int* cond_sample, const int8_t* cond, int num_intervals, int num_interactions)
{
  for (int s = 0; s < num_interactions; ++s) {
    int sum = 0;
    for (int k = 0; k < num_intervals; ++k) {
      sum += cond[s * num_intervals + k];
    }
    cond_sample[s] = (sum == num_intervals) ? -1 : sum;
  }
}

// calc_average_log_g (src/wl.cu:710-783), D3: shared[k][e] accumulates
// logG[w][e] / W for w = 0..W-1 in order, spectrum bins only.
WLO_API void wlo_calc_average_log_g(double* shared_logG,
                                    const int* len_histograms,
                                    const double* logG,
                                    const int* start,
                                    const int* end,
                                    const int8_t* spectrum,
                                    const int8_t* cond,
                                    const int* off_hist,
                                    const int* off_spec,
                                    const int64_t* off_shared,
                                    const int* cond_sample,
                                    int num_interactions,
                                    int walker_per_interval,
                                    int num_intervals,
                                    int total_len_histogram)
{
  (void)len_histograms;
  (void)total_len_histogram;
  const int W = walker_per_interval;
  for (int s = 0; s < num_interactions; ++s) {
    if (cond_sample[s] == -1) {
      continue;
    }
    for (int kk = 0; kk < num_intervals; ++kk) {
      const int k = s * num_intervals + kk;
      if (cond[k] != 1) {
        continue;
      }
      const int len = end[k] - start[k] + 1;
      for (int e = 0; e < len; ++e) {
        if (spectrum[off_spec[s] + start[k] + e - start[s * num_intervals]] != 1) {
          continue;
        }
        double acc = shared_logG[off_shared[k] + e];
        for (int w = 0; w < W; ++w) {
          acc += logG[off_hist[k * W + w] + e] / static_cast<double>(W);
        }
        shared_logG[off_shared[k] + e] = acc;
      }
    }
  }
}

// redistribute_g_values (src/wl.cu:969-1031): every bin of every walker of a
// flat interval takes the shared value.
WLO_API void wlo_redistribute_g_values(double* logG,
                                       const double* shared_logG,
                                       const int* len_histograms,
                                       const int* start,
                                       const int* end,
                                       const int* cond_sample,
                                       const int* off_hist,
                                       const int8_t* cond,
                                       const int64_t* off_shared,
                                       int num_intervals,
                                       int walker_per_interval,
                                       int num_interactions,
                                       int total_len_histogram)
{
  (void)len_histograms;
  (void)total_len_histogram;
  const int W = walker_per_interval;
  for (int s = 0; s < num_interactions; ++s) {
    if (cond_sample[s] == -1) {
      continue;
    }
    for (int kk = 0; kk < num_intervals; ++kk) {
      const int k = s * num_intervals + kk;
      if (cond[k] != 1) {
        continue;
      }
      const int len = end[k] - start[k] + 1;
      for (int w = 0; w < W; ++w) {
        for (int e = 0; e < len; ++e) {
          logG[off_hist[k * W + w] + e] = shared_logG[off_shared[k] + e];
        }
      }
    }
  }
}

// replica_exchange (src/wl.cu:810-890) with fisher_yates (785-808), D10, D13.
WLO_API void wlo_replica_exchange(int* off_lattice,
                                  int* E,
                                  int* idx,
                                  uint64_t* off_iter,
                                  const int* start,
                                  const int* end,
                                  const int* off_hist,
                                  const int* cond_sample,
                                  const double* logG,
                                  int even,
                                  int seed,
                                  int num_intervals,
                                  int walker_per_interactions,
                                  int num_interactions,
                                  int global_parity)
{
  const int I = num_intervals;
  const int W = walker_per_interactions / I;
  for (int k = 0; k < num_interactions * I; ++k) {
    const int kl = k % I;
    if (kl == I - 1) {
      continue;
    }
    const int par = global_parity ? (k % 2) : (kl % 2);
    if ((even && par != 0) || (!even && par == 0)) {
      continue;
    }
    const int s = k / I;
    if (cond_sample[s] == -1) {
      continue;
    }
    // Fisher-Yates by walker 0 of the interval (src/wl.cu:785-808).
    {
      const int64_t t0 = static_cast<int64_t>(k) * W;
      const uint32_t slot = static_cast<uint32_t>(t0 % walker_per_interactions);
      for (int i = W - 1; i > 0; --i) {
        const int r = pick(exchange_draw(seed + s, slot, off_iter[t0]), i + 1);
        std::swap(idx[t0 + i], idx[t0 + r]);
      }
    }
    for (int w = 0; w < W; ++w) {
      const int64_t t = static_cast<int64_t>(k) * W + w;
      const int64_t sw = static_cast<int64_t>(k + 1) * W + idx[t];
      if (E[t] > end[k + 1] || E[t] < start[k + 1]) {
        continue;
      }
      if (E[sw] > end[k] || E[sw] < start[k]) {
        continue;
      }
      const double giEX = logG[off_hist[t] + E[t] - start[k]];
      const double giEY = logG[off_hist[t] + E[sw] - start[k]];
      const double gjEY = logG[off_hist[sw] + E[sw] - start[k + 1]];
      const double gjEX = logG[off_hist[sw] + E[t] - start[k + 1]];
      const double p = accept_p((giEX - giEY) + (gjEY - gjEX));
      const uint32_t slot = static_cast<uint32_t>(t % walker_per_interactions);
      if (unit(exchange_draw(seed + s, slot, off_iter[t])) < p) {
        std::swap(off_lattice[t], off_lattice[sw]);
        std::swap(E[t], E[sw]);
      }
    }
  }
}

// check_histogram (src/wl.cu:892-967)
WLO_API void wlo_check_histogram(uint64_t* H,
                                 double* factor,
                                 int8_t* cond,
                                 const int* off_hist,
                                 const int* end,
                                 const int* start,
                                 const int8_t* spectrum,
                                 const int* off_spec,
                                 const int* cond_sample,
                                 double alpha,
                                 int num_walker_total,
                                 int walker_per_interactions,
                                 int num_intervals)
{
  const int W = walker_per_interactions / num_intervals;
  const int K = num_walker_total / W;
  for (int k = 0; k < K; ++k) {
    const int s = k / num_intervals;
    if (cond[k] == 1 || cond_sample[s] == -1) {
      continue;
    }
    const int len = end[k] - start[k] + 1;
    int finished = 0;
    for (int w = 0; w < W; ++w) {
      const int64_t t = static_cast<int64_t>(k) * W + w;
      uint64_t mn = std::numeric_limits<int>::max();
      uint64_t sum = 0;
      int cnt = 0;
      for (int i = 0; i < len; ++i) {
        if (spectrum[off_spec[s] + start[k] + i - start[s * num_intervals]] == 1) {
          const uint64_t h = H[off_hist[t] + i];
          mn = std::min(mn, h);
          sum += h;
          cnt += 1;
        }
      }
      if (cnt > 0) {
        const double mean = static_cast<double>(sum) / cnt;
        if (static_cast<double>(mn) >= alpha * mean) {
          finished += 1;
        }
      }
    }
    if (finished == W) {
      cond[k] = 1;
      for (int w = 0; w < W; ++w) {
        const int64_t t = static_cast<int64_t>(k) * W + w;
        for (int i = 0; i < len; ++i) {
          H[off_hist[t] + i] = 0;
        }
        factor[t] = sqrt_det(factor[t]);
      }
    }
  }
}

// wang_landau (src/wl.cu:1033-1175) with D2, D6, D12.
WLO_API void wlo_wang_landau(int8_t* lattice,
                             const int8_t* J,
                             uint64_t* H,
                             double* logG,
                             const double* factor,
                             int* E,
                             int* newE,
                             uint64_t* off_iter,
                             int* found,
                             const int* start,
                             const int* end,
                             const int* off_hist,
                             const int* off_lattice,
                             const int8_t* cond,
                             const int8_t* spectrum,
                             const int* off_spec,
                             const int* cond_sample,
                             int X,
                             int Y,
                             int num_iterations,
                             int num_lattices,
                             int num_intervals,
                             int walker_per_interactions,
                             int seed,
                             int boundary)
{
  const int N = X * Y;
  const int W = walker_per_interactions / num_intervals;
  for (int t = 0; t < num_lattices; ++t) {
    const int s = t / walker_per_interactions;
    const int k = t / W;
    if (cond_sample[s] == -1) {
      continue;
    }
    const int8_t* Js = J + static_cast<int64_t>(s) * 2 * N;
    int8_t* lat = lattice + off_lattice[t];
    const uint32_t slot = static_cast<uint32_t>(t % walker_per_interactions);
    const int emin = start[s * num_intervals];
    const int emax = end[s * num_intervals + num_intervals - 1];
    const double lnf = log_det(factor[t]);
    for (int it = 0; it < num_iterations; ++it) {
      uint32_t xs, xa;
      attempt_words(seed + s, slot, off_iter[t] & ATTEMPT_MASK, xs, xa);
      off_iter[t] += 1;
      const int sidx = pick(xs, N);
      const int i = sidx / Y, j = sidx % Y;
      const int en = flip_energy(lat, Js, X, Y, boundary, E[t], i, j);
      if (en < emin || en > emax || spectrum[off_spec[s] + en - emin] != 1) {
        found[t] = t + 1;
        newE[t] = en;
        break;
      }
      found[t] = 0;
      if (cond[k] == 0) {
        const int io = off_hist[t] + E[t] - start[k];
        if (en > end[k] || en < start[k]) {
          H[io] += 1;
          logG[io] += lnf;
        } else {
          const int in = off_hist[t] + en - start[k];
          const double p = accept_p(logG[io] - logG[in]);
          if (unit(xa) < p) {
            lat[sidx] *= -1;
            H[in] += 1;
            logG[in] += lnf;
            E[t] = en;
          } else {
            H[io] += 1;
            logG[io] += lnf;
          }
        }
      } else {
        if (en <= end[k] && en >= start[k]) {
          const int io = off_hist[t] + E[t] - start[k];
          const int in = off_hist[t] + en - start[k];
          const double p = accept_p(logG[io] - logG[in]);
          if (unit(xa) < p) {
            lat[sidx] *= -1;
            E[t] = en;
          }
        }
      }
    }
  }
}

// ------------------------------------------------------- intervals ---------

// interval_results (src/interval.cpp:5-41), float arithmetic as written there.
WLO_API int wlo_interval_results(int e_min,
                                 int e_max,
                                 unsigned int num_intervals,
                                 int num_walkers,
                                 float overlap,
                                 int* h_start,
                                 int* h_end,
                                 int64_t* len_histogram_over_all_walkers,
                                 int* len_interval)
{
  if (overlap > 1 || overlap < 0) {
    return -1;
  }
  const int e_range = e_max - e_min + 1;
  const int len = static_cast<int>(static_cast<float>(e_range) / (1.0F + overlap * static_cast<float>(num_intervals - 1)));
  const int step = static_cast<int>(overlap * static_cast<float>(len));
  int st = e_min;
  int64_t total = 0;
  for (unsigned int i = 0; i < num_intervals; ++i) {
    h_start[i] = st;
    if (i < num_intervals - 1) {
      h_end[i] = st + len - 1;
      total += static_cast<int64_t>(num_walkers) * len;
    } else {
      h_end[i] = e_max;
      total += static_cast<int64_t>(num_walkers) * (e_max - h_start[i] + 1);
    }
    if (h_start[i] >= h_end[i]) {
      return -2;
    }
    st += step;
  }
  *len_histogram_over_all_walkers = total;
  *len_interval = len;
  return 0;
}

// ------------------------------------------------------- mainrun loop ------

// One run of the mainrun loop (src/mainrun.cu:243-383) over S samples that
// are already set up: state arrays in the reference's flat layout.  Returns
// the number of loop iterations, or -1 on a "new energy" abort.
WLO_API int wlo_mainrun_loop(int8_t* lattice,
                             const int8_t* J,
                             uint64_t* H,
                             double* logG,
                             double* factor,
                             int* E,
                             uint64_t* off_iter,
                             int* off_lattice,
                             int* idx,
                             int8_t* cond,
                             int* cond_sample,
                             const int* start,
                             const int* end,
                             const int* off_hist,
                             const int8_t* spectrum,
                             const int* off_spec,
                             const int* len_histograms,
                             int S,
                             int I,
                             int W,
                             int X,
                             int Y,
                             int boundary,
                             int num_iterations,
                             double alpha,
                             double beta,
                             int exchange_every,
                             int seed,
                             int compat,
                             int64_t max_rounds,
                             int64_t* counter_io)
{
  const int total_walker = S * I * W;
  const int per_sample = I * W;
  std::vector<int64_t> off_shared(static_cast<size_t>(S) * I);
  int64_t size_shared = 0;
  for (int k = 0; k < S * I; ++k) {
    off_shared[k] = size_shared;
    size_shared += end[k] - start[k] + 1;
  }
  int total_len = 0;
  for (int s = 0; s < S; ++s) {
    total_len += len_histograms[s];
  }
  std::vector<double> shared(static_cast<size_t>(size_shared), 0.0);
  std::vector<int> newE(total_walker, 0), found(total_walker, 0);
  double max_f = 0;
  for (int t = 0; t < total_walker; ++t) {
    max_f = std::max(max_f, factor[t]);
  }
  const double ebeta = exp_det(beta);
  int64_t counter = (counter_io != nullptr) ? *counter_io : 1;
  int64_t rounds = 0;
  while (max_f - ebeta > 1e-10 && rounds < max_rounds) {
    wlo_wang_landau(lattice, J, H, logG, factor, E, newE.data(), off_iter, found.data(), start, end, off_hist,
                    off_lattice, cond, spectrum, off_spec, cond_sample, X, Y, num_iterations, total_walker, I,
                    per_sample, seed, boundary);
    for (int t = 0; t < total_walker; ++t) {
      if (found[t] != 0) {
        if (counter_io != nullptr) {
          *counter_io = counter;
        }
        return -1;
      }
    }
    wlo_check_histogram(H, factor, cond, off_hist, end, start, spectrum, off_spec, cond_sample, alpha, total_walker,
                        per_sample, I);
    wlo_calc_average_log_g(shared.data(), len_histograms, logG, start, end, spectrum, cond, off_hist, off_spec,
                           off_shared.data(), cond_sample, S, W, I, total_len);
    wlo_redistribute_g_values(logG, shared.data(), len_histograms, start, end, cond_sample, off_hist, cond,
                              off_shared.data(), I, W, S, total_len);
    std::fill(shared.begin(), shared.end(), 0.0);
    wlo_reset_d_cond(cond, factor, S * I, beta, W);
    wlo_check_interactions_finished(cond_sample, cond, I, S);
    max_f = 0;
    for (int t = 0; t < total_walker; ++t) {
      max_f = std::max(max_f, factor[t]);
    }
    if (counter % exchange_every == 0) {
      wlo_replica_exchange(off_lattice, E, idx, off_iter, start, end, off_hist, cond_sample, logG, 1, seed, I,
                           per_sample, S, compat);
      wlo_replica_exchange(off_lattice, E, idx, off_iter, start, end, off_hist, cond_sample, logG, 0, seed, I,
                           per_sample, S, compat);
    }
    if (!compat) {
      counter += 1;
    }
    rounds += 1;
  }
  if (counter_io != nullptr) {
    *counter_io = counter;
  }
  return static_cast<int>(rounds);
}

// ------------------------------------------------------- stitching ---------

namespace {

using imap = std::map<int, double>;

// find_stitching_keys (result_handling.cpp:115-143) with D7.
std::tuple<int, double> find_keys(const imap& a, const imap& b)
{
  int min_key = std::numeric_limits<int>::max();
  double min_diff = std::numeric_limits<double>::max();
  auto i1 = a.begin();
  auto i2 = b.begin();
  while (i1 != a.end() && i2 != b.end()) {
    if (i1->first == i2->first) {
      auto n1 = std::next(i1);
      auto n2 = std::next(i2);
      if (n1 != a.end() && n2 != b.end()) {
        const double d1 = (n1->second - i1->second) / (n1->first - i1->first);
        const double d2 = (n2->second - i2->second) / (n2->first - i2->first);
        const double dd = std::fabs(d1 - d2);
        if (dd < min_diff) {
          min_diff = dd;
          min_key = i1->first;
        }
      }
      ++i1;
      ++i2;
    } else if (i1->first < i2->first) {
      ++i1;
    } else {
      ++i2;
    }
  }
  return {min_key, min_diff};
}

imap filter_keys(const imap& m, int thr, bool less)
{
  imap out;
  for (const auto& [k, v] : m) {
    if ((less && k < thr) || (!less && k >= thr)) {
      out[k] = v;
    }
  }
  return out;
}

}  // namespace

// result_handling_stitched_histogram (result_handling.cpp:368-431) for one
// sample: logG_sample is the sample's histogram block (all walkers).  Writes
// the stitched, normalised (E, ln g) pairs; returns the count, or a negative
// error code.
WLO_API int wlo_stitch(const double* logG_sample,
                       const int* start,
                       const int* end,
                       int I,
                       int W,
                       int X,
                       int Y,
                       int* out_E,
                       double* out_lng,
                       int max_out)
{
  std::vector<imap> data(I);
  int64_t idx = 0;
  for (int i = 0; i < I; ++i) {
    const int len = end[i] - start[i] + 1;
    for (int k = 0; k < len; ++k) {
      const double v = logG_sample[idx + k];
      if (v != 0) {
        data[i][start[i] + k] = v;
      }
    }
    idx += static_cast<int64_t>(len) * W;
  }
  // rescale_by_minimum (:74-101)
  for (int i = 0; i < I; ++i) {
    double mn = std::numeric_limits<double>::max();
    for (const auto& kv : data[i]) {
      mn = std::min(mn, kv.second);
    }
    if (mn == std::numeric_limits<double>::max()) {
      mn = 0;
    }
    for (auto& kv : data[i]) {
      kv.second -= mn;
    }
  }
  // calculate_stitching_points (:157-191)
  std::vector<std::tuple<int, int>> keys;
  for (int i = 0; i < I - 1; ++i) {
    int abs_key = std::numeric_limits<int>::max();
    double min_der = std::numeric_limits<double>::max();
    int interval_index = std::numeric_limits<int>::max();
    for (int j = i + 1; j < I; ++j) {
      auto [mk, der] = find_keys(data[i], data[j]);
      if (mk < std::numeric_limits<int>::max()) {
        if (der < min_der) {
          abs_key = mk;
          min_der = der;
          interval_index = j;
        }
      } else {
        break;
      }
    }
    if (interval_index < std::numeric_limits<int>::max()) {
      keys.emplace_back(interval_index, abs_key);
    }
  }
  // keep keys not undercut by a later key (:383-396)
  std::vector<std::tuple<int, int>> real;
  for (size_t i = 0; i < keys.size(); ++i) {
    bool ok = true;
    for (size_t j = i + 1; j < keys.size(); ++j) {
      if (std::get<1>(keys[j]) <= std::get<1>(keys[i])) {
        ok = false;
      }
    }
    if (ok) {
      real.push_back(keys[i]);
    }
  }
  if (real.empty()) {
    return -2;
  }
  // smallest key per run of equal interval index (:399-421)
  std::vector<std::tuple<int, int>> smallest;
  int cur_key = std::get<0>(real[0]);
  int cur_min = std::get<1>(real[0]);
  for (const auto& [key, value] : real) {
    if (key != cur_key) {
      smallest.emplace_back(cur_key, cur_min);
      cur_key = key;
      cur_min = value;
    } else if (value < cur_min) {
      cur_min = value;
    }
  }
  smallest.emplace_back(cur_key, cur_min);
  // rescale_intervals_for_concatenation (:202-234) with D8
  for (size_t i = 0; i < smallest.size(); ++i) {
    const int e = std::get<1>(smallest[i]);
    const int nxt = std::get<0>(smallest[i]);
    const int prv = (i == 0) ? 0 : std::get<0>(smallest[i - 1]);
    auto ip = data[prv].find(e);
    auto in = data[nxt].find(e);
    if (ip == data[prv].end() || in == data[nxt].end()) {
      return -3;
    }
    const double shift = ip->second - in->second;
    for (auto& kv : data[nxt]) {
      kv.second += shift;
    }
  }
  // cut_overlapping_histogram_parts (:270-306)
  std::vector<imap> cut;
  {
    auto f = filter_keys(data[0], std::get<1>(smallest[0]), true);
    if (!f.empty()) {
      cut.push_back(f);
    }
  }
  for (size_t i = 1; i < smallest.size(); ++i) {
    auto f = filter_keys(data[std::get<0>(smallest[i - 1])], std::get<1>(smallest[i]), true);
    f = filter_keys(f, std::get<1>(smallest[i - 1]), false);
    if (!f.empty()) {
      cut.push_back(f);
    }
  }
  {
    auto f = filter_keys(data[std::get<0>(smallest.back())], std::get<1>(smallest.back()), false);
    if (!f.empty()) {
      cut.push_back(f);
    }
  }
  // rescale_map_values (:319-365): ln g + N ln2 - LSE(ln g)
  double mx = -std::numeric_limits<double>::infinity();
  for (const auto& m : cut) {
    for (const auto& kv : m) {
      mx = std::max(mx, kv.second);
    }
  }
  double se = 0;
  for (const auto& m : cut) {
    for (const auto& kv : m) {
      se += exp_det(kv.second - mx);
    }
  }
  const double offset = mx + log_det(se);
  const double l2 = log_det(2.0) * static_cast<double>(X) * static_cast<double>(Y);
  int n = 0;
  for (const auto& m : cut) {
    for (const auto& kv : m) {
      if (n >= max_out) {
        return -4;
      }
      out_E[n] = kv.first;
      out_lng[n] = kv.second + l2 - offset;
      ++n;
    }
  }
  return n;
}

// ln Z(beta) = LSE_E(ln g(E) - beta E).  Not in the reference: the step after
// its output, from the cited paper (README.md:198-200).
WLO_API double wlo_log_partition(const int* E, const double* lng, int n, double beta)
{
  double mx = -std::numeric_limits<double>::infinity();
  for (int i = 0; i < n; ++i) {
    mx = std::max(mx, lng[i] - beta * E[i]);
  }
  double se = 0;
  for (int i = 0; i < n; ++i) {
    se += exp_det((lng[i] - beta * E[i]) - mx);
  }
  return mx + log_det(se);
}

// Nishimori inverse temperature 1/2 ln((1-p)/p).
WLO_API double wlo_nishimori_beta(double p)
{
  return 0.5 * log_det((1.0 - p) / p);
}

// ----------------------------------------- task-parallel CPU baseline ------

// Runs `num_tasks` independent single-sample mainrun loops in parallel
// (OpenMP over tasks).  Task q owns the q-th block of every state array;
// keys[q] is its seed + sample.  Returns the total number of WL attempts
// performed in `attempts` and the number of failed tasks.
WLO_API int wlo_run_tasks(int num_tasks,
                          int8_t* lattice,
                          const int8_t* J,
                          uint64_t* H,
                          double* logG,
                          double* factor,
                          int* E,
                          uint64_t* off_iter,
                          int* off_lattice,
                          int* idx,
                          int8_t* cond,
                          int* cond_sample,
                          const int* start,
                          const int* end,
                          const int* off_hist,
                          const int8_t* spectrum,
                          const int* off_spec,
                          const int* len_histograms,
                          const int* keys,
                          int I,
                          int W,
                          int X,
                          int Y,
                          int boundary,
                          int num_iterations,
                          double alpha,
                          double beta,
                          int exchange_every,
                          int64_t max_rounds,
                          int64_t* rounds_out,
                          int num_threads)
{
  const int N = X * Y;
  const int per = I * W;
  int failed = 0;
#ifdef _OPENMP
  if (num_threads > 0) {
    omp_set_num_threads(num_threads);
  }
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : failed)
#endif
  for (int q = 0; q < num_tasks; ++q) {
    // Per-task views: offsets inside the task are task-local.
    const int64_t wbase = static_cast<int64_t>(q) * per;
    const int64_t hbase = off_hist[wbase];
    std::vector<int> loc_off_hist(per), loc_off_lat(per);
    for (int w = 0; w < per; ++w) {
      loc_off_hist[w] = static_cast<int>(off_hist[wbase + w] - hbase);
      loc_off_lat[w] = static_cast<int>(off_lattice[wbase + w] - wbase * N);
    }
    int zero = 0;
    const int r = wlo_mainrun_loop(lattice + wbase * N, J + static_cast<int64_t>(q) * 2 * N, H + hbase, logG + hbase,
                                   factor + wbase, E + wbase, off_iter + wbase, loc_off_lat.data(), idx + wbase,
                                   cond + static_cast<int64_t>(q) * I, cond_sample + q, start + static_cast<int64_t>(q) * I,
                                   end + static_cast<int64_t>(q) * I, loc_off_hist.data(), spectrum + off_spec[q], &zero,
                                   len_histograms + q, 1, I, W, X, Y, boundary, num_iterations, alpha, beta,
                                   exchange_every, keys[q], 0, max_rounds, nullptr);
    for (int w = 0; w < per; ++w) {
      off_lattice[wbase + w] = loc_off_lat[w] + static_cast<int>(wbase * N);
    }
    rounds_out[q] = r;
    if (r < 0) {
      failed += 1;
    }
  }
  return failed;
}

WLO_API int wlo_max_threads()
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
