/*
 * Host-side C entry points of libwl_amd: the domain logic around the kernels
 * (interval partition, prerun spectrum and seed lattices, stitching, log Z).
 * All pointers here are host pointers.  Status codes as in wl_amd.h.
 */
#ifndef WL_HOST_H_
#define WL_HOST_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* interval_results, src/interval.hpp:14-39 / src/interval.cpp:5-41.
 * Fills h_start/h_end (num_intervals entries each).  Returns 1 with
 * "Overlap is out of bounds." or "The interval arguments create empty
 * intervals." where the reference throws std::invalid_argument. */
int wl_interval_results(int e_min, int e_max, unsigned int num_intervals, int num_walkers, float overlap, int* h_start,
                        int* h_end, int64_t* len_histogram_over_all_walkers, int* len_interval);

/* get_energy_spectrum_information_from_prerun_results for one sample,
 * src/file_io.cpp:558-616.  hist: 0/1 flags over [-2XY, 2XY]; mask gets the
 * flags over [E_min, E_max] (len_mask entries).  Returns 2 if no energy is set. */
int wl_energy_spectrum(const uint8_t* hist, int len_hist, int X, int Y, int* E_min, int* E_max, int8_t* mask,
                       int* len_mask);

/* Spacing of the energies in a spectrum mask (4 periodic, 2 open; 0 for a
 * single energy or a null mask). */
int wl_energy_stride(const int8_t* mask, int len);

/* get_lattice_from_prerun_results for one sample, src/file_io.cpp:618-659.
 * out: I*W lattices of N spins.  Returns 2 if an interval has no lattice. */
int wl_select_seed_lattices(const int8_t* found, const int* found_energy, int num_found, int N, const int* start,
                            const int* end, int I, int W, int8_t* out);

/* result_handling_stitched_histogram without the file write,
 * src/result_handling.cpp:368-431: stitched, normalised (E, ln g) pairs of
 * one sample from its ln g block (reference layout, I intervals x W walkers). */
int wl_stitch_log_g(const double* logG_sample, const int* start, const int* end, int I, int W, int X, int Y,
                    int* out_E, double* out_lng, int max_out, int* n_out);

/* ln Z(beta) = log sum_E g(E) exp(-beta E); NaN when n > 0 and E or lng is
 * null.  No counterpart in the reference (its output stops at ln g,
 * README.md:198-200 cites the decoding step). */
double wl_log_partition(const int* E, const double* lng, int n, double beta);

/* Nishimori inverse temperature 1/2 ln((1-p)/p). */
double wl_nishimori_beta(double p);

/* Logical-class decision of num_samples samples from their ln Z
 * (num_samples x num_classes, class 0 = 'i', the sampled error's own coset).
 * decision[s] = first argmax (-1 if any ln Z is not finite); fail_weight[s] =
 * probability that a decoder choosing uniformly among the tied maxima does
 * not pick class 0 (0 or 1 without ties; NaN when undecided).  No counterpart
 * in the reference: the aggregation after decoding that README.md:198-200
 * cites (logical error curves). */
int wl_decide_classes(const double* log_z, int num_samples, int num_classes, int* decision, double* fail_weight);

/* Wilson score interval [lo, hi] of the failure rate failures/n at z
 * standard errors (z = 1.96 for 95 %).  No counterpart in the reference. */
int wl_wilson_interval(double failures, int n, double z, double* lo, double* hi);

#ifdef __cplusplus
}
#endif

#endif /* WL_HOST_H_ */
