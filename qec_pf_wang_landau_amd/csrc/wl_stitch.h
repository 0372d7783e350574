// Overlap-matched concatenation of per-interval ln g(E) into one normalised
// ln g(E) for a disorder sample, and log Z at inverse temperature beta.
//
// Follows result_handling_stitched_histogram (src/result_handling.cpp:368-431)
// step for step, on dense per-interval arrays instead of std::map (so the same
// code runs on the host and inside a device kernel):
//   1. walker 0 of each interval, entries with ln g != 0 (:26-60)
//   2. subtract the per-interval minimum (:74-101)
//   3. per interval i, the later interval j and shared energy whose forward
//      differences of ln g match best (:115-191); shared energies without a
//      successor in either interval are skipped (the reference dereferences
//      past the end there, :125-126)
//   4. keep keys not undercut by a later key, then the smallest key per run
//      of equal target interval (:383-421)
//   5. shift each target interval onto its predecessor at the key (:202-234),
//      looking the key up in the map that is searched (:209-218 compares with
//      another map's end)
//   6. cut the overlaps at the keys (:270-306)
//   7. ln g += N ln 2 - LSE(ln g) so that sum g = 2^N (:319-365)
// FP64 operations run in the reference's order; exp/log are the shared
// deterministic ones, so host and device give the same bits.
#pragma once

#include <cstdint>

#include "wl_math.h"

namespace wlamd {

struct stitch_work {
  double* vals;     // sum of interval lengths
  uint8_t* pres;    // same
  int* nxt;         // same: next present position in the interval, -1 if none
  int* key_iv;      // I entries
  int* key_e;       // I entries
};

// Returns the number of output entries, or -2 (no stitching key), -3 (key
// missing in an interval), -4 (output too small).
WL_HD int stitch_dense(const double* logG_sample, const int* start, const int* end, int I, int W, int X, int Y,
                       stitch_work wk, int* out_E, double* out_lng, int max_out)
{
  const int kIntMax = 0x7fffffff;
  const double kDblMax = 1.7976931348623157e308;
  // 1. walker 0 of every interval
  {
    int64_t idx = 0;
    int vo = 0;
    for (int i = 0; i < I; ++i) {
      const int len = end[i] - start[i] + 1;
      for (int k = 0; k < len; ++k) {
        const double v = logG_sample[idx + k];
        wk.vals[vo + k] = v;
        wk.pres[vo + k] = (v != 0.0) ? 1 : 0;
      }
      int nx = -1;
      for (int k = len - 1; k >= 0; --k) {
        wk.nxt[vo + k] = nx;
        if (wk.pres[vo + k]) {
          nx = k;
        }
      }
      idx += static_cast<int64_t>(len) * W;
      vo += len;
    }
  }
  // 2. minimum per interval
  {
    int vo = 0;
    for (int i = 0; i < I; ++i) {
      const int len = end[i] - start[i] + 1;
      double mn = kDblMax;
      for (int k = 0; k < len; ++k) {
        if (wk.pres[vo + k] && wk.vals[vo + k] < mn) {
          mn = wk.vals[vo + k];
        }
      }
      if (mn == kDblMax) {
        mn = 0.0;
      }
      for (int k = 0; k < len; ++k) {
        if (wk.pres[vo + k]) {
          wk.vals[vo + k] -= mn;
        }
      }
      vo += len;
    }
  }
  auto voff = [&](int i) {
    int o = 0;
    for (int q = 0; q < i; ++q) {
      o += end[q] - start[q] + 1;
    }
    return o;
  };
  // 3. stitching keys
  int nkeys = 0;
  for (int i = 0; i < I - 1; ++i) {
    int abs_key = kIntMax;
    double min_der = kDblMax;
    int target = kIntMax;
    const int oi = voff(i);
    for (int j = i + 1; j < I; ++j) {
      const int oj = voff(j);
      int min_key = kIntMax;
      double min_diff = kDblMax;
      const int lo = (start[i] > start[j]) ? start[i] : start[j];
      const int hi = (end[i] < end[j]) ? end[i] : end[j];
      for (int e = lo; e <= hi; ++e) {
        const int a = oi + e - start[i];
        const int b = oj + e - start[j];
        if (!wk.pres[a] || !wk.pres[b]) {
          continue;
        }
        const int na = wk.nxt[a];
        const int nb = wk.nxt[b];
        if (na < 0 || nb < 0) {
          continue;
        }
        const int ea = start[i] + na;
        const int eb = start[j] + nb;
        const double d1 = (wk.vals[oi + na] - wk.vals[a]) / static_cast<double>(ea - e);
        const double d2 = (wk.vals[oj + nb] - wk.vals[b]) / static_cast<double>(eb - e);
        double dd = d1 - d2;
        dd = (dd < 0.0) ? -dd : dd;
        if (dd < min_diff) {
          min_diff = dd;
          min_key = e;
        }
      }
      if (min_key < kIntMax) {
        if (min_diff < min_der) {
          abs_key = min_key;
          min_der = min_diff;
          target = j;
        }
      } else {
        break;
      }
    }
    if (target < kIntMax) {
      wk.key_iv[nkeys] = target;
      wk.key_e[nkeys] = abs_key;
      ++nkeys;
    }
  }
  // 4. keys not undercut by a later key; then the smallest per run
  int nreal = 0;
  for (int i = 0; i < nkeys; ++i) {
    bool ok = true;
    for (int j = i + 1; j < nkeys; ++j) {
      if (wk.key_e[j] <= wk.key_e[i]) {
        ok = false;
      }
    }
    if (ok) {
      wk.key_iv[nreal] = wk.key_iv[i];
      wk.key_e[nreal] = wk.key_e[i];
      ++nreal;
    }
  }
  if (nreal == 0) {
    return -2;
  }
  int nsm = 0;
  {
    int cur_iv = wk.key_iv[0];
    int cur_min = wk.key_e[0];
    for (int i = 0; i < nreal; ++i) {
      if (wk.key_iv[i] != cur_iv) {
        wk.key_iv[nsm] = cur_iv;
        wk.key_e[nsm] = cur_min;
        ++nsm;
        cur_iv = wk.key_iv[i];
        cur_min = wk.key_e[i];
      } else if (wk.key_e[i] < cur_min) {
        cur_min = wk.key_e[i];
      }
    }
    wk.key_iv[nsm] = cur_iv;
    wk.key_e[nsm] = cur_min;
    ++nsm;
  }
  // 5. shifts
  for (int i = 0; i < nsm; ++i) {
    const int e = wk.key_e[i];
    const int nx = wk.key_iv[i];
    const int pv = (i == 0) ? 0 : wk.key_iv[i - 1];
    if (e < start[pv] || e > end[pv] || e < start[nx] || e > end[nx]) {
      return -3;
    }
    const int a = voff(pv) + e - start[pv];
    const int onx = voff(nx);
    const int b = onx + e - start[nx];
    if (!wk.pres[a] || !wk.pres[b]) {
      return -3;
    }
    const double shift = wk.vals[a] - wk.vals[b];
    const int len = end[nx] - start[nx] + 1;
    for (int k = 0; k < len; ++k) {
      if (wk.pres[onx + k]) {
        wk.vals[onx + k] += shift;
      }
    }
  }
  // 6. cut: segment s takes interval src(s) over [lo(s), hi(s))
  int n = 0;
  for (int sgm = 0; sgm <= nsm; ++sgm) {
    const int src = (sgm == 0) ? 0 : wk.key_iv[sgm - 1];
    const int lo = (sgm == 0) ? -0x7fffffff : wk.key_e[sgm - 1];
    const int hi = (sgm == nsm) ? 0x7fffffff : wk.key_e[sgm];
    const int o = voff(src);
    const int len = end[src] - start[src] + 1;
    for (int k = 0; k < len; ++k) {
      const int e = start[src] + k;
      if (wk.pres[o + k] && e >= lo && e < hi) {
        if (n >= max_out) {
          return -4;
        }
        out_E[n] = e;
        out_lng[n] = wk.vals[o + k];
        ++n;
      }
    }
  }
  // 7. normalise
  double mx = -bits_dbl(0x7ff0000000000000ull);
  for (int q = 0; q < n; ++q) {
    mx = (out_lng[q] > mx) ? out_lng[q] : mx;
  }
  double se = 0.0;
  for (int q = 0; q < n; ++q) {
    se += wl_exp(out_lng[q] - mx);
  }
  const double offset = mx + wl_log(se);
  const double l2 = wl_log(2.0) * static_cast<double>(X) * static_cast<double>(Y);
  for (int q = 0; q < n; ++q) {
    out_lng[q] = out_lng[q] + l2 - offset;
  }
  return n;
}

// ln Z(beta) = LSE_E(ln g(E) - beta E).
WL_HD double log_partition(const int* E, const double* lng, int n, double beta)
{
  double mx = -bits_dbl(0x7ff0000000000000ull);
  for (int i = 0; i < n; ++i) {
    const double v = lng[i] - beta * E[i];
    mx = (v > mx) ? v : mx;
  }
  double se = 0.0;
  for (int i = 0; i < n; ++i) {
    se += wl_exp((lng[i] - beta * E[i]) - mx);
  }
  return mx + wl_log(se);
}

}  // namespace wlamd
