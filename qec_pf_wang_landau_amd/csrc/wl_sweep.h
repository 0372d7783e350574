// Device code of one walker's Wang-Landau sweep on LDS state, shared by the
// batched persistent kernel (wl_fused.hip) and the LDS-staged per-interval
// sweep of the round-graph path (wl_plain.hip).
//
// Spins and coupling signs are bit rows (one u32 per lattice row, Y <= 32;
// a set bit is spin -1 / an antiferromagnetic bond); ln g (f64) and H (u32)
// are per walker over the interval's reachable energies e0 + stride*b,
// walker-minor with stride W.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "wl_math.h"

namespace wlamd {

// Walker registers carried across attempts.
struct sweep_state {
  int e;       // energy
  int b;       // compact bin of e
  double g;    // ln g at b (mirrors the LDS value)
  bool alive;  // no energy outside the spectrum seen yet
  int bad_e;   // the first energy outside the spectrum, once alive is false
};

// Loop-invariant inputs of one walker's sweep.
struct sweep_ctx {
  uint32_t* L;            // lattice rows (bits) of the walker's slot
  const uint32_t* Jh;     // right-bond signs per row
  const uint32_t* Jv;     // down-bond signs per row
  const int8_t* mask;     // spectrum flags over [E_min, E_max]
  double* G;              // walker's ln g column (stride W)
  uint32_t* H;            // walker's histogram column (stride W)
  int X, Y, N;
  int emin, erange;       // E_min, E_max - E_min
  int sk, ek, e0, shift, W, nbm1;
  bool upd;               // interval still active
  double lnf;
};

// One Wang-Landau attempt (src/wl.cu:1076-1174) on LDS state, branch-free
// except for the rare exact acceptance fallback: both LDS reads that depend
// on the proposed energy (spectrum flag, ln g at the new bin) are issued
// together with clamped indices, and every effect is a select or an LDS
// atomic whose operand is zero when the effect does not apply.
template<int BOUNDARY>
__device__ __forceinline__ void attempt(const sweep_ctx& c, sweep_state& s, uint32_t xs, uint32_t xa)
{
  // site = floor(xs N / 2^32) with N = X Y; its row and column come from the
  // 64-bit product xs X: row = high word, column = floor(low word * Y / 2^32)
  // (xs X Y = row Y 2^32 + low Y, so floor(xs N / 2^32) = row Y + column).
  const uint64_t px = static_cast<uint64_t>(xs) * static_cast<uint32_t>(c.X);
  const int i = static_cast<int>(px >> 32);
  const int j = static_cast<int>((static_cast<uint64_t>(static_cast<uint32_t>(px)) * static_cast<uint32_t>(c.Y)) >> 32);
  const int ipp = (i + 1 < c.X) ? i + 1 : 0;
  const int inn = (i > 0) ? i - 1 : c.X - 1;
  const int jpp = (j + 1 < c.Y) ? j + 1 : 0;
  const int jnn = (j > 0) ? j - 1 : c.Y - 1;
  const uint32_t rc = c.L[i], ru = c.L[inn], rd = c.L[ipp];
  const uint32_t jh = c.Jh[i], jvu = c.Jv[inn], jvd = c.Jv[i];
  const uint32_t sbit = (rc >> j) & 1u;
  // mismatch (s_nb J_nb = -1) per neighbour
  const uint32_t mu = ((ru ^ jvu) >> j) & 1u;
  const uint32_t md = ((rd ^ jvd) >> j) & 1u;
  const uint32_t ml = ((rc >> jnn) ^ (jh >> jnn)) & 1u;
  const uint32_t mr = ((rc >> jpp) ^ (jh >> j)) & 1u;
  int sum;
  if (BOUNDARY == 1) {
    const int cu = i > 0, cd = i < c.X - 1, cl = j > 0, cr = j < c.Y - 1;
    sum = cu + cd + cl + cr - 2 * static_cast<int>(cu * mu + cd * md + cl * ml + cr * mr);
  } else {
    sum = 4 - 2 * static_cast<int>(mu + md + ml + mr);
  }
  const int neg = -static_cast<int>(sbit);              // 0 or -1
  const int en = s.e + (((2 * sum) ^ neg) - neg);        // E + 2 s sum, s = +1/-1
  const int rel = en - c.emin;
  const bool in_spec_range = static_cast<uint32_t>(rel) <= static_cast<uint32_t>(c.erange);
  const bool in_range = (en >= c.sk) & (en <= c.ek);
  // clamped so that the speculative read stays inside the walker's column
  const int b_new = in_range ? min(max((en - c.e0) >> c.shift, 0), c.nbm1) : s.b;
  const int8_t flag = c.mask[in_spec_range ? rel : 0];
  const double g_new = c.G[__mul24(b_new, c.W)];
  const bool valid = s.alive & in_spec_range & (flag == 1);
  const bool acc = valid & in_range & accept_fast(s.g - g_new, xa);
  atomicXor(&c.L[i], acc ? (1u << j) : 0u);
  s.e = acc ? en : s.e;
  s.b = acc ? b_new : s.b;
  s.g = acc ? g_new : s.g;
  const bool upd = valid & c.upd;
  s.g = upd ? s.g + c.lnf : s.g;
  const int at = __mul24(s.b, c.W);
  c.G[at] = s.g;  // unchanged value when !upd
  atomicAdd(&c.H[at], upd ? 1u : 0u);
  s.bad_e = (s.alive & !valid) ? en : s.bad_e;
  s.alive = valid;
}

// n attempts of one walker starting at attempt counter att: two attempts
// per Philox block, the next block evaluated while the current pair waits
// on LDS.  Stops early (alive = false) at an energy outside the spectrum;
// att then counts the attempts up to and including that one.
template<int BOUNDARY>
__device__ __forceinline__ void sweep(const sweep_ctx& cx, sweep_state& ws, const attempt_stream& rng, uint64_t& att,
                                      int n)
{
  int it = 0;
  uint64_t blk = att >> 1;
  philox4 cur = rng.eval(blk);
  if ((att & 1u) && it < n) {  // odd start: second half of the block
    attempt<BOUNDARY>(cx, ws, cur.v[2], cur.v[3]);
    ++att;
    ++it;
    ++blk;
    cur = rng.eval(blk);
  }
  for (; it + 2 <= n && ws.alive; it += 2) {
    const philox4 nx = rng.eval(blk + 1);
    attempt<BOUNDARY>(cx, ws, cur.v[0], cur.v[1]);
    const bool first_ok = ws.alive;
    attempt<BOUNDARY>(cx, ws, cur.v[2], cur.v[3]);  // no effect once dead
    att += first_ok ? 2 : 1;  // the attempt that met the new energy counts, later ones do not
    ++blk;
    cur = nx;
  }
  if (it < n && ws.alive) {
    attempt<BOUNDARY>(cx, ws, cur.v[0], cur.v[1]);
    ++att;
  }
}

}  // namespace wlamd
