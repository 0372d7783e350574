// ORACLE — test infrastructure only.
//
// Exact density of states g(E) of a 2D random-bond Ising model with the
// reference's coupling layout (src/wl.cu:77-87: J[i*Y+j] couples (i,j)-(i,j+1),
// J[N+i*Y+j] couples (i,j)-(i+1,j); E = -sum J s s').
//
//  * wlo_exact_dos_bruteforce: enumerate all 2^N spin configurations (N <= 26).
//  * wlo_exact_dos_transfer: row-to-row transfer matrix with energy
//    polynomials.  For periodic rows the first row is held fixed and the wrap
//    bond is closed at the end, so the cost is 2^Y * X * 4^Y * (4N+1).
//
// The transfer matrix is pinned against brute force on small lattices in
// tests/test_oracle.py; it then serves as the exact ln g(E) for <= 6x6.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#define WLO_API extern "C" __attribute__((visibility("default")))

namespace {

using u128 = unsigned __int128;

int energy_of(const int* s, const int8_t* J, int X, int Y, int boundary)
{
  const int N = X * Y;
  int e = 0;
  for (int i = 0; i < X; ++i) {
    for (int j = 0; j < Y; ++j) {
      const int a = s[i * Y + j];
      if (boundary == 0 || j + 1 < Y) {
        e += a * s[i * Y + (j + 1) % Y] * J[i * Y + j];
      }
      if (boundary == 0 || i + 1 < X) {
        e += a * s[((i + 1) % X) * Y + j] * J[N + i * Y + j];
      }
    }
  }
  return -e;
}

inline int spin(unsigned r, int j)
{
  return ((r >> j) & 1u) ? -1 : 1;
}

int row_energy(unsigned r, const int8_t* Jh_row, int Y, int boundary)
{
  int e = 0;
  for (int j = 0; j < Y; ++j) {
    if (boundary == 0 || j + 1 < Y) {
      e += spin(r, j) * spin(r, (j + 1) % Y) * Jh_row[j];
    }
  }
  return -e;
}

int bond_energy(unsigned r, unsigned rn, const int8_t* Jv_row, int Y)
{
  int e = 0;
  for (int j = 0; j < Y; ++j) {
    e += spin(r, j) * spin(rn, j) * Jv_row[j];
  }
  return -e;
}

}  // namespace

// counts[E + 2N] for E in [-2N, 2N]; counts has 4N+1 entries.
WLO_API int wlo_exact_dos_bruteforce(const int8_t* J, int X, int Y, int boundary, uint64_t* counts)
{
  const int N = X * Y;
  if (N > 26) {
    return -1;
  }
  std::memset(counts, 0, sizeof(uint64_t) * (4 * N + 1));
  std::vector<int> s(N);
  for (uint64_t c = 0; c < (1ull << N); ++c) {
    for (int i = 0; i < N; ++i) {
      s[i] = ((c >> i) & 1u) ? -1 : 1;
    }
    counts[energy_of(s.data(), J, X, Y, boundary) + 2 * N] += 1;
  }
  return 0;
}

// Exact ln g(E) (lng[E + 2N], -inf where g = 0) and, when every count fits,
// the counts themselves (counts may be null).  Returns 0, or -1 if Y > 12.
WLO_API int wlo_exact_dos_transfer(const int8_t* J, int X, int Y, int boundary, double* lng, uint64_t* counts)
{
  if (Y > 12 || X < 2) {
    return -1;
  }
  const int N = X * Y;
  const int L = 4 * N + 1;
  const unsigned R = 1u << Y;
  std::vector<u128> total(L, 0);
  std::vector<int> eh(static_cast<size_t>(X) * R);
  for (int i = 0; i < X; ++i) {
    for (unsigned r = 0; r < R; ++r) {
      eh[i * R + r] = row_energy(r, J + i * Y, Y, boundary);
    }
  }
  std::vector<int> ev(static_cast<size_t>(X) * R * R);
  for (int i = 0; i < X; ++i) {
    for (unsigned r = 0; r < R; ++r) {
      for (unsigned q = 0; q < R; ++q) {
        ev[(static_cast<size_t>(i) * R + r) * R + q] = bond_energy(r, q, J + N + i * Y, Y);
      }
    }
  }
  const unsigned first_rows = (boundary == 0) ? R : 1u;
  std::vector<u128> cur(static_cast<size_t>(R) * L), nxt(static_cast<size_t>(R) * L);
  for (unsigned r0 = 0; r0 < first_rows; ++r0) {
    std::fill(cur.begin(), cur.end(), 0);
    if (boundary == 0) {
      cur[static_cast<size_t>(r0) * L + eh[r0] + 2 * N] = 1;
    } else {
      for (unsigned r = 0; r < R; ++r) {
        cur[static_cast<size_t>(r) * L + eh[r] + 2 * N] = 1;
      }
    }
    for (int i = 1; i < X; ++i) {
      std::fill(nxt.begin(), nxt.end(), 0);
      for (unsigned r = 0; r < R; ++r) {
        const u128* src = &cur[static_cast<size_t>(r) * L];
        int lo = L, hi = -1;
        for (int e = 0; e < L; ++e) {
          if (src[e] != 0) {
            lo = (e < lo) ? e : lo;
            hi = e;
          }
        }
        if (hi < 0) {
          continue;
        }
        for (unsigned q = 0; q < R; ++q) {
          const int de = ev[(static_cast<size_t>(i - 1) * R + r) * R + q] + eh[i * R + q];
          u128* dst = &nxt[static_cast<size_t>(q) * L];
          for (int e = lo; e <= hi; ++e) {
            dst[e + de] += src[e];
          }
        }
      }
      cur.swap(nxt);
    }
    for (unsigned r = 0; r < R; ++r) {
      const int de = (boundary == 0) ? ev[(static_cast<size_t>(X - 1) * R + r) * R + r0] : 0;
      const u128* src = &cur[static_cast<size_t>(r) * L];
      for (int e = 0; e < L; ++e) {
        if (src[e] != 0) {
          total[e + de] += src[e];
        }
      }
    }
  }
  for (int e = 0; e < L; ++e) {
    if (total[e] == 0) {
      lng[e] = -INFINITY;
    } else {
      lng[e] = static_cast<double>(std::log(static_cast<long double>(total[e])));
    }
    if (counts != nullptr) {
      counts[e] = (total[e] >> 64) ? ~0ull : static_cast<uint64_t>(total[e]);
    }
  }
  return 0;
}
