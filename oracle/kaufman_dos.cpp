// ORACLE — test infrastructure only.
//
// Exact density of states g(E) of the ferromagnetic Ising model (every
// J = +1, E = -sum s s') on an X x Y torus with Y even, for lattices far
// beyond enumeration (16x16, 32x32), from Kaufman's partition function
// (B. Kaufman, Phys. Rev. 76, 1232 (1949)) -- the route by which Beale
// obtained the exact g(E) (P. D. Beale, Phys. Rev. Lett. 76, 78 (1996)).
// With m = X rows, n = Y columns and K the coupling:
//
//   Z = 1/2 (2 sinh 2K)^{mn/2} (Z1 + Z2 + Z3 + Z4),
//   Z1, Z2 = prod_{k=0}^{n-1} 2 cosh, 2 sinh (m gamma_{2k+1} / 2),
//   Z3, Z4 = prod_{k=0}^{n-1} 2 cosh, 2 sinh (m gamma_{2k} / 2),
//   cosh gamma_l = cosh 2K coth 2K - cos(pi l / n),  gamma_0 = 2K + ln tanh K.
//
// gamma_l = gamma_{2n-l}, so for n even the factors pair up (l odd in (0, n)
// for Z1, Z2; l even in (0, n) for Z3, Z4) and each pair is, with c = cosh
// gamma_l and z the root of z + 1/z = 2c of modulus >= 1,
//   (2 sinh 2K)^m (2 cosh(m gamma/2))^2 = (2 sinh 2K z)^m (1 + z^-m)^2
// (sinh: 1 - z^-m).  The two unpaired factors of Z3, Z4 (l = 0 and l = n),
// each with (2 sinh 2K)^{m/2}, are in v = e^{-2K}
//   ((1 - v)/v)^m +- (1 + v)^m   and   ((1 + v)/v)^m +- (1 - v)^m.
//
// Q(v) = v^N Z = sum_{j=0}^{2N} g(2j - 2N) v^j is a polynomial of degree 2N.
// Its coefficients come from discrete Fourier transforms of Q on circles
// |v| = rho (M >= 2N + 1 points, offset by half a step so that v = +-1, where
// sinh 2K = 0, is never met): each coefficient is read on the circle where
// its term is largest relative to the others (near its saddle point), which
// keeps its relative error near the rounding error of evaluating Q there
// (~1e-14 measured up to 16x4 and 10x6).  Coefficients below 1e9 are rounded
// to integers and must lie within 1e-3 of one, else the function reports
// failure.

#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#define WLO_API extern "C" __attribute__((visibility("default")))

namespace {

using cd = std::complex<double>;

// ln(a^m + sign b^m) without overflow
cd log_pow_sum(cd a, cd b, int m, double sign)
{
  if (std::abs(a) >= std::abs(b)) {
    return static_cast<double>(m) * std::log(a) + std::log(1.0 + sign * std::pow(b / a, m));
  }
  return static_cast<double>(m) * std::log(b) + std::log(sign + std::pow(a / b, m));
}

// ln Q(v) for the m x n torus (n even)
cd log_q(cd v, int m, int n, const std::vector<double>& cosines)
{
  const cd x = 1.0 / v;
  const cd C = 0.5 * (x + v);  // cosh 2K
  const cd s2 = x - v;         // 2 sinh 2K
  const cd s = 0.5 * s2;
  cd t[4] = {0.0, 0.0, 0.0, 0.0};
  const double dm = static_cast<double>(m);
  for (int l = 1; l < n; ++l) {
    const cd c = C * C / s - cosines[l];
    const cd r = std::sqrt(c * c - 1.0);
    cd z = c + r;
    if (std::abs(z) < 1.0) {
      z = c - r;
    }
    const cd zm = std::exp(-dm * std::log(z));
    const cd base = dm * std::log(s2 * z);
    const int a = (l & 1) ? 0 : 2;
    t[a] += base + 2.0 * std::log(1.0 + zm);
    t[a + 1] += base + 2.0 * std::log(1.0 - zm);
  }
  t[2] += log_pow_sum((1.0 - v) / v, 1.0 + v, m, 1.0) + log_pow_sum((1.0 + v) / v, 1.0 - v, m, 1.0);
  t[3] += log_pow_sum((1.0 - v) / v, 1.0 + v, m, -1.0) + log_pow_sum((1.0 + v) / v, 1.0 - v, m, -1.0);
  double top = t[0].real();
  for (int i = 1; i < 4; ++i) {
    top = std::max(top, t[i].real());
  }
  cd sum = 0.0;
  for (int i = 0; i < 4; ++i) {
    sum += std::exp(t[i] - top);
  }
  const double N = dm * n;
  return N * std::log(v) + top + std::log(0.5 * sum);
}

// in-place radix-2 DFT, F_j = sum_k q_k exp(-2 pi i j k / M)
void fft(std::vector<cd>& a)
{
  const size_t M = a.size();
  for (size_t i = 1, j = 0; i < M; ++i) {
    size_t bit = M >> 1;
    for (; j & bit; bit >>= 1) {
      j ^= bit;
    }
    j ^= bit;
    if (i < j) {
      std::swap(a[i], a[j]);
    }
  }
  for (size_t len = 2; len <= M; len <<= 1) {
    const double ang = -2.0 * M_PI / static_cast<double>(len);
    for (size_t i = 0; i < M; i += len) {
      for (size_t k = 0; k < len / 2; ++k) {
        const cd w = std::polar(1.0, ang * static_cast<double>(k));
        const cd u = a[i + k];
        const cd t = a[i + k + len / 2] * w;
        a[i + k] = u + t;
        a[i + k + len / 2] = u - t;
      }
    }
  }
}

}  // namespace

// lng[E + 2N] = ln g(E) for E in [-2N, 2N] (-inf where g = 0), for all J = +1
// on the X x Y torus.  Returns 0; -1 for an unsupported shape (Y odd, X < 3,
// Y < 4, N > 4096); -2 if a small coefficient is not close to an integer or
// an impossible energy's coefficient stands out of the rounding noise.
WLO_API int wlo_kaufman_dos(int X, int Y, double* lng)
{
  if (Y % 2 != 0 || X < 3 || Y < 4 || X * Y > 4096) {
    return -1;
  }
  const int m = X, n = Y, N = X * Y;
  std::vector<double> cosines(n);
  for (int l = 0; l < n; ++l) {
    cosines[l] = std::cos(M_PI * static_cast<double>(l) / n);
  }
  size_t M = 1;
  while (M < static_cast<size_t>(2 * N + 1)) {
    M <<= 1;
  }
  const double phi0 = M_PI / static_cast<double>(M);
  // radii from rho_min to 1/rho_min: a_0 = 2 dominates a_4 rho^4 = 2N rho^4 below N^(-1/4)
  const double lr_min = std::log(0.5 * std::pow(static_cast<double>(N), -0.25));
  const int R = static_cast<int>(std::ceil(-2.0 * lr_min / 0.05)) + 1;
  std::vector<double> best_dom(2 * N + 1, -1.0), best_val(2 * N + 1, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int ri = 0; ri < R; ++ri) {
    const double lr = lr_min + (-2.0 * lr_min) * ri / (R - 1);
    const double rho = std::exp(lr);
    std::vector<cd> lq(M);
    double top = -INFINITY;
    for (size_t k = 0; k < M; ++k) {
      const double ph = 2.0 * M_PI * static_cast<double>(k) / static_cast<double>(M) + phi0;
      lq[k] = log_q(std::polar(rho, ph), m, n, cosines);
      top = std::max(top, lq[k].real());
    }
    std::vector<cd> q(M);
    for (size_t k = 0; k < M; ++k) {
      q[k] = std::exp(lq[k] - top);
    }
    fft(q);
    double mx = 0.0;
    for (int j = 0; j <= 2 * N; ++j) {
      mx = std::max(mx, std::abs(q[j]));
    }
    for (int j = 0; j <= 2 * N; ++j) {
      const cd cj = q[j] * std::polar(1.0, -phi0 * j) / static_cast<double>(M);
      const double dom = std::abs(q[j]) / mx;
      const double lnv = (cj.real() > 0.0) ? std::log(cj.real()) + top - j * lr : -INFINITY;
#pragma omp critical
      if (dom > best_dom[j]) {
        best_dom[j] = dom;
        best_val[j] = lnv;
      }
    }
  }
  int rc = 0;
  for (int e = 0; e <= 4 * N; ++e) {
    lng[e] = -INFINITY;
  }
  // A flip changes E by 0, +-4 or +-8 on the 4-regular torus, so E = -2N mod
  // 4 and odd j never occur: their coefficients are rounding noise, and must
  // stay far below the dominance of any true coefficient (a self-check).
  // Even j with no states (E = -2N + 4, ...) are noise of the same size.
  constexpr double kNoise = 1e-8;
  for (int j = 0; j <= 2 * N; ++j) {
    if (j % 2 != 0 || best_dom[j] < kNoise) {
      if (j % 2 != 0 && best_dom[j] >= kNoise) {
        rc = -2;
      }
      continue;
    }
    double v = best_val[j];
    if (v < std::log(1e9)) {
      const double g = std::exp(v);
      const double r = std::nearbyint(g);
      if (g > 1e-3 && std::abs(g - r) > 1e-3) {
        rc = -2;
      }
      v = (r >= 1.0) ? std::log(r) : -INFINITY;
    }
    lng[2 * j] = v;
  }
  return rc;
}
