// coulomb_table.cpp -- the real-space Coulomb table of the fast pair form.
//
// g(s) = erfc(alpha sqrt(s)) / sqrt(s), s = r^2, on [2^kCoulTabELo, 2^e_hi)
// with 2^e_hi > cut_coulomb^2.  Each octave [2^e, 2^(e+1)) is split into
// kCoulTabM intervals of width 2^e / kCoulTabM, so the interval of s is read
// from the top bits of s as an FP32 number (exponent and the first
// log2(kCoulTabM) mantissa bits).  On each interval g is the degree-kCoulTabDeg
// polynomial in d = s - s_i that interpolates g at the Chebyshev nodes,
// computed in long double; the coefficients are stored in FP64.
// Pairs with r^2 < 0.01 inside the VDW cutoff are overlaps, whose energies
// are never used (VDW_Coulomb.cu:1299-1300), so the table starts just below.
#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.hpp"

namespace graspa {

int coulomb_table_octaves(double cut2)
{
  // 2^e_hi above cut2 (1 + 2^-20): r^2 < cut2 rounded to FP32 (relative
  // error 2^-24) stays inside the last octave, so lookups need no upper clamp
  const int e_hi = (int)std::floor(std::log2(cut2 * (1.0 + 0x1p-20))) + 1;
  return std::max(1, e_hi - kCoulTabELo);
}

static long double g_exact(long double alpha, long double s)
{
  const long double r = std::sqrt(s);
  return std::erfc(alpha * r) / r;
}

void build_coulomb_table(double alpha, double cut2, std::vector<double>& tab)
{
  constexpr int N = kCoulTabDeg + 1;
  const int n_oct = coulomb_table_octaves(cut2);
  tab.assign((size_t)n_oct * kCoulTabM * kCoulTabStride, 0.0);
  const long double pi = 3.141592653589793238462643383279502884L;
  long double u[N];
  for (int k = 0; k < N; ++k) u[k] = 0.5L * (1.0L + std::cos((2.0L * k + 1.0L) * pi / (2.0L * N)));
  for (int o = 0; o < n_oct; ++o) {
    const long double base = std::ldexp(1.0L, kCoulTabELo + o);
    const long double h = base / kCoulTabM;
    for (int j = 0; j < kCoulTabM; ++j) {
      const long double s0 = base + j * h;
      // Vandermonde system in u = d / h on [0, 1], Gaussian elimination with pivoting
      long double A[N][N + 1];
      for (int r = 0; r < N; ++r) {
        long double p = 1.0L;
        for (int c = 0; c < N; ++c) {
          A[r][c] = p;
          p *= u[r];
        }
        A[r][N] = g_exact(alpha, s0 + u[r] * h);
      }
      for (int c = 0; c < N; ++c) {
        int piv = c;
        for (int r = c + 1; r < N; ++r)
          if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
        for (int k = 0; k <= N; ++k) std::swap(A[c][k], A[piv][k]);
        for (int r = 0; r < N; ++r) {
          if (r == c) continue;
          const long double f = A[r][c] / A[c][c];
          for (int k = c; k <= N; ++k) A[r][k] -= f * A[c][k];
        }
      }
      double* t = &tab[((size_t)o * kCoulTabM + j) * kCoulTabStride];
      long double hk = 1.0L;
      for (int k = 0; k < N; ++k) {
        t[k] = (double)(A[k][N] / A[k][k] / hk);   // coefficient of d^k
        hk *= h;
      }
    }
  }
}

}  // namespace graspa

// Fills `out` with the table for (alpha, cut_coulomb^2); returns the number of
// doubles written, or the size needed when cap is too small.
extern "C" int graspa_coulomb_table(double alpha, double cut2, double* out, int cap)
{
  if (!(alpha > 0.0) || !(cut2 > 0.0)) return -1;
  std::vector<double> tab;
  graspa::build_coulomb_table(alpha, cut2, tab);
  if (!out || cap < (int)tab.size()) return (int)tab.size();
  for (size_t i = 0; i < tab.size(); ++i) out[i] = tab[i];
  return (int)tab.size();
}
