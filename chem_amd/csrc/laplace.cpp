// Minimax approximation of 1/x on [1, R] by an exponential sum
//     1/x ~ s(x) = sum_k w_k exp(-a_k x),  k = 1..n,
// the Laplace quadrature of SOS-AO-MP2 (the reference calls a Fortran
// Remez/Newton-Maehly code, src/math/laplace/src/lap_paraopt.F90 and
// lap_maehly.F90, through src/math/laplace/minimax.hpp:67-88; algorithm of
// D. Braess and W. Hackbusch, IMA J. Numer. Anal. 25, 685 (2005) and
// A. Takatsuka, S. Ten-no, W. Hackbusch, J. Chem. Phys. 129, 044112 (2008)).
//
// The error e(x) = 1/x - s(x) of the best approximation equioscillates at
// 2n+1 points 1 = x_0 < x_1 < ... < x_2n = R with e(x_j) = (-1)^j E.  Remez
// iteration: (1) with the alternation points fixed, solve the 2n+1 equations
// e(x_j) = (-1)^j E for (w, a, E) by damped Newton; (2) move the interior
// points to the extrema of the new error (Newton on e'(x) = 0).  Starting
// values come from tabulated solutions at nearby R; when the target R is far
// from the tabulated one the solution is continued geometrically in R.
// Arithmetic is long double.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "../../include/mchip.h"
#include "capi_common.hpp"

namespace mchip {
namespace {

using ld = long double;

struct ExpSum {
  int n;
  std::vector<ld> w, a;
  ld e(ld x) const
  {
    ld s = 0;
    for (int k = 0; k < n; ++k) s += w[k] * std::exp(-a[k] * x);
    return 1.0L / x - s;
  }
  ld de(ld x) const
  {
    ld s = 0;
    for (int k = 0; k < n; ++k) s += w[k] * a[k] * std::exp(-a[k] * x);
    return -1.0L / (x * x) + s;
  }
  ld d2e(ld x) const
  {
    ld s = 0;
    for (int k = 0; k < n; ++k) s += w[k] * a[k] * a[k] * std::exp(-a[k] * x);
    return 2.0L / (x * x * x) - s;
  }
};

// Solve A x = b (dense, partial pivoting), A is m x m row-major.
bool solve(std::vector<ld>& A, std::vector<ld>& b, int m)
{
  for (int k = 0; k < m; ++k) {
    int p = k;
    for (int i = k + 1; i < m; ++i)
      if (std::fabs(A[i * m + k]) > std::fabs(A[p * m + k])) p = i;
    if (A[p * m + k] == 0) return false;
    if (p != k) {
      for (int j = 0; j < m; ++j) std::swap(A[k * m + j], A[p * m + j]);
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < m; ++i) {
      const ld f = A[i * m + k] / A[k * m + k];
      for (int j = k; j < m; ++j) A[i * m + j] -= f * A[k * m + j];
      b[i] -= f * b[k];
    }
  }
  for (int i = m - 1; i >= 0; --i) {
    ld s = b[i];
    for (int j = i + 1; j < m; ++j) s -= A[i * m + j] * b[j];
    b[i] = s / A[i * m + i];
  }
  return true;
}

// Interior extrema of e on [1, R] by a log-spaced scan of e' sign changes.
std::vector<ld> scan_extrema(const ExpSum& s, ld R)
{
  std::vector<ld> xs;
  const int N = 20000;
  ld xp = 1.0L, dp = s.de(xp);
  for (int i = 1; i <= N; ++i) {
    const ld x = std::exp(std::log(R) * i / N);
    const ld d = s.de(x);
    if ((d > 0) != (dp > 0) && i < N) {
      // bisection + Newton polish inside [xp, x]
      ld lo = xp, hi = x;
      for (int it = 0; it < 80; ++it) {
        const ld mid = 0.5L * (lo + hi);
        if ((s.de(mid) > 0) == (dp > 0)) lo = mid; else hi = mid;
      }
      xs.push_back(0.5L * (lo + hi));
    }
    xp = x;
    dp = d;
  }
  return xs;
}

// One full Remez solve at fixed R from the given parameters.  Returns max |e|.
ld remez(ExpSum& s, ld R, int maxit, ld* equi_out)
{
  const int n = s.n, m = 2 * n + 1;
  std::vector<ld> xs = scan_extrema(s, R);
  if ((int)xs.size() != 2 * n - 1) throw std::runtime_error("minimax: alternation set not found");
  std::vector<ld> pts(m);
  ld E = 0;
  for (int it = 0; it < maxit; ++it) {
    pts[0] = 1.0L;
    for (int j = 0; j < 2 * n - 1; ++j) pts[j + 1] = xs[j];
    pts[m - 1] = R;
    // current levelled error estimate
    E = 0;
    for (int j = 0; j < m; ++j) E += ((j & 1) ? -1 : 1) * s.e(pts[j]);
    E /= m;
    // Newton on e(x_j) - (-1)^j E = 0 for (w, a, E)
    for (int nt = 0; nt < 30; ++nt) {
      std::vector<ld> A((size_t)m * m), r(m);
      ld rn = 0;
      for (int j = 0; j < m; ++j) {
        const ld x = pts[j], sg = (j & 1) ? -1 : 1;
        for (int k = 0; k < n; ++k) {
          const ld ex = std::exp(-s.a[k] * x);
          A[j * m + k] = -ex;                    // d e / d w_k
          A[j * m + n + k] = s.w[k] * x * ex;    // d e / d a_k
        }
        A[j * m + 2 * n] = -sg;                  // d/dE of -(sg E)
        r[j] = -(s.e(x) - sg * E);
        rn = std::max(rn, std::fabs(r[j]));
      }
      if (!solve(A, r, m)) throw std::runtime_error("minimax: singular Newton system");
      // damped step keeping exponents and weights positive
      ld lam = 1.0L;
      for (int k = 0; k < n; ++k) {
        if (s.a[k] + lam * r[n + k] <= 0) lam = std::min(lam, 0.5L * s.a[k] / std::fabs(r[n + k]));
        if (s.w[k] + lam * r[k] <= 0) lam = std::min(lam, 0.5L * s.w[k] / std::fabs(r[k]));
      }
      for (int k = 0; k < n; ++k) {
        s.w[k] += lam * r[k];
        s.a[k] += lam * r[n + k];
      }
      E += lam * r[2 * n];
      if (rn < 1e-17L * std::max<ld>(std::fabs(E), 1e-30L)) break;
    }
    // move interior points to the extrema of the new error
    std::vector<ld> nx = scan_extrema(s, R);
    if ((int)nx.size() != 2 * n - 1) throw std::runtime_error("minimax: alternation lost");
    ld mx = 0, mn = 1e300L;
    for (ld x : nx) {
      const ld v = std::fabs(s.e(x));
      mx = std::max(mx, v);
      mn = std::min(mn, v);
    }
    mx = std::max(mx, std::max(std::fabs(s.e(1.0L)), std::fabs(s.e(R))));
    mn = std::min(mn, std::min(std::fabs(s.e(1.0L)), std::fabs(s.e(R))));
    xs = nx;
    if (equi_out) *equi_out = (mx - mn) / mx;
    if ((mx - mn) <= 1e-9L * mx) return mx;
  }
  ld mx = 0;
  for (ld x : xs) mx = std::max(mx, std::fabs(s.e(x)));
  return mx;
}

}  // namespace
}  // namespace mchip

using namespace mchip;

extern "C" int mchip_laplace_minimax(int n, double R, double R0, const double* w0,
                                     const double* a0, int maxit, double* w, double* a,
                                     double* err, double* equi)
{
  MCHIP_TRY
  if (n < 1 || n > 40) throw std::runtime_error("minimax: number of points out of range");
  if (!(R > 1.0) || !(R0 > 1.0)) throw std::runtime_error("minimax: range must exceed 1");
  ExpSum s;
  s.n = n;
  s.w.assign(w0, w0 + n);
  s.a.assign(a0, a0 + n);
  // geometric continuation from R0 to R
  const ld ratio = std::log((ld)R / (ld)R0);
  const int nsteps = std::max(1, (int)std::ceil(std::fabs(ratio) / std::log(1.25L)));
  ld e = 0, eq = 0;
  for (int st = 1; st <= nsteps; ++st) {
    const ld Rs = (ld)R0 * std::exp(ratio * st / nsteps);
    e = remez(s, Rs, maxit, &eq);
  }
  for (int k = 0; k < n; ++k) {
    w[k] = (double)s.w[k];
    a[k] = (double)s.a[k];
  }
  *err = (double)e;
  if (equi) *equi = (double)eq;
  MCHIP_CATCH
}
