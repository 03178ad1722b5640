// McMurchie-Davidson Gaussian integral kernels, usable from host and device.
//
// The reference computes every AO integral with libcint (int1e_ovlp_sph,
// int1e_kin_sph, int1e_nuc_sph, int2c2e_sph, int3c2e_sph, int2e_sph; see
// src/ints/aofactory.cpp:380-430 and src/ints/integrals.cpp:250-370).  libcint
// is not available here, so the integrals are evaluated from scratch with the
// McMurchie-Davidson scheme (L. E. McMurchie, E. R. Davidson, J. Comput. Phys.
// 26, 218 (1978); notation of Helgaker, Jorgensen, Olsen, "Molecular
// Electronic-Structure Theory", Wiley 2000, sections 9.5 and 9.9):
//
//   E^{ij}_t   Hermite expansion coefficients of a Cartesian Gaussian product,
//   R^{n}_{tuv} Hermite Coulomb integrals from the Boys function,
//
// and transformed to real solid harmonics in libcint's order (p: x,y,z;
// l>=2: m=-l..l), normalised to one on the unit sphere.
//
// The Boys function F_m(T) uses a pretabulated grid with a 7th-order Taylor
// expansion about the nearest grid point followed by downward recursion
// (Helgaker et al. eq. 9.8.12-9.8.14); for T >= BOYS_TMAX it uses the
// asymptotic F_0 with upward recursion (eq. 9.8.9-9.8.10).
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MD_HD __host__ __device__ __forceinline__
#else
#define MD_HD inline
#endif

namespace mchip {
namespace md {

constexpr int LMAX = 3;                    // max angular momentum per shell (f)
constexpr int BOYS_MMAX = 4 * LMAX;        // max order needed for (ab|cd)
constexpr int BOYS_TAYLOR = 7;             // Taylor terms beyond F_m
constexpr int BOYS_NM = BOYS_MMAX + BOYS_TAYLOR + 1;
constexpr double BOYS_H = 0.05;            // grid spacing
constexpr double BOYS_TMAX = 40.0;         // table range [0, TMAX)
constexpr int BOYS_NT = 800;               // TMAX / H

constexpr double PI = 3.14159265358979323846264338327950288;

MD_HD int ncart(int l) { return (l + 1) * (l + 2) / 2; }
MD_HD int nsph(int l) { return 2 * l + 1; }

// Number of Hermite indices (t,u,v) with t+u+v <= L.
MD_HD constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

// Cartesian component k of shell l: exponents (i,j,k) with i+j+k=l, in the
// order x^l, x^{l-1}y, x^{l-1}z, x^{l-2}y^2, ... (i descending, then j).
MD_HD void cart_exps(int l, int k, int& ix, int& iy, int& iz)
{
  int n = 0;
  for (int a = l; a >= 0; --a) {
    for (int b = l - a; b >= 0; --b) {
      if (n == k) {
        ix = a;
        iy = b;
        iz = l - a - b;
        return;
      }
      ++n;
    }
  }
  ix = iy = iz = 0;
}

// Linear index of the Hermite triple (t,u,v), t+u+v <= L, ordered by total
// degree then (t desc, u desc); independent of L so tables can be shared.
MD_HD int herm_index(int t, int u, int v)
{
  int n = t + u + v;
  int base = n * (n + 1) * (n + 2) / 6;   // triples of degree < n
  // within degree n: t from n down to 0, u from n-t down to 0
  int r = n - t;                          // = u+v
  int off = r * (r + 1) / 2;              // entries with larger t
  return base + off + (r - u);
}

// ---------------------------------------------------------------------------
// Boys function
// ---------------------------------------------------------------------------

// Table layout: table[it * BOYS_NM + m] = F_m(it * BOYS_H).
// Computes F_0..F_mmax(T) into F.
MD_HD void boys(int mmax, double T, const double* table, double* F)
{
  if (T < BOYS_TMAX) {
    int it = (int)(T / BOYS_H + 0.5);
    if (it >= BOYS_NT) it = BOYS_NT - 1;
    double d = it * BOYS_H - T;           // expansion step (T0 - T)
    const double* row = table + it * BOYS_NM;
    // F_mmax(T) = sum_k F_{mmax+k}(T0) (T0-T)^k / k!
    double acc = 0.0;
    double fac = 1.0;
    double dk = 1.0;
    for (int k = 0; k <= BOYS_TAYLOR; ++k) {
      acc += row[mmax + k] * dk / fac;
      dk *= d;
      fac *= (double)(k + 1);
    }
    F[mmax] = acc;
    double e = exp(-T);
    for (int m = mmax - 1; m >= 0; --m) {
      F[m] = (2.0 * T * F[m + 1] + e) / (2.0 * m + 1.0);
    }
  }
  else {
    double e = exp(-T);
    F[0] = 0.5 * sqrt(PI / T);            // erf(sqrt(T)) = 1 to < 1e-18
    double inv2T = 0.5 / T;
    for (int m = 0; m < mmax; ++m) {
      F[m + 1] = ((2.0 * m + 1.0) * F[m] - e) * inv2T;
    }
  }
}

// ---------------------------------------------------------------------------
// Hermite expansion coefficients for one Cartesian direction
// ---------------------------------------------------------------------------
// E[i][j][t] for i<=la, j<=lb, t<=i+j, stored at E[(i*(LMAX+1)+j)*(2*LMAX+1)+t].
// XPA = P-A, XPB = P-B, oo2p = 1/(2p).  E^{00}_0 = 1 (the Gaussian product
// prefactor exp(-mu R_AB^2) is applied once for all three directions).
constexpr int ESTRIDE_T = 2 * LMAX + 1;
constexpr int ESIZE = (LMAX + 1) * (LMAX + 1) * ESTRIDE_T;

MD_HD int eidx(int i, int j, int t) { return (i * (LMAX + 1) + j) * ESTRIDE_T + t; }

MD_HD void hermite_E(int la, int lb, double XPA, double XPB, double oo2p, double* E)
{
  for (int n = 0; n < ESIZE; ++n) E[n] = 0.0;
  E[eidx(0, 0, 0)] = 1.0;
  // raise i with j = 0
  for (int i = 0; i < la; ++i) {
    for (int t = 0; t <= i + 1; ++t) {
      double v = XPA * E[eidx(i, 0, t)];
      if (t > 0) v += oo2p * E[eidx(i, 0, t - 1)];
      if (t + 1 <= i) v += (t + 1) * E[eidx(i, 0, t + 1)];
      E[eidx(i + 1, 0, t)] = v;
    }
  }
  // raise j for every i
  for (int i = 0; i <= la; ++i) {
    for (int j = 0; j < lb; ++j) {
      for (int t = 0; t <= i + j + 1; ++t) {
        double v = XPB * E[eidx(i, j, t)];
        if (t > 0) v += oo2p * E[eidx(i, j, t - 1)];
        if (t + 1 <= i + j) v += (t + 1) * E[eidx(i, j, t + 1)];
        E[eidx(i, j + 1, t)] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Hermite Coulomb integrals R^0_{tuv}(alpha, PC) for t+u+v <= L
// ---------------------------------------------------------------------------
// Output R[herm_index(t,u,v)].  Work space W must hold (L+1)*nherm(L) doubles.
// Recurrences (Helgaker eq. 9.9.18-9.9.20):
//   R^n_{000} = (-2 alpha)^n F_n(alpha |PC|^2)
//   R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X_PC R^{n+1}_{t,u,v}   (same for u,v)
MD_HD void hermite_R(int L, double alpha, double X, double Y, double Z,
                     const double* boys_table, double* W, double* R)
{
  double F[BOYS_MMAX + 1];
  double T = alpha * (X * X + Y * Y + Z * Z);
  boys(L, T, boys_table, F);
  const int NH = nherm(L);
  // W[n*NH + idx]: R^n for total degree <= L-n
  double m2a = -2.0 * alpha;
  double f = 1.0;
  for (int n = 0; n <= L; ++n) {
    W[n * NH + 0] = f * F[n];
    f *= m2a;
  }
  for (int deg = 1; deg <= L; ++deg) {
    for (int n = 0; n <= L - deg; ++n) {
      double* Rn = W + n * NH;
      const double* Rn1 = W + (n + 1) * NH;
      for (int t = deg; t >= 0; --t) {
        for (int u = deg - t; u >= 0; --u) {
          int v = deg - t - u;
          double val;
          if (t > 0) {
            val = X * Rn1[herm_index(t - 1, u, v)];
            if (t > 1) val += (t - 1) * Rn1[herm_index(t - 2, u, v)];
          }
          else if (u > 0) {
            val = Y * Rn1[herm_index(t, u - 1, v)];
            if (u > 1) val += (u - 1) * Rn1[herm_index(t, u - 2, v)];
          }
          else {
            val = Z * Rn1[herm_index(t, u, v - 1)];
            if (v > 1) val += (v - 1) * Rn1[herm_index(t, u, v - 2)];
          }
          Rn[herm_index(t, u, v)] = val;
        }
      }
    }
  }
  for (int i = 0; i < NH; ++i) R[i] = W[i];
}

// Same result as hermite_R with two level buffers of nherm(L) entries:
// level n (R^n for degree <= L-n) is built from level n+1, n = L..0.
// cur/nxt must each hold nherm(L) doubles; the result ends in R (== cur or nxt).
MD_HD const double* hermite_R2(int L, double alpha, double X, double Y, double Z,
                               const double* boys_table, double* cur, double* nxt)
{
  double F[BOYS_MMAX + 1];
  double T = alpha * (X * X + Y * Y + Z * Z);
  boys(L, T, boys_table, F);
  double m2a = -2.0 * alpha;
  double pw[BOYS_MMAX + 1];
  pw[0] = 1.0;
  for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * m2a;
  cur[0] = pw[L] * F[L];                 // level L: only R^L_000
  for (int n = L - 1; n >= 0; --n) {
    const int top = L - n;               // max degree at level n
    nxt[0] = pw[n] * F[n];
    for (int deg = 1; deg <= top; ++deg) {
      for (int t = deg; t >= 0; --t) {
        for (int u = deg - t; u >= 0; --u) {
          const int v = deg - t - u;
          double val;
          if (t > 0) {
            val = X * cur[herm_index(t - 1, u, v)];
            if (t > 1) val += (t - 1) * cur[herm_index(t - 2, u, v)];
          }
          else if (u > 0) {
            val = Y * cur[herm_index(t, u - 1, v)];
            if (u > 1) val += (u - 1) * cur[herm_index(t, u - 2, v)];
          }
          else {
            val = Z * cur[herm_index(t, u, v - 1)];
            if (v > 1) val += (v - 1) * cur[herm_index(t, u, v - 2)];
          }
          nxt[herm_index(t, u, v)] = val;
        }
      }
    }
    double* sw = cur;
    cur = nxt;
    nxt = sw;
  }
  return cur;
}

}  // namespace md
}  // namespace mchip
