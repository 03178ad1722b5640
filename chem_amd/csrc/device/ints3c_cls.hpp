// Three-centre integral kernels of one angular momentum class (la, lb, lc)
// for the GPU build of B(X|mu nu) (driver: ints_dev.hip).
//
// McMurchie-Davidson, as the host engine (ints/md.hpp, ints/ints_host.cpp),
// with the angular momenta fixed at compile time: one thread per (AO shell
// pair, auxiliary shell of degree lc), every recursion (Hermite E, Boys,
// Hermite R) and contraction unrolled with constant indices so that its arrays
// live in registers.  The auxiliary side uses that a real solid harmonic Y_lm
// is a harmonic polynomial: Y(r_C) exp(-g r_C^2) = (2g)^-l Y(d/dC)
// exp(-g r_C^2) (Hobson's theorem), so its Hermite expansion holds only the
// degree-l terms, with the solid harmonic's integer monomial weights.
#pragma once

#include <cstdint>

#include "../ints/md.hpp"
#include "dl_common.hpp"
#include "ints_dev.hpp"

namespace mchip {

using namespace md;

constexpr int AO_L = 2;   // max AO angular momentum on the GPU path

// Arguments of one class launch (one kernel argument, read through the
// kernel-argument segment).  ax_s[c] = (-1)^l coef (2 gamma)^-l folds the
// auxiliary contraction coefficient and the Hermite scale of the top degree;
// sinv[l][m] normalises solid harmonic m of degree l (its coefficient of
// monomial k is shp(l, m, k) * sinv[l][m]).
struct I3cArgs {
  const int* ao_np;
  const int* ao_po;
  const double* ao_O;
  const double* ao_ex;
  const double* ao_cf;
  const int* shl;          // auxiliary shells of this degree
  int nshl, nchunk;        // their number, blocks per task (64 shells each)
  const double* ax_g;
  const double* ax_O;
  const double* ax_s;
  const int* ax_x0;
  const int* ax_blk;
  const Task3* tasks;      // the class's (AO shell pair) tasks
  const int64_t* cp_off;
  const int* bsz;
  const uint8_t* mask;
  int nbb;
  const double* boys;
  double* tmp;             // canonical pair store [x][mu][nu]
  double sinv[LMAX + 1][2 * LMAX + 1];
};

namespace i3c {

constexpr int nh(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
constexpr int nc(int l) { return (l + 1) * (l + 2) / 2; }
// Hermite index of (t, u, v): by total degree, then t descending, u descending
constexpr int hix(int t, int u, int v)
{
  const int n = t + u + v, r = n - t;
  return n * (n + 1) * (n + 2) / 6 + r * (r + 1) / 2 + (r - u);
}
constexpr int hdeg(int i)
{
  int d = 0;
  while (nh(d) <= i) ++d;
  return d;
}
constexpr int hr(int i)   // u + v of Hermite index i
{
  const int o = i - nh(hdeg(i) - 1);
  int r = 0;
  while ((r + 1) * (r + 2) / 2 <= o) ++r;
  return r;
}
constexpr int ht(int i) { return hdeg(i) - hr(i); }
constexpr int hu(int i) { return hr(i) - (i - nh(hdeg(i) - 1) - hr(i) * (hr(i) + 1) / 2); }
constexpr int hv(int i) { return hdeg(i) - ht(i) - hu(i); }
// exponents of Cartesian component k of degree l (x^l, x^{l-1}y, x^{l-1}z, ...)
constexpr int cx(int l, int k)
{
  int n = 0;
  for (int a = l; a >= 0; --a)
    for (int b = l - a; b >= 0; --b) {
      if (n == k) return a;
      ++n;
    }
  return 0;
}
constexpr int cy(int l, int k)
{
  int n = 0;
  for (int a = l; a >= 0; --a)
    for (int b = l - a; b >= 0; --b) {
      if (n == k) return b;
      ++n;
    }
  return 0;
}
constexpr int cz(int l, int k) { return l - cx(l, k) - cy(l, k); }
// Integer coefficient of monomial k in the unnormalised real solid harmonic m
// of degree l, libcint's order (p: x, y, z; l >= 2: m = -l..l); the same
// shapes as the host table (ints_host.cpp, make_sph_table).
constexpr int shp(int l, int m, int k)
{
  const int x = cx(l, k), y = cy(l, k), z = cz(l, k);
  if (l == 0) return 1;
  if (l == 1) return (m == 0 && x == 1) || (m == 1 && y == 1) || (m == 2 && z == 1) ? 1 : 0;
  if (l == 2) {
    if (m == 0) return x == 1 && y == 1 ? 1 : 0;                   // xy
    if (m == 1) return y == 1 && z == 1 ? 1 : 0;                   // yz
    if (m == 2) return z == 2 ? 2 : (x == 2 || y == 2) ? -1 : 0;   // 2z2 - x2 - y2
    if (m == 3) return x == 1 && z == 1 ? 1 : 0;                   // xz
    return x == 2 ? 1 : y == 2 ? -1 : 0;                           // x2 - y2
  }
  if (m == 0) return x == 2 && y == 1 ? 3 : y == 3 ? -1 : 0;                             // y(3x2 - y2)
  if (m == 1) return x == 1 && y == 1 && z == 1 ? 1 : 0;                                 // xyz
  if (m == 2) return y == 1 && z == 2 ? 4 : (x == 2 && y == 1) || y == 3 ? -1 : 0;       // y(4z2 - x2 - y2)
  if (m == 3) return z == 3 ? 2 : (x == 2 && z == 1) || (y == 2 && z == 1) ? -3 : 0;     // z(2z2 - 3x2 - 3y2)
  if (m == 4) return x == 1 && z == 2 ? 4 : x == 3 || (x == 1 && y == 2) ? -1 : 0;       // x(4z2 - x2 - y2)
  if (m == 5) return x == 2 && z == 1 ? 1 : y == 2 && z == 1 ? -1 : 0;                  // z(x2 - y2)
  return x == 3 ? 1 : x == 1 && y == 2 ? -3 : 0;                                         // x(x2 - 3y2)
}
// auxiliary components per accumulator group: the group's Cartesian
// accumulators (group x ncart(la) x ncart(lb) doubles) stay within 24
constexpr int mgroup(int ncab, int nm) { return 24 / ncab < 1 ? 1 : (24 / ncab > nm ? nm : 24 / ncab); }

// Compile-time checks of the index helpers: the Hermite decode inverts the
// index for every triple up to degree 8, and the Cartesian exponents of
// every component sum to its degree and run in the order of md::cart_exps
// (x descending, then y descending); each solid harmonic's integer shape
// has at least one monomial and no monomial outside the degree.
constexpr bool hermite_index_ok()
{
  for (int i = 0; i < nh(8); ++i) {
    if (hix(ht(i), hu(i), hv(i)) != i) return false;
    if (ht(i) < 0 || hu(i) < 0 || hv(i) < 0 || ht(i) + hu(i) + hv(i) != hdeg(i)) return false;
  }
  return true;
}
constexpr bool cartesian_order_ok()
{
  for (int l = 0; l <= 4; ++l)
    for (int k = 0; k + 1 < nc(l); ++k) {
      if (cx(l, k) + cy(l, k) + cz(l, k) != l || cz(l, k) < 0) return false;
      if (cx(l, k) < cx(l, k + 1) || (cx(l, k) == cx(l, k + 1) && cy(l, k) <= cy(l, k + 1))) return false;
    }
  return true;
}
constexpr bool shapes_ok()
{
  for (int l = 0; l <= 3; ++l)
    for (int m = 0; m < 2 * l + 1; ++m) {
      int nz = 0;
      for (int k = 0; k < nc(l); ++k) nz += shp(l, m, k) != 0;
      if (nz == 0) return false;
    }
  return true;
}
static_assert(hermite_index_ok(), "Hermite index decode");
static_assert(cartesian_order_ok(), "Cartesian component order");
static_assert(shapes_ok(), "solid harmonic shapes");

}  // namespace i3c

// E^{ij}_t of one direction for i <= LA, j <= LB (ints/md.hpp, hermite_E)
template <int LA, int LB>
__device__ __forceinline__ void herm_E_t(double XPA, double XPB, double oo2p, double (&E)[(LA + 1) * (LB + 1) * (LA + LB + 1)])
{
  constexpr int TD = LA + LB + 1;
  auto ei = [](int i, int j, int t) { return (i * (LB + 1) + j) * TD + t; };
  E[0] = 1.0;
  static_for<LA>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    static_for<i + 2>([&](auto tt) {
      constexpr int t = decltype(tt)::value;
      double v = 0.0;
      if constexpr (t <= i) v = XPA * E[ei(i, 0, t)];
      if constexpr (t > 0) v += oo2p * E[ei(i, 0, t - 1)];
      if constexpr (t + 1 <= i) v += (t + 1) * E[ei(i, 0, t + 1)];
      E[ei(i + 1, 0, t)] = v;
    });
  });
  static_for<LA + 1>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    static_for<LB>([&](auto jj) {
      constexpr int j = decltype(jj)::value;
      static_for<i + j + 2>([&](auto tt) {
        constexpr int t = decltype(tt)::value;
        double v = 0.0;
        if constexpr (t <= i + j) v = XPB * E[ei(i, j, t)];
        if constexpr (t > 0) v += oo2p * E[ei(i, j, t - 1)];
        if constexpr (t + 1 <= i + j) v += (t + 1) * E[ei(i, j, t + 1)];
        E[ei(i, j + 1, t)] = v;
      });
    });
  });
}

// F_0..F_L(T) (ints/md.hpp, boys): Taylor expansion about the nearest grid
// point, then downward recursion; asymptotic form with upward recursion
// beyond the table
template <int L>
__device__ __forceinline__ void boys_t(double T, const double* __restrict__ tab, double (&F)[L + 1])
{
  if (T < BOYS_TMAX) {
    int it = (int)(T / BOYS_H + 0.5);
    if (it >= BOYS_NT) it = BOYS_NT - 1;
    const double d = it * BOYS_H - T;
    const double* row = tab + it * BOYS_NM + L;
    constexpr double inv_fact[8] = {1.0, 1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040};
    static_assert(BOYS_TAYLOR == 7, "Taylor terms");
    double acc = 0.0, dk = 1.0;
    static_for<BOYS_TAYLOR + 1>([&](auto kk) {
      constexpr int k = decltype(kk)::value;
      acc += row[k] * (dk * inv_fact[k]);
      dk *= d;
    });
    F[L] = acc;
    const double e = exp(-T);
    static_for<L>([&](auto qq) {
      constexpr int m = L - 1 - decltype(qq)::value;
      F[m] = (2.0 * T * F[m + 1] + e) * (1.0 / (2 * m + 1));
    });
  }
  else {
    const double e = exp(-T);
    F[0] = 0.5 * sqrt(PI / T);
    const double inv2T = 0.5 / T;
    static_for<L>([&](auto mm) {
      constexpr int m = decltype(mm)::value;
      F[m + 1] = ((2.0 * m + 1.0) * F[m] - e) * inv2T;
    });
  }
}

// R^0_{tuv}(alpha, PC) for t + u + v <= L in place (ints/md.hpp, hermite_R2):
// level n (degree <= L - n) from level n + 1, degrees descending, so that a
// value is overwritten only after the last read of it.
template <int L>
__device__ __forceinline__ void herm_R_t(double alpha, double X, double Y, double Z, const double* __restrict__ tab,
                                         double (&R)[i3c::nh(L)])
{
  using namespace i3c;
  double F[L + 1];
  boys_t<L>(alpha * (X * X + Y * Y + Z * Z), tab, F);
  const double m2a = -2.0 * alpha;
  double pw[L + 1];
  pw[0] = 1.0;
  static_for<L>([&](auto nn) { pw[decltype(nn)::value + 1] = pw[decltype(nn)::value] * m2a; });
  R[0] = pw[L] * F[L];
  static_for<L>([&](auto qq) {
    constexpr int n = L - 1 - decltype(qq)::value;
    constexpr int top = L - n;
    static_for<top>([&](auto dq) {
      constexpr int deg = top - decltype(dq)::value;
      static_for<nh(deg) - nh(deg - 1)>([&](auto ee) {
        constexpr int i = nh(deg - 1) + decltype(ee)::value;
        constexpr int t = ht(i), u = hu(i), v = hv(i);
        double val;
        if constexpr (t > 0) {
          val = X * R[hix(t - 1, u, v)];
          if constexpr (t > 1) val += (t - 1) * R[hix(t - 2, u, v)];
        }
        else if constexpr (u > 0) {
          val = Y * R[hix(t, u - 1, v)];
          if constexpr (u > 1) val += (u - 1) * R[hix(t, u - 2, v)];
        }
        else {
          val = Z * R[hix(t, u, v - 1)];
          if constexpr (v > 1) val += (v - 1) * R[hix(t, u, v - 2)];
        }
        R[i] = val;
      });
    });
    R[0] = pw[n] * F[n];
  });
}

// One class (LA, LB, LC): block = (task, chunk of 64 auxiliary shells of
// degree LC from shl).  ax_s[c] = (-1)^l coef (2 gamma)^-l folds the
// auxiliary contraction coefficient and the Hermite scale of the top degree.
template <int LA, int LB, int LC>
__global__ __launch_bounds__(64) void k_int3c2e_cls(const I3cArgs A)
{
  using namespace i3c;
  constexpr int LAB = LA + LB, L = LAB + LC;
  constexpr int NCA = nc(LA), NCB = nc(LB), NCC = nc(LC), NM = 2 * LC + 1;
  constexpr int MG = mgroup(NCA * NCB, NM), NG = (NM + MG - 1) / MG;
  constexpr int TD = LAB + 1, NE = (LA + 1) * (LB + 1) * TD;
  const int* __restrict__ ao_np = A.ao_np;
  const int* __restrict__ ao_po = A.ao_po;
  const double* __restrict__ ao_O = A.ao_O;
  const double* __restrict__ ao_ex = A.ao_ex;
  const double* __restrict__ ao_cf = A.ao_cf;
  const double* __restrict__ ax_O = A.ax_O;
  const double* __restrict__ boys_tab = A.boys;
  const int task = blockIdx.x / A.nchunk;
  const int ci = (blockIdx.x % A.nchunk) * 64 + threadIdx.x;
  if (ci >= A.nshl) return;
  const int c = A.shl[ci];
  const Task3 tk = A.tasks[task];
  if (!A.mask[((size_t)A.ax_blk[c] * A.nbb + tk.M) * A.nbb + tk.N]) return;
  const int a = tk.sa, b = tk.sb;
  const double Ax = ao_O[3 * a], Ay = ao_O[3 * a + 1], Az = ao_O[3 * a + 2];
  const double Bx = ao_O[3 * b], By = ao_O[3 * b + 1], Bz = ao_O[3 * b + 2];
  const double Cx = ax_O[3 * c], Cy = ax_O[3 * c + 1], Cz = ax_O[3 * c + 2];
  const double q = A.ax_g[c], sc = A.ax_s[c];
  const double AB2 = (Ax - Bx) * (Ax - Bx) + (Ay - By) * (Ay - By) + (Az - Bz) * (Az - Bz);
  const int pa0 = ao_po[a], pb0 = ao_po[b], npa = ao_np[a], npb = ao_np[b];
  const int BM = A.bsz[tk.M], BN = A.bsz[tk.N];
  double* base = A.tmp + A.cp_off[tk.cp];
  const int x0 = A.ax_x0[c];
  auto ei = [](int i, int j, int t) { return (i * (LB + 1) + j) * TD + t; };
  static_for<NG>([&](auto gg) {
    constexpr int m0 = decltype(gg)::value * MG;
    constexpr int MN = NM - m0 < MG ? NM - m0 : MG;
    double acc[MN][NCA][NCB];
    static_for<MN>([&](auto mm) {
      static_for<NCA>([&](auto ka) {
        static_for<NCB>([&](auto kb) { acc[mm][ka][kb] = 0.0; });
      });
    });
    for (int i = 0; i < npa; ++i) {
      const double al = ao_ex[pa0 + i], ca = ao_cf[pa0 + i];
      for (int j = 0; j < npb; ++j) {
        const double be = ao_ex[pb0 + j];
        const double p = al + be, mu = al * be / p;
        if (mu * AB2 > 41.4) continue;   // exp(-mu R^2) < 1e-18
        const double K = ca * ao_cf[pb0 + j] * exp(-mu * AB2);
        const double Px = (al * Ax + be * Bx) / p, Py = (al * Ay + be * By) / p, Pz = (al * Az + be * Bz) / p;
        const double oo2p = 0.5 / p;
        double Ex[NE], Ey[NE], Ez[NE];
        herm_E_t<LA, LB>(Px - Ax, Px - Bx, oo2p, Ex);
        herm_E_t<LA, LB>(Py - Ay, Py - By, oo2p, Ey);
        herm_E_t<LA, LB>(Pz - Az, Pz - Bz, oo2p, Ez);
        const double alpha = p * q / (p + q);
        double R[nh(L)];
        herm_R_t<L>(alpha, Px - Cx, Py - Cy, Pz - Cz, boys_tab, R);
        const double pref = 2.0 * PI * PI * sqrt(PI) / (p * q * sqrt(p + q)) * K * sc;
        static_for<MN>([&](auto mm) {
          constexpr int m = m0 + decltype(mm)::value;
          // S[tuv] = sum over the solid harmonic's monomials of R[tuv + k]
          double S[nh(LAB)];
          static_for<nh(LAB)>([&](auto ii) {
            constexpr int t = ht(ii), u = hu(ii), v = hv(ii);
            double x = 0.0;
            static_for<NCC>([&](auto kk) {
              constexpr int w = shp(LC, m, kk);
              if constexpr (w != 0)
                x += (double)w * R[hix(t + cx(LC, kk), u + cy(LC, kk), v + cz(LC, kk))];
            });
            S[ii] = x;
          });
          const double f = pref * A.sinv[LC][m];
          static_for<NCA>([&](auto ka) {
            constexpr int ax = cx(LA, ka), ay = cy(LA, ka), az = cz(LA, ka);
            static_for<NCB>([&](auto kb) {
              constexpr int bx = cx(LB, kb), by = cy(LB, kb), bz = cz(LB, kb);
              double r = 0.0;
              static_for<ax + bx + 1>([&](auto t) {
                static_for<ay + by + 1>([&](auto u) {
                  const double exy = Ex[ei(ax, bx, t)] * Ey[ei(ay, by, u)];
                  static_for<az + bz + 1>([&](auto v) {
                    r += exy * Ez[ei(az, bz, v)] * S[hix(t, u, v)];
                  });
                });
              });
              acc[mm][ka][kb] += f * r;
            });
          });
        });
      }
    }
    // spherical transform and store into the canonical pair block [x][mu][nu]
    static_for<MN>([&](auto mm) {
      constexpr int m = m0 + decltype(mm)::value;
      double* dst = base + (int64_t)(x0 + m) * BM * BN;
      static_for<2 * LA + 1>([&](auto sa) {
        static_for<2 * LB + 1>([&](auto sb) {
          double x = 0.0;
          static_for<NCA>([&](auto ka) {
            constexpr int wa = shp(LA, sa, ka);
            if constexpr (wa != 0) {
              double y = 0.0;
              static_for<NCB>([&](auto kb) {
                constexpr int wb = shp(LB, sb, kb);
                if constexpr (wb != 0) y += (double)wb * acc[mm][ka][kb];
              });
              x += (double)wa * y;
            }
          });
          x *= A.sinv[LA][sa] * A.sinv[LB][sb];
          dst[(tk.mu0 + sa) * BN + (tk.nu0 + sb)] = x;
          if (tk.M == tk.N) dst[(tk.nu0 + sb) * BN + (tk.mu0 + sa)] = x;
        });
      });
    });
  });
}


// Normalisations of the real solid harmonics (host table, checked against
// the integer shapes).
void solid_harmonic_norms(double (&inv)[LMAX + 1][2 * LMAX + 1]);

// One launch of class (LA, lb, lc), instantiated per LA in its own file.
void launch_int3c2e_la0(int lb, int lc, dim3 grid, hipStream_t st, const I3cArgs& a);
void launch_int3c2e_la1(int lb, int lc, dim3 grid, hipStream_t st, const I3cArgs& a);
void launch_int3c2e_la2(int lb, int lc, dim3 grid, hipStream_t st, const I3cArgs& a);

template <int LA>
void launch_int3c2e_la(int lb, int lc, dim3 grid, hipStream_t st, const I3cArgs& a)
{
#define MCHIP_I3C(LB, LC) \
  case LB * 4 + LC: hipLaunchKernelGGL((k_int3c2e_cls<LA, LB, LC>), grid, dim3(64), 0, st, a); return;
  static_assert(AO_L == 2 && LMAX == 3, "class table covers AO shells up to d, auxiliary shells up to f");
  switch (lb * 4 + lc) {
    MCHIP_I3C(0, 0) MCHIP_I3C(0, 1) MCHIP_I3C(0, 2) MCHIP_I3C(0, 3)
    MCHIP_I3C(1, 0) MCHIP_I3C(1, 1) MCHIP_I3C(1, 2) MCHIP_I3C(1, 3)
    MCHIP_I3C(2, 0) MCHIP_I3C(2, 1) MCHIP_I3C(2, 2) MCHIP_I3C(2, 3)
    default: throw std::runtime_error("GPU integrals: angular momentum class out of range");
  }
#undef MCHIP_I3C
}

}  // namespace mchip
