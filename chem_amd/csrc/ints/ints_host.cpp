// Host (OpenMP) evaluation of the AO integrals the DF Fock build consumes:
// overlap, kinetic, nuclear attraction, the Coulomb metric (X|Y), the
// three-centre integrals (X|mu nu) and the Schwarz factors.
//
// Reference counterparts: src/ints/aofactory.cpp (ao_overlap, ao_kinetic,
// ao_nuclear, ao_2c2e, ao_3c_fill, ao_schwarz, ao_3cschwarz) and the fill loops
// in src/ints/integrals.cpp:250-370 (3c) and 511-689 (Schwarz).
#include "ints_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>

#include "md.hpp"

namespace mchip {

using namespace md;

// ---------------------------------------------------------------------------
// normalisation (aofactory.cpp:273-310)
// ---------------------------------------------------------------------------
static double gaussian_int(int n, double a)
{
  double n1 = (n + 1) * 0.5;
  return std::tgamma(n1) / (2.0 * std::pow(a, n1));
}

void normalize_shell(Shell& s)
{
  const int l = s.l;
  const size_t np = s.alpha.size();
  std::vector<double> c(np);
  for (size_t i = 0; i < np; ++i)
    c[i] = s.coeff[i] / std::sqrt(gaussian_int(2 * l + 2, 2.0 * s.alpha[i]));
  double norm = 0.0;
  for (size_t p = 0; p < np; ++p)
    for (size_t q = 0; q < np; ++q)
      norm += c[p] * gaussian_int(2 * l + 2, s.alpha[p] + s.alpha[q]) * c[q];
  norm = 1.0 / std::sqrt(norm);
  for (size_t i = 0; i < np; ++i) s.coeff[i] = c[i] * norm;
}

// ---------------------------------------------------------------------------
// real solid harmonics, libcint order, unit norm on the sphere
// ---------------------------------------------------------------------------
namespace {

// integral of x^a y^b z^c over the unit sphere
double sphere_monomial(int a, int b, int c)
{
  if ((a & 1) || (b & 1) || (c & 1)) return 0.0;
  return 2.0 * std::tgamma((a + 1) * 0.5) * std::tgamma((b + 1) * 0.5) *
         std::tgamma((c + 1) * 0.5) / std::tgamma((a + b + c + 3) * 0.5);
}

int cart_index(int l, int ix, int iy, int iz)
{
  for (int k = 0; k < ncart(l); ++k) {
    int a, b, c;
    cart_exps(l, k, a, b, c);
    if (a == ix && b == iy && c == iz) return k;
  }
  return -1;
}

struct Term {
  double c;
  int x, y, z;
};

SphTable make_sph_table()
{
  SphTable t;
  std::memset(&t, 0, sizeof(t));
  // unnormalised shapes; signs and order as libcint (m = -l..l, p as x,y,z)
  std::vector<std::vector<std::vector<Term>>> shapes(LMAX + 1);
  shapes[0] = {{{1, 0, 0, 0}}};
  shapes[1] = {{{1, 1, 0, 0}}, {{1, 0, 1, 0}}, {{1, 0, 0, 1}}};
  shapes[2] = {
      {{1, 1, 1, 0}},                                  // xy
      {{1, 0, 1, 1}},                                  // yz
      {{2, 0, 0, 2}, {-1, 2, 0, 0}, {-1, 0, 2, 0}},    // 2z2 - x2 - y2
      {{1, 1, 0, 1}},                                  // xz
      {{1, 2, 0, 0}, {-1, 0, 2, 0}},                   // x2 - y2
  };
  shapes[3] = {
      {{3, 2, 1, 0}, {-1, 0, 3, 0}},                   // y(3x2 - y2)
      {{1, 1, 1, 1}},                                  // xyz
      {{4, 0, 1, 2}, {-1, 2, 1, 0}, {-1, 0, 3, 0}},    // y(4z2 - x2 - y2)
      {{2, 0, 0, 3}, {-3, 2, 0, 1}, {-3, 0, 2, 1}},    // z(2z2 - 3x2 - 3y2)
      {{4, 1, 0, 2}, {-1, 3, 0, 0}, {-1, 1, 2, 0}},    // x(4z2 - x2 - y2)
      {{1, 2, 0, 1}, {-1, 0, 2, 1}},                   // z(x2 - y2)
      {{1, 3, 0, 0}, {-3, 1, 2, 0}},                   // x(x2 - 3y2)
  };
  for (int l = 0; l <= LMAX; ++l) {
    for (int m = 0; m < nsph(l); ++m) {
      const auto& terms = shapes[l][m];
      double n2 = 0.0;
      for (const auto& a : terms)
        for (const auto& b : terms)
          n2 += a.c * b.c * sphere_monomial(a.x + b.x, a.y + b.y, a.z + b.z);
      double inv = 1.0 / std::sqrt(n2);
      for (const auto& a : terms) t.c[l][m][cart_index(l, a.x, a.y, a.z)] += a.c * inv;
    }
  }
  return t;
}

// F_m(T) by its convergent series F_m(T) = exp(-T) sum_k (2T)^k / prod_{j=0..k}(2m+2j+1)
long double boys_series(int m, long double T)
{
  long double term = 1.0L / (2 * m + 1);
  long double sum = term;
  for (int k = 1; k < 2000; ++k) {
    term *= 2.0L * T / (2 * m + 2 * k + 1);
    sum += term;
    if (term < sum * 1e-21L) break;
  }
  return std::exp(-T) * sum;
}

}  // namespace

const SphTable& sph_table()
{
  static SphTable t = make_sph_table();
  return t;
}

const std::vector<double>& boys_table()
{
  static std::vector<double> tab = [] {
    std::vector<double> v((size_t)BOYS_NT * BOYS_NM);
    for (int it = 0; it < BOYS_NT; ++it)
      for (int m = 0; m < BOYS_NM; ++m)
        v[(size_t)it * BOYS_NM + m] = (double)boys_series(m, (long double)it * BOYS_H);
    return v;
  }();
  return tab;
}

void boys_eval(int mmax, double T, double* F)
{
  boys(mmax, T, boys_table().data(), F);
}

// ---------------------------------------------------------------------------
// primitive pair data
// ---------------------------------------------------------------------------
namespace {

struct PrimPair {
  double p, P[3], K;            // K = c_a c_b exp(-mu R_AB^2)
  double Ex[ESIZE], Ey[ESIZE], Ez[ESIZE];
};

// Primitive pairs with exp(-mu R_AB^2) below e^-PRIM_CUT (about 1e-18) are
// skipped in the two-electron integrals.
constexpr double PRIM_CUT = 41.4;

void make_pairs(const Shell& a, const Shell& b, std::vector<PrimPair>& out)
{
  out.clear();
  double AB[3] = {a.O[0] - b.O[0], a.O[1] - b.O[1], a.O[2] - b.O[2]};
  double r2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
  for (size_t i = 0; i < a.alpha.size(); ++i) {
    for (size_t j = 0; j < b.alpha.size(); ++j) {
      PrimPair pp;
      double al = a.alpha[i], be = b.alpha[j];
      pp.p = al + be;
      double mu = al * be / pp.p;
      if (mu * r2 > PRIM_CUT) continue;
      pp.K = a.coeff[i] * b.coeff[j] * std::exp(-mu * r2);
      for (int d = 0; d < 3; ++d) pp.P[d] = (al * a.O[d] + be * b.O[d]) / pp.p;
      double oo2p = 0.5 / pp.p;
      hermite_E(a.l, b.l, pp.P[0] - a.O[0], pp.P[0] - b.O[0], oo2p, pp.Ex);
      hermite_E(a.l, b.l, pp.P[1] - a.O[1], pp.P[1] - b.O[1], oo2p, pp.Ey);
      hermite_E(a.l, b.l, pp.P[2] - a.O[2], pp.P[2] - b.O[2], oo2p, pp.Ez);
      out.push_back(pp);
    }
  }
}

// Spherical pair Hermite coefficients G[sa][sb][herm_index] of one primitive pair.
void pair_sph_hermite(int la, int lb, const PrimPair& pp, double* G)
{
  const SphTable& S = sph_table();
  const int L = la + lb;
  const int NH = nherm(L);
  const int na = nsph(la), nb = nsph(lb);
  std::fill(G, G + na * nb * NH, 0.0);
  for (int ka = 0; ka < ncart(la); ++ka) {
    int ax, ay, az;
    cart_exps(la, ka, ax, ay, az);
    for (int kb = 0; kb < ncart(lb); ++kb) {
      int bx, by, bz;
      cart_exps(lb, kb, bx, by, bz);
      double tmp[nherm(2 * LMAX)];
      for (int h = 0; h < NH; ++h) tmp[h] = 0.0;
      for (int t = 0; t <= ax + bx; ++t) {
        double ex = pp.Ex[eidx(ax, bx, t)];
        for (int u = 0; u <= ay + by; ++u) {
          double exy = ex * pp.Ey[eidx(ay, by, u)];
          for (int v = 0; v <= az + bz; ++v)
            tmp[herm_index(t, u, v)] += exy * pp.Ez[eidx(az, bz, v)];
        }
      }
      for (int sa = 0; sa < na; ++sa) {
        double ca = S.c[la][sa][ka];
        if (ca == 0.0) continue;
        for (int sb = 0; sb < nb; ++sb) {
          double cab = ca * S.c[lb][sb][kb];
          if (cab == 0.0) continue;
          double* g = G + (sa * nb + sb) * NH;
          for (int h = 0; h < NH; ++h) g[h] += cab * tmp[h];
        }
      }
    }
  }
}

// Hermite coefficients of a single contracted-free aux primitive, spherical
// component m, multiplied by (-1)^{t+u+v} (ket side of the Coulomb formula).
void aux_ket_hermite(int l, double gamma, double coef, double* H /* [nsph][nherm(l)] */)
{
  const SphTable& S = sph_table();
  const int NH = nherm(l);
  double E[ESIZE];
  hermite_E(l, 0, 0.0, 0.0, 0.5 / gamma, E);
  for (int m = 0; m < nsph(l); ++m) {
    double* h = H + m * NH;
    for (int i = 0; i < NH; ++i) h[i] = 0.0;
    for (int k = 0; k < ncart(l); ++k) {
      double c = S.c[l][m][k];
      if (c == 0.0) continue;
      int ix, iy, iz;
      cart_exps(l, k, ix, iy, iz);
      for (int t = 0; t <= ix; ++t)
        for (int u = 0; u <= iy; ++u)
          for (int v = 0; v <= iz; ++v) {
            double sgn = ((t + u + v) & 1) ? -1.0 : 1.0;
            h[herm_index(t, u, v)] +=
                sgn * coef * c * E[eidx(ix, 0, t)] * E[eidx(iy, 0, u)] * E[eidx(iz, 0, v)];
          }
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// one-electron integrals
// ---------------------------------------------------------------------------
void ints_1e(const Basis& bas, const std::vector<Charge>& charges,
             double* S_out, double* T_out, double* V_out)
{
  const int nbf = bas.nbf;
  const int ns = (int)bas.shells.size();
  const SphTable& SPH = sph_table();
  const double* btab = boys_table().data();
  // kinetic needs E up to lb+2; use a local recursion with wider index range
  constexpr int LB = LMAX + 2;
  constexpr int TS = 2 * LB + 1;
  auto eidx2 = [](int i, int j, int t) { return (i * (LB + 1) + j) * TS + t; };
  auto herm1d = [&](int la, int lb, double XPA, double XPB, double oo2p, double* E) {
    for (int n = 0; n < (LB + 1) * (LB + 1) * TS; ++n) E[n] = 0.0;
    E[eidx2(0, 0, 0)] = 1.0;
    for (int i = 0; i < la; ++i)
      for (int t = 0; t <= i + 1; ++t) {
        double v = XPA * E[eidx2(i, 0, t)];
        if (t > 0) v += oo2p * E[eidx2(i, 0, t - 1)];
        if (t + 1 <= i) v += (t + 1) * E[eidx2(i, 0, t + 1)];
        E[eidx2(i + 1, 0, t)] = v;
      }
    for (int i = 0; i <= la; ++i)
      for (int j = 0; j < lb; ++j)
        for (int t = 0; t <= i + j + 1; ++t) {
          double v = XPB * E[eidx2(i, j, t)];
          if (t > 0) v += oo2p * E[eidx2(i, j, t - 1)];
          if (t + 1 <= i + j) v += (t + 1) * E[eidx2(i, j, t + 1)];
          E[eidx2(i, j + 1, t)] = v;
        }
  };

#pragma omp parallel for schedule(dynamic)
  for (int ia = 0; ia < ns; ++ia) {
    std::vector<double> Ex((LB + 1) * (LB + 1) * TS), Ey(Ex.size()), Ez(Ex.size());
    double W[(2 * LMAX + 1) * nherm(2 * LMAX)], R[nherm(2 * LMAX)];
    for (int ib = 0; ib <= ia; ++ib) {
      const Shell& a = bas.shells[ia];
      const Shell& b = bas.shells[ib];
      const int la = a.l, lb = b.l;
      const int nca = ncart(la), ncb = ncart(lb);
      constexpr int NC2 = ((LMAX + 1) * (LMAX + 2) / 2) * ((LMAX + 1) * (LMAX + 2) / 2);
      double s_c[NC2], t_c[NC2], v_c[NC2];
      for (int k = 0; k < NC2; ++k) s_c[k] = t_c[k] = v_c[k] = 0.0;
      double AB[3] = {a.O[0] - b.O[0], a.O[1] - b.O[1], a.O[2] - b.O[2]};
      double r2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
      for (size_t i = 0; i < a.alpha.size(); ++i)
        for (size_t j = 0; j < b.alpha.size(); ++j) {
          double al = a.alpha[i], be = b.alpha[j], p = al + be, mu = al * be / p;
          double K = a.coeff[i] * b.coeff[j] * std::exp(-mu * r2);
          double P[3];
          for (int d = 0; d < 3; ++d) P[d] = (al * a.O[d] + be * b.O[d]) / p;
          herm1d(la, lb + 2, P[0] - a.O[0], P[0] - b.O[0], 0.5 / p, Ex.data());
          herm1d(la, lb + 2, P[1] - a.O[1], P[1] - b.O[1], 0.5 / p, Ey.data());
          herm1d(la, lb + 2, P[2] - a.O[2], P[2] - b.O[2], 0.5 / p, Ez.data());
          double s3 = std::pow(PI / p, 1.5);
          // nuclear attraction Hermite integrals per centre
          const int L = la + lb;
          std::vector<double> Rsum(nherm(L), 0.0);
          for (const auto& ch : charges) {
            hermite_R(L, p, P[0] - ch.xyz[0], P[1] - ch.xyz[1], P[2] - ch.xyz[2], btab, W, R);
            for (int h = 0; h < nherm(L); ++h) Rsum[h] += ch.Z * R[h];
          }
          double vpref = -2.0 * PI / p * K;
          for (int ka = 0; ka < nca; ++ka) {
            int ax, ay, az;
            cart_exps(la, ka, ax, ay, az);
            for (int kb = 0; kb < ncb; ++kb) {
              int bx, by, bz;
              cart_exps(lb, kb, bx, by, bz);
              double sx = Ex[eidx2(ax, bx, 0)], sy = Ey[eidx2(ay, by, 0)], sz = Ez[eidx2(az, bz, 0)];
              auto kin1 = [&](const std::vector<double>& E, int ai, int bj) {
                double v = -2.0 * be * be * E[eidx2(ai, bj + 2, 0)] + be * (2 * bj + 1) * E[eidx2(ai, bj, 0)];
                if (bj >= 2) v -= 0.5 * bj * (bj - 1) * E[eidx2(ai, bj - 2, 0)];
                return v;
              };
              double tx = kin1(Ex, ax, bx), ty = kin1(Ey, ay, by), tz = kin1(Ez, az, bz);
              s_c[ka * ncb + kb] += K * s3 * sx * sy * sz;
              t_c[ka * ncb + kb] += K * s3 * (tx * sy * sz + sx * ty * sz + sx * sy * tz);
              double v = 0.0;
              for (int t = 0; t <= ax + bx; ++t)
                for (int u = 0; u <= ay + by; ++u)
                  for (int w = 0; w <= az + bz; ++w)
                    v += Ex[eidx2(ax, bx, t)] * Ey[eidx2(ay, by, u)] * Ez[eidx2(az, bz, w)] *
                         Rsum[herm_index(t, u, w)];
              v_c[ka * ncb + kb] += vpref * v;
            }
          }
        }
      // cart -> sph
      const int oa = bas.shell_off[ia], ob = bas.shell_off[ib];
      for (int sa = 0; sa < nsph(la); ++sa)
        for (int sb = 0; sb < nsph(lb); ++sb) {
          double s = 0, t = 0, v = 0;
          for (int ka = 0; ka < nca; ++ka) {
            double ca = SPH.c[la][sa][ka];
            if (ca == 0.0) continue;
            for (int kb = 0; kb < ncb; ++kb) {
              double c = ca * SPH.c[lb][sb][kb];
              s += c * s_c[ka * ncb + kb];
              t += c * t_c[ka * ncb + kb];
              v += c * v_c[ka * ncb + kb];
            }
          }
          size_t i1 = (size_t)(oa + sa) * nbf + (ob + sb);
          size_t i2 = (size_t)(ob + sb) * nbf + (oa + sa);
          if (S_out) { S_out[i1] = s; S_out[i2] = s; }
          if (T_out) { T_out[i1] = t; T_out[i2] = t; }
          if (V_out) { V_out[i1] = v; V_out[i2] = v; }
        }
    }
  }
}

// ---------------------------------------------------------------------------
// two-centre Coulomb metric (X|Y)
// ---------------------------------------------------------------------------
void ints_2c2e(const Basis& aux, double* V_out)
{
  const int n = aux.nbf;
  const int ns = (int)aux.shells.size();
  const double* btab = boys_table().data();
#pragma omp parallel for schedule(dynamic)
  for (int ia = 0; ia < ns; ++ia) {
    double W[(2 * LMAX + 1) * nherm(2 * LMAX)], R[nherm(2 * LMAX)];
    double Ha[(2 * LMAX + 1) * nherm(LMAX)], Hc[(2 * LMAX + 1) * nherm(LMAX)];
    for (int ic = 0; ic <= ia; ++ic) {
      const Shell& a = aux.shells[ia];
      const Shell& c = aux.shells[ic];
      const int la = a.l, lc = c.l, L = la + lc;
      const int NHa = nherm(la), NHc = nherm(lc);
      double out[(2 * LMAX + 1) * (2 * LMAX + 1)] = {0};
      for (size_t i = 0; i < a.alpha.size(); ++i) {
        aux_ket_hermite(la, a.alpha[i], a.coeff[i], Ha);
        // bra side must not carry the (-1)^{t+u+v} sign
        for (int m = 0; m < nsph(la); ++m)
          for (int t = 0; t <= la; ++t)
            for (int u = 0; u <= la - t; ++u)
              for (int v = 0; v <= la - t - u; ++v)
                if ((t + u + v) & 1) Ha[m * NHa + herm_index(t, u, v)] *= -1.0;
        for (size_t k = 0; k < c.alpha.size(); ++k) {
          aux_ket_hermite(lc, c.alpha[k], c.coeff[k], Hc);
          double p = a.alpha[i], q = c.alpha[k];
          double alpha = p * q / (p + q);
          hermite_R(L, alpha, a.O[0] - c.O[0], a.O[1] - c.O[1], a.O[2] - c.O[2], btab, W, R);
          double pref = 2.0 * std::pow(PI, 2.5) / (p * q * std::sqrt(p + q));
          for (int ma = 0; ma < nsph(la); ++ma)
            for (int mc = 0; mc < nsph(lc); ++mc) {
              double s = 0.0;
              for (int t = 0; t <= la; ++t)
                for (int u = 0; u <= la - t; ++u)
                  for (int v = 0; v <= la - t - u; ++v) {
                    double ha = Ha[ma * NHa + herm_index(t, u, v)];
                    if (ha == 0.0) continue;
                    for (int t2 = 0; t2 <= lc; ++t2)
                      for (int u2 = 0; u2 <= lc - t2; ++u2)
                        for (int v2 = 0; v2 <= lc - t2 - u2; ++v2)
                          s += ha * Hc[mc * NHc + herm_index(t2, u2, v2)] *
                               R[herm_index(t + t2, u + u2, v + v2)];
                  }
              out[ma * (2 * LMAX + 1) + mc] += pref * s;
            }
        }
      }
      const int oa = aux.shell_off[ia], oc = aux.shell_off[ic];
      for (int ma = 0; ma < nsph(la); ++ma)
        for (int mc = 0; mc < nsph(lc); ++mc) {
          V_out[(size_t)(oa + ma) * n + (oc + mc)] = out[ma * (2 * LMAX + 1) + mc];
          V_out[(size_t)(oc + mc) * n + (oa + ma)] = out[ma * (2 * LMAX + 1) + mc];
        }
    }
  }
}

// ---------------------------------------------------------------------------
// Schwarz factors (integrals.cpp:511-689)
// ---------------------------------------------------------------------------
void schwarz_ao(const Basis& bas, double* Q)
{
  const int ns = (int)bas.shells.size();
  const double* btab = boys_table().data();
#pragma omp parallel for schedule(dynamic)
  for (int ia = 0; ia < ns; ++ia) {
    std::vector<PrimPair> pairs;
    std::vector<double> G;
    double W[(4 * LMAX + 1) * nherm(4 * LMAX)], R[nherm(4 * LMAX)];
    for (int ib = 0; ib <= ia; ++ib) {
      const Shell& a = bas.shells[ia];
      const Shell& b = bas.shells[ib];
      const int la = a.l, lb = b.l, L = la + lb, NH = nherm(L);
      const int na = nsph(la), nb = nsph(lb);
      make_pairs(a, b, pairs);
      const size_t npp = pairs.size();
      G.assign(npp * na * nb * NH, 0.0);
      for (size_t k = 0; k < npp; ++k) pair_sph_hermite(la, lb, pairs[k], G.data() + k * na * nb * NH);
      std::vector<double> diag(na * nb, 0.0);
      for (size_t k1 = 0; k1 < npp; ++k1)
        for (size_t k2 = 0; k2 < npp; ++k2) {
          const PrimPair& p1 = pairs[k1];
          const PrimPair& p2 = pairs[k2];
          double p = p1.p, q = p2.p, alpha = p * q / (p + q);
          hermite_R(2 * L, alpha, p1.P[0] - p2.P[0], p1.P[1] - p2.P[1], p1.P[2] - p2.P[2], btab, W, R);
          double pref = 2.0 * std::pow(PI, 2.5) / (p * q * std::sqrt(p + q)) * p1.K * p2.K;
          for (int s = 0; s < na * nb; ++s) {
            const double* g1 = G.data() + (k1 * na * nb + s) * NH;
            const double* g2 = G.data() + (k2 * na * nb + s) * NH;
            double acc = 0.0;
            for (int t = 0; t <= L; ++t)
              for (int u = 0; u <= L - t; ++u)
                for (int v = 0; v <= L - t - u; ++v) {
                  double x1 = g1[herm_index(t, u, v)];
                  if (x1 == 0.0) continue;
                  double inner = 0.0;
                  for (int t2 = 0; t2 <= L; ++t2)
                    for (int u2 = 0; u2 <= L - t2; ++u2)
                      for (int v2 = 0; v2 <= L - t2 - u2; ++v2) {
                        double sg = ((t2 + u2 + v2) & 1) ? -1.0 : 1.0;
                        inner += sg * g2[herm_index(t2, u2, v2)] * R[herm_index(t + t2, u + u2, v + v2)];
                      }
                  acc += x1 * inner;
                }
            diag[s] += pref * acc;
          }
        }
      double n = 0.0;
      for (double d : diag) n += std::fabs(d);
      Q[(size_t)ia * ns + ib] = std::sqrt(n);
      Q[(size_t)ib * ns + ia] = std::sqrt(n);
    }
  }
}

// ---------------------------------------------------------------------------
// four-centre integrals (ab|cd) (the reference's int2e_sph through libcint,
// used by its exact J/K for the atomic SCFs of the SAD guess, hfguess.cpp:142-452)
// ---------------------------------------------------------------------------
// (ab|cd) = sum_{pairs} 2 pi^{5/2} / (p q sqrt(p+q)) K_ab K_cd
//           sum_{tuv} G_ab[tuv] sum_{t'u'v'} (-1)^{t'+u'+v'} G_cd[t'u'v'] R_{t+t', u+u', v+v'}
// (Helgaker et al., eq. 9.9.33), spherical components, layout out[sa][sb][sc][sd].
void ints_4c2e_shell(const Basis& bas, int ia, int ib, int ic, int id, double* out)
{
  const double* btab = boys_table().data();
  const Shell& a = bas.shells[ia];
  const Shell& b = bas.shells[ib];
  const Shell& cs = bas.shells[ic];
  const Shell& ds = bas.shells[id];
  const int L1 = a.l + b.l, L2 = cs.l + ds.l, NH1 = nherm(L1), NH2 = nherm(L2);
  const int na = nsph(a.l), nb = nsph(b.l), nc = nsph(cs.l), nd = nsph(ds.l);
  std::vector<PrimPair> p1, p2;
  make_pairs(a, b, p1);
  make_pairs(cs, ds, p2);
  std::vector<double> G1(p1.size() * na * nb * NH1), G2(p2.size() * nc * nd * NH2);
  for (size_t k = 0; k < p1.size(); ++k) pair_sph_hermite(a.l, b.l, p1[k], G1.data() + k * na * nb * NH1);
  for (size_t k = 0; k < p2.size(); ++k) {
    double* g = G2.data() + k * nc * nd * NH2;
    pair_sph_hermite(cs.l, ds.l, p2[k], g);
    for (int s = 0; s < nc * nd; ++s)
      for (int t = 0; t <= L2; ++t)
        for (int u = 0; u <= L2 - t; ++u)
          for (int v = 0; v <= L2 - t - u; ++v)
            if ((t + u + v) & 1) g[s * NH2 + herm_index(t, u, v)] = -g[s * NH2 + herm_index(t, u, v)];
  }
  std::fill(out, out + (size_t)na * nb * nc * nd, 0.0);
  double W[(4 * LMAX + 1) * nherm(4 * LMAX)], R[nherm(4 * LMAX)];
  std::vector<double> T(NH1);
  for (size_t k1 = 0; k1 < p1.size(); ++k1)
    for (size_t k2 = 0; k2 < p2.size(); ++k2) {
      const PrimPair& q1 = p1[k1];
      const PrimPair& q2 = p2[k2];
      const double p = q1.p, q = q2.p, alpha = p * q / (p + q);
      hermite_R(L1 + L2, alpha, q1.P[0] - q2.P[0], q1.P[1] - q2.P[1], q1.P[2] - q2.P[2], btab, W, R);
      const double pref = 2.0 * std::pow(PI, 2.5) / (p * q * std::sqrt(p + q)) * q1.K * q2.K;
      for (int cd = 0; cd < nc * nd; ++cd) {
        const double* g2 = G2.data() + (k2 * nc * nd + cd) * NH2;
        // T[tuv] = sum_{t'u'v'} g2[t'u'v'] R[t+t', u+u', v+v']
        for (int t = 0; t <= L1; ++t)
          for (int u = 0; u <= L1 - t; ++u)
            for (int v = 0; v <= L1 - t - u; ++v) {
              double acc = 0.0;
              for (int t2 = 0; t2 <= L2; ++t2)
                for (int u2 = 0; u2 <= L2 - t2; ++u2)
                  for (int v2 = 0; v2 <= L2 - t2 - u2; ++v2)
                    acc += g2[herm_index(t2, u2, v2)] * R[herm_index(t + t2, u + u2, v + v2)];
              T[herm_index(t, u, v)] = acc;
            }
        for (int ab = 0; ab < na * nb; ++ab) {
          const double* g1 = G1.data() + (k1 * na * nb + ab) * NH1;
          double acc = 0.0;
          for (int h = 0; h < NH1; ++h) acc += g1[h] * T[h];
          out[(size_t)ab * nc * nd + cd] += pref * acc;
        }
      }
    }
}

// Every (mu nu|lambda sigma) of a (small) basis, row-major nbf^4.
void ints_4c2e_all(const Basis& bas, double* out)
{
  const int ns = (int)bas.shells.size(), n = bas.nbf;
  std::vector<int> off(ns);
  for (int i = 0; i < ns; ++i) off[i] = bas.shell_off[i];
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int ia = 0; ia < ns; ++ia)
    for (int ib = 0; ib < ns; ++ib) {
      std::vector<double> buf;
      for (int ic = 0; ic < ns; ++ic)
        for (int id = 0; id < ns; ++id) {
          const int na = nsph(bas.shells[ia].l), nb = nsph(bas.shells[ib].l), nc = nsph(bas.shells[ic].l),
                    nd = nsph(bas.shells[id].l);
          buf.assign((size_t)na * nb * nc * nd, 0.0);
          ints_4c2e_shell(bas, ia, ib, ic, id, buf.data());
          for (int sa = 0; sa < na; ++sa)
            for (int sb = 0; sb < nb; ++sb)
              for (int sc = 0; sc < nc; ++sc)
                for (int sd = 0; sd < nd; ++sd) {
                  const size_t i = off[ia] + sa, j = off[ib] + sb, k = off[ic] + sc, l = off[id] + sd;
                  out[((i * n + j) * n + k) * n + l] = buf[((sa * nb + sb) * nc + sc) * nd + sd];
                }
        }
    }
}

void schwarz_aux(const Basis& aux, double* Q)
{
  // diagonal of (X|X) summed per shell
  const int ns = (int)aux.shells.size();
  const double* btab = boys_table().data();
#pragma omp parallel for schedule(dynamic)
  for (int ia = 0; ia < ns; ++ia) {
    const Shell& a = aux.shells[ia];
    const int la = a.l, NH = nherm(la);
    double W[(2 * LMAX + 1) * nherm(2 * LMAX)], R[nherm(2 * LMAX)];
    double H1[(2 * LMAX + 1) * nherm(LMAX)], H2[(2 * LMAX + 1) * nherm(LMAX)];
    std::vector<double> diag(nsph(la), 0.0);
    for (size_t i = 0; i < a.alpha.size(); ++i)
      for (size_t k = 0; k < a.alpha.size(); ++k) {
        aux_ket_hermite(la, a.alpha[i], a.coeff[i], H1);
        aux_ket_hermite(la, a.alpha[k], a.coeff[k], H2);
        double p = a.alpha[i], q = a.alpha[k], alpha = p * q / (p + q);
        hermite_R(2 * la, alpha, 0.0, 0.0, 0.0, btab, W, R);
        double pref = 2.0 * std::pow(PI, 2.5) / (p * q * std::sqrt(p + q));
        for (int m = 0; m < nsph(la); ++m) {
          double s = 0.0;
          for (int t = 0; t <= la; ++t)
            for (int u = 0; u <= la - t; ++u)
              for (int v = 0; v <= la - t - u; ++v) {
                double sg = ((t + u + v) & 1) ? -1.0 : 1.0;   // undo ket sign on bra
                double h1 = sg * H1[m * NH + herm_index(t, u, v)];
                if (h1 == 0.0) continue;
                for (int t2 = 0; t2 <= la; ++t2)
                  for (int u2 = 0; u2 <= la - t2; ++u2)
                    for (int v2 = 0; v2 <= la - t2 - u2; ++v2)
                      s += h1 * H2[m * NH + herm_index(t2, u2, v2)] * R[herm_index(t + t2, u + u2, v + v2)];
              }
          diag[m] += pref * s;
        }
      }
    double n = 0.0;
    for (double d : diag) n += std::fabs(d);
    Q[ia] = std::sqrt(n);
  }
}

// ---------------------------------------------------------------------------
// three-centre integrals (X|mu nu) for one shell triple, spherical, layout
// out[m_c][sa][sb]
// ---------------------------------------------------------------------------
namespace {

struct AuxPrim {
  int l;
  double gamma, C[3];
  double H[(2 * LMAX + 1) * nherm(LMAX)];
};

void shell3c(const Shell& a, const Shell& b, const std::vector<PrimPair>& pairs,
             const AuxPrim* cprims, int ncprim, const double* btab, double* out)
{
  const SphTable& SPH = sph_table();
  const int la = a.l, lb = b.l, lc = cprims[0].l;
  const int lab = la + lb, L = lab + lc;
  const int nca = ncart(la), ncb = ncart(lb), nmc = nsph(lc);
  const int NHab = nherm(lab), NHc = nherm(lc);
  constexpr int NC = (LMAX + 1) * (LMAX + 2) / 2;
  double acc[(2 * LMAX + 1) * NC * NC];
  for (int i = 0; i < nmc * nca * ncb; ++i) acc[i] = 0.0;
  double W[(3 * LMAX + 1) * nherm(3 * LMAX)], R[nherm(3 * LMAX)];
  double Sm[(2 * LMAX + 1) * nherm(2 * LMAX)];
  for (int ic = 0; ic < ncprim; ++ic) {
    const AuxPrim& c = cprims[ic];
    for (const PrimPair& pp : pairs) {
      double p = pp.p, q = c.gamma, alpha = p * q / (p + q);
      hermite_R(L, alpha, pp.P[0] - c.C[0], pp.P[1] - c.C[1], pp.P[2] - c.C[2], btab, W, R);
      double pref = 2.0 * PI * PI * std::sqrt(PI) / (p * q * std::sqrt(p + q)) * pp.K;
      // S^m_{tuv} = sum_{t'u'v'} H^m_{t'u'v'} R_{t+t',u+u',v+v'}
      for (int m = 0; m < nmc; ++m) {
        const double* h = c.H + m * NHc;
        double* s = Sm + m * NHab;
        for (int t = 0; t <= lab; ++t)
          for (int u = 0; u <= lab - t; ++u)
            for (int v = 0; v <= lab - t - u; ++v) {
              double x = 0.0;
              for (int t2 = 0; t2 <= lc; ++t2)
                for (int u2 = 0; u2 <= lc - t2; ++u2)
                  for (int v2 = 0; v2 <= lc - t2 - u2; ++v2) {
                    double hv = h[herm_index(t2, u2, v2)];
                    if (hv != 0.0) x += hv * R[herm_index(t + t2, u + u2, v + v2)];
                  }
              s[herm_index(t, u, v)] = x * pref;
            }
      }
      for (int ka = 0; ka < nca; ++ka) {
        int ax, ay, az;
        cart_exps(la, ka, ax, ay, az);
        for (int kb = 0; kb < ncb; ++kb) {
          int bx, by, bz;
          cart_exps(lb, kb, bx, by, bz);
          double r[2 * LMAX + 1];
          for (int m = 0; m < nmc; ++m) r[m] = 0.0;
          for (int t = 0; t <= ax + bx; ++t) {
            double ex = pp.Ex[eidx(ax, bx, t)];
            for (int u = 0; u <= ay + by; ++u) {
              double exy = ex * pp.Ey[eidx(ay, by, u)];
              for (int v = 0; v <= az + bz; ++v) {
                double e = exy * pp.Ez[eidx(az, bz, v)];
                int hi = herm_index(t, u, v);
                for (int m = 0; m < nmc; ++m) r[m] += e * Sm[m * NHab + hi];
              }
            }
          }
          for (int m = 0; m < nmc; ++m) acc[(m * nca + ka) * ncb + kb] += r[m];
        }
      }
    }
  }
  const int nsa = nsph(la), nsb = nsph(lb);
  for (int m = 0; m < nmc; ++m)
    for (int sa = 0; sa < nsa; ++sa)
      for (int sb = 0; sb < nsb; ++sb) {
        double x = 0.0;
        for (int ka = 0; ka < nca; ++ka) {
          double ca = SPH.c[la][sa][ka];
          if (ca == 0.0) continue;
          for (int kb = 0; kb < ncb; ++kb) x += ca * SPH.c[lb][sb][kb] * acc[(m * nca + ka) * ncb + kb];
        }
        out[(m * nsa + sa) * nsb + sb] = x;
      }
}

}  // namespace

void aux_hermite_coeffs(int l, double gamma, double coef, double* H, int stride)
{
  double tmp[(2 * LMAX + 1) * nherm(LMAX)];
  aux_ket_hermite(l, gamma, coef, tmp);
  const int NH = nherm(l);
  for (int m = 0; m < nsph(l); ++m)
    for (int i = 0; i < NH; ++i) H[m * stride + i] = tmp[m * NH + i];
}

void ints_3c2e_shell(const Basis& ao, const Basis& aux, int ia, int ib, int ic, double* out)
{
  std::vector<PrimPair> pairs;
  make_pairs(ao.shells[ia], ao.shells[ib], pairs);
  const Shell& c = aux.shells[ic];
  std::vector<AuxPrim> cp(c.alpha.size());
  for (size_t k = 0; k < c.alpha.size(); ++k) {
    cp[k].l = c.l;
    cp[k].gamma = c.alpha[k];
    for (int d = 0; d < 3; ++d) cp[k].C[d] = c.O[d];
    aux_ket_hermite(c.l, c.alpha[k], c.coeff[k], cp[k].H);
  }
  shell3c(ao.shells[ia], ao.shells[ib], pairs, cp.data(), (int)cp.size(), boys_table().data(), out);
}

// ---------------------------------------------------------------------------
// Full B build: Schwarz block mask -> integrals -> block filter -> slabs
// (aoloader.cpp:341-377, aofactory.cpp:52-143, screening.cpp:23-30)
// ---------------------------------------------------------------------------
void build_B(const Basis& ao, const Basis& aux,
             const std::vector<int>& ao_blk_shells,    // shells per AO block
             const std::vector<int>& aux_blk_shells,   // shells per aux block
             const std::vector<uint8_t>& schwarz_mask, // [nxb][nbb][nbb]
             double filter_eps,
             BLayout& layout, std::vector<double>& data,
             std::vector<uint8_t>& final_mask)
{
  const int nbb = (int)ao_blk_shells.size();
  const int nxb = (int)aux_blk_shells.size();
  std::vector<int> ao_sh0(nbb + 1, 0), aux_sh0(nxb + 1, 0);
  for (int i = 0; i < nbb; ++i) ao_sh0[i + 1] = ao_sh0[i] + ao_blk_shells[i];
  for (int i = 0; i < nxb; ++i) aux_sh0[i + 1] = aux_sh0[i] + aux_blk_shells[i];
  std::vector<int> bsz(nbb), xsz(nxb), boff(nbb + 1, 0), xoff(nxb + 1, 0);
  for (int i = 0; i < nbb; ++i) {
    bsz[i] = 0;
    for (int s = ao_sh0[i]; s < ao_sh0[i + 1]; ++s) bsz[i] += nsph(ao.shells[s].l);
    boff[i + 1] = boff[i] + bsz[i];
  }
  for (int i = 0; i < nxb; ++i) {
    xsz[i] = 0;
    for (int s = aux_sh0[i]; s < aux_sh0[i + 1]; ++s) xsz[i] += nsph(aux.shells[s].l);
    xoff[i + 1] = xoff[i] + xsz[i];
  }
  const int naux = xoff[nxb];

  // canonical significant pairs (M <= N) from the Schwarz mask
  struct CPair { int m, n; int64_t off; };
  std::vector<CPair> cpairs;
  int64_t ctotal = 0;
  for (int n = 0; n < nbb; ++n)
    for (int m = 0; m <= n; ++m) {
      bool sig = false;
      for (int x = 0; x < nxb && !sig; ++x) sig = schwarz_mask[((size_t)x * nbb + m) * nbb + n] != 0;
      if (!sig) continue;
      cpairs.push_back({m, n, ctotal});
      ctotal += (int64_t)naux * bsz[m] * bsz[n];
    }
  // temp store per canonical pair: [x][mu][nu], mu in M, nu in N
  std::vector<double> tmp((size_t)ctotal, 0.0);

  // aux primitive data per aux shell
  std::vector<std::vector<AuxPrim>> auxp(aux.shells.size());
  for (size_t ic = 0; ic < aux.shells.size(); ++ic) {
    const Shell& c = aux.shells[ic];
    auxp[ic].resize(c.alpha.size());
    for (size_t k = 0; k < c.alpha.size(); ++k) {
      auxp[ic][k].l = c.l;
      auxp[ic][k].gamma = c.alpha[k];
      for (int d = 0; d < 3; ++d) auxp[ic][k].C[d] = c.O[d];
      aux_ket_hermite(c.l, c.alpha[k], c.coeff[k], auxp[ic][k].H);
    }
  }

  // shell-pair tasks
  struct Task { int cp, sa, sb; };
  std::vector<Task> tasks;
  for (int k = 0; k < (int)cpairs.size(); ++k) {
    int m = cpairs[k].m, n = cpairs[k].n;
    for (int sa = ao_sh0[m]; sa < ao_sh0[m + 1]; ++sa)
      for (int sb = ao_sh0[n]; sb < ao_sh0[n + 1]; ++sb) {
        if (m == n && sb < sa) continue;
        tasks.push_back({k, sa, sb});
      }
  }
  const double* btab = boys_table().data();
  std::vector<int> aux_shell_blk(aux.shells.size());
  for (int x = 0; x < nxb; ++x)
    for (int s = aux_sh0[x]; s < aux_sh0[x + 1]; ++s) aux_shell_blk[s] = x;

#pragma omp parallel
  {
    std::vector<PrimPair> pairs;
    double out[(2 * LMAX + 1) * (2 * LMAX + 1) * (2 * LMAX + 1)];
#pragma omp for schedule(dynamic, 4)
    for (size_t t = 0; t < tasks.size(); ++t) {
      const Task& tk = tasks[t];
      const CPair& cp = cpairs[tk.cp];
      const Shell& a = ao.shells[tk.sa];
      const Shell& b = ao.shells[tk.sb];
      make_pairs(a, b, pairs);
      const int na = nsph(a.l), nb = nsph(b.l);
      const int mu0 = ao.shell_off[tk.sa] - boff[cp.m];
      const int nu0 = ao.shell_off[tk.sb] - boff[cp.n];
      const int BM = bsz[cp.m], BN = bsz[cp.n];
      for (size_t ic = 0; ic < aux.shells.size(); ++ic) {
        int xb = aux_shell_blk[ic];
        if (!schwarz_mask[((size_t)xb * nbb + cp.m) * nbb + cp.n]) continue;
        const int nc = nsph(aux.shells[ic].l);
        shell3c(a, b, pairs, auxp[ic].data(), (int)auxp[ic].size(), btab, out);
        const int x0 = aux.shell_off[ic];
        for (int mc = 0; mc < nc; ++mc) {
          double* dst = tmp.data() + cp.off + (int64_t)(x0 + mc) * BM * BN;
          for (int sa = 0; sa < na; ++sa)
            for (int sb = 0; sb < nb; ++sb) {
              double v = out[(mc * na + sa) * nb + sb];
              dst[(mu0 + sa) * BN + (nu0 + sb)] = v;
              if (cp.m == cp.n) dst[(nu0 + sb) * BN + (mu0 + sa)] = v;
            }
        }
      }
    }
  }

  // block filter: drop (x,M,N) blocks with Frobenius norm < eps
  final_mask.assign((size_t)nxb * nbb * nbb, 0);
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < (int)cpairs.size(); ++k) {
    const CPair& cp = cpairs[k];
    const int BM = bsz[cp.m], BN = bsz[cp.n];
    for (int x = 0; x < nxb; ++x) {
      if (!schwarz_mask[((size_t)x * nbb + cp.m) * nbb + cp.n]) continue;
      double s = 0.0;
      const double* src = tmp.data() + cp.off + (int64_t)xoff[x] * BM * BN;
      for (int64_t i = 0; i < (int64_t)xsz[x] * BM * BN; ++i) s += src[i] * src[i];
      if (std::sqrt(s) >= filter_eps) {
        final_mask[((size_t)x * nbb + cp.m) * nbb + cp.n] = 1;
        final_mask[((size_t)x * nbb + cp.n) * nbb + cp.m] = 1;
      }
    }
  }

  layout.build(xsz, bsz, final_mask);
  data.assign((size_t)layout.total, 0.0);
  // scatter canonical blocks into slabs, both orientations
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < (int)cpairs.size(); ++k) {
    const CPair& cp = cpairs[k];
    const int m = cp.m, n = cp.n, BM = bsz[m], BN = bsz[n];
    int col_mn = layout.col_of(m, n);   // block M inside slab N
    int col_nm = layout.col_of(n, m);   // block N inside slab M
    if (col_mn < 0) continue;           // pair filtered out entirely
    double* slabN = data.data() + layout.slab_off[n];
    double* slabM = data.data() + layout.slab_off[m];
    const int ncN = layout.ncol[n], ncM = layout.ncol[m];
    for (int x = 0; x < nxb; ++x) {
      if (!final_mask[((size_t)x * nbb + m) * nbb + n]) continue;
      for (int xi = xoff[x]; xi < xoff[x + 1]; ++xi) {
        const double* src = tmp.data() + cp.off + (int64_t)xi * BM * BN;
        for (int mu = 0; mu < BM; ++mu)
          for (int nu = 0; nu < BN; ++nu) {
            double v = src[mu * BN + nu];
            slabN[((int64_t)xi * BN + nu) * ncN + col_mn + mu] = v;
            slabM[((int64_t)xi * BM + mu) * ncM + col_nm + nu] = v;
          }
      }
    }
  }
  (void)naux;
}

}  // namespace mchip
