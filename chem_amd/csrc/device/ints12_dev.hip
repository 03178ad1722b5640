// One-electron integrals (overlap, kinetic, nuclear attraction), the
// two-centre Coulomb metric (X|Y) and the AO Schwarz factors on the GPU, one
// kernel per angular momentum class, for the request set-up of the serving
// loop and the C-ABI.
//
// Same McMurchie-Davidson formulas as the host engine (ints_host.cpp, ints_1e,
// ints_2c2e and schwarz_ao; the reference's libcint calls int1e_ovlp_sph,
// int1e_kin_sph, int1e_nuc_sph and int2c2e_sph, aofactory.cpp:380-430, and its
// Schwarz factors, integrals.cpp:511-689), with the recursions
// of the three-centre class kernels (ints3c_cls.hpp): one thread per shell
// pair, everything unrolled with constant indices.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../ints/ints_host.hpp"
#include "dev.hpp"
#include "ints12_dev.hpp"
#include "ints3c_cls.hpp"

namespace mchip {

struct I1eArgs {
  const int* np;
  const int* po;
  const double* O;
  const double* ex;
  const double* cf;
  const int* off;          // first function of each shell
  const int* sha;          // shells of degree LA
  const int* shb;          // shells of degree LB
  int nsha, nshb;
  const double* chg;       // nchg x (Z, x, y, z)
  int nchg;
  const double* boys;
  double* S;
  double* T;
  double* V;
  int n;
  double sinv[LMAX + 1][2 * LMAX + 1];
};

// S, T and V of the shell pairs (a of degree LA, b of degree LB); LA == LB:
// the pairs with b at or below a in the shell list of that degree.
template <int LA, int LB>
__global__ __launch_bounds__(64) void k_int1e_cls(const I1eArgs A)
{
  using namespace i3c;
  constexpr int L = LA + LB, LB2 = LB + 2, TD = LA + LB2 + 1;
  constexpr int NCA = nc(LA), NCB = nc(LB), NE = (LA + 1) * (LB2 + 1) * TD;
  const int64_t idx = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (idx >= (int64_t)A.nsha * A.nshb) return;
  const int ia = (int)(idx / A.nshb), ib = (int)(idx % A.nshb);
  if (LA == LB && ib > ia) return;
  const int a = A.sha[ia], b = A.shb[ib];
  const double* __restrict__ ex = A.ex;
  const double* __restrict__ cf = A.cf;
  const double* __restrict__ chg = A.chg;
  const double Ax = A.O[3 * a], Ay = A.O[3 * a + 1], Az = A.O[3 * a + 2];
  const double Bx = A.O[3 * b], By = A.O[3 * b + 1], Bz = A.O[3 * b + 2];
  const double AB2 = (Ax - Bx) * (Ax - Bx) + (Ay - By) * (Ay - By) + (Az - Bz) * (Az - Bz);
  const int pa0 = A.po[a], pb0 = A.po[b], npa = A.np[a], npb = A.np[b];
  auto ei = [](int i, int j, int t) { return (i * (LB2 + 1) + j) * TD + t; };
  double sc[NCA][NCB], tc[NCA][NCB], vc[NCA][NCB];
  static_for<NCA>([&](auto ka) {
    static_for<NCB>([&](auto kb) { sc[ka][kb] = tc[ka][kb] = vc[ka][kb] = 0.0; });
  });
  for (int i = 0; i < npa; ++i) {
    const double al = ex[pa0 + i], ca = cf[pa0 + i];
    for (int j = 0; j < npb; ++j) {
      const double be = ex[pb0 + j];
      const double p = al + be, mu = al * be / p;
      const double K = ca * cf[pb0 + j] * exp(-mu * AB2);
      const double Px = (al * Ax + be * Bx) / p, Py = (al * Ay + be * By) / p, Pz = (al * Az + be * Bz) / p;
      const double oo2p = 0.5 / p;
      double Ex[NE], Ey[NE], Ez[NE];
      herm_E_t<LA, LB2>(Px - Ax, Px - Bx, oo2p, Ex);
      herm_E_t<LA, LB2>(Py - Ay, Py - By, oo2p, Ey);
      herm_E_t<LA, LB2>(Pz - Az, Pz - Bz, oo2p, Ez);
      const double s3 = (PI / p) * sqrt(PI / p);
      // nuclear attraction Hermite integrals summed over the charges
      double Rs[nh(L)];
      static_for<nh(L)>([&](auto h) { Rs[h] = 0.0; });
      for (int k = 0; k < A.nchg; ++k) {
        double R[nh(L)];
        herm_R_t<L>(p, Px - chg[4 * k + 1], Py - chg[4 * k + 2], Pz - chg[4 * k + 3], A.boys, R);
        const double Z = chg[4 * k];
        static_for<nh(L)>([&](auto h) { Rs[h] += Z * R[h]; });
      }
      const double vpref = -2.0 * PI / p * K;
      static_for<NCA>([&](auto ka) {
        constexpr int ax = cx(LA, ka), ay = cy(LA, ka), az = cz(LA, ka);
        static_for<NCB>([&](auto kb) {
          constexpr int bx = cx(LB, kb), by = cy(LB, kb), bz = cz(LB, kb);
          const double sx = Ex[ei(ax, bx, 0)], sy = Ey[ei(ay, by, 0)], sz = Ez[ei(az, bz, 0)];
          auto kin1 = [&](const double(&E)[NE], auto ai, auto bj) {
            constexpr int ia_ = decltype(ai)::value, jb = decltype(bj)::value;
            double v = -2.0 * be * be * E[ei(ia_, jb + 2, 0)] + be * (2 * jb + 1) * E[ei(ia_, jb, 0)];
            if constexpr (jb >= 2) v -= 0.5 * jb * (jb - 1) * E[ei(ia_, jb - 2, 0)];
            return v;
          };
          using IC = std::integral_constant<int, 0>;
          (void)IC{};
          const double tx = kin1(Ex, std::integral_constant<int, ax>{}, std::integral_constant<int, bx>{});
          const double ty = kin1(Ey, std::integral_constant<int, ay>{}, std::integral_constant<int, by>{});
          const double tz = kin1(Ez, std::integral_constant<int, az>{}, std::integral_constant<int, bz>{});
          sc[ka][kb] += K * s3 * sx * sy * sz;
          tc[ka][kb] += K * s3 * (tx * sy * sz + sx * ty * sz + sx * sy * tz);
          double v = 0.0;
          static_for<ax + bx + 1>([&](auto t) {
            static_for<ay + by + 1>([&](auto u) {
              static_for<az + bz + 1>([&](auto w) {
                v += Ex[ei(ax, bx, t)] * Ey[ei(ay, by, u)] * Ez[ei(az, bz, w)] * Rs[hix(t, u, w)];
              });
            });
          });
          vc[ka][kb] += vpref * v;
        });
      });
    }
  }
  // Cartesian -> spherical, both triangles
  const int oa = A.off[a], ob = A.off[b], n = A.n;
  static_for<2 * LA + 1>([&](auto sa) {
    static_for<2 * LB + 1>([&](auto sb) {
      double s = 0.0, t = 0.0, v = 0.0;
      static_for<NCA>([&](auto ka) {
        constexpr int wa = shp(LA, sa, ka);
        if constexpr (wa != 0) {
          double ys = 0.0, yt = 0.0, yv = 0.0;
          static_for<NCB>([&](auto kb) {
            constexpr int wb = shp(LB, sb, kb);
            if constexpr (wb != 0) {
              ys += (double)wb * sc[ka][kb];
              yt += (double)wb * tc[ka][kb];
              yv += (double)wb * vc[ka][kb];
            }
          });
          s += (double)wa * ys;
          t += (double)wa * yt;
          v += (double)wa * yv;
        }
      });
      const double f = A.sinv[LA][sa] * A.sinv[LB][sb];
      const int64_t i1 = (int64_t)(oa + sa) * n + (ob + sb), i2 = (int64_t)(ob + sb) * n + (oa + sa);
      A.S[i1] = A.S[i2] = s * f;
      A.T[i1] = A.T[i2] = t * f;
      A.V[i1] = A.V[i2] = v * f;
    });
  });
}

struct I2cArgs {
  const double* g;         // exponent per auxiliary shell
  const double* O;
  const double* s;         // (-1)^l coef (2 g)^-l per shell
  const int* off;
  const int* shx;          // shells of degree LX
  const int* shy;          // shells of degree LY
  int nshx, nshy;
  const double* boys;
  double* M;
  int n;
  double sinv[LMAX + 1][2 * LMAX + 1];
};

// (X|Y) of uncontracted auxiliary shells X of degree LX and Y of degree LY;
// both sides keep only their top-degree Hermite terms (the bra without the
// ket's (-1)^l).
template <int LX, int LY>
__global__ __launch_bounds__(64) void k_int2c2e_cls(const I2cArgs A)
{
  using namespace i3c;
  constexpr int L = LX + LY;
  const int64_t idx = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (idx >= (int64_t)A.nshx * A.nshy) return;
  const int ix = (int)(idx / A.nshy), iy = (int)(idx % A.nshy);
  if (LX == LY && iy > ix) return;
  const int x = A.shx[ix], y = A.shy[iy];
  const double p = A.g[x], q = A.g[y];
  const double alpha = p * q / (p + q);
  double R[nh(L)];
  herm_R_t<L>(alpha, A.O[3 * x] - A.O[3 * y], A.O[3 * x + 1] - A.O[3 * y + 1], A.O[3 * x + 2] - A.O[3 * y + 2],
              A.boys, R);
  // the bra factor without its (-1)^LX
  const double pref = 2.0 * PI * PI * sqrt(PI) / (p * q * sqrt(p + q)) * ((LX & 1) ? -A.s[x] : A.s[x]) * A.s[y];
  const int ox = A.off[x], oy = A.off[y], n = A.n;
  static_for<2 * LX + 1>([&](auto mx) {
    static_for<2 * LY + 1>([&](auto my) {
      double v = 0.0;
      static_for<nc(LX)>([&](auto kx) {
        constexpr int wx = shp(LX, mx, kx);
        if constexpr (wx != 0) {
          double yv = 0.0;
          static_for<nc(LY)>([&](auto ky) {
            constexpr int wy = shp(LY, my, ky);
            if constexpr (wy != 0)
              yv += (double)wy * R[hix(cx(LX, kx) + cx(LY, ky), cy(LX, kx) + cy(LY, ky), cz(LX, kx) + cz(LY, ky))];
          });
          v += (double)wx * yv;
        }
      });
      v *= pref * A.sinv[LX][mx] * A.sinv[LY][my];
      A.M[(int64_t)(ox + mx) * n + (oy + my)] = v;
      A.M[(int64_t)(oy + my) * n + (ox + mx)] = v;
    });
  });
}

struct SchArgs {
  const int* np;
  const int* po;
  const double* O;
  const double* ex;
  const double* cf;
  const int* sha;          // shells of degree LA
  const int* shb;          // shells of degree LB
  int nsha, nshb;
  const double* boys;
  double* diag;            // [pair][component]: (ab|ab) of one spherical component pair
  double sinv[LMAX + 1][2 * LMAX + 1];
};

// Diagonal (s s | s s) of one spherical component pair s = (sa, sb) of the
// shell pair (a of degree LA, b of degree LB; LA == LB: b at or below a):
// sum over primitive pairs k1, k2 of 2 pi^5/2 / (p q sqrt(p + q)) K1 K2
// sum_tuv G1[tuv] sum_t'u'v' (-1)^(t'+u'+v') G2[t'u'v'] R[t+t', u+u', v+v']
// with G the pair's Hermite coefficients of component s (ints_host.cpp,
// schwarz_ao).  Thread = (pair, component).
template <int LA, int LB>
__global__ __launch_bounds__(64) void k_schwarz_ao_cls(const SchArgs A)
{
  using namespace i3c;
  constexpr int L = LA + LB, TD = L + 1, NE = (LA + 1) * (LB + 1) * TD;
  constexpr int NSA = 2 * LA + 1, NSB = 2 * LB + 1, NS = NSA * NSB;
  const int64_t idx = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (idx >= (int64_t)A.nsha * A.nshb * NS) return;
  const int64_t pidx = idx / NS;
  const int sc = (int)(idx - pidx * NS);
  const int ia = (int)(pidx / A.nshb), ib = (int)(pidx % A.nshb);
  if (LA == LB && ib > ia) return;
  const int sa = sc / NSB, sb = sc - (sc / NSB) * NSB;
  const int a = A.sha[ia], b = A.shb[ib];
  const double* __restrict__ ex = A.ex;
  const double* __restrict__ cf = A.cf;
  const double Ax = A.O[3 * a], Ay = A.O[3 * a + 1], Az = A.O[3 * a + 2];
  const double Bx = A.O[3 * b], By = A.O[3 * b + 1], Bz = A.O[3 * b + 2];
  const double AB2 = (Ax - Bx) * (Ax - Bx) + (Ay - By) * (Ay - By) + (Az - Bz) * (Az - Bz);
  const int pa0 = A.po[a], pb0 = A.po[b], npa = A.np[a], npb = A.np[b];
  auto ei = [](int i, int j, int t) { return (i * (LB + 1) + j) * TD + t; };
  // G[tuv] of component (sa, sb) from the E coefficients of one primitive pair
  // (the integer shapes; the normalisations are applied at the end)
  auto hermite = [&](const double (&Ex)[NE], const double (&Ey)[NE], const double (&Ez)[NE], double (&G)[nh(L)]) {
    static_for<nh(L)>([&](auto h) { G[h] = 0.0; });
    static_for<NSA>([&](auto sA) {
      if (sA.value != sa) return;
      static_for<NSB>([&](auto sB) {
        if (sB.value != sb) return;
        static_for<nc(LA)>([&](auto ka) {
          constexpr int wa = shp(LA, decltype(sA)::value, ka);
          if constexpr (wa != 0) {
            constexpr int ax = cx(LA, ka), ay = cy(LA, ka), az = cz(LA, ka);
            static_for<nc(LB)>([&](auto kb) {
              constexpr int wb = shp(LB, decltype(sB)::value, kb);
              if constexpr (wb != 0) {
                constexpr int bx = cx(LB, kb), by = cy(LB, kb), bz = cz(LB, kb);
                static_for<ax + bx + 1>([&](auto t) {
                  static_for<ay + by + 1>([&](auto u) {
                    static_for<az + bz + 1>([&](auto v) {
                      G[hix(t, u, v)] += (double)(wa * wb) * Ex[ei(ax, bx, t)] * Ey[ei(ay, by, u)] * Ez[ei(az, bz, v)];
                    });
                  });
                });
              }
            });
          }
        });
      });
    });
  };
  double dsum = 0.0;
  for (int i1 = 0; i1 < npa; ++i1)
    for (int j1 = 0; j1 < npb; ++j1) {
      const double a1 = ex[pa0 + i1], b1 = ex[pb0 + j1];
      const double p = a1 + b1, mu1 = a1 * b1 / p;
      if (mu1 * AB2 > 41.4) continue;   // exp(-mu R^2) < 1e-18, as the host's pair list
      const double K1 = cf[pa0 + i1] * cf[pb0 + j1] * exp(-mu1 * AB2);
      const double P1x = (a1 * Ax + b1 * Bx) / p, P1y = (a1 * Ay + b1 * By) / p, P1z = (a1 * Az + b1 * Bz) / p;
      double G1[nh(L)];
      {
        double Ex[NE], Ey[NE], Ez[NE];
        herm_E_t<LA, LB>(P1x - Ax, P1x - Bx, 0.5 / p, Ex);
        herm_E_t<LA, LB>(P1y - Ay, P1y - By, 0.5 / p, Ey);
        herm_E_t<LA, LB>(P1z - Az, P1z - Bz, 0.5 / p, Ez);
        hermite(Ex, Ey, Ez, G1);
      }
      for (int i2 = 0; i2 < npa; ++i2)
        for (int j2 = 0; j2 < npb; ++j2) {
          const double a2 = ex[pa0 + i2], b2 = ex[pb0 + j2];
          const double q = a2 + b2, mu2 = a2 * b2 / q;
          if (mu2 * AB2 > 41.4) continue;
          const double K2 = cf[pa0 + i2] * cf[pb0 + j2] * exp(-mu2 * AB2);
          const double P2x = (a2 * Ax + b2 * Bx) / q, P2y = (a2 * Ay + b2 * By) / q, P2z = (a2 * Az + b2 * Bz) / q;
          double G2[nh(L)];
          {
            double Ex[NE], Ey[NE], Ez[NE];
            herm_E_t<LA, LB>(P2x - Ax, P2x - Bx, 0.5 / q, Ex);
            herm_E_t<LA, LB>(P2y - Ay, P2y - By, 0.5 / q, Ey);
            herm_E_t<LA, LB>(P2z - Az, P2z - Bz, 0.5 / q, Ez);
            hermite(Ex, Ey, Ez, G2);
          }
          double R[nh(2 * L)];
          herm_R_t<2 * L>(p * q / (p + q), P1x - P2x, P1y - P2y, P1z - P2z, A.boys, R);
          double acc = 0.0;
          static_for<nh(L)>([&](auto h1) {
            constexpr int t = ht(h1), u = hu(h1), v = hv(h1);
            double inner = 0.0;
            static_for<nh(L)>([&](auto h2) {
              constexpr int t2 = ht(h2), u2 = hu(h2), v2 = hv(h2);
              constexpr double sg = ((t2 + u2 + v2) & 1) ? -1.0 : 1.0;
              inner += sg * G2[h2] * R[hix(t + t2, u + u2, v + v2)];
            });
            acc += G1[h1] * inner;
          });
          dsum += 2.0 * PI * PI * sqrt(PI) / (p * q * sqrt(p + q)) * K1 * K2 * acc;
        }
    }
  const double f = A.sinv[LA][sa] * A.sinv[LB][sb];
  A.diag[pidx * NS + sc] = dsum * f * f;
}

// Q[a][b] = Q[b][a] = sqrt(sum over the components of |diag|) of the class's pairs
__global__ void k_schwarz_ao_reduce(const double* __restrict__ diag, int ns_comp, const int* __restrict__ sha,
                                    const int* __restrict__ shb, int nsha, int nshb, int same, int nshell,
                                    double* __restrict__ Q)
{
  const int64_t pidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx >= (int64_t)nsha * nshb) return;
  const int ia = (int)(pidx / nshb), ib = (int)(pidx % nshb);
  if (same && ib > ia) return;
  double n = 0.0;
  for (int s = 0; s < ns_comp; ++s) n += fabs(diag[pidx * ns_comp + s]);
  const int a = sha[ia], b = shb[ib];
  Q[(int64_t)a * nshell + b] = sqrt(n);
  Q[(int64_t)b * nshell + a] = sqrt(n);
}

namespace {

std::vector<std::vector<int>> shells_by_degree(const Basis& b, int lmax)
{
  std::vector<std::vector<int>> out(lmax + 1);
  for (int s = 0; s < (int)b.shells.size(); ++s) out[b.shells[s].l].push_back(s);
  return out;
}

unsigned blocks_of(int64_t threads)
{
  const int64_t nb = (threads + 63) / 64;
  if (nb >= ((int64_t)1 << 31)) throw std::runtime_error("GPU integrals: too many shell pairs");
  return (unsigned)nb;
}

}  // namespace

void ints_1e_device(Ctx& c, const Basis& bas, const std::vector<Charge>& charges, double* S, double* T, double* V)
{
  HIP_CHECK(hipSetDevice(c.device));
  const int n = bas.nbf;
  for (const auto& s : bas.shells)
    if (s.l > AO_L) throw std::runtime_error("GPU integrals: AO shells above d are not supported");
  std::vector<int> np, po, off;
  std::vector<double> O, ex, cf, chg;
  for (int s = 0; s < (int)bas.shells.size(); ++s) {
    const Shell& sh = bas.shells[s];
    np.push_back((int)sh.alpha.size());
    po.push_back((int)ex.size());
    off.push_back(bas.shell_off[s]);
    for (int d = 0; d < 3; ++d) O.push_back(sh.O[d]);
    ex.insert(ex.end(), sh.alpha.begin(), sh.alpha.end());
    cf.insert(cf.end(), sh.coeff.begin(), sh.coeff.end());
  }
  for (const auto& q : charges) {
    chg.push_back(q.Z);
    for (int d = 0; d < 3; ++d) chg.push_back(q.xyz[d]);
  }
  if (chg.empty()) chg.assign(4, 0.0);
  const auto byl = shells_by_degree(bas, AO_L);
  std::vector<int> shl, shl0(AO_L + 2, 0);
  for (int l = 0; l <= AO_L; ++l) {
    shl0[l] = (int)shl.size();
    shl.insert(shl.end(), byl[l].begin(), byl[l].end());
  }
  shl0[AO_L + 1] = (int)shl.size();
  DevBuf d_np = upload_vec(np, c.stream), d_po = upload_vec(po, c.stream), d_off = upload_vec(off, c.stream),
         d_O = upload_vec(O, c.stream), d_ex = upload_vec(ex, c.stream), d_cf = upload_vec(cf, c.stream),
         d_chg = upload_vec(chg, c.stream), d_shl = upload_vec(shl, c.stream),
         d_boys = upload_vec(boys_table(), c.stream);
  const size_t bytes = (size_t)n * n * sizeof(double);
  DevBuf dS(bytes), dT(bytes), dV(bytes);
  I1eArgs a{};
  solid_harmonic_norms(a.sinv);
  a.np = d_np.as<int>();
  a.po = d_po.as<int>();
  a.O = d_O.as<double>();
  a.ex = d_ex.as<double>();
  a.cf = d_cf.as<double>();
  a.off = d_off.as<int>();
  a.chg = d_chg.as<double>();
  a.nchg = (int)charges.size();
  a.boys = d_boys.as<double>();
  a.S = dS.as<double>();
  a.T = dT.as<double>();
  a.V = dV.as<double>();
  a.n = n;
  for (int la = 0; la <= AO_L; ++la)
    for (int lb = 0; lb <= la; ++lb) {
      a.sha = d_shl.as<int>() + shl0[la];
      a.shb = d_shl.as<int>() + shl0[lb];
      a.nsha = shl0[la + 1] - shl0[la];
      a.nshb = shl0[lb + 1] - shl0[lb];
      if (a.nsha == 0 || a.nshb == 0) continue;
      const dim3 grid(blocks_of((int64_t)a.nsha * a.nshb));
#define MCHIP_I1E(LA, LB) \
  if (la == LA && lb == LB) hipLaunchKernelGGL((k_int1e_cls<LA, LB>), grid, dim3(64), 0, c.stream, a);
      MCHIP_I1E(0, 0) MCHIP_I1E(1, 0) MCHIP_I1E(1, 1) MCHIP_I1E(2, 0) MCHIP_I1E(2, 1) MCHIP_I1E(2, 2)
#undef MCHIP_I1E
      HIP_CHECK(hipGetLastError());
    }
  d2h(S, dS.p, bytes, c.stream);
  d2h(T, dT.p, bytes, c.stream);
  d2h(V, dV.p, bytes, c.stream);
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

void ints_2c2e_device(Ctx& c, const Basis& aux, double* M)
{
  HIP_CHECK(hipSetDevice(c.device));
  const int n = aux.nbf, ns = (int)aux.shells.size();
  std::vector<double> g(ns), O(3 * ns), sc(ns);
  std::vector<int> off(ns);
  for (int s = 0; s < ns; ++s) {
    const Shell& sh = aux.shells[s];
    if (sh.alpha.size() != 1) throw std::runtime_error("GPU integrals: the auxiliary basis must be uncontracted");
    if (sh.l > LMAX) throw std::runtime_error("GPU integrals: auxiliary shells above f are not supported");
    g[s] = sh.alpha[0];
    off[s] = aux.shell_off[s];
    for (int d = 0; d < 3; ++d) O[3 * s + d] = sh.O[d];
    double v = sh.coeff[0];
    for (int i = 0; i < sh.l; ++i) v *= 0.5 / sh.alpha[0];
    sc[s] = (sh.l & 1) ? -v : v;
  }
  const auto byl = shells_by_degree(aux, LMAX);
  std::vector<int> shl, shl0(LMAX + 2, 0);
  for (int l = 0; l <= LMAX; ++l) {
    shl0[l] = (int)shl.size();
    shl.insert(shl.end(), byl[l].begin(), byl[l].end());
  }
  shl0[LMAX + 1] = (int)shl.size();
  DevBuf d_g = upload_vec(g, c.stream), d_O = upload_vec(O, c.stream), d_s = upload_vec(sc, c.stream),
         d_off = upload_vec(off, c.stream), d_shl = upload_vec(shl, c.stream),
         d_boys = upload_vec(boys_table(), c.stream);
  const size_t bytes = (size_t)n * n * sizeof(double);
  DevBuf dM(bytes);
  I2cArgs a{};
  solid_harmonic_norms(a.sinv);
  a.g = d_g.as<double>();
  a.O = d_O.as<double>();
  a.s = d_s.as<double>();
  a.off = d_off.as<int>();
  a.boys = d_boys.as<double>();
  a.M = dM.as<double>();
  a.n = n;
  for (int lx = 0; lx <= LMAX; ++lx)
    for (int ly = 0; ly <= lx; ++ly) {
      a.shx = d_shl.as<int>() + shl0[lx];
      a.shy = d_shl.as<int>() + shl0[ly];
      a.nshx = shl0[lx + 1] - shl0[lx];
      a.nshy = shl0[ly + 1] - shl0[ly];
      if (a.nshx == 0 || a.nshy == 0) continue;
      const dim3 grid(blocks_of((int64_t)a.nshx * a.nshy));
#define MCHIP_I2C(LX, LY) \
  if (lx == LX && ly == LY) hipLaunchKernelGGL((k_int2c2e_cls<LX, LY>), grid, dim3(64), 0, c.stream, a);
      MCHIP_I2C(0, 0) MCHIP_I2C(1, 0) MCHIP_I2C(1, 1) MCHIP_I2C(2, 0) MCHIP_I2C(2, 1) MCHIP_I2C(2, 2)
      MCHIP_I2C(3, 0) MCHIP_I2C(3, 1) MCHIP_I2C(3, 2) MCHIP_I2C(3, 3)
#undef MCHIP_I2C
      HIP_CHECK(hipGetLastError());
    }
  d2h(M, dM.p, bytes, c.stream);
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

void schwarz_ao_device(Ctx& c, const Basis& bas, double* Q)
{
  HIP_CHECK(hipSetDevice(c.device));
  const int ns = (int)bas.shells.size();
  for (const auto& sh : bas.shells)
    if (sh.l > AO_L) throw std::runtime_error("GPU integrals: AO shells above d are not supported");
  std::vector<int> np, po;
  std::vector<double> O, ex, cf;
  for (const auto& sh : bas.shells) {
    np.push_back((int)sh.alpha.size());
    po.push_back((int)ex.size());
    for (int d = 0; d < 3; ++d) O.push_back(sh.O[d]);
    ex.insert(ex.end(), sh.alpha.begin(), sh.alpha.end());
    cf.insert(cf.end(), sh.coeff.begin(), sh.coeff.end());
  }
  const auto byl = shells_by_degree(bas, AO_L);
  std::vector<int> shl, shl0(AO_L + 2, 0);
  for (int l = 0; l <= AO_L; ++l) {
    shl0[l] = (int)shl.size();
    shl.insert(shl.end(), byl[l].begin(), byl[l].end());
  }
  shl0[AO_L + 1] = (int)shl.size();
  if (ns == 0) return;
  DevBuf d_np = upload_vec(np, c.stream), d_po = upload_vec(po, c.stream), d_O = upload_vec(O, c.stream),
         d_ex = upload_vec(ex, c.stream), d_cf = upload_vec(cf, c.stream), d_shl = upload_vec(shl, c.stream),
         d_boys = upload_vec(boys_table(), c.stream);
  DevBuf dQ((size_t)ns * ns * sizeof(double));
  SchArgs a{};
  solid_harmonic_norms(a.sinv);
  a.np = d_np.as<int>();
  a.po = d_po.as<int>();
  a.O = d_O.as<double>();
  a.ex = d_ex.as<double>();
  a.cf = d_cf.as<double>();
  a.boys = d_boys.as<double>();
  for (int la = 0; la <= AO_L; ++la)
    for (int lb = 0; lb <= la; ++lb) {
      a.sha = d_shl.as<int>() + shl0[la];
      a.shb = d_shl.as<int>() + shl0[lb];
      a.nsha = shl0[la + 1] - shl0[la];
      a.nshb = shl0[lb + 1] - shl0[lb];
      if (a.nsha == 0 || a.nshb == 0) continue;
      const int nsc = (2 * la + 1) * (2 * lb + 1);
      const int64_t npair = (int64_t)a.nsha * a.nshb;
      DevBuf diag((size_t)npair * nsc * sizeof(double));
      a.diag = diag.as<double>();
      const dim3 grid(blocks_of(npair * nsc));
#define MCHIP_SCH(LA, LB) \
  if (la == LA && lb == LB) hipLaunchKernelGGL((k_schwarz_ao_cls<LA, LB>), grid, dim3(64), 0, c.stream, a);
      MCHIP_SCH(0, 0) MCHIP_SCH(1, 0) MCHIP_SCH(1, 1) MCHIP_SCH(2, 0) MCHIP_SCH(2, 1) MCHIP_SCH(2, 2)
#undef MCHIP_SCH
      HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(k_schwarz_ao_reduce, dim3(blocks_of(npair)), dim3(64), 0, c.stream, diag.as<double>(), nsc,
                         a.sha, a.shb, a.nsha, a.nshb, la == lb ? 1 : 0, ns, dQ.as<double>());
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipStreamSynchronize(c.stream));   // diag dies at the end of this scope
    }
  d2h(Q, dQ.p, dQ.bytes, c.stream);
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

}  // namespace mchip
