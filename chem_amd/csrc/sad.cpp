// Superposition of atomic densities (the reference's default SCF guess "SAD",
// hfmod::compute_guess, src/hf/hfguess.cpp:142-452).
//
// For every element present, the atom alone in its own shells of the
// molecule's basis runs a restricted SCF with fractional, spin-averaged
// occupations (desc::molecule with fractional(true) and spin_average(true),
// src/desc/molecule.cpp:37-93) and exact four-centre J and K: the core guess,
// no DIIS, convergence at RMS(FPS - SPF) < 1e-6 (the SAD_* defaults,
// src/hf/hfmod.hpp:50-53).  The atomic alpha densities are placed on the
// diagonal blocks of every atom of that element.  Host code (OpenMP in the
// integrals); the matrices are at most a few tens of functions.
#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mchip.h"
#include "capi_common.hpp"
#include "ints/ints_host.hpp"

namespace mchip {

// Occupations of a neutral atom with fractional, spin-averaged occupation
// (molecule.cpp:37-93): Z/2 alpha electrons; the orbitals of the closed
// shells below the atom (conf_orb: the noble-gas cores) hold one alpha
// electron each, the orbitals of the shell being filled share the rest
// evenly.  occ[i] is the factor applied to the i-th occupied orbital's
// coefficients (the square root of its alpha occupation).
void sad_occupations(int Z, int& nocc, std::vector<double>& occ)
{
  static const int conf_orb[] = {0, 2, 10, 18, 36, 54, 86, 118};
  if (Z < 1 || Z > 118) throw std::runtime_error("SAD: atomic number out of range");
  const double nele_a = 0.5 * Z;
  const int* end = conf_orb + 8;
  const int* lim = std::lower_bound(conf_orb, end, Z);
  if (Z < *lim) --lim;
  int ncore, nact;
  if (*lim == Z) {
    ncore = 0;
    nact = *lim / 2;
  }
  else {
    ncore = *lim / 2;
    nact = *(lim + 1) / 2 - ncore;
  }
  nocc = nact + ncore;
  const double frac = std::sqrt((nele_a - ncore) / nact);
  occ.assign(nocc, 1.0);
  for (int i = ncore; i < nocc; ++i) occ[i] = frac;
}

// Eigenvalues (ascending) and eigenvectors (columns of V, row-major n x n) of
// a small symmetric matrix by cyclic Jacobi rotations.
static void jacobi_eigen(std::vector<double> A, int n, std::vector<double>& w, std::vector<double>& V)
{
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double a = A[(size_t)i * n + j] * A[(size_t)i * n + j];
        tot += a;
        if (i != j) off += a;
      }
    if (off <= 1e-30 * std::max(tot, 1e-300)) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        const double theta = 0.5 * (aqq - app) / apq;
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return A[(size_t)a * n + a] < A[(size_t)b * n + b]; });
  w.resize(n);
  std::vector<double> Vs((size_t)n * n);
  for (int j = 0; j < n; ++j) {
    w[j] = A[(size_t)idx[j] * n + idx[j]];
    for (int i = 0; i < n; ++i) Vs[(size_t)i * n + j] = V[(size_t)i * n + idx[j]];
  }
  V.swap(Vs);
}

static void matmul(const std::vector<double>& A, const std::vector<double>& B, int n, std::vector<double>& C,
                   bool transA = false)
{
  C.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const double a = transA ? A[(size_t)k * n + i] : A[(size_t)i * n + k];
      if (a == 0.0) continue;
      for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += a * B[(size_t)k * n + j];
    }
}

struct AtomScf {
  std::vector<double> P;   // alpha density, row-major n x n
  double energy = 0.0;     // electronic energy tr(HP) + tr(FP)
  int iterations = 0;
};

// The atomic SCF of one atom (shells `bas`, nucleus Z at xyz): hfmod::compute
// (hfscf.cpp:326-552) with the core guess, exact J and K, no DIIS, restricted
// (equal alpha and beta counts).
static AtomScf atomic_scf(const Basis& bas, double Z, const double* xyz, double thr, int max_iter)
{
  const int n = bas.nbf;
  std::vector<double> S((size_t)n * n), T((size_t)n * n), V((size_t)n * n), H((size_t)n * n);
  std::vector<Charge> ch(1);
  ch[0].Z = Z;
  for (int d = 0; d < 3; ++d) ch[0].xyz[d] = xyz[d];
  ints_1e(bas, ch, S.data(), T.data(), V.data());
  for (size_t i = 0; i < H.size(); ++i) H[i] = T[i] + V[i];
  std::vector<double> eri((size_t)n * n * n * n);
  ints_4c2e_all(bas, eri.data());
  int nocc = 0;
  std::vector<double> occ;
  sad_occupations((int)std::lround(Z), nocc, occ);
  if (nocc > n) throw std::runtime_error("SAD: the atom's basis has fewer functions than occupied orbitals");
  // orthogonalizer X = U d U^T, d = 0 for eigenvalues < 0 else 1/sqrt
  // (orthogonalizer.cpp:12-45; its integer threshold is 0)
  std::vector<double> sw, SU;
  jacobi_eigen(S, n, sw, SU);
  std::vector<double> X((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += SU[(size_t)i * n + k] * (sw[k] < 0.0 ? 0.0 : 1.0 / std::sqrt(sw[k])) *
                                       SU[(size_t)j * n + k];
      X[(size_t)i * n + j] = s;
    }
  std::vector<double> P((size_t)n * n, 0.0), F(H), W1, W2;
  // diag_fock (hfdiagfock.cpp:12-211): F' = X^T F X, C = X C', occupied
  // columns scaled by the occupation factors, P = C_occ C_occ^T
  auto diag = [&](const std::vector<double>& Fm) {
    matmul(Fm, X, n, W1);
    matmul(X, W1, n, W2, true);
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        const double v = 0.5 * (W2[(size_t)i * n + j] + W2[(size_t)j * n + i]);
        W2[(size_t)i * n + j] = W2[(size_t)j * n + i] = v;
      }
    std::vector<double> e, Cp, C;
    jacobi_eigen(W2, n, e, Cp);
    matmul(X, Cp, n, C);
    std::fill(P.begin(), P.end(), 0.0);
    for (int k = 0; k < nocc; ++k) {
      const double f = occ[k] * occ[k];
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) P[(size_t)i * n + j] += f * C[(size_t)i * n + k] * C[(size_t)j * n + k];
    }
  };
  diag(H);
  AtomScf out;
  for (int iter = 0;; ++iter) {
    // F = H + J[2P] - K[P]
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double jv = 0.0, kv = 0.0;
        for (int k = 0; k < n; ++k)
          for (int l = 0; l < n; ++l) {
            const double p = P[(size_t)k * n + l];
            jv += eri[(((size_t)i * n + j) * n + k) * n + l] * p;
            kv += eri[(((size_t)i * n + k) * n + j) * n + l] * p;
          }
        F[(size_t)i * n + j] = H[(size_t)i * n + j] + 2.0 * jv - kv;
      }
    // e = F P S - S P F, E = tr(HP) + tr(FP)
    matmul(F, P, n, W1);
    matmul(W1, S, n, W2);
    double rms = 0.0, e = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double d = W2[(size_t)i * n + j] - W2[(size_t)j * n + i];
        rms += d * d;
        e += (H[(size_t)i * n + j] + F[(size_t)i * n + j]) * P[(size_t)j * n + i];
      }
    rms = std::sqrt(rms / ((double)n * n));
    out.energy = e;
    out.iterations = iter;
    if (rms < thr && iter > 0) break;
    if (iter > max_iter) throw std::runtime_error("SAD: atomic HF did not converge");
    diag(F);
  }
  out.P = P;
  return out;
}

// The SAD alpha density of a molecule (row-major nbf x nbf) from its basis
// (every atom's shells consecutive, in atom order).
void sad_density(const Basis& bas, const std::vector<double>& Z, const std::vector<double>& xyz, double* P_alpha,
                 double thr = 1e-6, int max_iter = 100)
{
  const int natom = (int)Z.size(), n = bas.nbf;
  std::vector<int> first(natom, -1), count(natom, 0);
  int prev = -1;
  for (size_t s = 0; s < bas.shells.size(); ++s) {
    const int a = bas.shells[s].atom;
    if (a < 0 || a >= natom) throw std::runtime_error("SAD: shell on an unknown atom");
    if (a != prev) {
      if (first[a] >= 0) throw std::runtime_error("SAD: the shells of an atom are not consecutive");
      first[a] = (int)s;
      prev = a;
    }
    ++count[a];
  }
  std::fill(P_alpha, P_alpha + (size_t)n * n, 0.0);
  std::map<int, AtomScf> dens;   // per element, from its first atom (hfguess.cpp:146-160)
  for (int a = 0; a < natom; ++a) {
    if (count[a] == 0) continue;
    const int z = (int)std::lround(Z[a]);
    if (!dens.count(z)) {
      Basis ab;
      for (int s = first[a]; s < first[a] + count[a]; ++s) ab.shells.push_back(bas.shells[s]);
      ab.finalize();
      dens[z] = atomic_scf(ab, Z[a], &xyz[3 * a], thr, max_iter);
    }
    const AtomScf& d = dens[z];
    const int off = bas.shell_off[first[a]];
    const int na = (int)std::lround(std::sqrt((double)d.P.size()));
    if (off + na > n) throw std::runtime_error("SAD: atom block outside the basis");
    for (int i = 0; i < na; ++i)
      for (int j = 0; j < na; ++j) P_alpha[(size_t)(off + i) * n + off + j] = d.P[(size_t)i * na + j];
  }
}

}  // namespace mchip

// ---------------------------------------------------------------------------
// C ABI (declared in include/mchip.h)
// ---------------------------------------------------------------------------
namespace mchip {
const Basis& basis_of(mchip_basis b);
}
using namespace mchip;

extern "C" int mchip_sad_guess(mchip_basis ao, int natom, const double* Z, const double* xyz, double* P_alpha)
{
  MCHIP_TRY
  if (natom <= 0) throw std::runtime_error("mchip_sad_guess: no atoms");
  std::vector<double> z(Z, Z + natom), x(xyz, xyz + 3 * natom);
  sad_density(basis_of(ao), z, x, P_alpha);
  MCHIP_CATCH
}

extern "C" int mchip_sad_atom(mchip_basis atom_basis, double Z, const double* xyz, double threshold, double* P,
                              double* energy, int* iterations, int* nocc, double* occ)
{
  MCHIP_TRY
  const Basis& b = basis_of(atom_basis);
  AtomScf r = atomic_scf(b, Z, xyz, threshold, 100);
  std::copy(r.P.begin(), r.P.end(), P);
  if (energy) *energy = r.energy;
  if (iterations) *iterations = r.iterations;
  int no = 0;
  std::vector<double> oc;
  sad_occupations((int)std::lround(Z), no, oc);
  if (nocc) *nocc = no;
  if (occ) std::copy(oc.begin(), oc.end(), occ);
  MCHIP_CATCH
}

extern "C" int mchip_int4c2e(mchip_basis b, double* out)
{
  MCHIP_TRY
  ints_4c2e_all(basis_of(b), out);
  MCHIP_CATCH
}
