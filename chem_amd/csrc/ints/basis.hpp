// Shells, contracted-function normalisation and real-solid-harmonic tables.
//
// Follows the reference's shell construction and normalisation:
//   * one shell per (angular momentum, coefficient column) of a BSE JSON
//     "electron_shells" entry, dropping primitives with |c| <= DBL_EPSILON
//     (src/desc/basis.cpp:304-336);
//   * primitive factor 1/sqrt(I(2l+2, 2a)) with I(n,a) = Gamma((n+1)/2)/(2 a^{(n+1)/2})
//     and renormalisation of the contracted radial function
//     (src/ints/aofactory.cpp:273-310).
// The angular factor lives in the Cartesian->spherical table (functions are
// unit-normalised real solid harmonics, libcint order).
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "md.hpp"

namespace mchip {

struct Shell {
  int l = 0;
  int atom = 0;                  // index of the atom the shell sits on
  std::array<double, 3> O{};     // centre (bohr)
  std::vector<double> alpha;     // exponents
  std::vector<double> coeff;     // normalised contraction coefficients
};

struct Basis {
  std::vector<Shell> shells;
  std::vector<int> shell_off;    // first function of each shell
  int nbf = 0;
  int max_l = 0;

  void finalize()
  {
    shell_off.resize(shells.size());
    nbf = 0;
    max_l = 0;
    for (size_t i = 0; i < shells.size(); ++i) {
      if (shells[i].l > md::LMAX)
        throw std::runtime_error("angular momentum above f is not supported");
      shell_off[i] = nbf;
      nbf += 2 * shells[i].l + 1;
      if (shells[i].l > max_l) max_l = shells[i].l;
    }
  }
};

// Normalise contraction coefficients in place (aofactory.cpp:273-310).
void normalize_shell(Shell& s);

// sph[l][m][k]: coefficient of Cartesian monomial k (md::cart_exps order)
// in unit-normalised real solid harmonic m (libcint order) of degree l.
struct SphTable {
  double c[md::LMAX + 1][2 * md::LMAX + 1][(md::LMAX + 1) * (md::LMAX + 2) / 2];
};
const SphTable& sph_table();

// Boys function table (md::BOYS_NT x md::BOYS_NM), computed once.
const std::vector<double>& boys_table();

}  // namespace mchip
