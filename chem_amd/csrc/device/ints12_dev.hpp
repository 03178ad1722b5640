// One-electron integrals and the two-centre Coulomb metric on the GPU.
#pragma once

#include <vector>

#include "../ints/basis.hpp"
#include "../ints/ints_host.hpp"
#include "dev.hpp"

namespace mchip {

// S, T (kinetic) and V (nuclear attraction of the charges), row-major nbf x
// nbf, into host arrays (same integrals as ints_1e).
void ints_1e_device(Ctx& c, const Basis& bas, const std::vector<Charge>& charges, double* S, double* T, double* V);

// (X|Y), row-major naux x naux, into a host array (same integrals as ints_2c2e).
void ints_2c2e_device(Ctx& c, const Basis& aux, double* M);

// AO Schwarz factors Q[a][b] = sqrt(sum over the shell pair's components of
// |(ab|ab)|), nshell x nshell, into a host array (same as schwarz_ao).
void schwarz_ao_device(Ctx& c, const Basis& bas, double* Q);

}  // namespace mchip
