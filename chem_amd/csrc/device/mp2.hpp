// SOS-AO-MP2 device driver.
#pragma once

#include <vector>

#include "dev.hpp"

namespace mchip {

struct Mp2Input {
  int nbf = 0, nocc = 0;
  const double* S = nullptr;     // host overlap, row-major
  const double* C = nullptr;     // host MO coefficients, row-major nbf x nbf (occupied first)
  const double* eps = nullptr;   // host orbital energies
  std::vector<double> omega, alpha;
  std::vector<int> bsz;          // AO block sizes
  int mo_split = 5;
  double filter_eps = 1e-10;
  // Z builder (mp::zmethod, z_builder.hpp): 0 LLMP_FULL, 1 LLMP_MEM (the same
  // contractions batched differently; one GPU path), 2 LL_FULL (B(X|ia) route
  // with a pivoted Cholesky factor of the virtual pseudo-density)
  int zmethod = 0;
};

// Z of one Laplace point from given factors (mp::Z::compute, z_builder.hpp):
//   zmethod 0/1: Z = LLMP_FULL_Z(L_o, P_v)   (z_full.cpp:69-397, z_mem.cpp:37-)
//   zmethod 2:   Z = LL_FULL_Z(L_o, L_v)     (z_full.cpp:428-692)
// Lo: device n x ro row-major; Pv: device n x n (zmethod 0/1); Lv: device
// n x rv row-major (zmethod 2).  Z: device naux x naux, full symmetric.
void z_point(Ctx& c, DevBT& bt, int zmethod, const double* Lo, int ro, const double* Pv, const double* Lv, int rv,
             double* Z);

// Pivoted Cholesky of the symmetric positive semidefinite N x N matrix in U
// (overwritten): returns L (N x rank, row-major, rows in the original order,
// columns in LMO position order) with U = L L^T to N * eps * max|diag|
// (math::pivinc_cd, piv_cd.cpp:125-483).
std::vector<double> pivoted_cholesky_dev(Ctx& c, DevBuf& U, int N, double filter_eps, int* rank_out,
                                         bool* neg_out = nullptr);

// Returns E = sum_k e_k (positive, unscaled); terms[k] = e_k; t_z = seconds
// spent in the Z builder (M2-M5) over all points.  Z_host, if not null,
// receives Z_k(X,Y) of every point (nlap x naux x naux, zero for a point whose
// occupied pseudo-density has rank 0).
double mp2_sos(Ctx& c, DevBT& bt, const Metric& met, const Mp2Input& in, std::vector<double>& terms,
               double* t_z, double* Z_host = nullptr);

// Same energy with the contractions of all Laplace points grouped and captured
// in one hipGraph, FP32 with FP64 accumulation.  t_z = seconds of the graph
// launch (M2-M5 of every point), t_prep = seconds of the FP64 per-point set-up
// (pseudo-densities, pivoted Cholesky, FP32 copies, autotuning).
double mp2_sos_batched(Ctx& c, DevBT& bt, const Metric& met, const Mp2Input& in, std::vector<double>& terms,
                       double* t_z, double* t_prep, double* Z_host = nullptr);

}  // namespace mchip
