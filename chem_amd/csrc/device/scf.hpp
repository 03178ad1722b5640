// Restricted HF driver on the GPU.
#pragma once

#include <functional>
#include <vector>

#include "dev.hpp"

namespace mchip {

struct ScfInput {
  int nbf = 0, nocc = 0;
  std::vector<int> bsz;          // AO block sizes (for the block filter)
  const double* S = nullptr;     // host, row-major
  const double* H = nullptr;     // host, T + V
  double enuc = 0.0;
  double threshold = 1e-6;
  int max_iter = 100;
  bool do_diis = true;
  int diis_start = 1, diis_min = 2, diis_max = 10;
  double filter_eps = 0.0;
  // SAD guess (hfguess.cpp:142-452): the alpha density of the first
  // iteration (row-major nbf x nbf; empty: the core guess).  Its first
  // exchange contracts every column of the density's pivoted Cholesky factor
  // (the SAD iteration of DFMO_K, jkdf.cpp:243-246).
  std::vector<double> guess_P;
};

struct ScfOutput {
  bool converged = false;
  int iterations = 0;
  double energy = 0.0;
  std::vector<double> C, eps, P, F;
  std::vector<double> history_e, history_rms, t_j, t_k;
  double energy_iter = 0.0;      // energy of the last iteration's own K
  double t_final_k = 0.0;        // seconds of the final FP64 K build (if any)
};

// K builder callback: (iteration, RMS of the previous iteration's error matrix
// (infinity before the first), P_dev, Cocc_dev row-major nbf x nocc, K_dev out).
// Returns true when this K comes from a different exchange operator than the
// previous iteration's (a precision switch): the SCF then restarts DIIS, whose
// stored Fock matrices and errors belong to the old operator.
using KBuildFn = std::function<bool(int, double, const double*, const double*, double*)>;

// Exchange builders of an SCF run for kmode (0 plain DFMO with the metric per
// build, 1 DFMO on the transformed B, 2 FP32, 3 BF16x3, 4 BF16x3 escalating to
// FP32 once the previous error RMS is <= c.escalate_rms or, below 1e-4, has not
// fallen below 0.8 x its lowest value for c.escalate_stall iterations, and from
// FP32 to FP64 after such a stall below 1e-6 (straight from BF16x3 to FP64 when
// the tensor has no FP32 copy); modes 2-4 get an FP64 final builder, with final_occ_nbatches occupied batches,
// 0 = occ_nbatches).  The builders capture c, bt and m by reference.
// tr(K P) of the FP64 exchange at occupied coefficients Cocc (device,
// row-major nbf x nocc), without K.
using EnergyFn = std::function<double(const double*)>;

// kenergy (optional): when the final FP64 exchange runs on the FP64 copy of
// L^-1 B in one occupied batch, the function that gives its tr(K P) directly
// (the reported energy needs K only through that trace); empty otherwise.
void select_k_builders(Ctx& c, DevBT& bt, const Metric* m, int kmode, int nocc, int occ_nbatches,
                       double threshold, KBuildFn& kb, KBuildFn& kfinal, int final_occ_nbatches = 0,
                       EnergyFn* kenergy = nullptr);

// final_k (optional): FP64 K builder used once at the converged orbitals to
// evaluate the reported energy when kbuild works in reduced precision;
// final_ek (optional, used instead of final_k): tr(K P) of that exchange.
// With final_ek the returned F holds the last iteration's exchange.
void scf_run(Ctx& c, DevBT& bt, const Metric* m, const ScfInput& in, ScfOutput& out,
             const KBuildFn& kbuild, const KBuildFn& final_k = KBuildFn(), const EnergyFn& final_ek = EnergyFn());

}  // namespace mchip
