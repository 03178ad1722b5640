// Host-side AO integral drivers (OpenMP).
#pragma once

#include <cstdint>
#include <vector>

#include "basis.hpp"
#include "btensor.hpp"

namespace mchip {

struct Charge {
  double Z;
  double xyz[3];
};

void boys_eval(int mmax, double T, double* F);

// Dense row-major nbf x nbf matrices; any output pointer may be null.
void ints_1e(const Basis& bas, const std::vector<Charge>& charges,
             double* S, double* T, double* V);

// Dense row-major naux x naux Coulomb metric (X|Y).
void ints_2c2e(const Basis& aux, double* V);

// Per-shell-pair Schwarz factors sqrt(sum |(ab|ab)|) (nshell x nshell, row-major)
// and per-aux-shell factors sqrt(sum |(X|X)|).
void schwarz_ao(const Basis& bas, double* Q);
void schwarz_aux(const Basis& aux, double* Q);

// Ket-side Hermite coefficients of one uncontracted aux primitive, times
// (-1)^{t+u+v}: H[m*stride + herm_index(t,u,v)].
void aux_hermite_coeffs(int l, double gamma, double coef, double* H, int stride);

// One shell triple (c | a b), layout out[m_c][sa][sb].
void ints_3c2e_shell(const Basis& ao, const Basis& aux, int ia, int ib, int ic, double* out);

// Four-centre (ab|cd) of one shell quartet, out[sa][sb][sc][sd], and every
// (mu nu|lambda sigma) of a small basis (row-major nbf^4).
void ints_4c2e_shell(const Basis& bas, int ia, int ib, int ic, int id, double* out);
void ints_4c2e_all(const Basis& bas, double* out);

// Screened, filtered, slab-laid-out B(X|mu nu).
void build_B(const Basis& ao, const Basis& aux,
             const std::vector<int>& ao_blk_shells,
             const std::vector<int>& aux_blk_shells,
             const std::vector<uint8_t>& schwarz_mask,
             double filter_eps,
             BLayout& layout, std::vector<double>& data,
             std::vector<uint8_t>& final_mask);

}  // namespace mchip
