// mchip_fock.hpp -- C++ builder layer over the mchip C ABI, with the names and
// call pattern of the reference's fock builders (src/fock/jkbuilder.hpp):
//
//   auto jbuilder = mchip::fock::create_j()
//                       .method(mchip::fock::jmethod::dfao)
//                       .eri3c2e_batched(b)          // mchip_dbt on the GPU
//                       .metric(m)                   // mchip_metric (V = L L^T)
//                       .build();
//   jbuilder->set_density_alpha(p);
//   jbuilder->compute_J();
//   auto J = jbuilder->get_J();
//
// Host matrices are dense row-major (mchip::fock::matrix), standing in for
// dbcsr::shared_matrix<double>; block structure lives in the B tensor.
// Failures are rethrown as std::runtime_error with the library's message, as
// the reference throws on every error.  Header-only; link with libmchip.so.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mchip.h"

namespace mchip {
namespace fock {

inline void check(int rc)
{
  if (rc != MCHIP_OK) throw std::runtime_error(mchip_last_error());
}

// jkbuilder.hpp:20-22.  Only the density-fitted builders are provided.
enum class jmethod { dfao };
enum class kmethod { dfao, dfmo, dfmem, dflmo };

// Precision of the exchange contractions: the reference's FP64 path, or the
// chosen forms on the transformed tensor (FP32, BF16 hi/lo split).
enum class precision { fp64, fp32, bf16x3 };

struct matrix {
  int rows = 0, cols = 0;
  std::vector<double> data;   // row-major
  matrix() = default;
  matrix(int r, int c) : rows(r), cols(c), data((size_t)r * c, 0.0) {}
  double* ptr() { return data.data(); }
  const double* ptr() const { return data.data(); }
};
using shared_matrix = std::shared_ptr<matrix>;

// JK_common (jkbuilder.hpp:62-150): densities and coefficients per spin.
class JK_common {
 protected:
  shared_matrix m_p_A, m_p_B, m_c_A, m_c_B;
  bool m_sym = true;

 public:
  virtual ~JK_common() = default;
  void set_density_alpha(shared_matrix p) { m_p_A = std::move(p); }
  void set_density_beta(shared_matrix p) { m_p_B = std::move(p); }
  void set_coeff_alpha(shared_matrix c) { m_c_A = std::move(c); }
  void set_coeff_beta(shared_matrix c) { m_c_B = std::move(c); }
  void set_sym(bool sym) { m_sym = sym; }
};

// J (jkbuilder.hpp:152-178)
class J : public JK_common {
 protected:
  shared_matrix m_J;

 public:
  virtual void init() = 0;
  virtual void compute_J() = 0;
  shared_matrix get_J() { return m_J; }
};

// K (jkbuilder.hpp:180-217); K is returned negated, as the reference does.
class K : public JK_common {
 protected:
  shared_matrix m_K_A, m_K_B;

 public:
  virtual void init() = 0;
  virtual void compute_K() = 0;
  shared_matrix get_K_A() { return m_K_A; }
  shared_matrix get_K_B() { return m_K_B; }
};

// DF_J::compute_J (jkdf.cpp:69-190): P_tot = 2 P_A (closed shell) or P_A + P_B.
class DF_J : public J {
  mchip_dbt m_b;
  mchip_metric m_metric;

 public:
  DF_J(mchip_dbt b, mchip_metric m) : m_b(b), m_metric(m) {}
  void init() override {}
  void compute_J() override
  {
    if (!m_p_A) throw std::runtime_error("DF_J: density alpha not set");
    const int n = m_p_A->rows;
    m_J = std::make_shared<matrix>(n, n);
    if (m_p_B) {
      matrix ptot(n, n);
      for (size_t i = 0; i < ptot.data.size(); ++i) ptot.data[i] = m_p_A->data[i] + m_p_B->data[i];
      check(mchip_dfj(m_b, m_metric, ptot.ptr(), 1.0, m_J->ptr()));
    }
    else {
      check(mchip_dfj(m_b, m_metric, m_p_A->ptr(), 2.0, m_J->ptr()));
    }
  }
};

// DFMO_K::compute_K (jkdf.cpp:221-462) from occupied coefficients per spin.
class DFMO_K : public K {
  mchip_dbt m_b;
  mchip_metric m_metric;
  int m_occ_nbatches;
  precision m_prec;

  void one(const shared_matrix& c, shared_matrix& k)
  {
    const int n = c->rows;
    k = std::make_shared<matrix>(n, n);
    switch (m_prec) {
      case precision::fp64:
        check(mchip_dfk_mo(m_b, m_metric, c->ptr(), c->cols, m_occ_nbatches, k->ptr()));
        break;
      case precision::fp32:
        check(mchip_dfk_mo_f32(m_b, c->ptr(), c->cols, m_occ_nbatches, k->ptr()));
        break;
      case precision::bf16x3:
        check(mchip_dfk_mo_bf16(m_b, c->ptr(), c->cols, m_occ_nbatches, k->ptr()));
        break;
    }
  }

 public:
  DFMO_K(mchip_dbt b, mchip_metric m, int occ_nbatches, precision p)
      : m_b(b), m_metric(m), m_occ_nbatches(occ_nbatches), m_prec(p)
  {
  }
  void init() override {}
  void compute_K() override
  {
    if (!m_c_A) throw std::runtime_error("DFMO_K: coefficients alpha not set");
    one(m_c_A, m_K_A);
    if (m_c_B) one(m_c_B, m_K_B);
  }
};

// DFAO_K (jkdf.cpp:510-648) and DFMEM_K (jkdf.cpp:719-859) from densities.
class DFAO_K : public K {
  mchip_dbt m_b;
  mchip_metric m_metric;
  int m_occ_nbatches;
  precision m_prec;

  void one(const shared_matrix& p, shared_matrix& k)
  {
    const int n = p->rows;
    k = std::make_shared<matrix>(n, n);
    check(mchip_dfk_ao(m_b, m_metric, p->ptr(), (int)m_prec, m_occ_nbatches, k->ptr()));
  }

 public:
  DFAO_K(mchip_dbt b, mchip_metric m, int occ_nbatches, precision p)
      : m_b(b), m_metric(m), m_occ_nbatches(occ_nbatches), m_prec(p)
  {
  }
  void init() override {}
  void compute_K() override
  {
    if (!m_p_A) throw std::runtime_error("DFAO_K: density alpha not set");
    one(m_p_A, m_K_A);
    if (m_p_B) one(m_p_B, m_K_B);
  }
};

// create_j_base (jkbuilder.hpp:621-690)
class create_j_base {
  jmethod c_method = jmethod::dfao;
  mchip_dbt c_b = nullptr;
  mchip_metric c_metric = nullptr;

 public:
  create_j_base& method(jmethod m) { c_method = m; return *this; }
  create_j_base& eri3c2e_batched(mchip_dbt b) { c_b = b; return *this; }
  create_j_base& metric(mchip_metric m) { c_metric = m; return *this; }
  std::shared_ptr<J> build()
  {
    if (!c_b) throw std::runtime_error("create_j: eri3c2e_batched is required");
    return std::make_shared<DF_J>(c_b, c_metric);
  }
};
inline create_j_base create_j() { return create_j_base(); }

// create_k_base (jkbuilder.hpp:692-790).  dflmo's symmetric branch is the
// occupied-factor exchange over all given columns, so it maps to DFMO_K.
class create_k_base {
  kmethod c_method = kmethod::dfmo;
  mchip_dbt c_b = nullptr;
  mchip_metric c_metric = nullptr;
  int c_occ_nbatches = 1;
  precision c_prec = precision::fp64;

 public:
  create_k_base& method(kmethod m) { c_method = m; return *this; }
  create_k_base& eri3c2e_batched(mchip_dbt b) { c_b = b; return *this; }
  create_k_base& metric(mchip_metric m) { c_metric = m; return *this; }
  create_k_base& occ_nbatches(int n) { c_occ_nbatches = n; return *this; }
  create_k_base& contraction_precision(precision p) { c_prec = p; return *this; }
  std::shared_ptr<K> build()
  {
    if (!c_b) throw std::runtime_error("create_k: eri3c2e_batched is required");
    switch (c_method) {
      case kmethod::dfmo:
      case kmethod::dflmo: return std::make_shared<DFMO_K>(c_b, c_metric, c_occ_nbatches, c_prec);
      case kmethod::dfao:
      case kmethod::dfmem: return std::make_shared<DFAO_K>(c_b, c_metric, c_occ_nbatches, c_prec);
    }
    throw std::runtime_error("create_k: unknown method");
  }
};
inline create_k_base create_k() { return create_k_base(); }

}  // namespace fock

// Z builders of SOS-AO-MP2 with the names and call pattern of
// src/mp/z_builder.hpp (mp::Z, zmethod, create_z):
//
//   auto z = mchip::mp::create_z().method(mchip::mp::zmethod::llmp_full)
//                .eri3c2e_batched(b).build();   // b: untransformed B (RI basis)
//   z->set_occ_coeff(locc).set_vir_density(pvir);    // LLMP builders
//   z->compute();
//   auto Z = z->zmat();                               // naux x naux
namespace mp {

using fock::matrix;
using fock::shared_matrix;

// z_builder.hpp:32-36
enum class zmethod { llmp_full, llmp_mem, ll_full };

// mp::Z (z_builder.hpp:43-116): occupied Cholesky factor L_occ (nbf x rank),
// virtual pseudo-density P_vir (LLMP builders) or its factor L_vir (LL_FULL).
class Z {
  mchip_dbt m_b;
  zmethod m_method;
  shared_matrix m_locc, m_lvir, m_pvir, m_zmat;

 public:
  Z(mchip_dbt b, zmethod m) : m_b(b), m_method(m) {}
  Z& set_occ_coeff(shared_matrix locc) { m_locc = std::move(locc); return *this; }
  Z& set_vir_coeff(shared_matrix lvir) { m_lvir = std::move(lvir); return *this; }
  Z& set_vir_density(shared_matrix pvir) { m_pvir = std::move(pvir); return *this; }
  void init() {}
  void compute()
  {
    if (!m_locc) throw std::runtime_error("Z: occupied coefficients not set");
    const int nbf = m_locc->rows;
    const int zm = m_method == zmethod::llmp_full ? 0 : m_method == zmethod::llmp_mem ? 1 : 2;
    if (zm == 2 && !m_lvir) throw std::runtime_error("Z: LL_FULL needs the virtual coefficients");
    if (zm != 2 && !m_pvir) throw std::runtime_error("Z: LLMP needs the virtual density");
    int dims[4] = {0, 0, 0, 0};
    int64_t nelem = 0;
    fock::check(mchip_dbt_info(m_b, dims, &nelem));
    const int naux = dims[0];
    if (!m_zmat || m_zmat->rows != naux) m_zmat = std::make_shared<matrix>(naux, naux);
    fock::check(mchip_z_compute(m_b, zm, nbf, m_locc->ptr(), m_locc->cols, m_pvir ? m_pvir->ptr() : nullptr,
                                m_lvir ? m_lvir->ptr() : nullptr, m_lvir ? m_lvir->cols : 0, m_zmat->ptr()));
  }
  shared_matrix zmat() { return m_zmat; }
};

// create_z (z_builder.hpp:324-406)
class create_z_base {
  zmethod c_method = zmethod::llmp_full;
  mchip_dbt c_b = nullptr;

 public:
  create_z_base& method(zmethod m) { c_method = m; return *this; }
  create_z_base& eri3c2e_batched(mchip_dbt b) { c_b = b; return *this; }
  std::shared_ptr<Z> build()
  {
    if (!c_b) throw std::runtime_error("create_z: eri3c2e_batched is required");
    return std::make_shared<Z>(c_b, c_method);
  }
};
inline create_z_base create_z() { return create_z_base(); }

}  // namespace mp
}  // namespace mchip
