// mchip_wfn.hpp -- C++ layer over include/mchip_wfn.h with the names and call
// pattern of the reference's output files (src/io/data_handler.hpp,
// src/desc/wfn.hpp, src/desc/molecule.hpp:322-352):
//
//   mchip::filio::data_handler out("run.h5", mchip::filio::create_mode::truncate);
//   mchip::desc::write_molecule("mol", *mol, out);
//   mchip::desc::write_hfwfn("hf", *wfn, out);
//   auto back = mchip::desc::read_hfwfn("hf", mol, w, out);
//
// The molecule type is any type with name(), charges(), coords(), nbf(),
// nocc_alpha(), nocc_beta() and mo_split() (mchip::desc::molecule has them).
// Header-only; link with libmchip_wfn.so.  Errors throw std::runtime_error
// with the reference's messages.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "mchip_fock.hpp"
#include "mchip_wfn.h"

namespace mchip {

namespace filio {

enum class create_mode { truncate, append };

inline void wfn_check(int rc)
{
  if (rc != 0) throw std::runtime_error(mchip_wfn_last_error());
}

// filio::data_handler (data_handler.hpp:147-373): the file behind a run's
// input or output; truncate creates it empty, append requires it to exist.
class data_handler {
  std::string m_file;

 public:
  data_handler(std::string filename, create_mode cmode) : m_file(std::move(filename))
  {
    if (cmode == create_mode::truncate) {
      wfn_check(mchip_wfn_create(m_file.c_str()));
    }
    else {
      struct stat st;
      if (stat(m_file.c_str(), &st) != 0)
        throw std::runtime_error("Datahandler: file " + m_file + " does not exist!");
    }
  }
  const std::string& filename() const { return m_file; }
};

}  // namespace filio

namespace desc {

using svector = std::shared_ptr<std::vector<double>>;

// desc::hf_wavefunction (wfn.hpp:21-111): beta parts null when restricted.
class hf_wavefunction {
  shared_matrix m_c_bo_A, m_c_bv_A, m_c_bo_B, m_c_bv_B;
  svector m_eps_occ_A, m_eps_occ_B, m_eps_vir_A, m_eps_vir_B;
  double m_scf_energy, m_nuc_energy, m_wfn_energy;

 public:
  hf_wavefunction(shared_matrix c_bo_A, shared_matrix c_bo_B, shared_matrix c_bv_A, shared_matrix c_bv_B,
                  svector eps_occ_A, svector eps_occ_B, svector eps_vir_A, svector eps_vir_B, double scf_energy,
                  double nuc_energy, double wfn_energy)
      : m_c_bo_A(std::move(c_bo_A)), m_c_bv_A(std::move(c_bv_A)), m_c_bo_B(std::move(c_bo_B)),
        m_c_bv_B(std::move(c_bv_B)), m_eps_occ_A(std::move(eps_occ_A)), m_eps_occ_B(std::move(eps_occ_B)),
        m_eps_vir_A(std::move(eps_vir_A)), m_eps_vir_B(std::move(eps_vir_B)), m_scf_energy(scf_energy),
        m_nuc_energy(nuc_energy), m_wfn_energy(wfn_energy)
  {
  }
  shared_matrix c_bo_A() const { return m_c_bo_A; }
  shared_matrix c_bv_A() const { return m_c_bv_A; }
  shared_matrix c_bo_B() const { return m_c_bo_B; }
  shared_matrix c_bv_B() const { return m_c_bv_B; }
  svector eps_occ_A() const { return m_eps_occ_A; }
  svector eps_vir_A() const { return m_eps_vir_A; }
  svector eps_occ_B() const { return m_eps_occ_B; }
  svector eps_vir_B() const { return m_eps_vir_B; }
  double scf_energy() const { return m_scf_energy; }
  double nuc_energy() const { return m_nuc_energy; }
  double wfn_energy() const { return m_wfn_energy; }
};
using shared_hf_wavefunction = std::shared_ptr<hf_wavefunction>;

// write_molecule (molecule.hpp:322-352); charge and multiplicity follow from
// the nuclear charges and the alpha / beta occupations
template <class Mol>
void write_molecule(const std::string& name, const Mol& mol, filio::data_handler& dh)
{
  double zsum = 0.0;
  for (double z : mol.charges()) zsum += z;
  const int charge = (int)std::lround(zsum) - mol.nocc_alpha() - mol.nocc_beta();
  const int mult = mol.nocc_alpha() - mol.nocc_beta() + 1;
  const auto& Z = mol.charges();
  filio::wfn_check(mchip_wfn_write_molecule(dh.filename().c_str(), name.c_str(), mol.name().c_str(), charge, mult,
                                            mol.mo_split(), (int)Z.size(), Z.data(), mol.coords().data()));
}

// write_hfwfn (wfn.hpp:266-289)
inline void write_hfwfn(const std::string& name, const hf_wavefunction& w, filio::data_handler& dh)
{
  if (!w.c_bo_A() && !w.c_bv_A()) throw std::runtime_error("write_hfwfn: no alpha coefficients");
  const int nbf = w.c_bo_A() ? w.c_bo_A()->rows : w.c_bv_A()->rows;
  auto cols = [&](const shared_matrix& m) {
    if (m && m->rows != nbf) throw std::runtime_error("write_hfwfn: coefficient matrices differ in rows");
    return m ? m->cols : 0;
  };
  auto mp = [](const shared_matrix& m) -> const double* { return m ? m->ptr() : nullptr; };
  auto vp = [](const svector& v) -> const double* { return v ? v->data() : nullptr; };
  auto vlen = [](const svector& v, int n, const char* what) {
    if (v && (int)v->size() != n) throw std::runtime_error(std::string("write_hfwfn: ") + what + " has the wrong length");
  };
  const int oa = cols(w.c_bo_A()), va = cols(w.c_bv_A()), ob = cols(w.c_bo_B()), vb = cols(w.c_bv_B());
  vlen(w.eps_occ_A(), oa, "eps_occ_A");
  vlen(w.eps_vir_A(), va, "eps_vir_A");
  vlen(w.eps_occ_B(), ob, "eps_occ_B");
  vlen(w.eps_vir_B(), vb, "eps_vir_B");
  filio::wfn_check(mchip_wfn_write_hfwfn(dh.filename().c_str(), name.c_str(), nbf, oa, va, mp(w.c_bo_A()),
                                         mp(w.c_bv_A()), vp(w.eps_occ_A()), vp(w.eps_vir_A()), ob, vb,
                                         mp(w.c_bo_B()), mp(w.c_bv_B()), vp(w.eps_occ_B()), vp(w.eps_vir_B()),
                                         w.scf_energy(), w.nuc_energy(), w.wfn_energy()));
}

// read_hfwfn (wfn.hpp:291-326): dimensions from the molecule (world is the
// reference's communicator argument, unused by a one-process read)
template <class Mol, class World>
shared_hf_wavefunction read_hfwfn(const std::string& name, const std::shared_ptr<Mol>& mol, const World&,
                                  filio::data_handler& dh)
{
  const int n = mol->nbf(), oa = mol->nocc_alpha(), ob = mol->nocc_beta();
  auto M = [&](int c, const char* nm) { return std::make_shared<matrix>(n, c, nm); };
  auto V = [](int len) { return std::make_shared<std::vector<double>>(len); };
  auto cboA = M(oa, "c_bo_A"), cbvA = M(n - oa, "c_bv_A"), cboB = M(ob, "c_bo_B"), cbvB = M(n - ob, "c_bv_B");
  auto eoA = V(oa), evA = V(n - oa), eoB = V(ob), evB = V(n - ob);
  int present = 0;
  double en[3];
  filio::wfn_check(mchip_wfn_read_hfwfn(dh.filename().c_str(), name.c_str(), n, oa, n - oa, ob, n - ob, cboA->ptr(),
                                        cbvA->ptr(), eoA->data(), evA->data(), cboB->ptr(), cbvB->ptr(), eoB->data(),
                                        evB->data(), &present, en));
  if (!(present & 1)) cboB = nullptr;
  if (!(present & 2)) cbvB = nullptr;
  if (!(present & 4)) eoB = nullptr;
  if (!(present & 8)) evB = nullptr;
  return std::make_shared<hf_wavefunction>(cboA, cboB, cbvA, cbvB, eoA, eoB, evA, evB, en[0], en[1], en[2]);
}

}  // namespace desc

}  // namespace mchip
