// mchip_fock.hpp -- C++ layer over the mchip C ABI with the names, factory
// arguments and call pattern of the reference's fock, ints and mp builders
// (src/fock/jkbuilder.hpp, src/ints/aoloader.hpp, src/mp/z_builder.hpp):
//
//   mchip::world w(0);                                   // one process, one GPU
//   auto mol = std::make_shared<mchip::desc::molecule>(...);
//   auto ao = mchip::ints::aoloader::create().set_world(w).set_molecule(mol).build();
//   mchip::fock::load_jints(jmethod::dfao, ints::metric::coulomb, *ao);
//   mchip::fock::load_kints(kmethod::dfmo, ints::metric::coulomb, *ao);
//   ao->compute();
//   auto j = mchip::fock::create_j().set_world(w).molecule(mol).method(jmethod::dfao)
//                .aoloader(*ao).print(0).metric(ints::metric::coulomb).build();
//   auto k = mchip::fock::create_k().set_world(w).molecule(mol).method(kmethod::dfmo)
//                .aoloader(*ao).print(0).metric(ints::metric::coulomb).occ_nbatches(3).build();
//   j->set_density_alpha(p);  k->set_coeff_alpha(c);
//   j->init(); k->init();  j->compute_J(); k->compute_K();
//   auto J = j->get_J(); auto K = k->get_K_A();          // K already negated
//
// Matrices are dense row-major (mchip::shared_matrix), standing in for
// dbcsr::shared_matrix<double>; J and K are stored full (the reference keeps
// the upper blocks of a symmetric DBCSR matrix).  Builders alias the caller's
// matrices and overwrite their own output on every compute, as the reference's
// do.  Every failure is a std::runtime_error carrying the library's message,
// as the reference throws on every error.  Header-only; link with libmchip.so.
#pragma once

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mchip.h"

namespace mchip {

inline void check(int rc)
{
  if (rc != MCHIP_OK) throw std::runtime_error(mchip_last_error());
}

// Dense row-major matrix (dbcsr::shared_matrix<double> in the reference).
struct matrix {
  std::string name;
  int rows = 0, cols = 0;
  std::vector<double> data;
  matrix() = default;
  matrix(int r, int c, std::string n = "") : name(std::move(n)), rows(r), cols(c), data((size_t)r * c, 0.0) {}
  double* ptr() { return data.data(); }
  const double* ptr() const { return data.data(); }
  double& operator()(int i, int j) { return data[(size_t)i * cols + j]; }
  double operator()(int i, int j) const { return data[(size_t)i * cols + j]; }
};
using shared_matrix = std::shared_ptr<matrix>;

// megalochem::world (src/megalochem.hpp:13-80): the reference's MPI
// communicator and DBCSR grid; here one process driving one GPU through a
// library context (its HIP stream and work arrays).
class world {
  struct holder {
    mchip_ctx c = nullptr;
    ~holder()
    {
      if (c) mchip_ctx_destroy(c);
    }
  };
  std::shared_ptr<holder> m_h;
  int m_device = 0;

 public:
  explicit world(int device = 0) : m_h(std::make_shared<holder>()), m_device(device)
  {
    check(mchip_ctx_create(device, &m_h->c));
  }
  mchip_ctx ctx() const { return m_h->c; }
  int device() const { return m_device; }
  int rank() const { return 0; }
  int size() const { return 1; }
};

// Wall-clock timers of a builder (util::mpi_time, printed by print_info).
class timers {
  std::string m_name;
  std::map<std::string, std::pair<double, int>> m_t;

 public:
  explicit timers(std::string name = "") : m_name(std::move(name)) {}
  template <class F>
  void time(const std::string& what, F&& f)
  {
    const auto t0 = std::chrono::steady_clock::now();
    f();
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    auto& e = m_t[what];
    e.first += s;
    e.second += 1;
  }
  double total(const std::string& what) const
  {
    auto it = m_t.find(what);
    return it == m_t.end() ? 0.0 : it->second.first;
  }
  void print_info() const
  {
    std::printf("Timings for %s\n", m_name.c_str());
    for (const auto& kv : m_t)
      std::printf("  %-28s %12.6f s  (%d calls)\n", kv.first.c_str(), kv.second.first, kv.second.second);
  }
};

namespace desc {

// dbcsr_common.hpp:216-230: n elements in blocks of `split`, the remainder last.
inline std::vector<int> split_range(int n, int split)
{
  std::vector<int> out;
  if (n <= 0) return out;
  const int nb = n % split == 0 ? n / split : n / split + 1;
  out.assign(nb, split);
  if (n % split) out.back() = n % split;
  return out;
}

// One basis set: shells as flat arrays (mchip_basis_create: raw contraction
// coefficients, centres in bohr, owning atom) and its blocking, as shell
// counts of consecutive shells per block (atomic blocks: one per centre,
// desc::cluster_basis, basis.cpp:52-80).
struct basis_set {
  std::vector<int> l, nprim, atom;
  std::vector<double> exps, coefs, centers;
  std::vector<int> block_shells;
};

// desc::molecule (src/desc/molecule.hpp:23-318): atoms, the orbital and
// auxiliary basis sets, electron counts and the block sizes of every tensor.
class molecule {
  struct bhold {
    mchip_basis b = nullptr;
    ~bhold()
    {
      if (b) mchip_basis_destroy(b);
    }
  };
  std::string m_name;
  std::vector<double> m_Z, m_xyz;
  basis_set m_ao, m_aux;
  int m_nocc_a, m_nocc_b, m_mo_split;
  std::shared_ptr<bhold> m_hao, m_haux;
  std::vector<int> m_b, m_x;
  int m_nbf = 0, m_naux = 0;

  static std::vector<int> blocks(mchip_basis h, const basis_set& bs)
  {
    // functions per block from the shells' angular momenta (spherical)
    std::vector<int> out;
    size_t s = 0;
    for (int nsh : bs.block_shells) {
      int n = 0;
      for (int i = 0; i < nsh; ++i, ++s) n += 2 * bs.l.at(s) + 1;
      out.push_back(n);
    }
    int nbf = 0;
    check(mchip_basis_nbf(h, &nbf));
    int tot = 0;
    for (int v : out) tot += v;
    if (s != bs.l.size() || tot != nbf) throw std::runtime_error("molecule: block shells do not cover the basis");
    return out;
  }

 public:
  class dims_t {
    const molecule* m;

   public:
    explicit dims_t(const molecule* mm) : m(mm) {}
    std::vector<int> b() const { return m->m_b; }
    std::vector<int> x() const { return m->m_x; }
    std::vector<int> oa() const { return split_range(m->m_nocc_a, m->m_mo_split); }
    std::vector<int> ob() const { return split_range(m->m_nocc_b, m->m_mo_split); }
    std::vector<int> va() const { return split_range(m->m_nbf - m->m_nocc_a, m->m_mo_split); }
    std::vector<int> vb() const { return split_range(m->m_nbf - m->m_nocc_b, m->m_mo_split); }
  };

  // Z: nuclear charges, xyz: natom x 3 coordinates in bohr
  molecule(std::string name, std::vector<double> Z, std::vector<double> xyz, basis_set ao, basis_set aux,
           int nocc_alpha, int nocc_beta, int mo_split = 5)
      : m_name(std::move(name)), m_Z(std::move(Z)), m_xyz(std::move(xyz)), m_ao(std::move(ao)),
        m_aux(std::move(aux)), m_nocc_a(nocc_alpha), m_nocc_b(nocc_beta), m_mo_split(mo_split),
        m_hao(std::make_shared<bhold>()), m_haux(std::make_shared<bhold>())
  {
    if (m_xyz.size() != 3 * m_Z.size()) throw std::runtime_error("molecule: xyz must hold 3 coordinates per atom");
    for (auto* p : {&m_ao, &m_aux}) {
      bhold& h = (p == &m_ao) ? *m_hao : *m_haux;
      check(mchip_basis_create((int)p->l.size(), p->l.data(), p->nprim.data(), p->exps.data(), p->coefs.data(),
                               p->centers.data(), p->atom.data(), &h.b));
    }
    m_b = blocks(m_hao->b, m_ao);
    m_x = blocks(m_haux->b, m_aux);
    check(mchip_basis_nbf(m_hao->b, &m_nbf));
    check(mchip_basis_nbf(m_haux->b, &m_naux));
  }
  const std::string& name() const { return m_name; }
  dims_t dims() const { return dims_t(this); }
  int nbf() const { return m_nbf; }
  int naux() const { return m_naux; }
  int nocc_alpha() const { return m_nocc_a; }
  int nocc_beta() const { return m_nocc_b; }
  int mo_split() const { return m_mo_split; }
  int natoms() const { return (int)m_Z.size(); }
  const std::vector<double>& charges() const { return m_Z; }
  const std::vector<double>& coords() const { return m_xyz; }
  mchip_basis c_basis() const { return m_hao->b; }
  mchip_basis c_dfbasis() const { return m_haux->b; }
  const basis_set& ao_set() const { return m_ao; }
  const basis_set& aux_set() const { return m_aux; }
  double nuclear_repulsion() const
  {
    double e = 0.0;
    for (size_t i = 0; i < m_Z.size(); ++i)
      for (size_t j = i + 1; j < m_Z.size(); ++j) {
        const double dx = m_xyz[3 * i] - m_xyz[3 * j], dy = m_xyz[3 * i + 1] - m_xyz[3 * j + 1],
                     dz = m_xyz[3 * i + 2] - m_xyz[3 * j + 2];
        e += m_Z[i] * m_Z[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
      }
    return e;
  }
};
using shared_molecule = std::shared_ptr<molecule>;

}  // namespace desc

namespace ints {

// ints::metric (src/ints/registry.hpp): only the Coulomb metric is on this path.
enum class metric { coulomb, erfc_coulomb, qr_fit, pari };

// ints::key (src/ints/aoloader.hpp:18-45), same order.
enum class key {
  pgrid2 = 0, pgrid3, pgrid4, ovlp_bb, ovlp_xx, pot_bb, kin_bb, ovlp_bb_inv, ovlp_xx_inv, coul_xbb, erfc_xbb,
  coul_xx, erfc_xx, coul_bbbb, coul_xx_inv, erfc_xx_inv, coul_xx_invsqrt, erfc_xx_invsqrt, dfit_coul_xbb,
  dfit_erfc_xbb, dfit_pari_xbb, dfit_qr_xbb, scr_xbb, qr_xbb, pari_xbb, NUM_KEYS
};

// The three-centre tensor B(X|mu nu) resident on the GPU (the reference's
// sbtensor<3> of coul_xbb), with its block mask.
struct btensor {
  mchip_dbt h = nullptr;
  int naux = 0, nbf = 0, nxb = 0, nbb = 0;
  explicit btensor(mchip_dbt t) : h(t)
  {
    int dims[4];
    int64_t ne = 0;
    check(mchip_dbt_info(h, dims, &ne));
    naux = dims[0];
    nbf = dims[1];
    nxb = dims[2];
    nbb = dims[3];
  }
  btensor(const btensor&) = delete;
  btensor& operator=(const btensor&) = delete;
  ~btensor()
  {
    if (h) mchip_dbt_destroy(h);
  }
};
using shared_btensor = std::shared_ptr<btensor>;

// The Coulomb metric's factors on the GPU: L^-1 (the reference's
// coul_xx_invsqrt is L^-T) and V^-1 (coul_xx_inv), aoloader::invert
// (aoloader.cpp:14-32).
struct metric_factors {
  mchip_metric h = nullptr;
  mchip_ctx ctx = nullptr;
  int naux = 0;
  metric_factors(mchip_metric m, mchip_ctx c, int n) : h(m), ctx(c), naux(n) {}
  metric_factors(const metric_factors&) = delete;
  metric_factors& operator=(const metric_factors&) = delete;
  ~metric_factors()
  {
    if (h) mchip_metric_destroy(h);
  }
  // host copies: V^-1 (symmetric) and L^-1 (row-major, lower triangular)
  shared_matrix inverse() const
  {
    auto v = std::make_shared<matrix>(naux, naux, "coul_xx_inv");
    check(mchip_metric_get(ctx, h, nullptr, v->ptr()));
    return v;
  }
  shared_matrix inverse_cholesky() const
  {
    std::vector<double> cm((size_t)naux * naux);
    check(mchip_metric_get(ctx, h, cm.data(), nullptr));
    auto l = std::make_shared<matrix>(naux, naux, "coul_xx_linv");
    for (int i = 0; i < naux; ++i)
      for (int j = 0; j < naux; ++j) (*l)(i, j) = cm[(size_t)j * naux + i];   // column-major -> row-major
    return l;
  }
};
using shared_metric = std::shared_ptr<metric_factors>;

// ints::key_registry (src/ints/registry.hpp): the objects an aoloader computed.
class key_registry {
  std::map<key, shared_matrix> m_mat;
  std::map<key, shared_btensor> m_bt;
  std::map<key, shared_metric> m_met;
  std::map<key, std::shared_ptr<std::vector<uint8_t>>> m_mask;

  template <class T>
  static const T& find(const std::map<key, T>& m, key k)
  {
    auto it = m.find(k);
    if (it == m.end())
      throw std::runtime_error("key_registry: key " + std::to_string((int)k) + " was not computed (request it)");
    return it->second;
  }

 public:
  template <class T>
  T get(key k) const;
  void insert(key k, shared_matrix m) { m_mat[k] = std::move(m); }
  void insert(key k, shared_btensor b) { m_bt[k] = std::move(b); }
  void insert(key k, shared_metric m) { m_met[k] = std::move(m); }
  void insert(key k, std::shared_ptr<std::vector<uint8_t>> m) { m_mask[k] = std::move(m); }
  bool present(key k) const { return m_mat.count(k) || m_bt.count(k) || m_met.count(k) || m_mask.count(k); }
};
template <>
inline shared_matrix key_registry::get<shared_matrix>(key k) const { return find(m_mat, k); }
template <>
inline shared_btensor key_registry::get<shared_btensor>(key k) const { return find(m_bt, k); }
template <>
inline shared_metric key_registry::get<shared_metric>(key k) const { return find(m_met, k); }
template <>
inline std::shared_ptr<std::vector<uint8_t>> key_registry::get<std::shared_ptr<std::vector<uint8_t>>>(key k) const
{
  return find(m_mask, k);
}

// ints::aoloader (src/ints/aoloader.hpp, aoloader.cpp:131-433): integrals
// requested by key, computed at compute() (one-electron matrices on the host,
// the Schwarz mask, B on the GPU, the metric factors on the GPU), kept in a
// registry.  Every Coulomb-metric key of the DF Fock and MP2 paths is served;
// keys of other metrics, exact four-centre integrals and the fitted-
// coefficient tensors of PARI/QR are rejected.  dfit_coul_xbb (DFAO_K's
// second tensor) is accepted: the DFAO exchange here contracts a factor of P
// with B and the metric and needs no fitted copy of B.
class aoloader {
  world m_world;
  desc::shared_molecule m_mol;
  int m_print = 0;
  double m_eps = 1e-10;
  bool m_keep[(int)key::NUM_KEYS] = {};
  bool m_req[(int)key::NUM_KEYS] = {};
  bool m_lowp[3] = {false, false, false};
  key_registry m_reg;
  timers m_time{"AO-loader"};

 public:
  class create_t {
    world* w = nullptr;
    desc::shared_molecule mol;
    int pr = 0;
    double eps = 1e-10;

   public:
    create_t& set_world(world& ww) { w = &ww; return *this; }
    create_t& set_molecule(desc::shared_molecule m) { mol = std::move(m); return *this; }
    create_t& print(int p) { pr = p; return *this; }
    // globals.block_threshold (megalochem_driver.cpp): Schwarz screening and
    // the post-fill block filter
    create_t& block_threshold(double e) { eps = e; return *this; }
    std::shared_ptr<aoloader> build()
    {
      if (!w) throw std::runtime_error("aoloader: required argument set_world not set");
      if (!mol) throw std::runtime_error("aoloader: required argument set_molecule not set");
      auto a = std::shared_ptr<aoloader>(new aoloader(*w, mol));
      a->m_print = pr;
      a->m_eps = eps;
      return a;
    }
  };
  static create_t create() { return create_t(); }

  aoloader(world w, desc::shared_molecule mol) : m_world(std::move(w)), m_mol(std::move(mol)) {}

  aoloader& request(key k, bool keep)
  {
    switch (k) {
      case key::ovlp_bb: case key::kin_bb: case key::pot_bb: case key::coul_xx: case key::coul_xx_inv:
      case key::coul_xx_invsqrt: case key::coul_xbb: case key::scr_xbb: case key::dfit_coul_xbb:
        break;
      default:
        throw std::runtime_error("aoloader: key " + std::to_string((int)k) +
                                 " is not served by this engine (Coulomb-metric DF keys only)");
    }
    m_req[(int)k] = true;
    m_keep[(int)k] = m_keep[(int)k] || keep;
    return *this;
  }
  // Chosen form: copies of L^-1 B (FP32, BF16 hi/lo, FP64) made with B, for
  // the reduced-precision exchanges (no counterpart in the reference).
  aoloader& lowp_copies(bool f32, bool bf16, bool f64)
  {
    m_lowp[0] = f32;
    m_lowp[1] = bf16;
    m_lowp[2] = f64;
    return *this;
  }

  void compute()
  {
    const desc::molecule& mol = *m_mol;
    const int nbf = mol.nbf(), naux = mol.naux();
    auto r = [&](key k) { return m_req[(int)k]; };
    if (r(key::ovlp_bb) || r(key::kin_bb) || r(key::pot_bb)) {
      m_time.time("one-electron integrals", [&] {
        auto S = std::make_shared<matrix>(nbf, nbf, "s_bb");
        auto T = std::make_shared<matrix>(nbf, nbf, "t_bb");
        auto V = std::make_shared<matrix>(nbf, nbf, "v_bb");
        check(mchip_int1e(mol.c_basis(), mol.natoms(), mol.charges().data(), mol.coords().data(), S->ptr(),
                          T->ptr(), V->ptr()));
        m_reg.insert(key::ovlp_bb, S);
        m_reg.insert(key::kin_bb, T);
        m_reg.insert(key::pot_bb, V);
      });
    }
    const bool need_b = r(key::coul_xbb) || r(key::dfit_coul_xbb) || r(key::scr_xbb);
    const bool need_v = r(key::coul_xx) || r(key::coul_xx_inv) || r(key::coul_xx_invsqrt) || r(key::dfit_coul_xbb);
    const auto& bs = mol.ao_set().block_shells;
    const auto& xs = mol.aux_set().block_shells;
    const int nbb = (int)bs.size(), nxb = (int)xs.size();
    if (need_b) {
      auto mask = std::make_shared<std::vector<uint8_t>>((size_t)nxb * nbb * nbb);
      m_time.time("Schwarz screening", [&] {
        check(mchip_schwarz_mask(mol.c_basis(), mol.c_dfbasis(), nbb, bs.data(), nxb, xs.data(), m_eps, mask->data(),
                                 nullptr, nullptr));
      });
      m_reg.insert(key::scr_xbb, mask);
      if (r(key::coul_xbb) || r(key::dfit_coul_xbb)) {
        mchip_dbt t = nullptr;
        m_time.time("three-centre integrals (GPU)", [&] {
          check(mchip_dbt_build_gpu(m_world.ctx(), mol.c_basis(), mol.c_dfbasis(), nbb, bs.data(), nxb, xs.data(),
                                    mask->data(), m_eps, &t));
          check(mchip_ctx_sync(m_world.ctx()));
        });
        m_reg.insert(key::coul_xbb, std::make_shared<btensor>(t));
      }
    }
    if (need_v) {
      auto M = std::make_shared<matrix>(naux, naux, "coul_xx");
      m_time.time("two-centre integrals", [&] { check(mchip_int2c2e(mol.c_dfbasis(), M->ptr())); });
      m_reg.insert(key::coul_xx, M);
      if (r(key::coul_xx_inv) || r(key::coul_xx_invsqrt) || r(key::dfit_coul_xbb)) {
        mchip_metric h = nullptr;
        m_time.time("metric factors (GPU)", [&] { check(mchip_metric_create(m_world.ctx(), naux, M->ptr(), &h)); });
        auto f = std::make_shared<metric_factors>(h, m_world.ctx(), naux);
        m_reg.insert(key::coul_xx_inv, f);
        m_reg.insert(key::coul_xx_invsqrt, f);
      }
    }
    if ((m_lowp[0] || m_lowp[1] || m_lowp[2]) && m_reg.present(key::coul_xbb) && m_reg.present(key::coul_xx_inv)) {
      auto b = m_reg.get<shared_btensor>(key::coul_xbb);
      auto f = m_reg.get<shared_metric>(key::coul_xx_inv);
      m_time.time("copies of L^-1 B", [&] { check(mchip_dbt_make_lowp(b->h, f->h, m_lowp[0], m_lowp[1], m_lowp[2])); });
    }
  }
  const key_registry& get_registry() const { return m_reg; }
  const world& get_world() const { return m_world; }
  desc::shared_molecule get_molecule() const { return m_mol; }
  void print_info() const { m_time.print_info(); }
};

}  // namespace ints

namespace fock {

// jkbuilder.hpp:20-60
enum class jmethod { exact, dfao };
enum class kmethod { exact, dfao, dfmo, dfmem, dfrobust, dflmo };

inline jmethod str_to_jmethod(const std::string& s)
{
  if (s == "exact") return jmethod::exact;
  if (s == "dfao") return jmethod::dfao;
  throw std::runtime_error("Invalid jmethod");
}

inline kmethod str_to_kmethod(const std::string& s)
{
  if (s == "exact") return kmethod::exact;
  if (s == "dfao") return kmethod::dfao;
  if (s == "dfmo") return kmethod::dfmo;
  if (s == "dfmem") return kmethod::dfmem;
  if (s == "dfrobust") return kmethod::dfrobust;
  if (s == "dflmo") return kmethod::dflmo;
  throw std::runtime_error("Invalid kmethod");
}

// Precision of the exchange contractions (no counterpart in the reference):
// the reference's FP64, or the chosen forms on the copies of L^-1 B.
enum class precision { fp64, fp32, bf16x3 };

// JK_common (jkbuilder.hpp:62-150)
class JK_common {
 protected:
  world m_world;
  desc::shared_molecule m_mol;
  int m_print;
  timers TIME;
  shared_matrix m_p_A, m_p_B, m_c_A, m_c_B;
  shared_matrix m_u_A, m_u_B, m_v_A, m_v_B;
  bool m_SAD_iter = false;
  int m_SAD_rank = 0;
  bool m_sym = true;

 public:
  JK_common(world w, desc::shared_molecule mol, int print, std::string name)
      : m_world(std::move(w)), m_mol(std::move(mol)), m_print(print), TIME(std::move(name))
  {
  }
  virtual ~JK_common() = default;
  void set_density_alpha(shared_matrix& p) { m_p_A = p; }
  void set_density_beta(shared_matrix& p) { m_p_B = p; }
  void set_coeff_alpha(shared_matrix& c) { m_c_A = c; }
  void set_coeff_beta(shared_matrix& c) { m_c_B = c; }
  void set_coeff_left_alpha(shared_matrix& u) { m_u_A = u; }
  void set_coeff_right_alpha(shared_matrix& v) { m_v_A = v; }
  void set_coeff_left_beta(shared_matrix& u) { m_u_B = u; }
  void set_coeff_right_beta(shared_matrix& v) { m_v_B = v; }
  void set_sym(bool sym) { m_sym = sym; }
  void set_SAD(bool SAD, int rank)
  {
    m_SAD_iter = SAD;
    m_SAD_rank = rank;
  }
  void print_info() { TIME.print_info(); }
};

// J (jkbuilder.hpp:152-178)
class J : public JK_common {
 protected:
  shared_matrix m_J;
  void init_base() { m_J = std::make_shared<matrix>(m_mol->nbf(), m_mol->nbf(), "J_bb"); }

 public:
  using JK_common::JK_common;
  virtual void compute_J() = 0;
  virtual void init() = 0;
  shared_matrix get_J() { return m_J; }
};

// K (jkbuilder.hpp:180-217).  K_B exists only if a beta density was set
// before init(), as in K::init_base (jkcommon.cpp:41-56).  K is returned
// negated, as the reference does.
class K : public JK_common {
 protected:
  shared_matrix m_K_A, m_K_B;
  void init_base()
  {
    const int n = m_mol->nbf();
    m_K_A = std::make_shared<matrix>(n, n, "K_bb_A");
    if (m_p_B) m_K_B = std::make_shared<matrix>(n, n, "K_bb_B");
  }
  void need(const shared_matrix& m, const char* what) const
  {
    if (!m) throw std::runtime_error(std::string(what) + " not set");
  }

 public:
  using JK_common::JK_common;
  virtual void compute_K() = 0;
  virtual void init() = 0;
  shared_matrix get_K_A() { return m_K_A; }
  shared_matrix get_K_B() { return m_K_B; }
};

// DF_J::compute_J (jkdf.cpp:69-190): P_tot = 2 P_A, or P_A + P_B when a beta
// density is set.
class DF_J : public J {
  ints::shared_btensor m_b;
  ints::shared_metric m_v;

 public:
  DF_J(world w, desc::shared_molecule mol, int print, ints::shared_btensor b, ints::shared_metric v)
      : J(std::move(w), std::move(mol), print, "DF_J"), m_b(std::move(b)), m_v(std::move(v))
  {
  }
  void init() override { init_base(); }
  void compute_J() override
  {
    if (!m_p_A) throw std::runtime_error("DF_J: density alpha not set");
    if (!m_J) init_base();
    TIME.time("compute_J", [&] {
      if (m_p_B) {
        matrix ptot(m_p_A->rows, m_p_A->cols);
        for (size_t i = 0; i < ptot.data.size(); ++i) ptot.data[i] = m_p_A->data[i] + m_p_B->data[i];
        check(mchip_dfj(m_b->h, m_v->h, ptot.ptr(), 1.0, m_J->ptr()));
      }
      else {
        check(mchip_dfj(m_b->h, m_v->h, m_p_A->ptr(), 2.0, m_J->ptr()));
      }
    });
  }
};

// DFMO_K::compute_K (jkdf.cpp:221-462): the occupied columns of the MO
// coefficients (dims().oa() / ob(); all columns in a SAD iteration,
// jkdf.cpp:243-246), K = -(L^-1 B C)^T (L^-1 B C), in occupied batches.
class DFMO_K : public K {
  ints::shared_btensor m_b;
  ints::shared_metric m_v;
  int m_occ_nbatches;
  precision m_prec;

  void one(const shared_matrix& c, int nocc, shared_matrix& k)
  {
    if (nocc > c->cols) throw std::runtime_error("DFMO_K: fewer coefficient columns than occupied orbitals");
    std::vector<double> co((size_t)c->rows * nocc);
    for (int i = 0; i < c->rows; ++i)
      for (int j = 0; j < nocc; ++j) co[(size_t)i * nocc + j] = (*c)(i, j);
    switch (m_prec) {
      case precision::fp64: check(mchip_dfk_mo(m_b->h, m_v->h, co.data(), nocc, m_occ_nbatches, k->ptr())); break;
      case precision::fp32: check(mchip_dfk_mo_f32(m_b->h, co.data(), nocc, m_occ_nbatches, k->ptr())); break;
      case precision::bf16x3: check(mchip_dfk_mo_bf16(m_b->h, co.data(), nocc, m_occ_nbatches, k->ptr())); break;
    }
  }

 public:
  DFMO_K(world w, desc::shared_molecule mol, int print, ints::shared_btensor b, ints::shared_metric v,
         int occ_nbatches, precision p)
      : K(std::move(w), std::move(mol), print, "DFMO_K"), m_b(std::move(b)), m_v(std::move(v)),
        m_occ_nbatches(occ_nbatches), m_prec(p)
  {
  }
  void init() override { init_base(); }
  void compute_K() override
  {
    need(m_c_A, "DFMO_K: coefficients alpha");
    if (!m_K_A) init_base();
    TIME.time("compute_K", [&] {
      one(m_c_A, m_SAD_iter ? m_c_A->cols : m_mol->nocc_alpha(), m_K_A);
      if (m_K_B) {
        need(m_c_B, "DFMO_K: coefficients beta");
        one(m_c_B, m_SAD_iter ? m_c_B->cols : m_mol->nocc_beta(), m_K_B);
      }
    });
  }
};

// DFAO_K (jkdf.cpp:510-648) and DFMEM_K (jkdf.cpp:719-859): K from the
// density.  P is factorised on the GPU (pivoted Cholesky) and contracted like
// an occupied factor: the same K, term by term, as the fitted-coefficient and
// the batched-inverse contractions.  P must be positive semidefinite.
class DFAO_K : public K {
  ints::shared_btensor m_b;
  ints::shared_metric m_v;
  int m_occ_nbatches;
  precision m_prec;

 public:
  DFAO_K(world w, desc::shared_molecule mol, int print, ints::shared_btensor b, ints::shared_metric v,
         int occ_nbatches, precision p, const char* name = "DFAO_K")
      : K(std::move(w), std::move(mol), print, name), m_b(std::move(b)), m_v(std::move(v)),
        m_occ_nbatches(occ_nbatches), m_prec(p)
  {
  }
  void init() override { init_base(); }
  void compute_K() override
  {
    need(m_p_A, "DFAO_K: density alpha");
    if (!m_K_A) init_base();
    TIME.time("compute_K", [&] {
      check(mchip_dfk_ao(m_b->h, m_v->h, m_p_A->ptr(), (int)m_prec, m_occ_nbatches, m_K_A->ptr()));
      if (m_K_B) check(mchip_dfk_ao(m_b->h, m_v->h, m_p_B->ptr(), (int)m_prec, m_occ_nbatches, m_K_B->ptr()));
    });
  }
};

class DFMEM_K : public DFAO_K {
 public:
  DFMEM_K(world w, desc::shared_molecule mol, int print, ints::shared_btensor b, ints::shared_metric v,
          precision p)
      : DFAO_K(std::move(w), std::move(mol), print, std::move(b), std::move(v), 1, p, "DFMEM_K")
  {
  }
};

// DFLMO_K (jkdf.cpp:890-1345).  set_sym(true): K = -sum ht (V^-1 ht) with
// ht = B C over every column of the given C (the reference contracts all of
// them, :912,918).  set_sym(false): the two-sided exchange with the left
// (u, nbf x nocc) and right (v, nocc x nbf) coefficients, K not symmetric.
class DFLMO_K : public K {
  ints::shared_btensor m_b;
  ints::shared_metric m_v;
  int m_occ_nbatches;

 public:
  DFLMO_K(world w, desc::shared_molecule mol, int print, ints::shared_btensor b, ints::shared_metric v,
          int occ_nbatches)
      : K(std::move(w), std::move(mol), print, "DFLMO_K"), m_b(std::move(b)), m_v(std::move(v)),
        m_occ_nbatches(occ_nbatches)
  {
  }
  void init() override { init_base(); }
  void compute_K() override
  {
    if (!m_K_A) init_base();
    TIME.time("compute_K", [&] {
      auto sym = [&](const shared_matrix& c, shared_matrix& k) {
        check(mchip_dfk_mo(m_b->h, m_v->h, c->ptr(), c->cols, m_occ_nbatches, k->ptr()));
      };
      auto asym = [&](const shared_matrix& u, const shared_matrix& v, shared_matrix& k) {
        if (u->cols != v->rows) throw std::runtime_error("DFLMO_K: left and right coefficients disagree in width");
        check(mchip_dfk_lmo_uv(m_b->h, m_v->h, u->ptr(), v->ptr(), u->cols, m_occ_nbatches, 0, k->ptr()));
      };
      if (m_sym) {
        need(m_c_A, "DFLMO_K: coefficients alpha");
        sym(m_c_A, m_K_A);
        if (m_K_B) {
          need(m_c_B, "DFLMO_K: coefficients beta");
          sym(m_c_B, m_K_B);
        }
      }
      else {
        need(m_u_A, "DFLMO_K: left coefficients alpha");
        need(m_v_A, "DFLMO_K: right coefficients alpha");
        asym(m_u_A, m_v_A, m_K_A);
        if (m_K_B) {
          need(m_u_B, "DFLMO_K: left coefficients beta");
          need(m_v_B, "DFLMO_K: right coefficients beta");
          asym(m_u_B, m_v_B, m_K_B);
        }
      }
    });
  }
};

// load_jints / load_kints (jkbuilder.hpp:497-614): the keys each method needs.
inline void load_jints(jmethod jmet, ints::metric metr, ints::aoloader& ao)
{
  if (jmet != jmethod::dfao) throw std::runtime_error("load_jints: only the density-fitted J is on this path");
  if (metr != ints::metric::coulomb) throw std::runtime_error("load_jints: only the Coulomb metric is on this path");
  ao.request(ints::key::scr_xbb, true);
  ao.request(ints::key::coul_xx, false);
  ao.request(ints::key::coul_xx_inv, true);
  ao.request(ints::key::coul_xbb, true);
}

inline void load_kints(kmethod kmet, ints::metric metr, ints::aoloader& ao)
{
  if (kmet == kmethod::exact || kmet == kmethod::dfrobust)
    throw std::runtime_error("load_kints: exact and robust exchange are not on this path");
  if (metr != ints::metric::coulomb) {
    if (kmet == kmethod::dfmo) throw std::runtime_error("DFMO with non coulomb metric disabled.");
    throw std::runtime_error("load_kints: only the Coulomb metric is on this path");
  }
  ao.request(ints::key::scr_xbb, true);
  ao.request(ints::key::coul_xx, false);
  ao.request(ints::key::coul_xbb, true);
  switch (kmet) {
    case kmethod::dfao:
      ao.request(ints::key::coul_xx_inv, false);
      ao.request(ints::key::dfit_coul_xbb, true);
      break;
    case kmethod::dfmo: ao.request(ints::key::coul_xx_invsqrt, true); break;
    default: ao.request(ints::key::coul_xx_inv, true); break;
  }
}

namespace detail {
struct base_args {
  world* w = nullptr;
  desc::shared_molecule mol;
  ints::aoloader* ao = nullptr;
  int print = 0;
  bool has_metric = false;
  ints::metric metr = ints::metric::coulomb;
  void required(const char* who) const
  {
    if (!w) throw std::runtime_error(std::string(who) + ": required argument set_world not set");
    if (!mol) throw std::runtime_error(std::string(who) + ": required argument molecule not set");
    if (!ao) throw std::runtime_error(std::string(who) + ": required argument aoloader not set");
    if (!has_metric) throw std::runtime_error(std::string(who) + ": required argument metric not set");
    if (metr != ints::metric::coulomb)
      throw std::runtime_error(std::string(who) + ": only the Coulomb metric is on this path");
  }
};
}  // namespace detail

// create_j_base (jkbuilder.hpp:621-690)
class create_j_base {
  detail::base_args a;
  bool has_method = false;
  jmethod c_method = jmethod::dfao;

 public:
  create_j_base& set_world(world& w) { a.w = &w; return *this; }
  create_j_base& molecule(desc::shared_molecule m) { a.mol = std::move(m); return *this; }
  create_j_base& method(jmethod m) { c_method = m; has_method = true; return *this; }
  create_j_base& aoloader(ints::aoloader& ao) { a.ao = &ao; return *this; }
  create_j_base& print(int n) { a.print = n; return *this; }
  create_j_base& metric(ints::metric m) { a.metr = m; a.has_metric = true; return *this; }
  std::shared_ptr<J> build()
  {
    if (!has_method) throw std::runtime_error("create_j: required argument method not set");
    a.required("create_j");
    if (c_method != jmethod::dfao) throw std::runtime_error("create_j: only the density-fitted J is on this path");
    const auto& reg = a.ao->get_registry();
    return std::make_shared<DF_J>(*a.w, a.mol, a.print, reg.get<ints::shared_btensor>(ints::key::coul_xbb),
                                  reg.get<ints::shared_metric>(ints::key::coul_xx_inv));
  }
};
inline create_j_base create_j() { return create_j_base(); }

// create_k_base (jkbuilder.hpp:692-841); occ_nbatches defaults to 1 as in the
// factory (:718).  contraction_precision picks the chosen form's exchange on
// the copies of L^-1 B (no counterpart in the reference).
class create_k_base {
  detail::base_args a;
  bool has_method = false;
  kmethod c_method = kmethod::dfmo;
  int c_occ_nbatches = 1;
  precision c_prec = precision::fp64;

 public:
  create_k_base& set_world(world& w) { a.w = &w; return *this; }
  create_k_base& molecule(desc::shared_molecule m) { a.mol = std::move(m); return *this; }
  create_k_base& method(kmethod m) { c_method = m; has_method = true; return *this; }
  create_k_base& aoloader(ints::aoloader& ao) { a.ao = &ao; return *this; }
  create_k_base& print(int n) { a.print = n; return *this; }
  create_k_base& metric(ints::metric m) { a.metr = m; a.has_metric = true; return *this; }
  create_k_base& occ_nbatches(int n) { c_occ_nbatches = n; return *this; }
  create_k_base& contraction_precision(precision p) { c_prec = p; return *this; }
  std::shared_ptr<K> build()
  {
    if (!has_method) throw std::runtime_error("create_k: required argument method not set");
    a.required("create_k");
    const auto& reg = a.ao->get_registry();
    auto b = reg.get<ints::shared_btensor>(ints::key::coul_xbb);
    switch (c_method) {
      case kmethod::dfmo:
        return std::make_shared<DFMO_K>(*a.w, a.mol, a.print, b, reg.get<ints::shared_metric>(ints::key::coul_xx_invsqrt),
                                        c_occ_nbatches, c_prec);
      case kmethod::dfao:
        return std::make_shared<DFAO_K>(*a.w, a.mol, a.print, b, reg.get<ints::shared_metric>(ints::key::coul_xx_inv),
                                        c_occ_nbatches, c_prec);
      case kmethod::dfmem:
        return std::make_shared<DFMEM_K>(*a.w, a.mol, a.print, b, reg.get<ints::shared_metric>(ints::key::coul_xx_inv),
                                         c_prec);
      case kmethod::dflmo:
        if (c_prec != precision::fp64) throw std::runtime_error("create_k: DFLMO runs in FP64");
        return std::make_shared<DFLMO_K>(*a.w, a.mol, a.print, b, reg.get<ints::shared_metric>(ints::key::coul_xx_inv),
                                         c_occ_nbatches);
      default: throw std::runtime_error("create_k: exact and robust exchange are not on this path");
    }
  }
};
inline create_k_base create_k() { return create_k_base(); }

}  // namespace fock

// Z builders of SOS-AO-MP2 (src/mp/z_builder.hpp): Z_XY =
// sum_{mu nu} b3(X|mu nu) B(Y|mu nu) from the occupied factor L_occ (nbf x
// rank) and the virtual pseudo-density P_vir (LLMP builders) or its factor
// L_vir (LL_FULL).  FP64.
namespace mp {

// z_builder.hpp:21-36
enum class zmethod { llmp_full, llmp_mem, ll_full };
inline zmethod str_to_zmethod(const std::string& s)
{
  if (s == "llmp_full") return zmethod::llmp_full;
  if (s == "llmp_mem") return zmethod::llmp_mem;
  if (s == "ll_full") return zmethod::ll_full;
  throw std::runtime_error("Invalid zbuilder mathod.");
}

// mp::Z (z_builder.hpp:43-116)
class Z {
  world m_world;
  desc::shared_molecule m_mol;
  int m_print;
  timers TIME;
  ints::shared_btensor m_b;
  zmethod m_method;
  shared_matrix m_pocc, m_pvir, m_locc, m_lvir, m_zmat;

 public:
  Z(world w, desc::shared_molecule mol, int print, ints::shared_btensor b, zmethod m)
      : m_world(std::move(w)), m_mol(std::move(mol)), m_print(print),
        TIME(m == zmethod::llmp_full ? "LLMP_FULL_Z" : m == zmethod::llmp_mem ? "LLMP_MEM_Z" : "LL_FULL_Z"),
        m_b(std::move(b)), m_method(m)
  {
  }
  Z& set_occ_density(shared_matrix& p) { m_pocc = p; return *this; }
  Z& set_vir_density(shared_matrix& p) { m_pvir = p; return *this; }
  Z& set_occ_coeff(shared_matrix& l) { m_locc = l; return *this; }
  Z& set_vir_coeff(shared_matrix& l) { m_lvir = l; return *this; }
  void init() {}
  void compute()
  {
    if (!m_locc) throw std::runtime_error("Z: occupied coefficients not set");
    const int zm = m_method == zmethod::llmp_full ? 0 : m_method == zmethod::llmp_mem ? 1 : 2;
    if (zm == 2 && !m_lvir) throw std::runtime_error("Z: LL_FULL needs the virtual coefficients");
    if (zm != 2 && !m_pvir) throw std::runtime_error("Z: LLMP needs the virtual density");
    const int naux = m_b->naux;
    if (!m_zmat || m_zmat->rows != naux) m_zmat = std::make_shared<matrix>(naux, naux, "zmat");
    TIME.time("compute", [&] {
      check(mchip_z_compute(m_b->h, zm, m_locc->rows, m_locc->ptr(), m_locc->cols,
                            m_pvir ? m_pvir->ptr() : nullptr, m_lvir ? m_lvir->ptr() : nullptr,
                            m_lvir ? m_lvir->cols : 0, m_zmat->ptr()));
    });
  }
  shared_matrix zmat() { return m_zmat; }
  void print_info() { TIME.print_info(); }
};

// load_zints (z_builder.hpp:286-317)
inline void load_zints(zmethod, ints::metric metr, ints::aoloader& ao)
{
  if (metr != ints::metric::coulomb) throw std::runtime_error("Invalid metric for z builder");
  ao.request(ints::key::scr_xbb, true);
  ao.request(ints::key::coul_xx, false);
  ao.request(ints::key::coul_xx_inv, true);
  ao.request(ints::key::coul_xbb, true);
}

// create_z (z_builder.hpp:324-406)
class create_z_base {
  world* c_w = nullptr;
  desc::shared_molecule c_mol;
  int c_print = 0;
  bool has_method = false, has_metric = false;
  zmethod c_method = zmethod::llmp_full;
  ints::aoloader* c_ao = nullptr;
  ints::metric c_metric = ints::metric::coulomb;

 public:
  create_z_base& set_world(world& w) { c_w = &w; return *this; }
  create_z_base& set_molecule(desc::shared_molecule m) { c_mol = std::move(m); return *this; }
  create_z_base& print(int n) { c_print = n; return *this; }
  create_z_base& method(zmethod m) { c_method = m; has_method = true; return *this; }
  create_z_base& aoloader(ints::aoloader& ao) { c_ao = &ao; return *this; }
  create_z_base& metric(ints::metric m) { c_metric = m; has_metric = true; return *this; }
  std::shared_ptr<Z> build()
  {
    if (!c_w) throw std::runtime_error("create_z: required argument set_world not set");
    if (!c_mol) throw std::runtime_error("create_z: required argument set_molecule not set");
    if (!has_method) throw std::runtime_error("create_z: required argument method not set");
    if (!c_ao) throw std::runtime_error("create_z: required argument aoloader not set");
    if (!has_metric) throw std::runtime_error("create_z: required argument metric not set");
    if (c_metric != ints::metric::coulomb) throw std::runtime_error("Invalid metric for z builder.");
    return std::make_shared<Z>(*c_w, c_mol, c_print, c_ao->get_registry().get<ints::shared_btensor>(ints::key::coul_xbb),
                               c_method);
  }
};
inline create_z_base create_z() { return create_z_base(); }

}  // namespace mp
}  // namespace mchip
