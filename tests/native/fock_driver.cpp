// Test driver of the C++ builder layer (include/mchip_fock.hpp): reads a
// molecule, its basis sets and the densities / coefficients to contract from
// an input file, runs every J, K and Z builder the way the reference's SCF,
// ADC and MP2 drivers call them (create_j / create_k / create_z over an
// aoloader), and writes each result to an output file for the test to check
// against the oracle.
//
// File format (both directions): records of
//   int32 name length, name, char type ('d' double, 'i' int32), int64 count, data.
//
// Usage: fock_driver <input> <output>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "mchip_fock.hpp"

namespace {

struct record {
  char type = 'd';
  std::vector<double> d;
  std::vector<int> i;
};

std::map<std::string, record> read_file(const char* path)
{
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  std::map<std::string, record> out;
  while (true) {
    int32_t nl = 0;
    if (!f.read(reinterpret_cast<char*>(&nl), sizeof(nl))) break;
    std::string name(nl, '\0');
    f.read(&name[0], nl);
    record r;
    f.read(&r.type, 1);
    int64_t n = 0;
    f.read(reinterpret_cast<char*>(&n), sizeof(n));
    if (r.type == 'd') {
      r.d.resize(n);
      f.read(reinterpret_cast<char*>(r.d.data()), n * sizeof(double));
    }
    else {
      r.i.resize(n);
      f.read(reinterpret_cast<char*>(r.i.data()), n * sizeof(int32_t));
    }
    if (!f) throw std::runtime_error("truncated record " + name);
    out[name] = std::move(r);
  }
  return out;
}

class writer {
  std::ofstream f;

 public:
  explicit writer(const char* path) : f(path, std::ios::binary)
  {
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
  }
  void put(const std::string& name, const std::vector<double>& v)
  {
    const int32_t nl = (int32_t)name.size();
    f.write(reinterpret_cast<const char*>(&nl), sizeof(nl));
    f.write(name.data(), nl);
    const char t = 'd';
    f.write(&t, 1);
    const int64_t n = (int64_t)v.size();
    f.write(reinterpret_cast<const char*>(&n), sizeof(n));
    f.write(reinterpret_cast<const char*>(v.data()), n * sizeof(double));
  }
  void put(const std::string& name, const mchip::shared_matrix& m) { put(name, m->data); }
};

mchip::desc::basis_set basis(const std::map<std::string, record>& in, const std::string& p)
{
  mchip::desc::basis_set b;
  b.l = in.at(p + ".l").i;
  b.nprim = in.at(p + ".nprim").i;
  b.atom = in.at(p + ".atom").i;
  b.block_shells = in.at(p + ".block_shells").i;
  b.exps = in.at(p + ".exps").d;
  b.coefs = in.at(p + ".coefs").d;
  b.centers = in.at(p + ".centers").d;
  return b;
}

mchip::shared_matrix mat(const std::map<std::string, record>& in, const std::string& name, int rows, int cols)
{
  const auto& d = in.at(name).d;
  if ((int64_t)d.size() != (int64_t)rows * cols) throw std::runtime_error("matrix " + name + " has the wrong size");
  auto m = std::make_shared<mchip::matrix>(rows, cols, name);
  m->data = d;
  return m;
}

}  // namespace

int main(int argc, char** argv)
{
  using namespace mchip;
  if (argc != 3) {
    std::fprintf(stderr, "usage: fock_driver <input> <output>\n");
    return 2;
  }
  try {
    const auto in = read_file(argv[1]);
    const int nocc = in.at("nocc").i.at(0);
    const int nu = in.at("nu").i.at(0);
    const int rank_o = in.at("rank_o").i.at(0), rank_v = in.at("rank_v").i.at(0);
    world w(0);
    auto mol = std::make_shared<desc::molecule>("water cluster", in.at("Z").d, in.at("xyz").d, basis(in, "ao"),
                                                basis(in, "jk"), nocc, nocc);
    const int n = mol->nbf();
    writer out(argv[2]);

    // the SCF driver's builders (hfmod::compute, hfscf.cpp:183-276)
    auto ao = ints::aoloader::create().set_world(w).set_molecule(mol).print(0).build();
    fock::load_jints(fock::jmethod::dfao, ints::metric::coulomb, *ao);
    for (auto km : {fock::kmethod::dfmo, fock::kmethod::dfao, fock::kmethod::dfmem, fock::kmethod::dflmo})
      fock::load_kints(km, ints::metric::coulomb, *ao);
    ao->request(ints::key::ovlp_bb, true);   // the SCF's own request (hfmod)
    ao->compute();
    const auto& reg = ao->get_registry();
    auto f = reg.get<ints::shared_metric>(ints::key::coul_xx_inv);
    out.put("Vinv", f->inverse());
    out.put("Linv", f->inverse_cholesky());
    out.put("S", reg.get<shared_matrix>(ints::key::ovlp_bb));

    auto P = mat(in, "P", n, n);
    auto C = mat(in, "C", n, n);   // all MO columns, occupied first
    auto Co = mat(in, "Cocc", n, nocc);

    auto jb = fock::create_j().set_world(w).molecule(mol).method(fock::jmethod::dfao).aoloader(*ao).print(0)
                  .metric(ints::metric::coulomb).build();
    jb->set_density_alpha(P);
    jb->init();
    jb->compute_J();
    out.put("J", jb->get_J());

    auto kbuild = [&](fock::kmethod km, int nb) {
      return fock::create_k().set_world(w).molecule(mol).method(km).aoloader(*ao).print(0)
          .metric(ints::metric::coulomb).occ_nbatches(nb).build();
    };
    auto kmo = kbuild(fock::kmethod::dfmo, 3);
    kmo->set_coeff_alpha(C);
    kmo->init();
    kmo->compute_K();
    out.put("K_dfmo", kmo->get_K_A());
    if (kmo->get_K_B()) throw std::runtime_error("K_B without a beta density set before init");

    for (auto km : {fock::kmethod::dfao, fock::kmethod::dfmem}) {
      auto k = kbuild(km, 1);
      k->set_density_alpha(P);
      k->init();
      k->compute_K();
      out.put(km == fock::kmethod::dfao ? "K_dfao" : "K_dfmem", k->get_K_A());
    }

    // DFLMO: symmetric over the given columns, then the two-sided form of the
    // ADC drivers (set_sym(false), u / v coefficients)
    auto kl = kbuild(fock::kmethod::dflmo, 2);
    kl->set_coeff_alpha(Co);
    kl->init();
    kl->compute_K();
    out.put("K_dflmo", kl->get_K_A());
    auto U = mat(in, "U", n, nu);
    auto V = mat(in, "V", nu, n);
    auto ku = kbuild(fock::kmethod::dflmo, 1);
    ku->set_sym(false);
    ku->set_coeff_left_alpha(U);
    ku->set_coeff_right_alpha(V);
    ku->init();
    ku->compute_K();
    out.put("K_uv", ku->get_K_A());

    // SAD iteration: DFMO over every column of the given coefficients
    auto Cs = mat(in, "Csad", n, nu);
    auto ks = kbuild(fock::kmethod::dfmo, 1);
    ks->set_SAD(true, nu);
    ks->set_coeff_alpha(Cs);
    ks->init();
    ks->compute_K();
    out.put("K_sad", ks->get_K_A());

    // a beta density set before init: K_B exists and is built from the beta coefficients
    auto Cb = mat(in, "Cb", n, n);
    auto kb = kbuild(fock::kmethod::dfmo, 1);
    kb->set_density_beta(P);
    kb->set_coeff_alpha(C);
    kb->set_coeff_beta(Cb);
    kb->init();
    kb->compute_K();
    if (!kb->get_K_B()) throw std::runtime_error("K_B missing with a beta density set before init");
    out.put("K_A_uhf", kb->get_K_A());
    out.put("K_B_uhf", kb->get_K_B());

    // required factory arguments are checked like CHECK_REQUIRED does
    double missing = 0.0;
    try {
      fock::create_k().set_world(w).molecule(mol).method(fock::kmethod::dfmo).metric(ints::metric::coulomb).build();
    }
    catch (const std::runtime_error& e) {
      if (std::strstr(e.what(), "aoloader")) missing = 1.0;
    }
    out.put("missing_aoloader_caught", std::vector<double>{missing});

    // the MP2 driver's Z builders (mpmod::compute, mp_compute.cpp:135-353) on the RI basis
    auto molri = std::make_shared<desc::molecule>("water cluster (ri)", in.at("Z").d, in.at("xyz").d,
                                                  basis(in, "ao"), basis(in, "ri"), nocc, nocc);
    auto ari = ints::aoloader::create().set_world(w).set_molecule(molri).build();
    mp::load_zints(mp::zmethod::llmp_full, ints::metric::coulomb, *ari);
    ari->compute();
    auto Lo = mat(in, "Locc", n, rank_o);
    auto Pv = mat(in, "Pvir", n, n);
    auto Lv = mat(in, "Lvir", n, rank_v);
    for (auto zm : {mp::zmethod::llmp_full, mp::zmethod::llmp_mem, mp::zmethod::ll_full}) {
      auto z = mp::create_z().set_world(w).set_molecule(molri).print(0).method(zm).aoloader(*ari)
                   .metric(ints::metric::coulomb).build();
      z->set_occ_coeff(Lo);
      if (zm == mp::zmethod::ll_full)
        z->set_vir_coeff(Lv);
      else
        z->set_vir_density(Pv);
      z->init();
      z->compute();
      out.put(zm == mp::zmethod::llmp_full ? "Z_llmp_full" : zm == mp::zmethod::llmp_mem ? "Z_llmp_mem" : "Z_ll_full",
              z->zmat());
      z->print_info();
    }
    jb->print_info();
    kmo->print_info();
    ao->print_info();
    std::printf("fock_driver: done\n");
    return 0;
  }
  catch (const std::exception& e) {
    std::fprintf(stderr, "fock_driver: %s\n", e.what());
    return 1;
  }
}
