// Test driver of the C++ wavefunction-file layer (include/mchip_wfn.hpp): the
// reference's write_molecule / write_hfwfn / read_hfwfn call pattern on a
// molecule-shaped type, values fixed by formula so the test can check the file.
//
// Usage: wfn_driver <file.h5>
#include <cmath>
#include <cstdio>
#include <string>

#include "mchip_wfn.hpp"

namespace {

// the accessors read_hfwfn and write_molecule use, as desc::molecule has them
struct small_molecule {
  std::vector<double> Z{8, 1, 1}, xyz{0, 0, 0, 1.43, 0, 1.1, -1.43, 0, 1.1};
  const std::string& name() const
  {
    static const std::string n = "water";
    return n;
  }
  const std::vector<double>& charges() const { return Z; }
  const std::vector<double>& coords() const { return xyz; }
  int nbf() const { return 24; }
  int nocc_alpha() const { return 5; }
  int nocc_beta() const { return 5; }
  int mo_split() const { return 5; }
};

double value(int i, int j) { return std::sin(0.37 * i + 1.91 * j) / (1 + i + j); }

}  // namespace

int main(int argc, char** argv)
{
  using namespace mchip;
  if (argc != 2) {
    std::fprintf(stderr, "usage: wfn_driver <file>\n");
    return 2;
  }
  // the real molecule type satisfies the same interface
  auto wm = &desc::write_molecule<desc::molecule>;
  (void)wm;
  try {
    auto mol = std::make_shared<small_molecule>();
    filio::data_handler out(argv[1], filio::create_mode::truncate);
    desc::write_molecule("mol", *mol, out);
    const int n = 24, o = 5;
    auto co = std::make_shared<matrix>(n, o, "c_bo_A"), cv = std::make_shared<matrix>(n, n - o, "c_bv_A");
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < o; ++j) (*co)(i, j) = value(i, j);
      for (int j = 0; j < n - o; ++j) (*cv)(i, j) = value(i, j + o);
    }
    auto eo = std::make_shared<std::vector<double>>(o), ev = std::make_shared<std::vector<double>>(n - o);
    for (int j = 0; j < n; ++j) (j < o ? (*eo)[j] : (*ev)[j - o]) = -1.0 + 0.1 * j;
    desc::hf_wavefunction w(co, nullptr, cv, nullptr, eo, nullptr, ev, nullptr, -85.25, 9.125, -76.125);
    desc::write_hfwfn("hf", w, out);
    filio::data_handler in(argv[1], filio::create_mode::append);
    struct no_world {} wd;
    auto back = desc::read_hfwfn("hf", mol, wd, in);
    if (back->c_bo_B() || back->eps_vir_B()) throw std::runtime_error("beta parts read from a restricted file");
    if (back->c_bo_A()->data != co->data || back->c_bv_A()->data != cv->data || *back->eps_occ_A() != *eo ||
        *back->eps_vir_A() != *ev || back->wfn_energy() != -76.125)
      throw std::runtime_error("round trip differs");
    bool caught = false;
    try {
      filio::data_handler missing(std::string(argv[1]) + ".none", filio::create_mode::append);
    }
    catch (const std::runtime_error& e) {
      caught = std::string(e.what()).find("does not exist!") != std::string::npos;
    }
    if (!caught) throw std::runtime_error("missing file not reported");
    std::printf("wfn_driver: done\n");
    return 0;
  }
  catch (const std::exception& e) {
    std::fprintf(stderr, "wfn_driver: %s\n", e.what());
    return 1;
  }
}
