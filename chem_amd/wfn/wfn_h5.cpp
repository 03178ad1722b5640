// HDF5 wavefunction files in the reference's layout (include/mchip_wfn.h):
// the datasets and groups that filio::data_handler, write_molecule,
// write_hfwfn and read_hfwfn create and read (src/io/data_handler.hpp,
// src/desc/molecule.hpp:322-352, src/desc/wfn.hpp:188-326), on the serial
// HDF5 C library: one process writes, so the MPI-IO file driver of the
// reference is not needed for an identical file layout.
#include <hdf5.h>

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mchip_wfn.h"

namespace {

thread_local std::string g_err;

// Owner of one HDF5 identifier, closed by the matching H5*close.
struct hid {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  hid(hid_t i, herr_t (*c)(hid_t), const std::string& what) : id(i), close(c)
  {
    if (id < 0) throw std::runtime_error(what);
  }
  hid(const hid&) = delete;
  hid& operator=(const hid&) = delete;
  ~hid()
  {
    if (id >= 0 && close) close(id);
  }
  operator hid_t() const { return id; }
};

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

hid open_file(const char* path, bool rw)
{
  if (!path) throw std::runtime_error("null file name");
  return hid(H5Fopen(path, rw ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose,
             std::string("Datahandler: file ") + path + " does not exist!");
}

void create_group(hid_t f, const std::string& g)
{
  if (H5Lexists(f, g.c_str(), H5P_DEFAULT) > 0)
    throw std::runtime_error("group " + g + " exists already in the file");
  hid grp(H5Gcreate2(f, g.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "cannot create group " + g);
}

// data_handler::write: a new dataset of the file type, written from memory
void write(hid_t f, const std::string& name, hid_t filetype, hid_t memtype, const void* data,
           const std::vector<hsize_t>& dims)
{
  hid space(H5Screate_simple((int)dims.size(), dims.data(), nullptr), H5Sclose, "dataspace for " + name);
  hid dset(H5Dcreate2(f, name.c_str(), filetype, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
           "cannot create dataset " + name);
  if (H5Dwrite(dset, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    throw std::runtime_error("cannot write dataset " + name);
}

void write_f64(hid_t f, const std::string& name, const double* data, const std::vector<hsize_t>& dims)
{
  write(f, name, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, data, dims);
}

void write_i32(hid_t f, const std::string& name, int v) { write(f, name, H5T_STD_I32LE, H5T_NATIVE_INT, &v, {1}); }

// the reference's fixed 128-byte strings (data_handler.hpp:158-160, 236-270)
void write_string(hid_t f, const std::string& name, const std::string& s)
{
  if (s.size() > 127) throw std::runtime_error("Datahandler: string exceeds size limit.");
  hid t(H5Tcopy(H5T_C_S1), H5Tclose, "string type");
  H5Tset_size(t, 128);
  char buf[128];
  std::memset(buf, 0, sizeof(buf));
  std::memcpy(buf, s.data(), s.size());
  write(f, name, t, t, buf, {1});
}

bool exists(hid_t f, const std::string& name) { return H5Lexists(f, name.c_str(), H5P_DEFAULT) > 0; }

// data_handler::read into out, with the dataset's dimensions checked against dims
void read_f64(hid_t f, const std::string& name, double* out, const std::vector<hsize_t>& dims, const char* what)
{
  hid dset(H5Dopen2(f, name.c_str(), H5P_DEFAULT), H5Dclose, "Read hfwfn: could not find " + name + ".");
  hid space(H5Dget_space(dset), H5Sclose, "dataspace of " + name);
  const int rank = H5Sget_simple_extent_ndims(space);
  std::vector<hsize_t> d(rank > 0 ? rank : 0);
  if (rank > 0) H5Sget_simple_extent_dims(space, d.data(), nullptr);
  if (d != dims) throw std::runtime_error(std::string("Read hfwfn: incompatible ") + what + " dimensions.");
  if (H5Dread(dset, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
    throw std::runtime_error("cannot read dataset " + name);
}

}  // namespace

#define WFN_TRY try {
#define WFN_CATCH                 \
  return 0;                       \
  }                               \
  catch (const std::exception& e) \
  {                               \
    g_err = e.what();             \
    return -1;                    \
  }

extern "C" const char* mchip_wfn_last_error(void) { return g_err.c_str(); }

extern "C" int mchip_wfn_create(const char* path)
{
  WFN_TRY
  quiet();
  if (!path) throw std::runtime_error("null file name");
  hid f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, std::string("cannot create ") + path);
  WFN_CATCH
}

extern "C" int mchip_wfn_write_molecule(const char* path, const char* tag, const char* name, int charge, int mult,
                                        int mo_split, int natom, const double* Z, const double* xyz)
{
  WFN_TRY
  quiet();
  if (!tag || !name || natom < 0 || (natom > 0 && (!Z || !xyz))) throw std::runtime_error("write_molecule: bad arguments");
  hid f = open_file(path, true);
  const std::string g(tag);
  create_group(f, g);
  write_string(f, g + "/name", name);
  write_i32(f, g + "/charge", charge);
  write_i32(f, g + "/mult", mult);
  write_i32(f, g + "/mo_split", mo_split);
  std::vector<double> a((size_t)natom * 4);
  for (int i = 0; i < natom; ++i) {
    a[4 * i] = Z[i];
    for (int d = 0; d < 3; ++d) a[4 * i + 1 + d] = xyz[3 * i + d];
  }
  write_f64(f, g + "/atoms", a.data(), {(hsize_t)natom, 4});
  WFN_CATCH
}

extern "C" int mchip_wfn_write_hfwfn(const char* path, const char* tag, int nbf, int nocc_a, int nvir_a,
                                     const double* c_bo_A, const double* c_bv_A, const double* eps_occ_A,
                                     const double* eps_vir_A, int nocc_b, int nvir_b, const double* c_bo_B,
                                     const double* c_bv_B, const double* eps_occ_B, const double* eps_vir_B,
                                     double scf_energy, double nuc_energy, double wfn_energy)
{
  WFN_TRY
  quiet();
  if (!tag || nbf <= 0 || nocc_a < 0 || nvir_a < 0 || nocc_b < 0 || nvir_b < 0)
    throw std::runtime_error("write_hfwfn: bad arguments");
  hid f = open_file(path, true);
  const std::string g(tag);
  create_group(f, g);
  // write_mat skips absent matrices; hfmod makes none for zero columns
  auto mat = [&](const char* n, const double* p, int cols) {
    if (p && cols > 0) write_f64(f, g + "/" + n, p, {(hsize_t)nbf, (hsize_t)cols});
  };
  auto vec = [&](const char* n, const double* p, int len) {
    if (p) write_f64(f, g + "/" + n, p, {(hsize_t)len});
  };
  mat("c_bo_A", c_bo_A, nocc_a);
  mat("c_bo_B", c_bo_B, nocc_b);
  mat("c_bv_A", c_bv_A, nvir_a);
  mat("c_bv_B", c_bv_B, nvir_b);
  vec("eps_occ_A", eps_occ_A, nocc_a);
  vec("eps_vir_A", eps_vir_A, nvir_a);
  vec("eps_occ_B", eps_occ_B, nocc_b);
  vec("eps_vir_B", eps_vir_B, nvir_b);
  write_f64(f, g + "/scf_energy", &scf_energy, {1});
  write_f64(f, g + "/nuc_energy", &nuc_energy, {1});
  write_f64(f, g + "/wfn_energy", &wfn_energy, {1});
  WFN_CATCH
}

extern "C" int mchip_wfn_read_hfwfn(const char* path, const char* tag, int nbf, int nocc_a, int nvir_a, int nocc_b,
                                    int nvir_b, double* c_bo_A, double* c_bv_A, double* eps_occ_A,
                                    double* eps_vir_A, double* c_bo_B, double* c_bv_B, double* eps_occ_B,
                                    double* eps_vir_B, int* present, double* energies)
{
  WFN_TRY
  quiet();
  if (!tag || nbf <= 0 || !c_bo_A || !c_bv_A || !eps_occ_A || !eps_vir_A || !energies)
    throw std::runtime_error("read_hfwfn: bad arguments");
  hid f = open_file(path, false);
  const std::string g = std::string(tag) + "/";
  read_f64(f, g + "c_bo_A", c_bo_A, {(hsize_t)nbf, (hsize_t)nocc_a}, "matrix");
  read_f64(f, g + "c_bv_A", c_bv_A, {(hsize_t)nbf, (hsize_t)nvir_a}, "matrix");
  read_f64(f, g + "eps_occ_A", eps_occ_A, {(hsize_t)nocc_a}, "vector");
  read_f64(f, g + "eps_vir_A", eps_vir_A, {(hsize_t)nvir_a}, "vector");
  int bits = 0;
  struct opt {
    const char* n;
    double* p;
    std::vector<hsize_t> d;
    const char* what;
  } beta[4] = {{"c_bo_B", c_bo_B, {(hsize_t)nbf, (hsize_t)nocc_b}, "matrix"},
               {"c_bv_B", c_bv_B, {(hsize_t)nbf, (hsize_t)nvir_b}, "matrix"},
               {"eps_occ_B", eps_occ_B, {(hsize_t)nocc_b}, "vector"},
               {"eps_vir_B", eps_vir_B, {(hsize_t)nvir_b}, "vector"}};
  for (int i = 0; i < 4; ++i) {
    if (!exists(f, g + beta[i].n)) continue;
    bits |= 1 << i;
    if (beta[i].p) read_f64(f, g + beta[i].n, beta[i].p, beta[i].d, beta[i].what);
  }
  const char* en[3] = {"scf_energy", "nuc_energy", "wfn_energy"};
  for (int i = 0; i < 3; ++i) read_f64(f, g + en[i], energies + i, {1}, "scalar");
  if (present) *present = bits;
  WFN_CATCH
}
