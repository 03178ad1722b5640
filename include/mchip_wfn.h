/* HDF5 wavefunction files in the reference's layout (libmchip_wfn.so, host
 * only, linked against the HDF5 C library).
 *
 * The reference's driver keeps one output file per run (filio::data_handler,
 * src/io/data_handler.hpp:162-215) and writes, per job tag, a group with
 *   molecule  <tag>/name (128-byte string), charge, mult, mo_split (int32 [1]),
 *             atoms (float64 [natoms][4]: Z, x, y, z)      molecule.hpp:322-352
 *   hfwfn     <tag>/c_bo_A, c_bv_A [, c_bo_B, c_bv_B]  (float64 [nbf][n], row
 *             major), eps_occ_A, eps_vir_A [, _B] (float64 [n]), scf_energy,
 *             nuc_energy, wfn_energy (float64 [1])          wfn.hpp:188-289
 * and its MP2 driver reads an SCF result back with read_hfwfn
 * (wfn.hpp:291-326).  Matrices with no columns (nocc or nvir 0) and absent
 * beta parts are not written, as in write_mat.  Every function returns 0 on
 * success and -1 on error, with the message in mchip_wfn_last_error(). */
#ifndef MCHIP_WFN_H
#define MCHIP_WFN_H

#ifdef __cplusplus
extern "C" {
#endif

const char* mchip_wfn_last_error(void);

/* Create (truncate) the file: data_handler(filename, create_mode::truncate)
 * (data_handler.hpp:162-180).  The writers below open an existing file. */
int mchip_wfn_create(const char* path);

/* write_molecule (molecule.hpp:322-352).  xyz: natom x 3, bohr. */
int mchip_wfn_write_molecule(const char* path, const char* tag, const char* name, int charge, int mult,
                             int mo_split, int natom, const double* Z, const double* xyz);

/* write_hfwfn (wfn.hpp:266-289).  c_bo: nbf x nocc, c_bv: nbf x nvir, row
 * major.  Beta pointers NULL for a restricted wavefunction (the reference's
 * RHF writes no _B datasets).  wfn_energy = scf_energy + nuc_energy in the
 * reference (hfscf.cpp:541-544); scf_energy is the electronic energy. */
int mchip_wfn_write_hfwfn(const char* path, const char* tag, int nbf, int nocc_a, int nvir_a, const double* c_bo_A,
                          const double* c_bv_A, const double* eps_occ_A, const double* eps_vir_A, int nocc_b,
                          int nvir_b, const double* c_bo_B, const double* c_bv_B, const double* eps_occ_B,
                          const double* eps_vir_B, double scf_energy, double nuc_energy, double wfn_energy);

/* read_hfwfn (wfn.hpp:291-326).  The alpha datasets are required ("Read
 * hfwfn: could not find <name>."), the beta ones optional; each matrix must
 * have the molecule's dimensions ("Read hfwfn: incompatible matrix
 * dimensions.").  present (optional) receives bit 0 c_bo_B, 1 c_bv_B,
 * 2 eps_occ_B, 3 eps_vir_B for the beta parts found (their outputs are left
 * untouched otherwise; beta output pointers may be NULL when nothing is
 * wanted).  energies[3] = scf, nuc, wfn. */
int mchip_wfn_read_hfwfn(const char* path, const char* tag, int nbf, int nocc_a, int nvir_a, int nocc_b, int nvir_b,
                         double* c_bo_A, double* c_bv_A, double* eps_occ_A, double* eps_vir_A, double* c_bo_B,
                         double* c_bv_B, double* eps_occ_B, double* eps_vir_B, int* present, double* energies);

#ifdef __cplusplus
}
#endif

#endif
