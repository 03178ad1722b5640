// C ABI: basis sets, AO integrals and the host B container.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "../../include/mchip.h"
#include "capi_common.hpp"
#include "ints/ints_host.hpp"

using namespace mchip;

struct mchip_basis_s {
  Basis basis;
};

struct mchip_btensor_s {
  BLayout layout;
  std::vector<double> data;
};

namespace mchip {
thread_local std::string g_last_error;
void set_error(const std::string& s) { g_last_error = s; }
}  // namespace mchip

extern "C" const char* mchip_last_error(void) { return g_last_error.c_str(); }

namespace mchip {
const Basis& basis_of(mchip_basis b) { return b->basis; }
const BLayout& btensor_layout(mchip_btensor t) { return t->layout; }
const double* btensor_data(mchip_btensor t) { return t->data.data(); }
}  // namespace mchip

extern "C" int mchip_basis_create(int nshell, const int* l, const int* nprim,
                                  const double* exps, const double* coefs,
                                  const double* centers, const int* atom,
                                  mchip_basis* out)
{
  MCHIP_TRY
  auto b = std::make_unique<mchip_basis_s>();
  size_t off = 0;
  for (int i = 0; i < nshell; ++i) {
    Shell s;
    s.l = l[i];
    s.atom = atom[i];
    for (int d = 0; d < 3; ++d) s.O[d] = centers[3 * i + d];
    s.alpha.assign(exps + off, exps + off + nprim[i]);
    s.coeff.assign(coefs + off, coefs + off + nprim[i]);
    off += nprim[i];
    normalize_shell(s);
    b->basis.shells.push_back(std::move(s));
  }
  b->basis.finalize();
  *out = b.release();
  MCHIP_CATCH
}

extern "C" int mchip_basis_destroy(mchip_basis b)
{
  delete b;
  return MCHIP_OK;
}

extern "C" int mchip_basis_nbf(mchip_basis b, int* nbf)
{
  MCHIP_TRY
  *nbf = b->basis.nbf;
  MCHIP_CATCH
}

extern "C" int mchip_basis_norm_coefs(mchip_basis b, double* coefs)
{
  MCHIP_TRY
  size_t off = 0;
  for (const auto& s : b->basis.shells) {
    std::memcpy(coefs + off, s.coeff.data(), s.coeff.size() * sizeof(double));
    off += s.coeff.size();
  }
  MCHIP_CATCH
}

extern "C" int mchip_boys(int mmax, double T, double* F)
{
  MCHIP_TRY
  if (mmax < 0 || mmax > md::BOYS_MMAX) throw std::runtime_error("mchip_boys: mmax out of range");
  boys_eval(mmax, T, F);
  MCHIP_CATCH
}

extern "C" int mchip_int1e(mchip_basis b, int natom, const double* Z, const double* xyz,
                           double* S, double* T, double* V)
{
  MCHIP_TRY
  std::vector<Charge> ch(natom);
  for (int i = 0; i < natom; ++i) {
    ch[i].Z = Z[i];
    for (int d = 0; d < 3; ++d) ch[i].xyz[d] = xyz[3 * i + d];
  }
  ints_1e(b->basis, ch, S, T, V);
  MCHIP_CATCH
}

extern "C" int mchip_int2c2e(mchip_basis aux, double* V)
{
  MCHIP_TRY
  ints_2c2e(aux->basis, V);
  MCHIP_CATCH
}

extern "C" int mchip_schwarz(mchip_basis ao, mchip_basis aux, double* Qao, double* Qaux)
{
  MCHIP_TRY
  if (Qao) schwarz_ao(ao->basis, Qao);
  if (Qaux) schwarz_aux(aux->basis, Qaux);
  MCHIP_CATCH
}

extern "C" int mchip_schwarz_mask(mchip_basis ao, mchip_basis aux, int nbb, const int* ao_blk_shells, int nxb,
                                  const int* aux_blk_shells, double eps, uint8_t* mask, double* norm_mn,
                                  double* norm_x)
{
  MCHIP_TRY
  const int nsa = (int)ao->basis.shells.size(), nsx = (int)aux->basis.shells.size();
  std::vector<int> boff(nbb + 1, 0), xoff(nxb + 1, 0);
  for (int i = 0; i < nbb; ++i) boff[i + 1] = boff[i] + ao_blk_shells[i];
  for (int i = 0; i < nxb; ++i) xoff[i + 1] = xoff[i] + aux_blk_shells[i];
  if (boff[nbb] != nsa || xoff[nxb] != nsx)
    throw std::runtime_error("mchip_schwarz_mask: block shell counts do not cover the basis");
  std::vector<double> Qao((size_t)nsa * nsa), Qx(nsx);
  schwarz_ao(ao->basis, Qao.data());
  schwarz_aux(aux->basis, Qx.data());
  // Frobenius block norms (sums in shell order, row by row)
  std::vector<double> nmn((size_t)nbb * nbb), nx(nxb);
  for (int M = 0; M < nbb; ++M)
    for (int N = 0; N < nbb; ++N) {
      double s = 0.0;
      for (int i = boff[M]; i < boff[M + 1]; ++i)
        for (int j = boff[N]; j < boff[N + 1]; ++j) s += Qao[(size_t)i * nsa + j] * Qao[(size_t)i * nsa + j];
      nmn[(size_t)M * nbb + N] = std::sqrt(s);
    }
  for (int X = 0; X < nxb; ++X) {
    double s = 0.0;
    for (int i = xoff[X]; i < xoff[X + 1]; ++i) s += Qx[i] * Qx[i];
    nx[X] = std::sqrt(s);
  }
  // keep (X, M, N) iff (float)(|Q_X| |Q_MN|) > eps: the product in double,
  // rounded to float, compared in double (screening.cpp:23-30)
  for (int X = 0; X < nxb; ++X)
    for (int MN = 0; MN < nbb * nbb; ++MN)
      mask[(size_t)X * nbb * nbb + MN] = (double)(float)(nx[X] * nmn[MN]) > eps ? 1 : 0;
  if (norm_mn) std::copy(nmn.begin(), nmn.end(), norm_mn);
  if (norm_x) std::copy(nx.begin(), nx.end(), norm_x);
  MCHIP_CATCH
}

extern "C" int mchip_int3c2e_shell(mchip_basis ao, mchip_basis aux, int ia, int ib, int ic,
                                   double* out)
{
  MCHIP_TRY
  const int nsa = (int)ao->basis.shells.size(), nsx = (int)aux->basis.shells.size();
  if (ia < 0 || ib < 0 || ic < 0 || ia >= nsa || ib >= nsa || ic >= nsx)
    throw std::runtime_error("mchip_int3c2e_shell: shell index out of range");
  ints_3c2e_shell(ao->basis, aux->basis, ia, ib, ic, out);
  MCHIP_CATCH
}

extern "C" int mchip_btensor_build(mchip_basis ao, mchip_basis aux,
                                   int nbb, const int* ao_blk_shells,
                                   int nxb, const int* aux_blk_shells,
                                   const uint8_t* schwarz_mask, double filter_eps,
                                   mchip_btensor* out)
{
  MCHIP_TRY
  std::vector<int> bsh(ao_blk_shells, ao_blk_shells + nbb);
  std::vector<int> xsh(aux_blk_shells, aux_blk_shells + nxb);
  int sb = 0, sx = 0;
  for (int v : bsh) sb += v;
  for (int v : xsh) sx += v;
  if (sb != (int)ao->basis.shells.size() || sx != (int)aux->basis.shells.size())
    throw std::runtime_error("mchip_btensor_build: block shell counts do not cover the basis");
  std::vector<uint8_t> msk(schwarz_mask, schwarz_mask + (size_t)nxb * nbb * nbb);
  auto t = std::make_unique<mchip_btensor_s>();
  std::vector<uint8_t> fmask;
  build_B(ao->basis, aux->basis, bsh, xsh, msk, filter_eps, t->layout, t->data, fmask);
  *out = t.release();
  MCHIP_CATCH
}

extern "C" int mchip_btensor_from_dense(int nbb, const int* bsz, int nxb, const int* xsz,
                                        const uint8_t* mask, const double* dense,
                                        mchip_btensor* out)
{
  MCHIP_TRY
  auto t = std::make_unique<mchip_btensor_s>();
  std::vector<int> b(bsz, bsz + nbb), x(xsz, xsz + nxb);
  std::vector<uint8_t> m(mask, mask + (size_t)nxb * nbb * nbb);
  t->layout.build(x, b, m);
  const BLayout& L = t->layout;
  t->data.assign((size_t)L.total, 0.0);
  const int nbf = L.nbf;
  for (int n = 0; n < nbb; ++n) {
    double* slab = t->data.data() + L.slab_off[n];
    for (int p = L.pair_ptr[n]; p < L.pair_ptr[n + 1]; ++p) {
      const int mb = L.pair_m[p], col = L.pair_col[p];
      for (int xb = 0; xb < nxb; ++xb) {
        if (!L.kept(xb, mb, n)) continue;
        for (int xi = L.xoff[xb]; xi < L.xoff[xb + 1]; ++xi)
          for (int nu = 0; nu < L.bsz[n]; ++nu)
            for (int mu = 0; mu < L.bsz[mb]; ++mu)
              slab[((int64_t)xi * L.bsz[n] + nu) * L.ncol[n] + col + mu] =
                  dense[((int64_t)xi * nbf + L.boff[mb] + mu) * nbf + L.boff[n] + nu];
      }
    }
  }
  *out = t.release();
  MCHIP_CATCH
}

extern "C" int mchip_btensor_destroy(mchip_btensor t)
{
  delete t;
  return MCHIP_OK;
}

extern "C" int mchip_btensor_info(mchip_btensor t, int* dims, int64_t* nelem)
{
  MCHIP_TRY
  dims[0] = t->layout.naux;
  dims[1] = t->layout.nbf;
  dims[2] = t->layout.nxb;
  dims[3] = t->layout.nbb;
  *nelem = t->layout.total;
  MCHIP_CATCH
}

extern "C" int mchip_btensor_mask(mchip_btensor t, uint8_t* mask)
{
  MCHIP_TRY
  std::memcpy(mask, t->layout.mask.data(), t->layout.mask.size());
  MCHIP_CATCH
}

extern "C" int mchip_btensor_to_dense(mchip_btensor t, double* dense)
{
  MCHIP_TRY
  const BLayout& L = t->layout;
  const int nbf = L.nbf;
  std::memset(dense, 0, sizeof(double) * (size_t)L.naux * nbf * nbf);
  for (int n = 0; n < L.nbb; ++n) {
    const double* slab = t->data.data() + L.slab_off[n];
    for (int p = L.pair_ptr[n]; p < L.pair_ptr[n + 1]; ++p) {
      const int mb = L.pair_m[p], col = L.pair_col[p];
      for (int xi = 0; xi < L.naux; ++xi)
        for (int nu = 0; nu < L.bsz[n]; ++nu)
          for (int mu = 0; mu < L.bsz[mb]; ++mu)
            dense[((int64_t)xi * nbf + L.boff[mb] + mu) * nbf + L.boff[n] + nu] =
                slab[((int64_t)xi * L.bsz[n] + nu) * L.ncol[n] + col + mu];
    }
  }
  MCHIP_CATCH
}

extern "C" int mchip_btensor_dense_rows(mchip_btensor t, int x0, int x1, double* dense)
{
  MCHIP_TRY
  const BLayout& L = t->layout;
  if (x0 < 0 || x1 > L.naux || x0 > x1) throw std::runtime_error("mchip_btensor_dense_rows: bad range");
  const int nbf = L.nbf;
  std::memset(dense, 0, sizeof(double) * (size_t)(x1 - x0) * nbf * nbf);
  for (int n = 0; n < L.nbb; ++n) {
    const double* slab = t->data.data() + L.slab_off[n];
    for (int p = L.pair_ptr[n]; p < L.pair_ptr[n + 1]; ++p) {
      const int mb = L.pair_m[p], col = L.pair_col[p];
      for (int xi = x0; xi < x1; ++xi)
        for (int nu = 0; nu < L.bsz[n]; ++nu)
          for (int mu = 0; mu < L.bsz[mb]; ++mu)
            dense[((int64_t)(xi - x0) * nbf + L.boff[mb] + mu) * nbf + L.boff[n] + nu] =
                slab[((int64_t)xi * L.bsz[n] + nu) * L.ncol[n] + col + mu];
    }
  }
  MCHIP_CATCH
}
