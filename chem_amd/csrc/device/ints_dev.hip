// GPU evaluation of the screened three-centre tensor B(X|mu nu).
//
// One kernel per angular momentum class (la, lb, lc) (ints3c_cls.hpp): one
// thread per (AO shell pair, auxiliary shell of degree lc) of a block triple
// the Schwarz mask keeps (screening.cpp:23-30 / aofactory.cpp:52-143),
// writing the canonical pair store [x][mu][nu]; then FP64 block norms, the
// block filter (aoloader.cpp:373) decided on the host, and a scatter into the
// pair-slab layout (both orientations).  AO shells up to d, auxiliary shells
// up to f, uncontracted auxiliary basis.  A wave works on one AO shell pair
// and 64 auxiliary shells of one degree.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../ints/ints_host.hpp"
#include "dev.hpp"
#include "ints3c_cls.hpp"
#include "ints_dev.hpp"
#include "kernels.hpp"

namespace mchip {

// norm2[cp * nxb + xb] = sum of squares of rows xoff[xb]..xoff[xb+1] of pair cp
__global__ void k_pair_block_norm2(const double* __restrict__ tmp, const int64_t* __restrict__ cp_off,
                                   const int* __restrict__ cp_size, const int* __restrict__ xoff, int nxb,
                                   double* __restrict__ norm2)
{
  const int cp = blockIdx.x, xb = blockIdx.y;   // grid = (pairs, aux blocks)
  const int64_t per = cp_size[cp];
  const double* src = tmp + cp_off[cp] + (int64_t)xoff[xb] * per;
  const int64_t n = (int64_t)(xoff[xb + 1] - xoff[xb]) * per;
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += src[i] * src[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) norm2[(int64_t)cp * nxb + xb] = red[0];
}

// copy kept blocks of canonical pair cp into slab N (block M at col_mn) and
// slab M (block N at col_nm); grid = (pairs, naux chunks)
__global__ void k_pair_scatter_slabs(const double* __restrict__ tmp, const int64_t* __restrict__ cp_off,
                                     const int* __restrict__ cp_m, const int* __restrict__ cp_n,
                                     const int* __restrict__ col_mn, const int* __restrict__ col_nm,
                                     const int* __restrict__ bsz, const int* __restrict__ ncol,
                                     const int64_t* __restrict__ slab_off, const uint8_t* __restrict__ keep,
                                     const int* __restrict__ x_blk, int nxb, int naux, int xper,
                                     double* __restrict__ out)
{
  const int cp = blockIdx.x;
  if (col_mn[cp] < 0) return;
  const int m = cp_m[cp], n = cp_n[cp];
  const int BM = bsz[m], BN = bsz[n];
  const int x0 = blockIdx.y * xper, x1 = min(naux, x0 + xper);
  double* slabN = out + slab_off[n];
  double* slabM = out + slab_off[m];
  const int ncN = ncol[n], ncM = ncol[m];
  for (int xi = x0; xi < x1; ++xi) {
    if (!keep[(size_t)cp * nxb + x_blk[xi]]) continue;
    const double* src = tmp + cp_off[cp] + (int64_t)xi * BM * BN;
    for (int e = threadIdx.x; e < BM * BN; e += blockDim.x) {
      const int mu = e / BN, nu = e % BN;
      const double v = src[e];
      slabN[((int64_t)xi * BN + nu) * ncN + col_mn[cp] + mu] = v;
      slabM[((int64_t)xi * BM + mu) * ncM + col_nm[cp] + nu] = v;
    }
  }
}

// sinv[l][m] with c_sph[l][m][k] = shp(l, m, k) * sinv[l][m]: read off a
// monomial whose integer weight is a power of two (the division is exact),
// and the integer shapes checked against the host table.
void solid_harmonic_norms(double (&inv)[LMAX + 1][2 * LMAX + 1])
{
  const SphTable& S = sph_table();
  for (int l = 0; l <= LMAX; ++l)
    for (int m = 0; m < 2 * LMAX + 1; ++m) inv[l][m] = 0.0;
  for (int l = 0; l <= LMAX; ++l)
    for (int m = 0; m < 2 * l + 1; ++m) {
      bool found = false;
      for (int k = 0; k < i3c::nc(l) && !found; ++k) {
        const int w = i3c::shp(l, m, k);
        if (w == 1 || w == -1 || w == 2 || w == -2 || w == 4 || w == -4) {
          inv[l][m] = S.c[l][m][k] / w;
          found = true;
        }
      }
      if (!found) throw std::runtime_error("GPU integrals: solid harmonic normalisation");
      for (int k = 0; k < i3c::nc(l); ++k)
        if (std::fabs(i3c::shp(l, m, k) * inv[l][m] - S.c[l][m][k]) > 1e-15 * std::fabs(inv[l][m]))
          throw std::runtime_error("GPU integrals: solid harmonic shapes disagree with the host table");
    }
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
void build_B_device(Ctx& c, const Basis& ao, const Basis& aux, const std::vector<int>& ao_blk_shells,
                    const std::vector<int>& aux_blk_shells, const std::vector<uint8_t>& schwarz_mask,
                    double filter_eps, DevBT& t)
{
  HIP_CHECK(hipSetDevice(c.device));
  const int nbb = (int)ao_blk_shells.size(), nxb = (int)aux_blk_shells.size();
  for (const auto& s : ao.shells)
    if (s.l > AO_L) throw std::runtime_error("GPU integrals: AO shells above d are not supported");
  for (const auto& s : aux.shells)
    if (s.alpha.size() != 1) throw std::runtime_error("GPU integrals: the auxiliary basis must be uncontracted");
  std::vector<int> ao_sh0(nbb + 1, 0), aux_sh0(nxb + 1, 0);
  for (int i = 0; i < nbb; ++i) ao_sh0[i + 1] = ao_sh0[i] + ao_blk_shells[i];
  for (int i = 0; i < nxb; ++i) aux_sh0[i + 1] = aux_sh0[i] + aux_blk_shells[i];
  std::vector<int> bsz(nbb), xsz(nxb), boff(nbb + 1, 0), xoff(nxb + 1, 0);
  for (int i = 0; i < nbb; ++i) {
    bsz[i] = 0;
    for (int s = ao_sh0[i]; s < ao_sh0[i + 1]; ++s) bsz[i] += nsph(ao.shells[s].l);
    boff[i + 1] = boff[i] + bsz[i];
  }
  for (int i = 0; i < nxb; ++i) {
    xsz[i] = 0;
    for (int s = aux_sh0[i]; s < aux_sh0[i + 1]; ++s) xsz[i] += nsph(aux.shells[s].l);
    xoff[i + 1] = xoff[i] + xsz[i];
  }
  const int naux = xoff[nxb];
  // canonical significant pairs
  std::vector<int> cpm, cpn, cpsize;
  std::vector<int64_t> cpoff;
  int64_t ctotal = 0;
  for (int n = 0; n < nbb; ++n)
    for (int m = 0; m <= n; ++m) {
      bool sig = false;
      for (int x = 0; x < nxb && !sig; ++x) sig = schwarz_mask[((size_t)x * nbb + m) * nbb + n] != 0;
      if (!sig) continue;
      cpm.push_back(m);
      cpn.push_back(n);
      cpsize.push_back(bsz[m] * bsz[n]);
      cpoff.push_back(ctotal);
      ctotal += (int64_t)naux * bsz[m] * bsz[n];
    }
  const int ncp = (int)cpm.size();
  // tasks grouped by class (la, lb): cls_t0[la * (AO_L + 1) + lb] .. next
  constexpr int NAB = (AO_L + 1) * (AO_L + 1);
  std::vector<std::vector<Task3>> by_ab(NAB);
  for (int k = 0; k < ncp; ++k) {
    const int m = cpm[k], n = cpn[k];
    for (int sa = ao_sh0[m]; sa < ao_sh0[m + 1]; ++sa)
      for (int sb = ao_sh0[n]; sb < ao_sh0[n + 1]; ++sb) {
        if (m == n && sb < sa) continue;
        by_ab[ao.shells[sa].l * (AO_L + 1) + ao.shells[sb].l].push_back(
            Task3{sa, sb, k, ao.shell_off[sa] - boff[m], ao.shell_off[sb] - boff[n], m, n, 0});
      }
  }
  std::vector<Task3> tasks;
  std::vector<int64_t> cls_t0(NAB + 1, 0);
  for (int k = 0; k < NAB; ++k) {
    cls_t0[k] = (int64_t)tasks.size();
    tasks.insert(tasks.end(), by_ab[k].begin(), by_ab[k].end());
  }
  cls_t0[NAB] = (int64_t)tasks.size();
  // shell arrays
  std::vector<int> aonp, aopo;
  std::vector<double> aoO, aoex, aocf;
  for (const auto& s : ao.shells) {
    aonp.push_back((int)s.alpha.size());
    aopo.push_back((int)aoex.size());
    for (int d = 0; d < 3; ++d) aoO.push_back(s.O[d]);
    aoex.insert(aoex.end(), s.alpha.begin(), s.alpha.end());
    aocf.insert(aocf.end(), s.coeff.begin(), s.coeff.end());
  }
  const int nsx = (int)aux.shells.size();
  std::vector<int> axx0(nsx), axblk(nsx);
  std::vector<double> axg(nsx), axO(3 * nsx), axs(nsx);
  std::vector<int> shl;                       // auxiliary shells by degree, basis order within
  std::vector<int> shl0(LMAX + 2, 0);
  for (int x = 0; x < nxb; ++x)
    for (int s = aux_sh0[x]; s < aux_sh0[x + 1]; ++s) axblk[s] = x;
  for (int s = 0; s < nsx; ++s) {
    const Shell& sh = aux.shells[s];
    if (sh.l > LMAX) throw std::runtime_error("GPU integrals: auxiliary shells above f are not supported");
    axg[s] = sh.alpha[0];
    axx0[s] = aux.shell_off[s];
    for (int d = 0; d < 3; ++d) axO[3 * s + d] = sh.O[d];
    // (-1)^l coef (2 gamma)^-l: the top-degree Hermite coefficients of the
    // solid harmonic Gaussian without its integer monomial weights
    double v = sh.coeff[0];
    for (int i = 0; i < sh.l; ++i) v *= 0.5 / sh.alpha[0];
    axs[s] = (sh.l & 1) ? -v : v;
  }
  for (int l = 0; l <= LMAX; ++l) {
    shl0[l] = (int)shl.size();
    for (int s = 0; s < nsx; ++s)
      if (aux.shells[s].l == l) shl.push_back(s);
  }
  shl0[LMAX + 1] = (int)shl.size();
  I3cArgs ia{};
  solid_harmonic_norms(ia.sinv);
  const auto& btab = boys_table();
  DevBuf d_aonp = upload_vec(aonp, c.stream), d_aopo = upload_vec(aopo, c.stream),
         d_aoO = upload_vec(aoO, c.stream), d_aoex = upload_vec(aoex, c.stream),
         d_aocf = upload_vec(aocf, c.stream), d_axg = upload_vec(axg, c.stream),
         d_axO = upload_vec(axO, c.stream), d_axs = upload_vec(axs, c.stream),
         d_axx0 = upload_vec(axx0, c.stream), d_axblk = upload_vec(axblk, c.stream),
         d_shl = upload_vec(shl, c.stream), d_tasks = upload_vec(tasks, c.stream),
         d_cpoff = upload_vec(cpoff, c.stream), d_bsz = upload_vec(bsz, c.stream),
         d_mask = upload_vec(schwarz_mask, c.stream), d_boys = upload_vec(btab, c.stream),
         d_cpsize = upload_vec(cpsize, c.stream), d_xoff = upload_vec(xoff, c.stream);
  DevBuf tmp((size_t)std::max<int64_t>(ctotal, 1) * sizeof(double));
  hipLaunchKernelGGL(k_fill, dim3(gs_blocks(std::max<int64_t>(ctotal, 1))), dim3(256), 0, c.stream,
                     tmp.as<double>(), std::max<int64_t>(ctotal, 1), 0.0);
  for (int ab = 0; ab < NAB; ++ab) {
    const int la = ab / (AO_L + 1), lb = ab % (AO_L + 1);
    const int64_t nt = cls_t0[ab + 1] - cls_t0[ab];
    for (int lc = 0; lc <= LMAX; ++lc) {
      const int nshl = shl0[lc + 1] - shl0[lc];
      if (nt == 0 || nshl == 0) continue;
      const int nchunk = (nshl + 63) / 64;
      const int64_t nblocks = nt * nchunk;
      // launch in slices of whole tasks: a dispatch holds fewer than 2^32
      // work-items (64 per block), and each slice starts at a task boundary
      const int64_t slice = std::max<int64_t>(1, ((int64_t)1 << 25) / nchunk) * nchunk;
      for (int64_t b0 = 0; b0 < nblocks; b0 += slice) {
        const int64_t nb = std::min(slice, nblocks - b0);
        I3cArgs a = ia;
        a.ao_np = d_aonp.as<int>();
        a.ao_po = d_aopo.as<int>();
        a.ao_O = d_aoO.as<double>();
        a.ao_ex = d_aoex.as<double>();
        a.ao_cf = d_aocf.as<double>();
        a.shl = d_shl.as<int>() + shl0[lc];
        a.nshl = nshl;
        a.nchunk = nchunk;
        a.ax_g = d_axg.as<double>();
        a.ax_O = d_axO.as<double>();
        a.ax_s = d_axs.as<double>();
        a.ax_x0 = d_axx0.as<int>();
        a.ax_blk = d_axblk.as<int>();
        a.tasks = d_tasks.as<Task3>() + cls_t0[ab] + b0 / nchunk;
        a.cp_off = d_cpoff.as<int64_t>();
        a.bsz = d_bsz.as<int>();
        a.mask = d_mask.as<uint8_t>();
        a.nbb = nbb;
        a.boys = d_boys.as<double>();
        a.tmp = tmp.as<double>();
        const dim3 grid((unsigned)nb);
        if (la == 0) launch_int3c2e_la0(lb, lc, grid, c.stream, a);
        else if (la == 1) launch_int3c2e_la1(lb, lc, grid, c.stream, a);
        else launch_int3c2e_la2(lb, lc, grid, c.stream, a);
        HIP_CHECK(hipGetLastError());
      }
    }
  }
  // block norms -> final mask (host)
  DevBuf norm2((size_t)std::max(1, ncp * nxb) * sizeof(double));
  if (ncp)
    hipLaunchKernelGGL(k_pair_block_norm2, dim3(ncp, nxb), dim3(256), 0, c.stream, tmp.as<double>(),
                       d_cpoff.as<int64_t>(), d_cpsize.as<int>(), d_xoff.as<int>(), nxb, norm2.as<double>());
  std::vector<double> hn((size_t)ncp * nxb);
  if (ncp) d2h(hn.data(), norm2.p, hn.size() * sizeof(double), c.stream);
  HIP_CHECK(hipStreamSynchronize(c.stream));
  std::vector<uint8_t> fmask((size_t)nxb * nbb * nbb, 0), keep((size_t)std::max(1, ncp) * nxb, 0);
  for (int k = 0; k < ncp; ++k)
    for (int x = 0; x < nxb; ++x) {
      if (!schwarz_mask[((size_t)x * nbb + cpm[k]) * nbb + cpn[k]]) continue;
      if (std::sqrt(hn[(size_t)k * nxb + x]) >= filter_eps) {
        keep[(size_t)k * nxb + x] = 1;
        fmask[((size_t)x * nbb + cpm[k]) * nbb + cpn[k]] = 1;
        fmask[((size_t)x * nbb + cpn[k]) * nbb + cpm[k]] = 1;
      }
    }
  BLayout L;
  L.build(xsz, bsz, fmask);
  std::vector<int> colmn(std::max(1, ncp), -1), colnm(std::max(1, ncp), -1);
  for (int k = 0; k < ncp; ++k) {
    colmn[k] = L.col_of(cpm[k], cpn[k]);
    colnm[k] = L.col_of(cpn[k], cpm[k]);
  }
  std::vector<int> xblk(naux);
  for (int x = 0; x < nxb; ++x)
    for (int i = xoff[x]; i < xoff[x + 1]; ++i) xblk[i] = x;
  std::vector<int64_t> so(L.slab_off.begin(), L.slab_off.end());
  DevBuf d_cpm = upload_vec(cpm, c.stream), d_cpn = upload_vec(cpn, c.stream),
         d_colmn = upload_vec(colmn, c.stream), d_colnm = upload_vec(colnm, c.stream),
         d_ncol = upload_vec(L.ncol, c.stream), d_so = upload_vec(so, c.stream),
         d_keep = upload_vec(keep, c.stream), d_xblk = upload_vec(xblk, c.stream);
  DevBuf out((size_t)std::max<int64_t>(L.total, 1) * sizeof(double));
  hipLaunchKernelGGL(k_fill, dim3(gs_blocks(std::max<int64_t>(L.total, 1))), dim3(256), 0, c.stream,
                     out.as<double>(), std::max<int64_t>(L.total, 1), 0.0);
  const int xper = 64;
  if (ncp)
    hipLaunchKernelGGL(k_pair_scatter_slabs, dim3(ncp, (naux + xper - 1) / xper), dim3(256), 0, c.stream,
                       tmp.as<double>(), d_cpoff.as<int64_t>(), d_cpm.as<int>(), d_cpn.as<int>(),
                       d_colmn.as<int>(), d_colnm.as<int>(), d_bsz.as<int>(), d_ncol.as<int>(),
                       d_so.as<int64_t>(), d_keep.as<uint8_t>(), d_xblk.as<int>(), nxb, naux, xper,
                       out.as<double>());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c.stream));
  tmp.release();
  dbt_adopt(c, L, std::move(out), t);
}

}  // namespace mchip
