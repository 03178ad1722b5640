// GPU evaluation of the screened three-centre tensor B(X|mu nu).
//
// Same McMurchie-Davidson formulation as the host engine (ints/md.hpp,
// ints/ints_host.cpp): one thread per (AO shell pair, auxiliary shell) of a
// block triple the Schwarz mask keeps (screening.cpp:23-30 /
// aofactory.cpp:52-143), writing the canonical pair store [x][mu][nu];
// then FP64 block norms, the block filter (aoloader.cpp:373) decided on the
// host, and a scatter into the pair-slab layout (both orientations).
// AO shells up to d, auxiliary shells up to f, uncontracted auxiliary basis.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <set>
#include <vector>

#include "../ints/ints_host.hpp"
#include "../ints/md.hpp"
#include "dev.hpp"
#include "ints_dev.hpp"
#include "kernels.hpp"

namespace mchip {

using namespace md;

constexpr int AO_L = 2;                          // max AO angular momentum on the GPU path
constexpr int AO_NC = (AO_L + 1) * (AO_L + 2) / 2;
constexpr int AX_NM = 2 * LMAX + 1;
constexpr int EDIM = (AO_L + 1) * (AO_L + 1) * (2 * AO_L + 1);

__constant__ double c_sph[LMAX + 1][2 * LMAX + 1][(LMAX + 1) * (LMAX + 2) / 2];

__device__ __forceinline__ int eix(int i, int j, int t) { return (i * (AO_L + 1) + j) * (2 * AO_L + 1) + t; }

__device__ void herm_E_ao(int la, int lb, double XPA, double XPB, double oo2p, double* E)
{
  for (int n = 0; n < EDIM; ++n) E[n] = 0.0;
  E[0] = 1.0;
  for (int i = 0; i < la; ++i)
    for (int t = 0; t <= i + 1; ++t) {
      double v = XPA * E[eix(i, 0, t)];
      if (t > 0) v += oo2p * E[eix(i, 0, t - 1)];
      if (t + 1 <= i) v += (t + 1) * E[eix(i, 0, t + 1)];
      E[eix(i + 1, 0, t)] = v;
    }
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j < lb; ++j)
      for (int t = 0; t <= i + j + 1; ++t) {
        double v = XPB * E[eix(i, j, t)];
        if (t > 0) v += oo2p * E[eix(i, j, t - 1)];
        if (t + 1 <= i + j) v += (t + 1) * E[eix(i, j, t + 1)];
        E[eix(i, j + 1, t)] = v;
      }
}

__global__ __launch_bounds__(64) void k_int3c2e(
    const int* __restrict__ ao_l, const int* __restrict__ ao_np, const int* __restrict__ ao_po,
    const double* __restrict__ ao_O, const double* __restrict__ ao_ex, const double* __restrict__ ao_cf,
    const int* __restrict__ ax_l, const double* __restrict__ ax_g, const double* __restrict__ ax_O,
    const double* __restrict__ ax_H, const int* __restrict__ ax_x0, const int* __restrict__ ax_blk,
    int nsh_aux, int nchunk, const Task3* __restrict__ tasks, const int64_t* __restrict__ cp_off,
    const int* __restrict__ bsz, const uint8_t* __restrict__ mask, int nbb,
    const double* __restrict__ boys_tab, double* __restrict__ tmp)
{
  const int task = blockIdx.x / nchunk;
  const int c = (blockIdx.x % nchunk) * 64 + threadIdx.x;
  if (c >= nsh_aux) return;
  const Task3 tk = tasks[task];
  const int xb = ax_blk[c];
  if (!mask[((size_t)xb * nbb + tk.M) * nbb + tk.N]) return;
  const int a = tk.sa, b = tk.sb;
  const int la = ao_l[a], lb = ao_l[b], lc = ax_l[c];
  const int lab = la + lb, L = lab + lc;
  const int nca = ncart(la), ncb = ncart(lb), nmc = nsph(lc);
  const double* A = ao_O + 3 * a;
  const double* Bc = ao_O + 3 * b;
  const double* C = ax_O + 3 * c;
  const double gam = ax_g[c];
  const double* H = ax_H + (size_t)c * AX_NM * nherm(LMAX);
  const int NHc = nherm(lc);
  double acc[AX_NM * AO_NC * AO_NC];
  for (int i = 0; i < nmc * nca * ncb; ++i) acc[i] = 0.0;
  double Ex[EDIM], Ey[EDIM], Ez[EDIM];
  double r0[nherm(2 * AO_L + LMAX)], r1[nherm(2 * AO_L + LMAX)];
  double Sm[nherm(2 * AO_L)];
  const double AB2 = (A[0] - Bc[0]) * (A[0] - Bc[0]) + (A[1] - Bc[1]) * (A[1] - Bc[1]) +
                     (A[2] - Bc[2]) * (A[2] - Bc[2]);
  const int pa0 = ao_po[a], pb0 = ao_po[b];
  for (int i = 0; i < ao_np[a]; ++i) {
    const double al = ao_ex[pa0 + i];
    for (int j = 0; j < ao_np[b]; ++j) {
      const double be = ao_ex[pb0 + j];
      const double p = al + be, mu = al * be / p;
      if (mu * AB2 > 41.4) continue;                 // exp(-mu R^2) < 1e-18
      const double K = ao_cf[pa0 + i] * ao_cf[pb0 + j] * exp(-mu * AB2);
      double P[3];
      for (int d = 0; d < 3; ++d) P[d] = (al * A[d] + be * Bc[d]) / p;
      const double oo2p = 0.5 / p;
      herm_E_ao(la, lb, P[0] - A[0], P[0] - Bc[0], oo2p, Ex);
      herm_E_ao(la, lb, P[1] - A[1], P[1] - Bc[1], oo2p, Ey);
      herm_E_ao(la, lb, P[2] - A[2], P[2] - Bc[2], oo2p, Ez);
      const double q = gam, alpha = p * q / (p + q);
      const double* R = hermite_R2(L, alpha, P[0] - C[0], P[1] - C[1], P[2] - C[2], boys_tab, r0, r1);
      const double pref = 2.0 * PI * PI * sqrt(PI) / (p * q * sqrt(p + q)) * K;
      for (int m = 0; m < nmc; ++m) {
        const double* h = H + m * nherm(LMAX);
        for (int t = 0; t <= lab; ++t)
          for (int u = 0; u <= lab - t; ++u)
            for (int v = 0; v <= lab - t - u; ++v) {
              double x = 0.0;
              for (int t2 = 0; t2 <= lc; ++t2)
                for (int u2 = 0; u2 <= lc - t2; ++u2)
                  for (int v2 = 0; v2 <= lc - t2 - u2; ++v2) {
                    const double hv = h[herm_index(t2, u2, v2)];
                    if (hv != 0.0) x += hv * R[herm_index(t + t2, u + u2, v + v2)];
                  }
              Sm[herm_index(t, u, v)] = x * pref;
            }
        for (int ka = 0; ka < nca; ++ka) {
          int ax, ay, az;
          cart_exps(la, ka, ax, ay, az);
          for (int kb = 0; kb < ncb; ++kb) {
            int bx, by, bz;
            cart_exps(lb, kb, bx, by, bz);
            double r = 0.0;
            for (int t = 0; t <= ax + bx; ++t) {
              const double ex = Ex[eix(ax, bx, t)];
              for (int u = 0; u <= ay + by; ++u) {
                const double exy = ex * Ey[eix(ay, by, u)];
                for (int v = 0; v <= az + bz; ++v) r += exy * Ez[eix(az, bz, v)] * Sm[herm_index(t, u, v)];
              }
            }
            acc[(m * nca + ka) * ncb + kb] += r;
          }
        }
      }
    }
  }
  (void)NHc;
  // spherical transform and store into the canonical pair block [x][mu][nu]
  const int nsa = nsph(la), nsb = nsph(lb);
  const int BM = bsz[tk.M], BN = bsz[tk.N];
  double* base = tmp + cp_off[tk.cp];
  const int x0 = ax_x0[c];
  for (int m = 0; m < nmc; ++m)
    for (int sa = 0; sa < nsa; ++sa)
      for (int sb = 0; sb < nsb; ++sb) {
        double x = 0.0;
        for (int ka = 0; ka < nca; ++ka) {
          const double ca = c_sph[la][sa][ka];
          if (ca == 0.0) continue;
          for (int kb = 0; kb < ncb; ++kb) x += ca * c_sph[lb][sb][kb] * acc[(m * nca + ka) * ncb + kb];
        }
        double* dst = base + (int64_t)(x0 + m) * BM * BN;
        dst[(tk.mu0 + sa) * BN + (tk.nu0 + sb)] = x;
        if (tk.M == tk.N) dst[(tk.nu0 + sb) * BN + (tk.mu0 + sa)] = x;
      }
}

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
  std::vector<Task3> tasks;
  for (int k = 0; k < ncp; ++k) {
    const int m = cpm[k], n = cpn[k];
    for (int sa = ao_sh0[m]; sa < ao_sh0[m + 1]; ++sa)
      for (int sb = ao_sh0[n]; sb < ao_sh0[n + 1]; ++sb) {
        if (m == n && sb < sa) continue;
        tasks.push_back(Task3{sa, sb, k, ao.shell_off[sa] - boff[m], ao.shell_off[sb] - boff[n], m, n, 0});
      }
  }
  // shell arrays
  std::vector<int> aol, aonp, aopo;
  std::vector<double> aoO, aoex, aocf;
  for (const auto& s : ao.shells) {
    aol.push_back(s.l);
    aonp.push_back((int)s.alpha.size());
    aopo.push_back((int)aoex.size());
    for (int d = 0; d < 3; ++d) aoO.push_back(s.O[d]);
    aoex.insert(aoex.end(), s.alpha.begin(), s.alpha.end());
    aocf.insert(aocf.end(), s.coeff.begin(), s.coeff.end());
  }
  const int nsx = (int)aux.shells.size();
  std::vector<int> axl(nsx), axx0(nsx), axblk(nsx);
  std::vector<double> axg(nsx), axO(3 * nsx), axH((size_t)nsx * AX_NM * nherm(LMAX), 0.0);
  for (int x = 0; x < nxb; ++x)
    for (int s = aux_sh0[x]; s < aux_sh0[x + 1]; ++s) axblk[s] = x;
  for (int s = 0; s < nsx; ++s) {
    const Shell& sh = aux.shells[s];
    axl[s] = sh.l;
    axg[s] = sh.alpha[0];
    axx0[s] = aux.shell_off[s];
    for (int d = 0; d < 3; ++d) axO[3 * s + d] = sh.O[d];
    aux_hermite_coeffs(sh.l, sh.alpha[0], sh.coeff[0], axH.data() + (size_t)s * AX_NM * nherm(LMAX), nherm(LMAX));
  }
  {
    // the constant table is written once per device (request workers share it)
    static std::mutex mu;
    static std::set<int> done;
    std::lock_guard<std::mutex> lk(mu);
    if (!done.count(c.device)) {
      const SphTable& S = sph_table();
      HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_sph), &S.c[0][0][0], sizeof(S.c)));
      done.insert(c.device);
    }
  }
  const auto& btab = boys_table();
  DevBuf d_aol = upload_vec(aol, c.stream), d_aonp = upload_vec(aonp, c.stream),
         d_aopo = upload_vec(aopo, c.stream), d_aoO = upload_vec(aoO, c.stream),
         d_aoex = upload_vec(aoex, c.stream), d_aocf = upload_vec(aocf, c.stream),
         d_axl = upload_vec(axl, c.stream), d_axg = upload_vec(axg, c.stream),
         d_axO = upload_vec(axO, c.stream), d_axH = upload_vec(axH, c.stream),
         d_axx0 = upload_vec(axx0, c.stream), d_axblk = upload_vec(axblk, c.stream),
         d_tasks = upload_vec(tasks, c.stream), d_cpoff = upload_vec(cpoff, c.stream),
         d_bsz = upload_vec(bsz, c.stream), d_mask = upload_vec(schwarz_mask, c.stream),
         d_boys = upload_vec(btab, c.stream), d_cpsize = upload_vec(cpsize, c.stream),
         d_xoff = upload_vec(xoff, c.stream);
  DevBuf tmp((size_t)std::max<int64_t>(ctotal, 1) * sizeof(double));
  hipLaunchKernelGGL(k_fill, dim3(gs_blocks(std::max<int64_t>(ctotal, 1))), dim3(256), 0, c.stream,
                     tmp.as<double>(), std::max<int64_t>(ctotal, 1), 0.0);
  const int nchunk = (nsx + 63) / 64;
  const int64_t nblocks = (int64_t)tasks.size() * nchunk;
  // launch in slices of whole tasks: a dispatch holds fewer than 2^32
  // work-items (64 per block), and each slice starts at a task boundary
  const int64_t slice = std::max<int64_t>(1, ((int64_t)1 << 25) / nchunk) * nchunk;
  for (int64_t b0 = 0; b0 < nblocks; b0 += slice) {
    const int64_t nb = std::min(slice, nblocks - b0);
    const int64_t t0 = b0 / nchunk;
    hipLaunchKernelGGL(k_int3c2e, dim3((unsigned)nb), dim3(64), 0, c.stream, d_aol.as<int>(),
                       d_aonp.as<int>(), d_aopo.as<int>(), d_aoO.as<double>(), d_aoex.as<double>(),
                       d_aocf.as<double>(), d_axl.as<int>(), d_axg.as<double>(), d_axO.as<double>(),
                       d_axH.as<double>(), d_axx0.as<int>(), d_axblk.as<int>(), nsx, nchunk,
                       d_tasks.as<Task3>() + t0, d_cpoff.as<int64_t>(), d_bsz.as<int>(),
                       d_mask.as<uint8_t>(), nbb, d_boys.as<double>(), tmp.as<double>());
    HIP_CHECK(hipGetLastError());
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
