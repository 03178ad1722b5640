// Restricted Hartree-Fock loop on the GPU (FP64).
//
// Mirrors hf::hfmod::compute (src/hf/hfscf.cpp:326-552):
//   one_electron   :131-181   X = orthogonalizer(S) (orthogonalizer.cpp:14-45,
//                             eigenvalues < 0 -> 0, else 1/sqrt), H = T + V,
//                             block filter(eps) of S, X, H
//   guess "core"   hfguess.cpp:127-141; "SAD" hfguess.cpp:142-452 (density
//                  from the host's atomic SCFs, first exchange on its
//                  pivoted Cholesky factor)
//   form_fock      :228-276   F = H + J + K
//   compute_errmat :298-324   e = F P S - S P F  (formed as A - A^T, A = F P S)
//   energy         :278-296   E = tr(H P) + tr(F P)
//   stop           :430       RMS(e) < threshold and iter > 0
//   DIIS           diis.hpp:44-215 (drop the largest-norm error when full)
//   diag_fock      hfdiagfock.cpp:12-211  F' = X^T F X, eig, C = X C',
//                             P = C_occ C_occ^T
// All matrices stay in HBM; only the DIIS B-matrix solve and the convergence
// test run on the host.
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <limits>
#include <memory>
#include <stdexcept>

#include "dfock.hpp"
#include "kernels.hpp"
#include "mp2.hpp"
#include "scf.hpp"

namespace mchip {

__global__ void k_block_filter(double* M, int n, const int* boff, int nb, double eps)
{
  const int bi = blockIdx.x, bj = blockIdx.y;
  const int r0 = boff[bi], r1 = boff[bi + 1], c0 = boff[bj], c1 = boff[bj + 1];
  __shared__ double red[256];
  double s = 0.0;
  const int w = c1 - c0;
  const int cnt = (r1 - r0) * w;
  for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
    const double v = M[(int64_t)(r0 + e / w) * n + c0 + e % w];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  const bool drop = sqrt(red[0]) < eps;
  if (drop)
    for (int e = threadIdx.x; e < cnt; e += blockDim.x) M[(int64_t)(r0 + e / w) * n + c0 + e % w] = 0.0;
}

// partial dot products, fixed-order final sum on the host side of the call
__global__ void k_dot_partial(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                              double* __restrict__ part)
{
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += a[i] * b[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out = A - A^T  (row-major n x n)
__global__ void k_antisym(const double* __restrict__ A, int n, double* __restrict__ out)
{
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx / n), c = (int)(idx % n);
  out[idx] = A[idx] - A[(int64_t)c * n + r];
}

// C_occ row-major (n x nocc) from column-major eigenvectors (n x n)
__global__ void k_take_occ(const double* __restrict__ Ccm, int n, int nocc, double* __restrict__ Cocc)
{
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * nocc) return;
  const int mu = (int)(idx / nocc), i = (int)(idx % nocc);
  Cocc[idx] = Ccm[(int64_t)i * n + mu];
}

// d = 0 for w < 0, else 1/sqrt(w); Ud = U * diag(d) (column-major)
__global__ void k_scale_cols(const double* __restrict__ U, const double* __restrict__ w, int n,
                             double* __restrict__ Ud)
{
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int col = (int)(idx / n);
  const double ev = w[col];
  const double d = (ev < 0.0) ? 0.0 : 1.0 / sqrt(ev);
  Ud[idx] = U[idx] * d;
}

static double dot_dev(Ctx& c, const double* a, const double* b, int64_t n, DevBuf& part)
{
  const int nb = 256;
  part.ensure(nb * sizeof(double));
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(256), 0, c.stream, a, b, n, part.as<double>());
  double h[256];
  d2h(h, part.p, nb * sizeof(double), c.stream);
  double s = 0.0;
  for (int i = 0; i < nb; ++i) s += h[i];
  return s;
}

// row-major C = A * B (all n x n): column-major view computes B^T A^T = (AB)^T
static void gemm_rm(Ctx& c, int n, const double* A, const double* B, double* C, double alpha = 1.0,
                    double beta = 0.0, bool transA = false, bool transB = false)
{
  // row-major C = op(A) op(B)  <=>  column-major C^T = op(B)^T op(A)^T
  RB_CHECK(rocblas_dgemm(c.blas, transB ? rocblas_operation_transpose : rocblas_operation_none,
                         transA ? rocblas_operation_transpose : rocblas_operation_none, n, n, n,
                         &alpha, B, n, A, n, &beta, C, n));
}

// Solve the (n+1) DIIS system by Gaussian elimination with partial pivoting.
static std::vector<double> diis_solve(const std::vector<double>& Bm, int n)
{
  const int m = n + 1;
  std::vector<double> A((size_t)m * m, 0.0), x(m, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) A[i * m + j] = Bm[i * n + j];
  for (int i = 0; i < n; ++i) {
    A[n * m + i] = -1.0;
    A[i * m + n] = -1.0;
  }
  A[n * m + n] = 0.0;
  std::vector<double> rhs(m, 0.0);
  rhs[n] = -1.0;
  for (int k = 0; k < m; ++k) {
    int piv = k;
    for (int i = k + 1; i < m; ++i)
      if (std::fabs(A[i * m + k]) > std::fabs(A[piv * m + k])) piv = i;
    if (piv != k) {
      for (int j = 0; j < m; ++j) std::swap(A[k * m + j], A[piv * m + j]);
      std::swap(rhs[k], rhs[piv]);
    }
    const double d = A[k * m + k];
    if (d == 0.0) continue;
    for (int i = k + 1; i < m; ++i) {
      const double f = A[i * m + k] / d;
      if (f == 0.0) continue;
      for (int j = k; j < m; ++j) A[i * m + j] -= f * A[k * m + j];
      rhs[i] -= f * rhs[k];
    }
  }
  for (int i = m - 1; i >= 0; --i) {
    double s = rhs[i];
    for (int j = i + 1; j < m; ++j) s -= A[i * m + j] * x[j];
    x[i] = (A[i * m + i] != 0.0) ? s / A[i * m + i] : 0.0;
  }
  x.resize(n);
  return x;
}

struct DiisState {
  std::deque<DevBuf> err, trial;
  std::vector<double> B;   // row-major nerr x nerr
  std::vector<double> coeffs;
  std::vector<DevBuf> spare;   // dropped vectors, reused so that iterations allocate nothing
  DevBuf take(size_t bytes)
  {
    if (!spare.empty()) {
      DevBuf b = std::move(spare.back());
      spare.pop_back();
      return b;
    }
    return DevBuf(bytes);
  }
};

void scf_run(Ctx& c, DevBT& bt, const Metric* m, const ScfInput& in, ScfOutput& out,
             const KBuildFn& kbuild, const KBuildFn& final_k, const EnergyFn& final_ek)
{
  const int n = in.nbf, nocc = in.nocc;
  const int64_t nn = (int64_t)n * n;
  const size_t mb = (size_t)nn * sizeof(double);
  auto alloc = [&](DevBuf& b) { b.alloc(mb); };
  DevBuf S, H, X, F, P, J, K, E, W1, W2, Cev, Cocc, w, part, boff;
  for (DevBuf* b : {&S, &H, &X, &F, &P, &J, &K, &E, &W1, &W2, &Cev}) alloc(*b);
  Cocc.alloc((size_t)n * std::max(nocc, 1) * sizeof(double));
  w.alloc(n * sizeof(double));
  DevBuf ework(n * sizeof(double));
  h2d(S.p, in.S, mb, c.stream);
  h2d(H.p, in.H, mb, c.stream);
  std::vector<int> hboff(in.bsz.size() + 1, 0);
  for (size_t i = 0; i < in.bsz.size(); ++i) hboff[i + 1] = hboff[i] + in.bsz[i];
  boff = upload_vec(hboff, c.stream);
  const int nb = (int)in.bsz.size();
  int* info = c.info.as<int>();

  // X = U diag(d) U^T from eig(S)
  HIP_CHECK(hipMemcpyAsync(W1.p, S.p, mb, hipMemcpyDeviceToDevice, c.stream));
  RB_CHECK(rocsolver_dsyevd(c.blas, rocblas_evect_original, rocblas_fill_upper, n, W1.as<double>(), n,
                            w.as<double>(), ework.as<double>(), info));
  hipLaunchKernelGGL(k_scale_cols, dim3(nblk(nn, 256)), dim3(256), 0, c.stream, W1.as<double>(),
                     w.as<double>(), n, W2.as<double>());
  {
    const double one = 1.0, zero = 0.0;   // column-major X = Ud U^T (symmetric)
    RB_CHECK(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_transpose, n, n, n, &one,
                           W2.as<double>(), n, W1.as<double>(), n, &zero, X.as<double>(), n));
  }
  if (in.filter_eps > 0 && nb > 0) {
    for (DevBuf* b : {&S, &X, &H})
      hipLaunchKernelGGL(k_block_filter, dim3(nb, nb), dim3(256), 0, c.stream, b->as<double>(), n,
                         boff.as<int>(), nb, in.filter_eps);
  }

  auto diag = [&](const double* Fm) {
    // F' = X^T F X ; eig ; C = X C' ; P = C_occ C_occ^T
    gemm_rm(c, n, Fm, X.as<double>(), W1.as<double>());                       // F X
    gemm_rm(c, n, X.as<double>(), W1.as<double>(), W2.as<double>(), 1.0, 0.0, true, false);  // X^T F X
    RB_CHECK(rocsolver_dsyevd(c.blas, rocblas_evect_original, rocblas_fill_upper, n, W2.as<double>(), n,
                              w.as<double>(), ework.as<double>(), info));
    // column-major eigenvectors V (in W2): C_cm = X_cm * V  (X symmetric)
    const double one = 1.0, zero = 0.0;
    RB_CHECK(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                           X.as<double>(), n, W2.as<double>(), n, &zero, Cev.as<double>(), n));
    if (nocc > 0) {
      hipLaunchKernelGGL(k_take_occ, dim3(nblk((int64_t)n * nocc, 256)), dim3(256), 0, c.stream,
                         Cev.as<double>(), n, nocc, Cocc.as<double>());
      // P = C_occ C_occ^T: column-major view of row-major Cocc (n x nocc) is nocc x n = Cocc^T
      RB_CHECK(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, n, n, nocc,
                             &one, Cocc.as<double>(), nocc, Cocc.as<double>(), nocc, &zero,
                             P.as<double>(), n));
    }
    else {
      hipLaunchKernelGGL(k_fill, dim3(gs_blocks(nn)), dim3(256), 0, c.stream, P.as<double>(), nn, 0.0);
    }
  };

  // first density: the core guess, or the SAD density with its Cholesky
  // factor for the first exchange (hfguess.cpp:423-452)
  DevBuf Lsad;
  int sad_rank = 0;
  if (!in.guess_P.empty()) {
    if ((int64_t)in.guess_P.size() != nn) throw std::runtime_error("SCF: the guess density is not nbf x nbf");
    h2d(P.p, in.guess_P.data(), mb, c.stream);
    DevBuf U(mb);
    HIP_CHECK(hipMemcpyAsync(U.p, P.p, mb, hipMemcpyDeviceToDevice, c.stream));
    std::vector<double> hL = pivoted_cholesky_dev(c, U, n, in.filter_eps, &sad_rank);
    Lsad.alloc((size_t)n * std::max(sad_rank, 1) * sizeof(double));
    if (sad_rank > 0) h2d(Lsad.p, hL.data(), hL.size() * sizeof(double), c.stream);
  }
  else {
    diag(H.as<double>());
  }
  const double* t64 = bt.data64t.p ? bt.data64t.as<double>() : nullptr;

  DiisState diis;
  double energy = 0.0;
  int iter = 0;
  out.history_e.clear();
  out.history_rms.clear();
  out.t_j.clear();
  out.t_k.clear();
  bool converged = false;
  double prev_rms = std::numeric_limits<double>::infinity();
  while (true) {
    // form_fock
    HIP_CHECK(hipEventRecord(c.ev0, c.stream));
    dfj(c, bt, m, P.as<double>(), J.as<double>(), 2.0);
    HIP_CHECK(hipEventRecord(c.ev1, c.stream));
    HIP_CHECK(hipEventSynchronize(c.ev1));
    float tj = 0.f;
    HIP_CHECK(hipEventElapsedTime(&tj, c.ev0, c.ev1));
    HIP_CHECK(hipEventRecord(c.ev0, c.stream));
    if (iter == 0 && !in.guess_P.empty()) {
      // the SAD iteration: every column of the guess density's factor, FP64
      dfk_mo(c, bt, m, Lsad.as<double>(), sad_rank, sad_rank, K.as<double>(), 1, t64);
    }
    else if (kbuild(iter, prev_rms, P.as<double>(), Cocc.as<double>(), K.as<double>())) {
      for (auto& b : diis.err) diis.spare.push_back(std::move(b));
      for (auto& b : diis.trial) diis.spare.push_back(std::move(b));
      diis.err.clear();
      diis.trial.clear();
      diis.B.clear();
      diis.coeffs.clear();
    }
    HIP_CHECK(hipEventRecord(c.ev1, c.stream));
    HIP_CHECK(hipEventSynchronize(c.ev1));
    float tk = 0.f;
    HIP_CHECK(hipEventElapsedTime(&tk, c.ev0, c.ev1));
    out.t_j.push_back(tj * 1e-3);
    out.t_k.push_back(tk * 1e-3);
    HIP_CHECK(hipMemcpyAsync(F.p, H.p, mb, hipMemcpyDeviceToDevice, c.stream));
    hipLaunchKernelGGL(k_axpby, dim3(gs_blocks(nn)), dim3(256), 0, c.stream, W1.as<double>(), J.as<double>(), nn, 1.0, 0.0);
    hipLaunchKernelGGL(k_axpby, dim3(gs_blocks(nn)), dim3(256), 0, c.stream, W1.as<double>(), K.as<double>(), nn, 1.0, 1.0);
    hipLaunchKernelGGL(k_axpby, dim3(gs_blocks(nn)), dim3(256), 0, c.stream, F.as<double>(), W1.as<double>(), nn, 1.0, 1.0);
    // error matrix
    gemm_rm(c, n, F.as<double>(), P.as<double>(), W1.as<double>());     // F P
    gemm_rm(c, n, W1.as<double>(), S.as<double>(), W2.as<double>());    // F P S
    hipLaunchKernelGGL(k_antisym, dim3(nblk(nn, 256)), dim3(256), 0, c.stream, W2.as<double>(), n, E.as<double>());
    // energy and RMS
    const double e1 = dot_dev(c, H.as<double>(), P.as<double>(), nn, part);
    const double e2 = dot_dev(c, F.as<double>(), P.as<double>(), nn, part);
    energy = e1 + e2;
    const double rms = std::sqrt(dot_dev(c, E.as<double>(), E.as<double>(), nn, part) / ((double)n * n));
    out.history_e.push_back(energy + in.enuc);
    out.history_rms.push_back(rms);
    prev_rms = rms;
    if (rms < in.threshold && iter > 0) {
      converged = true;
      break;
    }
    if (iter > in.max_iter) break;
    if (in.do_diis && iter >= in.diis_start) {
      const bool reduce = (int)diis.err.size() >= in.diis_max;
      DevBuf ec = diis.take(mb), fc = diis.take(mb);
      HIP_CHECK(hipMemcpyAsync(ec.p, E.p, mb, hipMemcpyDeviceToDevice, c.stream));
      HIP_CHECK(hipMemcpyAsync(fc.p, F.p, mb, hipMemcpyDeviceToDevice, c.stream));
      diis.err.push_back(std::move(ec));
      int max_pos = 0;
      double max_norm = -1.0;
      for (size_t i = 0; i < diis.err.size(); ++i) {
        const double nrm = dot_dev(c, diis.err[i].as<double>(), diis.err[i].as<double>(), nn, part);
        if (nrm > max_norm) {
          max_norm = nrm;
          max_pos = (int)i;
        }
      }
      int nold = (int)diis.trial.size();
      if (reduce) {
        diis.spare.push_back(std::move(diis.err[max_pos]));
        diis.err.erase(diis.err.begin() + max_pos);
      }
      diis.trial.push_back(std::move(fc));
      if (reduce) {
        diis.spare.push_back(std::move(diis.trial[max_pos]));
        diis.trial.erase(diis.trial.begin() + max_pos);
      }
      const int nerr = (int)diis.err.size();
      std::vector<double> Bn((size_t)nerr * nerr, 0.0);
      // carry over the old B-matrix without row/column max_pos
      std::vector<int> keep;
      for (int i = 0; i < nold; ++i)
        if (!reduce || i != max_pos) keep.push_back(i);
      for (int i = 0; i < nerr - 1; ++i)
        for (int j = 0; j < nerr - 1; ++j) Bn[i * nerr + j] = diis.B[keep[i] * nold + keep[j]];
      for (int i = 0; i < nerr; ++i) {
        const double v = dot_dev(c, diis.err[i].as<double>(), diis.err[nerr - 1].as<double>(), nn, part);
        Bn[(nerr - 1) * nerr + i] = v;
        Bn[i * nerr + (nerr - 1)] = v;
      }
      diis.B = Bn;
      diis.coeffs = diis_solve(diis.B, nerr);
      if ((int)diis.trial.size() >= in.diis_min) {
        hipLaunchKernelGGL(k_fill, dim3(gs_blocks(nn)), dim3(256), 0, c.stream, F.as<double>(), nn, 0.0);
        for (int i = 0; i < nerr; ++i)
          hipLaunchKernelGGL(k_axpby, dim3(gs_blocks(nn)), dim3(256), 0, c.stream, F.as<double>(),
                             diis.trial[i].as<double>(), nn, diis.coeffs[i], 1.0);
      }
    }
    diag(F.as<double>());
    ++iter;
  }
  HIP_CHECK(hipStreamSynchronize(c.stream));
  out.energy_iter = energy + in.enuc;
  if (final_ek) {
    // tr(K P) of the FP64 exchange at the final orbitals; J is already FP64 at
    // this density: E = tr(H P) + tr((H + J + K) P)
    HIP_CHECK(hipEventRecord(c.ev0, c.stream));
    const double ek = final_ek(Cocc.as<double>());
    HIP_CHECK(hipEventRecord(c.ev1, c.stream));
    HIP_CHECK(hipEventSynchronize(c.ev1));
    float tk = 0.f;
    HIP_CHECK(hipEventElapsedTime(&tk, c.ev0, c.ev1));
    out.t_final_k = tk * 1e-3;
    energy = 2.0 * dot_dev(c, H.as<double>(), P.as<double>(), nn, part) +
             dot_dev(c, J.as<double>(), P.as<double>(), nn, part) + ek;
  }
  else if (final_k) {
    // FP64 exchange at the final orbitals; J is already FP64 at this density
    HIP_CHECK(hipEventRecord(c.ev0, c.stream));
    final_k(iter, prev_rms, P.as<double>(), Cocc.as<double>(), K.as<double>());
    HIP_CHECK(hipEventRecord(c.ev1, c.stream));
    HIP_CHECK(hipEventSynchronize(c.ev1));
    float tk = 0.f;
    HIP_CHECK(hipEventElapsedTime(&tk, c.ev0, c.ev1));
    out.t_final_k = tk * 1e-3;
    HIP_CHECK(hipMemcpyAsync(F.p, H.p, mb, hipMemcpyDeviceToDevice, c.stream));
    hipLaunchKernelGGL(k_axpby, dim3(gs_blocks(nn)), dim3(256), 0, c.stream, F.as<double>(), J.as<double>(), nn, 1.0, 1.0);
    hipLaunchKernelGGL(k_axpby, dim3(gs_blocks(nn)), dim3(256), 0, c.stream, F.as<double>(), K.as<double>(), nn, 1.0, 1.0);
    energy = dot_dev(c, H.as<double>(), P.as<double>(), nn, part) + dot_dev(c, F.as<double>(), P.as<double>(), nn, part);
  }
  out.converged = converged;
  out.iterations = iter;
  out.energy = energy + in.enuc;
  out.eps.resize(n);
  d2h(out.eps.data(), w.p, n * sizeof(double), c.stream);
  out.C.resize(nn);
  // return C row-major (n x n): transpose of the column-major eigenvector matrix
  std::vector<double> tmp(nn);
  d2h(tmp.data(), Cev.p, mb, c.stream);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out.C[(size_t)i * n + j] = tmp[(size_t)j * n + i];
  out.P.resize(nn);
  d2h(out.P.data(), P.p, mb, c.stream);
  out.F.resize(nn);
  d2h(out.F.data(), F.p, mb, c.stream);
}

void select_k_builders(Ctx& c, DevBT& bt, const Metric* m, int kmode, int nocc, int occ_nbatches,
                       double threshold, KBuildFn& kb, KBuildFn& kfinal, int final_occ_nbatches,
                       EnergyFn* kenergy)
{
  if (kenergy) *kenergy = EnergyFn();
  if (final_occ_nbatches <= 0) final_occ_nbatches = occ_nbatches;
  if (kmode == 0 && !m) throw std::runtime_error("SCF: plain K needs the metric");
  if (kmode >= 2 && !bt.transformed && !m)
    throw std::runtime_error("SCF: the final FP64 exchange on the untransformed B needs the metric");
  if (kmode == 1 && !bt.transformed) throw std::runtime_error("SCF: B is not transformed");
  DevBT* t = &bt;
  Ctx* cp = &c;
  const Metric* km = bt.transformed ? nullptr : m;
  KBuildFn k64 = [cp, t, km, nocc, occ_nbatches](int, double, const double*, const double* Cocc, double* K) {
    dfk_mo(*cp, *t, km, Cocc, nocc, nocc, K, occ_nbatches);
    return false;
  };
  kfinal = KBuildFn();
  if (kmode >= 2) {   // the final FP64 exchange of the low-precision modes
    // on the FP64 L^-1 B copy when it exists ((L^-1 B) C instead of L^-1 (B C))
    const double* t64 = bt.data64t.p ? bt.data64t.as<double>() : nullptr;
    k64 = [cp, t, km, nocc, final_occ_nbatches, t64](int, double, const double*, const double* Cocc, double* K) {
      dfk_mo(*cp, *t, km, Cocc, nocc, nocc, K, final_occ_nbatches, t64);
      return false;
    };
    if (kenergy && t64 && final_occ_nbatches <= 1 && nocc > 0)
      *kenergy = [cp, t, nocc](const double* Cocc) { return dfk_energy_tilde64(*cp, *t, Cocc, nocc, nocc); };
    else if (kenergy && !t64 && km && nocc > 0)
      *kenergy = [cp, t, km, nocc, final_occ_nbatches](const double* Cocc) {
        return dfk_energy_metric(*cp, *t, *km, Cocc, nocc, nocc, final_occ_nbatches);
      };
  }
  if (kmode == 0 || kmode == 1) {
    kb = k64;
  }
  else if (kmode == 2) {
    if (!bt.lowp_tilde || !bt.data32.p) throw std::runtime_error("SCF: the FP32 mode needs the FP32 copy of L^-1 B");
    kb = [cp, t, nocc, occ_nbatches](int, double, const double*, const double* Cocc, double* K) {
      dfk_mo_f32(*cp, *t, Cocc, nocc, nocc, K, occ_nbatches);
      return false;
    };
    kfinal = k64;
  }
  else if (kmode == 3) {
    if (!bt.lowp_tilde || !bt.data_hi.p) throw std::runtime_error("SCF: the BF16 mode needs the BF16 split copy of L^-1 B");
    kb = [cp, t, nocc, occ_nbatches](int, double, const double*, const double* Cocc, double* K) {
      dfk_mo_bf16(*cp, *t, Cocc, nocc, nocc, K, occ_nbatches);
      return false;
    };
    kfinal = k64;
  }
  else if (kmode == 4) {
    if (!bt.lowp_tilde || !bt.data_hi.p)
      throw std::runtime_error("SCF: the escalation mode needs the BF16 split copy of L^-1 B");
    // without the FP32 copy the FP32 stage is skipped (BF16x3 -> FP64)
    const bool have32 = bt.data32.p != nullptr;
    // stage 0 BF16x3, 1 FP32, 2 FP64: a stage ends at the RMS trigger (BF16x3
    // only) or after `stall` iterations without progress once the error RMS is
    // below the stage's arming level (early SCF iterations also stall, far
    // above any precision's noise floor)
    struct Esc {
      int stage = 0;
      double best = std::numeric_limits<double>::infinity();
      int flat = 0;
    };
    auto st = std::make_shared<Esc>();
    const double tau = c.escalate_rms;
    const int stall = c.escalate_stall;
    kb = [cp, t, nocc, occ_nbatches, tau, stall, st, k64, have32](int it, double prev_rms, const double* P,
                                                                  const double* Cocc, double* K) {
      bool switched = false;
      constexpr double arm[2] = {1e-4, 1e-6};
      if (st->stage < 2 && std::isfinite(prev_rms)) {
        bool next = st->stage == 0 && tau > 0.0 && prev_rms <= tau;
        if (prev_rms >= arm[st->stage]) {
          st->flat = 0;
        }
        else if (prev_rms < 0.8 * st->best) {
          st->flat = 0;
        }
        else if (stall > 0 && ++st->flat >= stall) {
          next = true;
        }
        st->best = std::min(st->best, prev_rms);
        if (next) {
          st->stage += (st->stage == 0 && !have32) ? 2 : 1;
          st->best = std::numeric_limits<double>::infinity();
          st->flat = 0;
          switched = true;
        }
      }
      if (st->stage == 0) dfk_mo_bf16(*cp, *t, Cocc, nocc, nocc, K, occ_nbatches);
      else if (st->stage == 1) dfk_mo_f32(*cp, *t, Cocc, nocc, nocc, K, occ_nbatches);
      else k64(it, prev_rms, P, Cocc, K);
      return switched;
    };
    kfinal = k64;
  }
  else {
    throw std::runtime_error("SCF: unknown K mode");
  }
}

}  // namespace mchip
