// SOS-AO-MP2 on the GPU (FP64): mp::mpmod::compute (src/mp/mp_compute.cpp:
// 27-407) with the LLMP_FULL_Z builder (src/mp/z_full.cpp:69-397).
//
// Per Laplace point (omega, alpha):
//   M0  C_o' = C_o e^{+alpha eps/2}, C_v' = C_v e^{-alpha eps/2};
//       P~_o = omega^{1/4} (L_S^-1 C_o')(L_S^-1 C_o')^T, P_v = omega^{1/4} C_v' C_v'^T
//   M1  L~ = pivoted Cholesky of P~_o (piv_cd.cpp:125-483, device kernels,
//       columns ordered by LMO position), L_o = L_S L~, block filter
//   M2  b(X,nu,i)  = sum_mu B(X|mu nu) L_o(mu,i)        k_slab_gemm
//   M3  b2(X,N,i)  = sum_nu P_v(N,nu) b(X,nu,i)         rocBLAS strided batch
//   M4  b3(X|M N)  = sum_i L_o(M,i) b2(X,N,i) on B's pair pattern   k_slab_out_gemm
//   M5  Z(X,Y)     = sum_{mu nu} b3(X|mu nu) B(Y|mu nu)  k_slabseg_gemm (upper tiles)
//   M6  e = tr(Z~ Z~), Z~ = Z V^-1
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <numeric>

#include "dfock.hpp"
#include "kernels.hpp"
#include "mp2.hpp"

namespace mchip {

// ---- pivoted Cholesky kernels (row-major N x N symmetric U) -----------------
// out[0] = index of the first maximum |U_kk| over k >= I, out[1] = U_kk
__global__ void k_pcd_pivot(const double* __restrict__ U, int N, int I, double* __restrict__ out)
{
  __shared__ double bv[1024];
  __shared__ int bi[1024];
  double best = -1.0;
  int idx = N;
  for (int k = I + threadIdx.x; k < N; k += blockDim.x) {
    const double v = fabs(U[(int64_t)k * N + k]);
    if (v > best || (v == best && k < idx)) {
      best = v;
      idx = k;
    }
  }
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = idx;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      const double v = bv[threadIdx.x + off];
      const int j = bi[threadIdx.x + off];
      if (v > bv[threadIdx.x] || (v == bv[threadIdx.x] && j < bi[threadIdx.x])) {
        bv[threadIdx.x] = v;
        bi[threadIdx.x] = j;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = (double)bi[0];
    out[1] = (bi[0] < N) ? U[(int64_t)bi[0] * N + bi[0]] : 0.0;
  }
}

__global__ void k_swap_rows(double* __restrict__ A, int ncols, int ld, int r1, int r2)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncols) return;
  const double t = A[(int64_t)r1 * ld + k];
  A[(int64_t)r1 * ld + k] = A[(int64_t)r2 * ld + k];
  A[(int64_t)r2 * ld + k] = t;
}

__global__ void k_swap_cols(double* __restrict__ A, int nrows, int ld, int c1, int c2)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nrows) return;
  const double t = A[(int64_t)k * ld + c1];
  A[(int64_t)k * ld + c1] = A[(int64_t)k * ld + c2];
  A[(int64_t)k * ld + c2] = t;
}

// L[k][I] = U[k][I] / sqrt(piv) for k >= I (row-major L, leading dim ldl)
__global__ void k_pcd_column(const double* __restrict__ U, int N, int I, double piv,
                             double* __restrict__ L, int ldl)
{
  const int k = I + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const double s = sqrt(piv);
  L[(int64_t)k * ldl + I] = (k == I) ? s : U[(int64_t)k * N + I] / s;
}

// U[a][b] -= U[a][I] U[b][I] / piv for a, b > I
__global__ void k_pcd_update(double* __restrict__ U, int N, int I, double piv)
{
  const int a = I + 1 + blockIdx.y;
  const int b = I + 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= N || b >= N) return;
  U[(int64_t)a * N + b] -= U[(int64_t)a * N + I] * U[(int64_t)b * N + I] / piv;
}

// ---- small helpers ---------------------------------------------------------
// X (row-major rows x cols) column j scaled by exp(sgn * 0.5 * alpha * e[j]) * fac
__global__ void k_scale_cols_exp(double* __restrict__ X, int rows, int cols, const double* __restrict__ e,
                                 double sgn_alpha_half, double fac)
{
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)rows * cols) return;
  const int j = (int)(idx % cols);
  X[idx] *= exp(sgn_alpha_half * e[j]) * fac;
}

// zero (rblk x cblk) blocks of row-major M (rows x cols) with Frobenius norm < eps
__global__ void k_block_filter_rc(double* M, int cols, const int* roff, const int* coff, int ncb, double eps)
{
  const int bi = blockIdx.x / ncb, bj = blockIdx.x % ncb;
  const int r0 = roff[bi], r1 = roff[bi + 1], c0 = coff[bj], c1 = coff[bj + 1];
  const int w = c1 - c0, cnt = (r1 - r0) * w;
  __shared__ double red[256];
  double s = 0.0;
  for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
    const double v = M[(int64_t)(r0 + e / w) * cols + c0 + e % w];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (sqrt(red[0]) < eps)
    for (int e = threadIdx.x; e < cnt; e += blockDim.x) M[(int64_t)(r0 + e / w) * cols + c0 + e % w] = 0.0;
}

// part[b] = partial sums of sum_{x,y} Z[x][y] Z[y][x]
__global__ void k_trace_sq(const double* __restrict__ Z, int n, double* __restrict__ part)
{
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)n * n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i / n), y = (int)(i % n);
    s += Z[i] * Z[(int64_t)y * n + x];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ---- pivoted Cholesky driver -------------------------------------------------
// Decomposes the row-major N x N matrix in U (destroyed); returns L (host,
// N x rank row-major) in the original row order with LMO-ordered columns.
std::vector<double> pivoted_cholesky_dev(Ctx& c, DevBuf& U, int N, double filter_eps, int* rank_out)
{
  DevBuf L((size_t)N * N * sizeof(double)), piv(2 * sizeof(double));
  hipLaunchKernelGGL(k_fill, dim3(gs_blocks((int64_t)N * N)), dim3(256), 0, c.stream, L.as<double>(),
                     (int64_t)N * N, 0.0);
  std::vector<int> perm(N);
  std::iota(perm.begin(), perm.end(), 0);
  double hp[2];
  hipLaunchKernelGGL(k_pcd_pivot, dim3(1), dim3(1024), 0, c.stream, U.as<double>(), N, 0, piv.as<double>());
  HIP_CHECK(hipMemcpyAsync(hp, piv.p, sizeof(hp), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  const double thresh = N * 2.220446049250313e-16 * std::fabs(hp[1]);
  int rank = N;
  for (int I = 0; I < N; ++I) {
    hipLaunchKernelGGL(k_pcd_pivot, dim3(1), dim3(1024), 0, c.stream, U.as<double>(), N, I, piv.as<double>());
    HIP_CHECK(hipMemcpyAsync(hp, piv.p, sizeof(hp), hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    const int j = (int)hp[0];
    const double pv = hp[1];
    if (j != I && j < N) {
      hipLaunchKernelGGL(k_swap_rows, dim3(nblk(N, 256)), dim3(256), 0, c.stream, U.as<double>(), N, N, I, j);
      hipLaunchKernelGGL(k_swap_cols, dim3(nblk(N, 256)), dim3(256), 0, c.stream, U.as<double>(), N, N, I, j);
      if (I > 0)
        hipLaunchKernelGGL(k_swap_rows, dim3(nblk(I, 256)), dim3(256), 0, c.stream, L.as<double>(), I, N, I, j);
      std::swap(perm[I], perm[j]);
    }
    if (pv < 0.0 && std::fabs(pv) > thresh) throw std::runtime_error("Negative Pivot element. CD not possible.");
    if (std::fabs(pv) < thresh) {
      rank = I;
      break;
    }
    hipLaunchKernelGGL(k_pcd_column, dim3(nblk(N - I, 256)), dim3(256), 0, c.stream, U.as<double>(), N, I, pv,
                       L.as<double>(), N);
    if (I + 1 < N)
      hipLaunchKernelGGL(k_pcd_update, dim3(nblk(N - I - 1, 256), N - I - 1), dim3(256), 0, c.stream,
                         U.as<double>(), N, I, pv);
  }
  HIP_CHECK(hipGetLastError());
  std::vector<double> hL((size_t)N * N);
  HIP_CHECK(hipMemcpyAsync(hL.data(), L.p, hL.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  // rows back to the original order
  std::vector<double> Lo((size_t)N * rank, 0.0);
  for (int r = 0; r < N; ++r)
    for (int k = 0; k < rank; ++k) Lo[(size_t)perm[r] * rank + k] = hL[(size_t)r * N + k];
  // LMO position ordering of the columns (piv_cd.cpp:485-560)
  std::vector<double> pos(rank);
  for (int k = 0; k < rank; ++k) {
    int first = N - 1, last = 0;
    bool any = false;
    for (int r = 0; r < N; ++r)
      if (Lo[(size_t)r * rank + k] > filter_eps) {
        if (!any) first = r;
        last = r;
        any = true;
      }
    pos[k] = any ? 0.5 * (first + last) : (double)(N - 1);
  }
  std::vector<int> order(rank);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pos[a] < pos[b]; });
  std::vector<double> out((size_t)N * rank);
  for (int r = 0; r < N; ++r)
    for (int k = 0; k < rank; ++k) out[(size_t)r * rank + k] = Lo[(size_t)r * rank + order[k]];
  *rank_out = rank;
  return out;
}

static std::vector<int> split_range(int n, int split)
{
  std::vector<int> v;
  if (n <= 0) return v;
  const int nb = (n % split == 0) ? n / split : n / split + 1;
  v.assign(nb, split);
  if (n % split) v.back() = n % split;
  return v;
}

double mp2_sos(Ctx& c, DevBT& bt, const Metric& met, const Mp2Input& in, std::vector<double>& terms,
               double* t_z)
{
  const BLayout& L = bt.L;
  if (bt.transformed) throw std::runtime_error("MP2 needs the untransformed three-centre tensor");
  const int n = in.nbf, no = in.nocc, nv = n - in.nocc, naux = L.naux;
  if (n != L.nbf || met.naux != naux) throw std::runtime_error("MP2: dimensions do not match B / metric");
  const size_t nn = (size_t)n * n;
  DevBuf LS(nn * sizeof(double)), Co((size_t)n * std::max(no, 1) * sizeof(double)),
      Cv((size_t)n * std::max(nv, 1) * sizeof(double)), Wo((size_t)n * std::max(no, 1) * sizeof(double)),
      Wv((size_t)n * std::max(nv, 1) * sizeof(double)), Po(nn * sizeof(double)), Pv(nn * sizeof(double)),
      eo(std::max(no, 1) * sizeof(double)), ev(std::max(nv, 1) * sizeof(double));
  // occupied / virtual coefficients (row-major n x no / n x nv)
  std::vector<double> hco((size_t)n * no), hcv((size_t)n * nv);
  for (int r = 0; r < n; ++r) {
    for (int i = 0; i < no; ++i) hco[(size_t)r * no + i] = in.C[(size_t)r * n + i];
    for (int a = 0; a < nv; ++a) hcv[(size_t)r * nv + a] = in.C[(size_t)r * n + no + a];
  }
  HIP_CHECK(hipMemcpyAsync(LS.p, in.S, nn * sizeof(double), hipMemcpyHostToDevice, c.stream));
  if (no) HIP_CHECK(hipMemcpyAsync(eo.p, in.eps, no * sizeof(double), hipMemcpyHostToDevice, c.stream));
  if (nv) HIP_CHECK(hipMemcpyAsync(ev.p, in.eps + no, nv * sizeof(double), hipMemcpyHostToDevice, c.stream));
  // S = L_S L_S^T (column-major lower), L_S^-1
  DevBuf LSinv(nn * sizeof(double));
  int* info = c.info.as<int>();
  RB_CHECK(rocsolver_dpotrf(c.blas, rocblas_fill_lower, n, LS.as<double>(), n, info));
  hipLaunchKernelGGL(k_zero_upper_cm, dim3(nblk((int64_t)nn, 256)), dim3(256), 0, c.stream, LS.as<double>(), n);
  HIP_CHECK(hipMemcpyAsync(LSinv.p, LS.p, nn * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  RB_CHECK(rocsolver_dtrtri(c.blas, rocblas_fill_lower, rocblas_diagonal_non_unit, n, LSinv.as<double>(), n, info));
  hipLaunchKernelGGL(k_zero_upper_cm, dim3(nblk((int64_t)nn, 256)), dim3(256), 0, c.stream, LSinv.as<double>(), n);
  // block offsets for the filters
  std::vector<int> boff(in.bsz.size() + 1, 0);
  for (size_t i = 0; i < in.bsz.size(); ++i) boff[i + 1] = boff[i] + in.bsz[i];
  DevBuf d_boff = upload_vec(boff, c.stream);
  const int nbb = (int)in.bsz.size();
  // Z builder workspaces
  DevBuf Zp, Zm(nn == 0 ? 8 : (size_t)naux * naux * sizeof(double)), Zt((size_t)naux * naux * sizeof(double)),
      part(256 * sizeof(double));
  DevBuf b3((size_t)std::max<int64_t>(L.total, 1) * sizeof(double));
  const double one = 1.0, zero = 0.0;
  terms.clear();
  double total = 0.0;
  double tz = 0.0;
  for (size_t k = 0; k < in.omega.size(); ++k) {
    const double w = in.omega[k], al = in.alpha[k];
    const double w4 = std::pow(w, 0.25);
    // M0: scaled coefficients
    if (no) {
      HIP_CHECK(hipMemcpyAsync(Wo.p, hco.data(), (size_t)n * no * sizeof(double), hipMemcpyHostToDevice, c.stream));
      hipLaunchKernelGGL(k_scale_cols_exp, dim3(nblk((int64_t)n * no, 256)), dim3(256), 0, c.stream,
                         Wo.as<double>(), n, no, eo.as<double>(), 0.5 * al, 1.0);
      // c_ortho = L_S^-1 C_o' : column-major (no x n) view times L_S^-T
      RB_CHECK(rocblas_dtrmm(c.blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                             rocblas_diagonal_non_unit, no, n, &one, LSinv.as<double>(), n, Wo.as<double>(), no,
                             Co.as<double>(), no));
      RB_CHECK(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, n, n, no, &w4,
                             Co.as<double>(), no, Co.as<double>(), no, &zero, Po.as<double>(), n));
    }
    if (nv) {
      HIP_CHECK(hipMemcpyAsync(Wv.p, hcv.data(), (size_t)n * nv * sizeof(double), hipMemcpyHostToDevice, c.stream));
      hipLaunchKernelGGL(k_scale_cols_exp, dim3(nblk((int64_t)n * nv, 256)), dim3(256), 0, c.stream,
                         Wv.as<double>(), n, nv, ev.as<double>(), -0.5 * al, 1.0);
      RB_CHECK(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, n, n, nv, &w4,
                             Wv.as<double>(), nv, Wv.as<double>(), nv, &zero, Pv.as<double>(), n));
    }
    // M1: pivoted Cholesky of P~_o, L_o = L_S L~
    int rank = 0;
    std::vector<double> hLt = pivoted_cholesky_dev(c, Po, n, in.filter_eps, &rank);
    if (rank == 0) {
      terms.push_back(0.0);
      continue;
    }
    DevBuf Lt((size_t)n * rank * sizeof(double)), Lo((size_t)n * rank * sizeof(double));
    HIP_CHECK(hipMemcpyAsync(Lt.p, hLt.data(), hLt.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    // column-major (rank x n) view: Lo^T = Lt^T L_S^T
    RB_CHECK(rocblas_dtrmm(c.blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                           rocblas_diagonal_non_unit, rank, n, &one, LS.as<double>(), n, Lt.as<double>(), rank,
                           Lo.as<double>(), rank));
    std::vector<int> u = split_range(rank, in.mo_split), uoff(u.size() + 1, 0);
    for (size_t i = 0; i < u.size(); ++i) uoff[i + 1] = uoff[i] + u[i];
    DevBuf d_uoff = upload_vec(uoff, c.stream);
    if (in.filter_eps > 0) {
      hipLaunchKernelGGL(k_block_filter_rc, dim3(nbb * (int)u.size()), dim3(256), 0, c.stream, Lo.as<double>(),
                         rank, d_boff.as<int>(), d_uoff.as<int>(), (int)u.size(), in.filter_eps);
      hipLaunchKernelGGL(k_block_filter_rc, dim3(nbb * nbb), dim3(256), 0, c.stream, Pv.as<double>(), n,
                         d_boff.as<int>(), d_boff.as<int>(), nbb, in.filter_eps);
    }
    HIP_CHECK(hipEventRecord(c.ev0, c.stream));
    // M2: b = B . L_o   [X][nu][i]
    const size_t bsize = (size_t)naux * n * rank;
    c.scratch_a.ensure(bsize * sizeof(double));
    c.scratch_b.ensure(bsize * sizeof(double));
    double* bb = c.scratch_a.as<double>();
    double* b2 = c.scratch_b.as<double>();
    slab_gemm(c, bt, Lo.as<double>(), rank, rank, bb, rank, 1.0);
    // M3: b2_X = P_v b_X for every X (column-major rank x n blocks)
    RB_CHECK(rocblas_dgemm_strided_batched(c.blas, rocblas_operation_none, rocblas_operation_none, rank, n, n,
                                           &one, bb, rank, (rocblas_stride)n * rank, Pv.as<double>(), n, 0, &zero,
                                           b2, rank, (rocblas_stride)n * rank, naux));
    // M4: b3 on the pair slabs
    slab_out_gemm(c, bt, b2, rank, rank, Lo.as<double>(), rank, b3.as<double>());
    // M5: Z = sum_slabs b3 B^T (symmetric)
    slab_syrk(c, bt, b3.as<double>(), bt.data64.as<double>(), Zm.as<double>(), Zp);
    HIP_CHECK(hipEventRecord(c.ev1, c.stream));
    HIP_CHECK(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    tz += ms * 1e-3;
    // M6: Z~ = Z V^-1, e = sum Z~_xy Z~_yx
    RB_CHECK(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_none, naux, naux, naux, &one,
                           met.Vinv.as<double>(), naux, Zm.as<double>(), naux, &zero, Zt.as<double>(), naux));
    hipLaunchKernelGGL(k_trace_sq, dim3(256), dim3(256), 0, c.stream, Zt.as<double>(), naux, part.as<double>());
    double hpart[256];
    HIP_CHECK(hipMemcpyAsync(hpart, part.p, sizeof(hpart), hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    double e = 0.0;
    for (double v : hpart) e += v;
    terms.push_back(e);
    total += e;
  }
  HIP_CHECK(hipGetLastError());
  if (t_z) *t_z = tz;
  return total;
}

}  // namespace mchip
