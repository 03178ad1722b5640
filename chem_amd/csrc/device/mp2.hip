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
// LLMP_MEM_Z (z_mem.cpp) runs the same contractions in memory batches; with B
// resident in HBM it is the same GPU path.  LL_FULL_Z (z_full.cpp:428-692)
// replaces M3-M5 by a pivoted Cholesky factor L_v of the orthogonalised
// virtual pseudo-density (mp_compute.cpp:309-336):
//   b(X,i,a) = sum_N L_v(N,a) b(X,N,i)   rocBLAS strided batch
//   Z(X,Y)   = sum_ia b(X,i,a) b(Y,i,a)   rocBLAS GEMM
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>

#include "dfock.hpp"
#include "kernels.hpp"
#include "linalg.hpp"
#include "mp2.hpp"

namespace mchip {

// ---- pivoted Cholesky kernels (row-major N x N symmetric U) -----------------
// Right-looking blocked form (LAPACK dpstrf's scheme): a panel of PCD_NB
// columns is factorised by one workgroup, then the trailing matrix takes the
// panel's rank-PCD_NB update.  The updated diagonal d drives the pivot choice
// (largest |d_k|, first index on ties, as piv_cd.cpp:125-483).  All state
// (d, the permutation, rank, stop and error flags) lives on the device: the
// host reads it once per panel.
constexpr int PCD_NB = 32;

struct PcdState {
  int rank;      // columns factorised so far
  int done;      // 1: the factorisation has stopped
  int neg;       // 1: stopped at a negative pivot above the threshold
  int pad;
  double thresh; // N eps |largest diagonal|, set at the first pivot
};

__global__ void k_pcd_init(const double* __restrict__ U, int N, double* __restrict__ d, int* __restrict__ perm,
                           PcdState* __restrict__ st)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < N) {
    d[k] = U[(int64_t)k * N + k];
    perm[k] = k;
  }
  if (k == 0) {
    st->rank = 0;
    st->done = N == 0 ? 1 : 0;
    st->neg = 0;
    st->pad = 0;
    st->thresh = 0.0;
  }
}

// Columns J0 .. J0 + PCD_NB - 1 (one workgroup of 1024 threads).  U holds the
// Schur complement updated through column J0 - 1; a column's panel terms
// are subtracted here, the trailing update of the panel comes after.
__global__ __launch_bounds__(1024) void k_pcd_panel(double* __restrict__ U, int N, double* __restrict__ L,
                                                    double* __restrict__ d, int* __restrict__ perm, int J0,
                                                    PcdState* __restrict__ st)
{
  __shared__ double bv[1024];
  __shared__ int bi[1024];
  __shared__ double s_thresh;
  const int tid = threadIdx.x;
  if (st->done) return;
  if (tid == 0) s_thresh = st->thresh;
  __syncthreads();
  const int J1 = min(J0 + PCD_NB, N);
  for (int j = J0; j < J1; ++j) {
    double best = -1.0;
    int idx = N;
    for (int k = j + tid; k < N; k += 1024) {
      const double v = fabs(d[k]);
      if (v > best || (v == best && k < idx)) {
        best = v;
        idx = k;
      }
    }
    bv[tid] = best;
    bi[tid] = idx;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
      if (tid < off) {
        const double v = bv[tid + off];
        const int i2 = bi[tid + off];
        if (v > bv[tid] || (v == bv[tid] && i2 < bi[tid])) {
          bv[tid] = v;
          bi[tid] = i2;
        }
      }
      __syncthreads();
    }
    const int p = bi[0];
    const double pv = d[p];
    if (j == 0 && tid == 0) s_thresh = N * 2.220446049250313e-16 * fabs(pv);
    __syncthreads();
    const double thresh = s_thresh;
    if (fabs(pv) < thresh || pv < 0.0) {   // rank reached, or a negative pivot: stop
      if (tid == 0) {
        st->rank = j;
        st->done = 1;
        st->neg = fabs(pv) < thresh ? 0 : 1;
        st->thresh = thresh;
      }
      return;
    }
    if (p != j) {
      for (int k = tid; k < N; k += 1024) {   // rows j and p of U
        const double t = U[(int64_t)j * N + k];
        U[(int64_t)j * N + k] = U[(int64_t)p * N + k];
        U[(int64_t)p * N + k] = t;
      }
      __syncthreads();
      for (int k = tid; k < N; k += 1024) {   // columns j and p of U
        const double t = U[(int64_t)k * N + j];
        U[(int64_t)k * N + j] = U[(int64_t)k * N + p];
        U[(int64_t)k * N + p] = t;
      }
      for (int k = tid; k < j; k += 1024) {   // rows j and p of the factor so far
        const double t = L[(int64_t)j * N + k];
        L[(int64_t)j * N + k] = L[(int64_t)p * N + k];
        L[(int64_t)p * N + k] = t;
      }
      if (tid == 0) {
        const double t = d[j];
        d[j] = d[p];
        d[p] = t;
        const int q = perm[j];
        perm[j] = perm[p];
        perm[p] = q;
      }
      __syncthreads();
    }
    const double sq = sqrt(d[j]);
    for (int r = j + tid; r < N; r += 1024) {
      if (r == j) {
        L[(int64_t)j * N + j] = sq;
      }
      else {
        double v = U[(int64_t)r * N + j];
        for (int i = J0; i < j; ++i) v -= L[(int64_t)r * N + i] * L[(int64_t)j * N + i];
        v /= sq;
        L[(int64_t)r * N + j] = v;
        d[r] -= v * v;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    st->rank = J1;
    st->thresh = s_thresh;
    if (J1 == N) st->done = 1;
  }
}

// U[a][b] -= sum over the panel's columns of L[a][i] L[b][i], a, b >= J0 + PCD_NB
// (16 x 16 tiles of U, the panel rows of both sides staged in LDS)
__global__ __launch_bounds__(256) void k_pcd_trailing(double* __restrict__ U, int N, const double* __restrict__ L,
                                                      int J0, const PcdState* __restrict__ st)
{
  if (st->done) return;
  __shared__ double la[16][PCD_NB + 1], lb[16][PCD_NB + 1];
  const int J1 = J0 + PCD_NB;
  const int a0 = J1 + blockIdx.y * 16, b0 = J1 + blockIdx.x * 16;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  for (int e = threadIdx.x; e < 16 * PCD_NB; e += 256) {
    const int r = e / PCD_NB, i = e % PCD_NB;
    la[r][i] = a0 + r < N ? L[(int64_t)(a0 + r) * N + J0 + i] : 0.0;
    lb[r][i] = b0 + r < N ? L[(int64_t)(b0 + r) * N + J0 + i] : 0.0;
  }
  __syncthreads();
  const int a = a0 + ty, b = b0 + tx;
  if (a >= N || b >= N) return;
  double v = 0.0;
#pragma unroll 8
  for (int i = 0; i < PCD_NB; ++i) v += la[ty][i] * lb[tx][i];
  U[(int64_t)a * N + b] -= v;
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
// Throws at a negative pivot above the threshold (piv_cd.cpp's "Negative
// Pivot element. CD not possible."); *neg_out, when given, receives that
// case instead of the exception.
std::vector<double> pivoted_cholesky_dev(Ctx& c, DevBuf& U, int N, double filter_eps, int* rank_out, bool* neg_out)
{
  DevBuf L((size_t)std::max(N, 1) * std::max(N, 1) * sizeof(double)), d((size_t)std::max(N, 1) * sizeof(double)),
      perm_d((size_t)std::max(N, 1) * sizeof(int)), st(sizeof(PcdState));
  hipLaunchKernelGGL(k_fill, dim3(gs_blocks((int64_t)N * N)), dim3(256), 0, c.stream, L.as<double>(),
                     (int64_t)N * N, 0.0);
  hipLaunchKernelGGL(k_pcd_init, dim3(nblk(std::max(N, 1), 256)), dim3(256), 0, c.stream, U.as<double>(), N,
                     d.as<double>(), perm_d.as<int>(), st.as<PcdState>());
  PcdState hs{};
  for (int J0 = 0; J0 < N; J0 += PCD_NB) {
    hipLaunchKernelGGL(k_pcd_panel, dim3(1), dim3(1024), 0, c.stream, U.as<double>(), N, L.as<double>(),
                       d.as<double>(), perm_d.as<int>(), J0, st.as<PcdState>());
    const int rest = N - (J0 + PCD_NB);
    if (rest > 0)
      hipLaunchKernelGGL(k_pcd_trailing, dim3(nblk(rest, 16), nblk(rest, 16)), dim3(256), 0, c.stream,
                         U.as<double>(), N, L.as<double>(), J0, st.as<PcdState>());
    HIP_CHECK(hipGetLastError());
    d2h(&hs, st.p, sizeof(hs), c.stream);   // one read of the state per panel
    if (hs.done) break;
  }
  if (N == 0) hs.rank = 0;
  if (hs.neg) {
    if (neg_out) {
      *neg_out = true;
      *rank_out = 0;
      return {};
    }
    throw std::runtime_error("Negative Pivot element. CD not possible.");
  }
  if (neg_out) *neg_out = false;
  const int rank = hs.rank;
  std::vector<double> hL((size_t)N * N);
  std::vector<int> perm(N);
  if (N > 0) {
    d2h(hL.data(), L.p, hL.size() * sizeof(double), c.stream);
    d2h(perm.data(), perm_d.p, perm.size() * sizeof(int), c.stream);
  }
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

// Shared set-up of both MP2 drivers: Cholesky factor of S and its inverse,
// orbital energies and coefficients, block offsets for the filters.
struct Mp2Setup {
  int n = 0, no = 0, nv = 0, nbb = 0;
  DevBuf LS, LSinv, eo, ev, d_boff, Wo, Wv, Co, Po;
  std::vector<double> hco, hcv;
};

static void mp2_setup(Ctx& c, const Mp2Input& in, Mp2Setup& s)
{
  const int n = in.nbf, no = in.nocc, nv = n - in.nocc;
  s.n = n;
  s.no = no;
  s.nv = nv;
  const size_t nn = (size_t)n * n;
  s.LS.alloc(nn * sizeof(double));
  s.LSinv.alloc(nn * sizeof(double));
  s.eo.alloc(std::max(no, 1) * sizeof(double));
  s.ev.alloc(std::max(nv, 1) * sizeof(double));
  s.Wo.alloc((size_t)n * std::max(no, 1) * sizeof(double));
  s.Wv.alloc((size_t)n * std::max(nv, 1) * sizeof(double));
  s.Co.alloc((size_t)n * std::max(no, 1) * sizeof(double));
  s.Po.alloc(nn * sizeof(double));
  s.hco.resize((size_t)n * no);
  s.hcv.resize((size_t)n * nv);
  for (int r = 0; r < n; ++r) {
    for (int i = 0; i < no; ++i) s.hco[(size_t)r * no + i] = in.C[(size_t)r * n + i];
    for (int a = 0; a < nv; ++a) s.hcv[(size_t)r * nv + a] = in.C[(size_t)r * n + no + a];
  }
  h2d(s.LS.p, in.S, nn * sizeof(double), c.stream);
  if (no) h2d(s.eo.p, in.eps, no * sizeof(double), c.stream);
  if (nv) h2d(s.ev.p, in.eps + no, nv * sizeof(double), c.stream);
  // S = L_S L_S^T (column-major lower), L_S^-1
  if (chol_lower(c, s.LS.as<double>(), n, n) != 0)
    throw std::runtime_error("MP2: the overlap matrix is not positive definite");
  hipLaunchKernelGGL(k_zero_upper_cm, dim3(nblk((int64_t)nn, 256)), dim3(256), 0, c.stream, s.LS.as<double>(), n);
  HIP_CHECK(hipMemcpyAsync(s.LSinv.p, s.LS.p, nn * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  trtri_lower(c, s.LSinv.as<double>(), n, n);
  hipLaunchKernelGGL(k_zero_upper_cm, dim3(nblk((int64_t)nn, 256)), dim3(256), 0, c.stream, s.LSinv.as<double>(), n);
  std::vector<int> boff(in.bsz.size() + 1, 0);
  for (size_t i = 0; i < in.bsz.size(); ++i) boff[i + 1] = boff[i] + in.bsz[i];
  s.d_boff = upload_vec(boff, c.stream);
  s.nbb = (int)in.bsz.size();
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

// Pseudo-density factors of Laplace point k (mp_compute.cpp:232-302):
// P~_o = w^1/4 (L_S^-1 C_o')(...)^T factorised by pivoted Cholesky, L_o = L_S L~
// (n x rank row-major, in Lo), and the filtered virtual pseudo-density P_v
// (n x n, in Pv).  Returns the rank.
static int mp2_point_factors(Ctx& c, Mp2Setup& s, const Mp2Input& in, size_t k, DevBuf& Lo, DevBuf& Pv,
                             DevBuf* Lv = nullptr, int* rank_v = nullptr)
{
  const int n = s.n, no = s.no, nv = s.nv;
  const size_t nn = (size_t)n * n;
  const double w = in.omega[k], al = in.alpha[k];
  const double w4 = std::pow(w, 0.25);
  const double one = 1.0, zero = 0.0;
  Pv.alloc(nn * sizeof(double));
  if (no) {
    h2d(s.Wo.p, s.hco.data(), (size_t)n * no * sizeof(double), c.stream);
    hipLaunchKernelGGL(k_scale_cols_exp, dim3(nblk((int64_t)n * no, 256)), dim3(256), 0, c.stream,
                       s.Wo.as<double>(), n, no, s.eo.as<double>(), 0.5 * al, 1.0);
    // c_ortho = L_S^-1 C_o' : column-major (no x n) view times L_S^-T
    RB_CHECK(rocblas_dtrmm(c.blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                           rocblas_diagonal_non_unit, no, n, &one, s.LSinv.as<double>(), n, s.Wo.as<double>(), no,
                           s.Co.as<double>(), no));
    RB_CHECK(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, n, n, no, &w4,
                           s.Co.as<double>(), no, s.Co.as<double>(), no, &zero, s.Po.as<double>(), n));
  }
  if (nv) {
    h2d(s.Wv.p, s.hcv.data(), (size_t)n * nv * sizeof(double), c.stream);
    hipLaunchKernelGGL(k_scale_cols_exp, dim3(nblk((int64_t)n * nv, 256)), dim3(256), 0, c.stream,
                       s.Wv.as<double>(), n, nv, s.ev.as<double>(), -0.5 * al, 1.0);
    RB_CHECK(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, n, n, nv, &w4,
                           s.Wv.as<double>(), nv, s.Wv.as<double>(), nv, &zero, Pv.as<double>(), n));
  }
  else {
    hipLaunchKernelGGL(k_fill, dim3(gs_blocks((int64_t)nn)), dim3(256), 0, c.stream, Pv.as<double>(), (int64_t)nn, 0.0);
  }
  // LL_FULL_Z: L_v = L_S CD(P~_v), P~_v = omega^{1/4} (L_S^-1 C_v')(L_S^-1 C_v')^T
  // (mp_compute.cpp:309-336); made before P_v is filtered, from the same C_v'
  if (Lv && rank_v) {
    *rank_v = 0;
    if (nv) {
      DevBuf Cv((size_t)n * nv * sizeof(double)), Pvt(nn * sizeof(double));
      RB_CHECK(rocblas_dtrmm(c.blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                             rocblas_diagonal_non_unit, nv, n, &one, s.LSinv.as<double>(), n, s.Wv.as<double>(), nv,
                             Cv.as<double>(), nv));
      RB_CHECK(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, n, n, nv, &w4,
                             Cv.as<double>(), nv, Cv.as<double>(), nv, &zero, Pvt.as<double>(), n));
      int rv = 0;
      std::vector<double> hLv = pivoted_cholesky_dev(c, Pvt, n, in.filter_eps, &rv);
      *rank_v = rv;
      if (rv > 0) {
        DevBuf Lvt((size_t)n * rv * sizeof(double));
        Lv->alloc((size_t)n * rv * sizeof(double));
        h2d(Lvt.p, hLv.data(), hLv.size() * sizeof(double), c.stream);
        RB_CHECK(rocblas_dtrmm(c.blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                               rocblas_diagonal_non_unit, rv, n, &one, s.LS.as<double>(), n, Lvt.as<double>(), rv,
                               Lv->as<double>(), rv));
        if (in.filter_eps > 0) {
          std::vector<int> r = split_range(rv, in.mo_split), roff(r.size() + 1, 0);
          for (size_t i = 0; i < r.size(); ++i) roff[i + 1] = roff[i] + r[i];
          DevBuf d_roff = upload_vec(roff, c.stream);
          hipLaunchKernelGGL(k_block_filter_rc, dim3(s.nbb * (int)r.size()), dim3(256), 0, c.stream,
                             Lv->as<double>(), rv, s.d_boff.as<int>(), d_roff.as<int>(), (int)r.size(),
                             in.filter_eps);
        }
        HIP_CHECK(hipStreamSynchronize(c.stream));   // host factor and offsets die here
      }
    }
  }
  // M1: pivoted Cholesky of P~_o, L_o = L_S L~
  int rank = 0;
  std::vector<double> hLt = pivoted_cholesky_dev(c, s.Po, n, in.filter_eps, &rank);
  if (rank == 0) return 0;
  DevBuf Lt((size_t)n * rank * sizeof(double));
  Lo.alloc((size_t)n * rank * sizeof(double));
  h2d(Lt.p, hLt.data(), hLt.size() * sizeof(double), c.stream);
  // column-major (rank x n) view: Lo^T = Lt^T L_S^T
  RB_CHECK(rocblas_dtrmm(c.blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                         rocblas_diagonal_non_unit, rank, n, &one, s.LS.as<double>(), n, Lt.as<double>(), rank,
                         Lo.as<double>(), rank));
  std::vector<int> u = split_range(rank, in.mo_split), uoff(u.size() + 1, 0);
  for (size_t i = 0; i < u.size(); ++i) uoff[i + 1] = uoff[i] + u[i];
  DevBuf d_uoff = upload_vec(uoff, c.stream);
  if (in.filter_eps > 0) {
    hipLaunchKernelGGL(k_block_filter_rc, dim3(s.nbb * (int)u.size()), dim3(256), 0, c.stream, Lo.as<double>(),
                       rank, s.d_boff.as<int>(), d_uoff.as<int>(), (int)u.size(), in.filter_eps);
    hipLaunchKernelGGL(k_block_filter_rc, dim3(s.nbb * s.nbb), dim3(256), 0, c.stream, Pv.as<double>(), n,
                       s.d_boff.as<int>(), s.d_boff.as<int>(), s.nbb, in.filter_eps);
  }
  HIP_CHECK(hipStreamSynchronize(c.stream));   // host factor and offsets die here
  return rank;
}

// e = sum_xy Z~_xy Z~_yx with Z~ = V^-1 Z (mp_compute.cpp:361-389), FP64.
static double mp2_point_energy(Ctx& c, const Metric& met, const double* Z, DevBuf& Zt, DevBuf& part)
{
  const int naux = met.naux;
  const double one = 1.0, zero = 0.0;
  RB_CHECK(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_none, naux, naux, naux, &one,
                         met.Vinv.as<double>(), naux, Z, naux, &zero, Zt.as<double>(), naux));
  hipLaunchKernelGGL(k_trace_sq, dim3(256), dim3(256), 0, c.stream, Zt.as<double>(), naux, part.as<double>());
  double hpart[256];
  d2h(hpart, part.p, sizeof(hpart), c.stream);
  double e = 0.0;
  for (double v : hpart) e += v;
  return e;
}

void z_point(Ctx& c, DevBT& bt, int zmethod, const double* Lo, int ro, const double* Pv, const double* Lv, int rv,
             double* Z)
{
  const BLayout& L = bt.L;
  const int n = L.nbf, naux = L.naux;
  const double one = 1.0, zero = 0.0;
  if (zmethod < 0 || zmethod > 2) throw std::runtime_error("Z builder: unknown method");
  if (ro == 0 || (zmethod == 2 && rv == 0)) {
    hipLaunchKernelGGL(k_fill, dim3(gs_blocks((int64_t)naux * naux)), dim3(256), 0, c.stream, Z,
                       (int64_t)naux * naux, 0.0);
    return;
  }
  // M2: b = B . L_o   [X][nu][i]
  const size_t bsize = (size_t)naux * n * ro;
  c.scratch_a.ensure(bsize * sizeof(double));
  double* bb = c.scratch_a.as<double>();
  slab_gemm(c, bt, Lo, ro, ro, bb, ro, 1.0);
  if (zmethod == 2) {
    // b(X,i,a) = sum_N L_v(N,a) b(X,N,i)  ("Na, XiN -> Xia"): per X the
    // column-major (rv x ro) product L_v^T b_X
    const size_t csize = (size_t)naux * ro * rv;
    c.scratch_b.ensure(csize * sizeof(double));
    double* cc = c.scratch_b.as<double>();
    RB_CHECK(rocblas_dgemm_strided_batched(c.blas, rocblas_operation_none, rocblas_operation_transpose, rv, ro, n,
                                           &one, Lv, rv, 0, bb, ro, (rocblas_stride)n * ro, &zero, cc, rv,
                                           (rocblas_stride)ro * rv, naux));
    // Z(X,Y) = sum_ia b(X,i,a) b(Y,i,a)  ("Mia, Nia -> MN")
    RB_CHECK(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, naux, naux, ro * rv, &one,
                           cc, ro * rv, cc, ro * rv, &zero, Z, naux));
    return;
  }
  // M3: b2_X = P_v b_X for every X (column-major ro x n blocks)
  c.scratch_b.ensure(bsize * sizeof(double));
  double* b2 = c.scratch_b.as<double>();
  RB_CHECK(rocblas_dgemm_strided_batched(c.blas, rocblas_operation_none, rocblas_operation_none, ro, n, n, &one, bb,
                                         ro, (rocblas_stride)n * ro, Pv, n, 0, &zero, b2, ro,
                                         (rocblas_stride)n * ro, naux));
  // M4: b3 on the pair slabs
  c.scratch_c.ensure((size_t)std::max<int64_t>(L.total, 1) * sizeof(double));
  double* b3 = c.scratch_c.as<double>();
  slab_out_gemm(c, bt, b2, ro, ro, Lo, ro, b3);
  // M5: Z = sum_slabs b3 B^T (symmetric)
  slab_syrk(c, bt, b3, bt.data64.as<double>(), Z, c.scratch_partial);
}

double mp2_sos(Ctx& c, DevBT& bt, const Metric& met, const Mp2Input& in, std::vector<double>& terms,
               double* t_z, double* Z_host)
{
  const BLayout& L = bt.L;
  if (bt.transformed) throw std::runtime_error("MP2 needs the untransformed three-centre tensor");
  const int n = in.nbf, naux = L.naux;
  if (n != L.nbf || met.naux != naux) throw std::runtime_error("MP2: dimensions do not match B / metric");
  if (in.zmethod < 0 || in.zmethod > 2) throw std::runtime_error("MP2: unknown Z method");
  Mp2Setup s;
  mp2_setup(c, in, s);
  DevBuf Zm((size_t)naux * naux * sizeof(double)), Zt((size_t)naux * naux * sizeof(double)),
      part(256 * sizeof(double));
  terms.clear();
  double total = 0.0, tz = 0.0;
  const bool ll = in.zmethod == 2;
  for (size_t k = 0; k < in.omega.size(); ++k) {
    DevBuf Lo, Pv, Lv;
    int rv = 0;
    const int rank = mp2_point_factors(c, s, in, k, Lo, Pv, ll ? &Lv : nullptr, ll ? &rv : nullptr);
    if (rank == 0 || (ll && rv == 0)) {
      if (Z_host) std::fill(Z_host + k * (size_t)naux * naux, Z_host + (k + 1) * (size_t)naux * naux, 0.0);
      terms.push_back(0.0);
      continue;
    }
    HIP_CHECK(hipEventRecord(c.ev0, c.stream));
    z_point(c, bt, in.zmethod, Lo.as<double>(), rank, Pv.as<double>(), ll ? Lv.as<double>() : nullptr, rv,
            Zm.as<double>());
    HIP_CHECK(hipEventRecord(c.ev1, c.stream));
    HIP_CHECK(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    tz += ms * 1e-3;
    if (Z_host) d2h(Z_host + k * (size_t)naux * naux, Zm.p, (size_t)naux * naux * sizeof(double), c.stream);
    const double e = mp2_point_energy(c, met, Zm.as<double>(), Zt, part);
    terms.push_back(e);
    total += e;
  }
  HIP_CHECK(hipGetLastError());
  if (t_z) *t_z = tz;
  return total;
}

// FP32 copy of a row-major (rows x w) FP64 block into columns [c0, c0 + w) of
// a row-major FP32 matrix with leading dimension ld.
__global__ void k_cols_to_f32(const double* __restrict__ in, int rows, int w, float* __restrict__ out, int ld,
                              int c0)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows * w) return;
  const int r = (int)(i / w), col = (int)(i - (int64_t)r * w);
  out[(int64_t)r * ld + c0 + col] = (float)in[i];
}

// All Laplace points at once (LLMP_FULL_Z, z_full.cpp:69-397, for every point
// of mp_compute.cpp:232-389).  The pseudo-densities and their pivoted
// Cholesky factors are made per point in FP64; then the four contractions of
// all points go out as grouped launches captured in one hipGraph, in FP32
// with FP64 accumulation, on B~ = L^-1 B (V = L L^T; FP32 copy made slab by
// slab in FP64 from B):
//   M2  b[X][nu][i]     = sum_mu B~(X|mu nu) L_all(mu, i)  one slab GEMM over the
//                                                         concatenated columns of all points,
//                                                         direct-to-LDS on the padded FP32 copy
//   M3  b2_g[X]         = P_v,g b_g[X]                     grouped over (point, X), P_v as an FP32 hi/lo pair,
//                                                         direct-to-LDS on b transposed per X
//   M4  b3_g(X|M N)     = sum_i L_g(M,i) b2_g(X,i,N)        grouped over points, B's pattern,
//                                                         direct-to-LDS (rank blocks padded to 16)
//   M5  Z'_g(X,Y)       = sum b3_g(X|mu nu) B~(Y|mu nu)      grouped over points, FP64-flushed
// so Z'_g = L^-1 Z_g L^-T and e_g = tr(Z_g V^-1 Z_g V^-1) = sum_XY Z'_g(X,Y)^2
// (FP64, after the graph).  The reference's form applies V^-1 after Z: Z then
// carries the small-eigenvalue directions of V that V^-1 amplifies, and the
// rounding of FP32 products alone moves e_g by ~1e-5 relative; in the
// orthonormal auxiliary basis e_g is a well-conditioned function of Z'.
// Every FP32 MFMA chain is cut into FP64 running sums after 16 products
// (64 in M5).  The pipeline depths of M2 and M3 are picked by timing both
// candidates on the live data.
double mp2_sos_batched(Ctx& c, DevBT& bt, const Metric& met, const Mp2Input& in, std::vector<double>& terms,
                       double* t_z, double* t_prep, double* Z_host)
{
  const BLayout& L = bt.L;
  if (bt.transformed) throw std::runtime_error("MP2 needs the untransformed three-centre tensor");
  const int n = in.nbf, naux = L.naux;
  if (n != L.nbf || met.naux != naux) throw std::runtime_error("MP2: dimensions do not match B / metric");
  if (in.zmethod == 2) throw std::runtime_error("MP2: the grouped FP32 graph implements the LLMP Z builders");
  const int npt = (int)in.omega.size();
  const auto tp0 = std::chrono::steady_clock::now();
  Mp2Setup s;
  mp2_setup(c, in, s);
  std::vector<DevBuf> Lo(npt), Pv(npt);
  // each point's rank block padded with zero columns to a multiple of 16
  // (whole 64-byte contraction steps for M4's direct-to-LDS loads)
  std::vector<int> rank(npt), rank16(npt), pcol(npt + 1, 0);
  for (int g = 0; g < npt; ++g) {
    rank[g] = mp2_point_factors(c, s, in, g, Lo[g], Pv[g]);
    rank16[g] = (rank[g] + 15) / 16 * 16;
    pcol[g + 1] = pcol[g] + rank16[g];
  }
  const int R = pcol[npt];
  terms.assign(npt, 0.0);
  if (R == 0) {
    if (Z_host) std::fill(Z_host, Z_host + (size_t)npt * naux * naux, 0.0);
    if (t_z) *t_z = 0.0;
    if (t_prep) *t_prep = 0.0;
    return 0.0;
  }
  // FP32 operands: all L_o side by side (n x R), the P_v (n x n each, rows
  // padded to ldn, a multiple of 16), B~
  const int ldn = (n + 15) / 16 * 16;
  const size_t nl = (size_t)n * ldn;
  if (naux > 65535) throw std::runtime_error("MP2: the grouped graph handles at most 65535 aux functions");
  // P_v as an FP32 pair (hi at [g], lo at [npt + g]): P_v b cancels heavily,
  // and a single FP32 rounding of P_v alone costs ~1e-5 of max|Z|
  DevBuf Lall((size_t)n * R * sizeof(float)), Pv32((size_t)2 * npt * nl * sizeof(float)),
      B32((size_t)std::max<int64_t>(L.total, 1) * sizeof(float));
  HIP_CHECK(hipMemsetAsync(Lall.p, 0, Lall.bytes, c.stream));   // the padding columns (before any capture)
  for (int g = 0; g < npt; ++g) {
    if (rank[g] == 0) continue;
    hipLaunchKernelGGL(k_cols_to_f32, dim3(nblk((int64_t)n * rank[g], 256)), dim3(256), 0, c.stream,
                       Lo[g].as<double>(), n, rank[g], Lall.as<float>(), R, pcol[g]);
    hipLaunchKernelGGL(k_to_f32x2_rows, dim3(gs_blocks((int64_t)nl)), dim3(256), 0, c.stream, Pv[g].as<double>(),
                       n, ldn, Pv32.as<float>() + (size_t)g * nl, Pv32.as<float>() + (size_t)(npt + g) * nl);
  }
  HIP_CHECK(hipGetLastError());
  if (L.total) tilde_to_f32(c, bt, met, B32.as<float>());
  // the padded FP32 copy of B~ for M2's direct-to-LDS slab GEMM (made here
  // unless the tensor holds one; released at the end if made here)
  const bool had32 = bt.data32.p != nullptr && bt.lowp_tilde;
  const bool had_lowp = bt.lowp_tilde;
  if (!had32) dbt_make_f32_from(c, bt, B32.as<float>());
  DevBuf G2((size_t)std::max<int64_t>(dbt_padded_cols(bt) * R, 1) * sizeof(float));
  const size_t bsize = (size_t)naux * n * R;
  // b transposed per aux function, rows of ldn: the second operand of M3
  DevBuf bb(bsize * sizeof(float)), b2(bsize * sizeof(float)), bT((size_t)naux * R * ldn * sizeof(float)),
      b3((size_t)npt * std::max<int64_t>(L.total, 1) * sizeof(float));
  std::vector<int> pw(rank16.begin(), rank16.end());
  std::vector<int> pc(pcol.begin(), pcol.end() - 1);
  DevBuf d_pcol = upload_vec(pc, c.stream), d_pw = upload_vec(pw, c.stream);
  // M3 tiles (point, m0, first column): 128 x 128 over the columns (X, c) of
  // each point, column index X rank + c
  std::vector<int4> t3;
  for (int g = 0; g < npt; ++g)
    for (int64_t col = 0; col < (int64_t)naux * rank16[g]; col += 128)
      for (int m0 = 0; m0 < n; m0 += 128) t3.push_back(int4{g, m0, (int)col, 0});
  if ((int64_t)naux * R >= ((int64_t)1 << 31)) throw std::runtime_error("MP2: too many (aux, rank) columns");
  // M4 tiles (slab, m0, c0, point)
  std::vector<Tile> t4;
  for (int g = 0; g < npt; ++g) {
    if (rank[g] == 0) continue;
    for (int nb = 0; nb < L.nbb; ++nb) {
      const int64_t rows = (int64_t)naux * L.bsz[nb];
      for (int64_t m0 = 0; m0 < rows; m0 += SG_BM)
        for (int c0 = 0; c0 < L.ncol[nb]; c0 += 128) t4.push_back(Tile{nb, (int)m0, c0, g});
    }
  }
  while (t4.size() % 8) t4.push_back(Tile{-1, 0, 0, 0});   // whole XCD segments (empty tiles return)
  // M5: upper 128 x 128 tiles of Z, slabs grouped in chunks of similar length
  const int ntz = (naux + SG_BM - 1) / SG_BM;
  std::vector<int2> t5;
  for (int i = 0; i < ntz; ++i)
    for (int j = i; j < ntz; ++j) t5.push_back(int2{i, j});
  const int ntile = (int)t5.size();
  std::vector<int64_t> soff(L.nbb), slen(L.nbb);
  int64_t tot = 0;
  for (int sb = 0; sb < L.nbb; ++sb) {
    soff[sb] = L.slab_off[sb];
    slen[sb] = (int64_t)L.bsz[sb] * L.ncol[sb];
    tot += slen[sb];
  }
  const int want = std::max(1, std::min(L.nbb, 2048 / std::max(1, ntile * npt) + 1));
  std::vector<int> cs(1, 0);
  int64_t acc = 0;
  for (int sb = 0; sb < L.nbb; ++sb) {
    acc += slen[sb];
    if ((int)cs.size() < want && acc * want >= tot * (int64_t)cs.size() && sb + 1 < L.nbb) cs.push_back(sb + 1);
  }
  cs.push_back(L.nbb);
  const int nch = (int)cs.size() - 1;
  DevBuf d_t3 = upload_vec(t3, c.stream), d_t4 = upload_vec(t4, c.stream), d_t5 = upload_vec(t5, c.stream),
         d_soff = upload_vec(soff, c.stream), d_slen = upload_vec(slen, c.stream), d_cs = upload_vec(cs, c.stream);
  DevBuf partial((size_t)npt * nch * naux * naux * sizeof(double)), Z((size_t)npt * naux * naux * sizeof(double));
  // M2 on the padded FP32 copy; its pipeline depth is timed on the live
  // operands (the first call also makes the tile table)
  int m2_depth = 2;
  {
    float best = 1e30f;
    for (int depth : {2, 4}) {
      slab_gemm_f32_f64sum(c, bt, Lall.as<float>(), R, R, G2.as<float>(), bb.as<float>(), R, depth);
      HIP_CHECK(hipEventRecord(c.ev0, c.stream));
      slab_gemm_f32_f64sum(c, bt, Lall.as<float>(), R, R, G2.as<float>(), bb.as<float>(), R, depth);
      HIP_CHECK(hipEventRecord(c.ev1, c.stream));
      HIP_CHECK(hipEventSynchronize(c.ev1));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
      if (ms < best) {
        best = ms;
        m2_depth = depth;
      }
    }
  }
  // M3: b transposed (rows of the contraction index), then the grouped
  // direct-to-LDS GEMM; its pipeline depth is timed on the live operands
  const dim3 tgrid((unsigned)((R + 31) / 32), (unsigned)((ldn + 31) / 32), (unsigned)naux);
  auto m3 = [&](int depth) {
    if (depth == 2)
      hipLaunchKernelGGL(k_group_apply_f32_dl<2>, dim3((unsigned)t3.size()), dim3(512), 0, c.stream,
                         Pv32.as<float>(), npt, bT.as<float>(), b2.as<float>(), n, ldn, R, naux, d_pcol.as<int>(),
                         d_pw.as<int>(), d_t3.as<int4>());
    else
      hipLaunchKernelGGL(k_group_apply_f32_dl<4>, dim3((unsigned)t3.size()), dim3(512), 0, c.stream,
                         Pv32.as<float>(), npt, bT.as<float>(), b2.as<float>(), n, ldn, R, naux, d_pcol.as<int>(),
                         d_pw.as<int>(), d_t3.as<int4>());
  };
  int m3_depth = 2;
  {
    hipLaunchKernelGGL(k_transpose_pad_f32, tgrid, dim3(256), 0, c.stream, bb.as<float>(), n, R, ldn,
                       bT.as<float>());
    float best = 1e30f;
    for (int depth : {2, 4}) {
      m3(depth);
      HIP_CHECK(hipEventRecord(c.ev0, c.stream));
      m3(depth);
      HIP_CHECK(hipEventRecord(c.ev1, c.stream));
      HIP_CHECK(hipEventSynchronize(c.ev1));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
      if (ms < best) {
        best = ms;
        m3_depth = depth;
      }
    }
  }
  // M4 on the direct-to-LDS pipeline; its depth is timed on the live operands
  // (b2 from the tuning runs above)
  auto m4 = [&](int depth) {
    if (depth == 2)
      hipLaunchKernelGGL(k_slab_out_gemm_f32_dl<2>, dim3((unsigned)t4.size()), dim3(512), 0, c.stream,
                         b2.as<float>(), R, bt.slab_off.as<int64_t>(), bt.ncol.as<int>(), bt.bsz.as<int>(),
                         bt.boff.as<int>(), bt.colmap.as<int>(), bt.colmap_off.as<int>(), d_t4.as<Tile>(), naux, n,
                         Lall.as<float>(), b3.as<float>(), d_pcol.as<int>(), d_pw.as<int>(), (int64_t)L.total);
    else
      hipLaunchKernelGGL(k_slab_out_gemm_f32_dl<4>, dim3((unsigned)t4.size()), dim3(512), 0, c.stream,
                         b2.as<float>(), R, bt.slab_off.as<int64_t>(), bt.ncol.as<int>(), bt.bsz.as<int>(),
                         bt.boff.as<int>(), bt.colmap.as<int>(), bt.colmap_off.as<int>(), d_t4.as<Tile>(), naux, n,
                         Lall.as<float>(), b3.as<float>(), d_pcol.as<int>(), d_pw.as<int>(), (int64_t)L.total);
  };
  int m4_depth = 2;
  if (!t4.empty()) {
    float best = 1e30f;
    for (int depth : {2, 4}) {
      m4(depth);
      HIP_CHECK(hipEventRecord(c.ev0, c.stream));
      m4(depth);
      HIP_CHECK(hipEventRecord(c.ev1, c.stream));
      HIP_CHECK(hipEventSynchronize(c.ev1));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
      if (ms < best) {
        best = ms;
        m4_depth = depth;
      }
    }
  }
  HIP_CHECK(hipStreamSynchronize(c.stream));
  if (t_prep) *t_prep = std::chrono::duration<double>(std::chrono::steady_clock::now() - tp0).count();
  // capture M2..M5 and the chunk reductions in one graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  HIP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  slab_gemm_f32_f64sum(c, bt, Lall.as<float>(), R, R, G2.as<float>(), bb.as<float>(), R, m2_depth);
  hipLaunchKernelGGL(k_transpose_pad_f32, tgrid, dim3(256), 0, c.stream, bb.as<float>(), n, R, ldn, bT.as<float>());
  m3(m3_depth);
  if (!t4.empty()) m4(m4_depth);
  hipLaunchKernelGGL(k_slabseg_f32, dim3((unsigned)(ntile * nch * npt)), dim3(512), 0, c.stream, b3.as<float>(),
                     (int64_t)L.total, B32.as<float>(), d_soff.as<int64_t>(), d_slen.as<int64_t>(), d_cs.as<int>(),
                     nch, naux, d_t5.as<int2>(), ntile, partial.as<double>());
  for (int g = 0; g < npt; ++g)
    hipLaunchKernelGGL(k_reduce_sym<double>, dim3(nblk((int64_t)naux * naux, 256)), dim3(256), 0, c.stream,
                       partial.as<double>() + (size_t)g * nch * naux * naux, nch, naux, 0, 1.0,
                       Z.as<double>() + (size_t)g * naux * naux);
  const hipError_t cap_err = hipGetLastError();
  HIP_CHECK(hipStreamEndCapture(c.stream, &graph));
  HIP_CHECK(cap_err);
  HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  HIP_CHECK(hipEventRecord(c.ev0, c.stream));
  HIP_CHECK(hipGraphLaunch(exec, c.stream));
  HIP_CHECK(hipEventRecord(c.ev1, c.stream));
  HIP_CHECK(hipEventSynchronize(c.ev1));
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  if (t_z) *t_z = ms * 1e-3;
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  if (!had32) {   // the tensor as the caller left it
    bt.data32.release();
    bt.lowp_tilde = had_lowp;
    bt.f32_plan.clear();
  }
  // e_g = sum Z'_g(X,Y)^2 in FP64; Z_g = L Z'_g L^T only when asked for
  DevBuf part(256 * sizeof(double)), Lf, Zg64;
  if (Z_host) {
    Lf.alloc((size_t)naux * naux * sizeof(double));
    Zg64.alloc((size_t)naux * naux * sizeof(double));
    HIP_CHECK(hipMemcpyAsync(Lf.p, met.Linv.p, (size_t)naux * naux * sizeof(double), hipMemcpyDeviceToDevice,
                             c.stream));
    trtri_lower(c, Lf.as<double>(), naux, naux);
  }
  double total = 0.0;
  for (int g = 0; g < npt; ++g) {
    double* Zg = Z_host ? Z_host + (size_t)g * naux * naux : nullptr;
    if (rank[g] == 0) {
      if (Zg) std::fill(Zg, Zg + (size_t)naux * naux, 0.0);
      continue;
    }
    const double* Zp = Z.as<double>() + (size_t)g * naux * naux;
    hipLaunchKernelGGL(k_trace_sq, dim3(256), dim3(256), 0, c.stream, Zp, naux, part.as<double>());
    double hpart[256];
    d2h(hpart, part.p, sizeof(hpart), c.stream);
    double e = 0.0;
    for (double v : hpart) e += v;
    terms[g] = e;
    total += e;
    if (Zg) {
      // Z = L Z' L^T (symmetric, so row/column order does not matter)
      const double one = 1.0;
      RB_CHECK(rocblas_dtrmm(c.blas, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                             rocblas_diagonal_non_unit, naux, naux, &one, Lf.as<double>(), naux, Zp, naux,
                             Zg64.as<double>(), naux));
      RB_CHECK(rocblas_dtrmm(c.blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                             rocblas_diagonal_non_unit, naux, naux, &one, Lf.as<double>(), naux,
                             Zg64.as<double>(), naux, Zg64.as<double>(), naux));
      d2h(Zg, Zg64.p, (size_t)naux * naux * sizeof(double), c.stream);
    }
  }
  return total;
}

}  // namespace mchip
