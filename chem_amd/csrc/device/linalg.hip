// Dense FP64 Cholesky factorisation and triangular inverse (the metric
// V = L L^T and L^-1 of aoloader.cpp:14-32 / LLT.cpp; the overlap factor of
// mp_compute.cpp:167-178).
//
// Blocked right-looking Cholesky and backward blocked inverse with block size
// 64: the diagonal blocks by one workgroup each in LDS (k_chol_diag,
// k_trti_diag), the panels by k_trmm_panel and the trailing updates by the
// library's FP64 MFMA GEMM (kernels_dgemm.hip), every product out of place or
// staged through LDS.  No vendor BLAS or solver runs on this path.
#include <algorithm>
#include <cmath>

#include "kernels.hpp"
#include "linalg.hpp"

namespace mchip {

namespace {
constexpr int NB = 64;
}

// Unblocked Cholesky of the kb x kb diagonal block at (k0, k0) of A
// (column-major, lda), written back in place (lower part), and the inverse of
// its factor into Winv (column-major, ld NB, upper part zero).  info receives
// the 1-based global index of the first non-positive pivot if it is still 0.
__global__ __launch_bounds__(256) void k_chol_diag(double* __restrict__ A, int lda, int k0, int kb,
                                                   double* __restrict__ Winv, int* __restrict__ info)
{
  __shared__ double L[NB][NB + 1];   // L[row][col]
  __shared__ int bad;
  const int tid = threadIdx.x;
  if (tid == 0) bad = 0;
  for (int idx = tid; idx < kb * kb; idx += blockDim.x) {
    const int i = idx % kb, j = idx / kb;
    L[i][j] = (i >= j) ? A[(int64_t)(k0 + j) * lda + (k0 + i)] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < kb; ++j) {
    const double d = L[j][j];
    if (!(d > 0.0)) {
      if (tid == 0) bad = j + 1;
      __syncthreads();
      break;   // d is read from LDS by every thread: all leave together
    }
    const double s = sqrt(d);
    __syncthreads();   // every thread has read L[j][j]
    if (tid == j) L[j][j] = s;
    for (int i = j + 1 + tid; i < kb; i += blockDim.x) L[i][j] /= s;
    __syncthreads();
    // trailing update of the block: L[i][l] -= L[i][j] L[l][j], j < l <= i
    const int m = kb - j - 1;
    for (int idx = tid; idx < m * m; idx += blockDim.x) {
      const int i = j + 1 + idx % m, l = j + 1 + idx / m;
      if (l <= i) L[i][l] -= L[i][j] * L[l][j];
    }
    __syncthreads();
  }
  if (bad) {
    if (tid == 0 && *info == 0) *info = k0 + bad;
    return;
  }
  for (int idx = tid; idx < kb * kb; idx += blockDim.x) {
    const int i = idx % kb, j = idx / kb;
    if (i >= j) A[(int64_t)(k0 + j) * lda + (k0 + i)] = L[i][j];
  }
  // inverse of the lower-triangular block, one column per thread (forward
  // substitution): X[c][c] = 1 / L[c][c], X[i][c] = -sum_{c<=l<i} L[i][l] X[l][c] / L[i][i]
  __shared__ double X[NB][NB + 1];
  if (tid < NB) {
    const int cc = tid;
    for (int i = 0; i < NB; ++i) X[i][cc] = 0.0;
    if (cc < kb) {
      X[cc][cc] = 1.0 / L[cc][cc];
      for (int i = cc + 1; i < kb; ++i) {
        double s = 0.0;
        for (int l = cc; l < i; ++l) s += L[i][l] * X[l][cc];
        X[i][cc] = -s / L[i][i];
      }
    }
    for (int i = 0; i < NB; ++i) Winv[(int64_t)cc * NB + i] = X[i][cc];
  }
}

// In-place inverse of the lower-triangular jb x jb diagonal block at (j0, j0).
__global__ __launch_bounds__(64) void k_trti_diag(double* __restrict__ A, int lda, int j0, int jb)
{
  __shared__ double L[NB][NB + 1];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < jb * jb; idx += blockDim.x) {
    const int i = idx % jb, j = idx / jb;
    L[i][j] = (i >= j) ? A[(int64_t)(j0 + j) * lda + (j0 + i)] : 0.0;
  }
  __syncthreads();
  __shared__ double X[NB][NB + 1];
  const int cc = tid;
  if (cc >= jb) return;
  X[cc][cc] = 1.0 / L[cc][cc];
  for (int i = cc + 1; i < jb; ++i) {
    double s = 0.0;
    for (int l = cc; l < i; ++l) s += L[i][l] * X[l][cc];
    X[i][cc] = -s / L[i][i];
  }
  for (int i = cc; i < jb; ++i) A[(int64_t)(j0 + cc) * lda + (j0 + i)] = X[i][cc];
}

int chol_lower(Ctx& c, double* A, int n, int lda)
{
  if (n <= 0) return 0;
  DevBuf winv((size_t)NB * NB * sizeof(double));
  int* info = c.info.as<int>();
  HIP_CHECK(hipMemsetAsync(info, 0, sizeof(int), c.stream));
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = std::min(NB, n - k0);
    hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(256), 0, c.stream, A, lda, k0, kb, winv.as<double>(), info);
    HIP_CHECK(hipGetLastError());
    const int m = n - k0 - kb;
    if (m <= 0) break;
    double* A21 = A + (int64_t)k0 * lda + (k0 + kb);
    double* A22 = A + (int64_t)(k0 + kb) * lda + (k0 + kb);
    // A21 <- A21 L11^-T (the panel is read into LDS before it is overwritten)
    trmm_panel(c.stream, m, kb, A21, lda, winv.as<double>(), NB, true, 1.0, A21, lda);
    // A22 <- A22 - A21 A21^T: the row-major view of the column-major A22 is
    // its transpose, and the update is symmetric
    dgemm_rm(c.stream, false, false, m, m, kb, A21, lda, A21, lda, A22, lda, -1.0, 1.0, 0);
  }
  int h_info = 0;
  d2h(&h_info, info, sizeof(int), c.stream);
  return h_info;
}

void trtri_lower(Ctx& c, double* A, int n, int lda)
{
  if (n <= 0) return;
  DevBuf tmp((size_t)n * NB * sizeof(double));
  const int last = (n - 1) / NB * NB;
  for (int j0 = last; j0 >= 0; j0 -= NB) {
    const int jb = std::min(NB, n - j0);
    hipLaunchKernelGGL(k_trti_diag, dim3(1), dim3(64), 0, c.stream, A, lda, j0, jb);
    HIP_CHECK(hipGetLastError());
    const int m = n - j0 - jb;
    if (m <= 0) continue;
    const double* A11 = A + (int64_t)j0 * lda + j0;
    double* A21 = A + (int64_t)j0 * lda + (j0 + jb);
    const double* A22 = A + (int64_t)(j0 + jb) * lda + (j0 + jb);
    // T = L22^-1 A21 (A22 already inverted, lower triangular), column-major
    // m x jb in tmp: row-major T^T(c, r) = sum_k A21(k, c) L22^-1(r, k)
    double* T = tmp.as<double>();
    dgemm_rm(c.stream, true, false, jb, m, m, A21, lda, A22, lda, T, m, 1.0, 0.0, DG_TRI_B_UPPER);
    // A21 <- -T L11^-1 (A11 inverted above)
    trmm_panel(c.stream, m, jb, T, m, A11, lda, false, -1.0, A21, lda);
  }
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

}  // namespace mchip
