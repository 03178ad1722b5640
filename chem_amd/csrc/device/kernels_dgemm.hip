// Dense FP64 GEMMs of the metric and the L^-1 B set-up, on FP64 MFMA:
//
//   k_dgemm       C = alpha op(A) op(B) + beta C (row-major C), each operand
//                 read either contraction-contiguous or output-contiguous, with
//                 an optional triangular structure that limits the contraction
//                 range of a tile (zeros of a triangular factor are skipped);
//   k_trmm_panel  X W for a panel X of at most 64 columns and a triangular
//                 64 x 64 W (the diagonal-block steps of the blocked Cholesky
//                 and triangular inverse), in place or out of place.
//
// The metric's Cholesky factor, L^-1, V^-1 = L^-T L^-1 and the slabs of
// L^-1 B (aoloader.cpp:14-32, LLT.cpp:8-109) all run on these.
#include "gemm_mfma.hpp"
#include "kernels.hpp"

namespace mchip {

namespace {
constexpr int DG_BM = 128, DG_BN = 128;
constexpr int DG_LDA = DG_BM + 8, DG_LDB = DG_BN + 8;   // LDS rows: [k][m], [k][n]
}  // namespace

// Tiles are numbered m fastest (tiles sharing a column block of B are
// neighbours) and grouped per XCD (grid padded to a multiple of 8).
template <bool AK, bool BK>
__global__ __launch_bounds__(256) void k_dgemm(int64_t M, int64_t N, int K, const double* __restrict__ A,
                                               int64_t lda, const double* __restrict__ B, int64_t ldb,
                                               double* __restrict__ C, int64_t ldc, double alpha, double beta,
                                               int tri, int ntm, int ntiles)
{
  __shared__ double As[2][GM_BK * DG_LDA];
  __shared__ double Bs[2][GM_BK * DG_LDB];
  const int tile = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (tile >= ntiles) return;
  const int64_t m0 = (int64_t)(tile % ntm) * DG_BM;
  const int64_t n0 = (int64_t)(tile / ntm) * DG_BN;
  // contraction range of this tile
  int64_t k_lo = 0, ke = K;
  if ((tri & DG_TRI_A_LOWER) && m0 + DG_BM < ke) ke = m0 + DG_BM;   // A(m,k) = 0 for k > m
  if ((tri & DG_TRI_A_UPPER) && m0 > k_lo) k_lo = m0;               // A(m,k) = 0 for k < m
  if ((tri & DG_TRI_B_LOWER) && n0 > k_lo) k_lo = n0;               // B(k,n) = 0 for k < n
  if ((tri & DG_TRI_B_UPPER) && n0 + DG_BN < ke) ke = n0 + DG_BN;   // B(k,n) = 0 for k > n
  // steps start on a multiple of the step; elements before k_lo are masked
  const int64_t kb = k_lo / GM_BK * GM_BK;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * 64, wn0 = (wave & 1) * 64;
  Mfma<double>::v4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<double>::v4{0.0, 0.0, 0.0, 0.0};
  // staging: 8 elements of A and 8 of B per thread and step.
  // contraction-contiguous operand: thread (r = tid >> 4, k = tid & 15), rows r + 16 j
  // output-contiguous operand: thread (c = tid & 127, k = tid >> 7), k + 2 j
  double ra[8], rb[8];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      int64_t m, k;
      if constexpr (AK) {
        m = m0 + (tid >> 4) + 16 * j;
        k = k0 + (tid & 15);
        ra[j] = (m < M && k >= k_lo && k < ke) ? A[m * lda + k] : 0.0;
      }
      else {
        m = m0 + (tid & 127);
        k = k0 + (tid >> 7) + 2 * j;
        ra[j] = (m < M && k >= k_lo && k < ke) ? A[k * lda + m] : 0.0;
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      int64_t n, k;
      if constexpr (BK) {
        n = n0 + (tid >> 4) + 16 * j;
        k = k0 + (tid & 15);
        rb[j] = (n < N && k >= k_lo && k < ke) ? B[n * ldb + k] : 0.0;
      }
      else {
        n = n0 + (tid & 127);
        k = k0 + (tid >> 7) + 2 * j;
        rb[j] = (n < N && k >= k_lo && k < ke) ? B[k * ldb + n] : 0.0;
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if constexpr (AK)
        As[buf][(tid & 15) * DG_LDA + (tid >> 4) + 16 * j] = ra[j];
      else
        As[buf][((tid >> 7) + 2 * j) * DG_LDA + (tid & 127)] = ra[j];
      if constexpr (BK)
        Bs[buf][(tid & 15) * DG_LDB + (tid >> 4) + 16 * j] = rb[j];
      else
        Bs[buf][((tid >> 7) + 2 * j) * DG_LDB + (tid & 127)] = rb[j];
    }
  };
  const int nsteps = ke > kb ? (int)((ke - kb + GM_BK - 1) / GM_BK) : 0;
  if (nsteps > 0) {
    load(kb);
    store(0);
  }
  __syncthreads();
  for (int s = 0; s < nsteps; ++s) {
    const int cur = s & 1;
    if (s + 1 < nsteps) load(kb + (int64_t)(s + 1) * GM_BK);
    mfma_step<double, 4, 4, DG_LDA, DG_LDB>(As[cur], Bs[cur], wm0, wn0, lane, acc);
    if (s + 1 < nsteps) store(cur ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t r = m0 + wm0 + 16 * i + Mfma<double>::row(lane, e);
        const int64_t c = n0 + wn0 + 16 * j + (lane & 15);
        if (r < M && c < N) {
          double* p = C + r * ldc + c;
          const double v = alpha * acc[i][j][e];
          *p = beta == 0.0 ? v : v + beta * *p;
        }
      }
}

template __global__ void k_dgemm<true, true>(int64_t, int64_t, int, const double*, int64_t, const double*, int64_t,
                                             double*, int64_t, double, double, int, int, int);
template __global__ void k_dgemm<true, false>(int64_t, int64_t, int, const double*, int64_t, const double*, int64_t,
                                              double*, int64_t, double, double, int, int, int);
template __global__ void k_dgemm<false, true>(int64_t, int64_t, int, const double*, int64_t, const double*, int64_t,
                                              double*, int64_t, double, double, int, int, int);
template __global__ void k_dgemm<false, false>(int64_t, int64_t, int, const double*, int64_t, const double*, int64_t,
                                               double*, int64_t, double, double, int, int, int);

// Y(r, c) = alpha sum_k X(r, k) Wop(k, c) for rows r < m and columns c < w
// (w <= 64): X and Y column-major (leading dimensions ldx, ldy; Y may be X),
// W lower triangular, column-major with leading dimension ldw;
// Wop = W^T if trans, else W.  Each workgroup reads its 64 rows of X into LDS
// before it writes any of Y, so Y = X is safe.
__global__ __launch_bounds__(256) void k_trmm_panel(int m, int w, const double* X, int ldx, const double* __restrict__ W,
                                                    int ldw, int trans, double alpha, double* Y, int ldy)
{
  __shared__ double xs[64][65];   // [row][k]
  __shared__ double ws[64][65];   // Wop[k][c]
  const int r0 = blockIdx.x * 64;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < 64 * w; idx += 256) {
    const int rr = idx & 63, k = idx >> 6;
    xs[rr][k] = (r0 + rr < m) ? X[(int64_t)k * ldx + r0 + rr] : 0.0;
  }
  for (int idx = tid; idx < w * w; idx += 256) {
    const int a = idx % w, b = idx / w;      // W(a, b), column-major
    const double v = (a >= b) ? W[(int64_t)b * ldw + a] : 0.0;
    if (trans)
      ws[b][a] = v;   // Wop(k = b, c = a) = W(a, b)
    else
      ws[a][b] = v;   // Wop(k = a, c = b) = W(a, b)
  }
  __syncthreads();
  for (int idx = tid; idx < 64 * w; idx += 256) {
    const int rr = idx & 63, c = idx >> 6;
    double s = 0.0;
    for (int k = 0; k < w; ++k) s += xs[rr][k] * ws[k][c];
    if (r0 + rr < m) Y[(int64_t)c * ldy + r0 + rr] = alpha * s;
  }
}

void dgemm_rm(hipStream_t st, bool ak, bool bk, int64_t M, int64_t N, int K, const double* A, int64_t lda,
              const double* B, int64_t ldb, double* C, int64_t ldc, double alpha, double beta, int tri)
{
  if (M <= 0 || N <= 0) return;
  const int64_t ntm = (M + DG_BM - 1) / DG_BM, ntn = (N + DG_BN - 1) / DG_BN;
  const int64_t nt = ntm * ntn;
  if (nt >= ((int64_t)1 << 31) - 8) throw std::runtime_error("dgemm: too many tiles");
  const unsigned grid = (unsigned)((nt + 7) / 8 * 8);
#define MCHIP_DG(AKV, BKV)                                                                                  \
  hipLaunchKernelGGL((k_dgemm<AKV, BKV>), dim3(grid), dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, alpha, \
                     beta, tri, (int)ntm, (int)nt)
  if (ak && bk) MCHIP_DG(true, true);
  else if (ak) MCHIP_DG(true, false);
  else if (bk) MCHIP_DG(false, true);
  else MCHIP_DG(false, false);
#undef MCHIP_DG
  HIP_CHECK(hipGetLastError());
}

void trmm_panel(hipStream_t st, int m, int w, const double* X, int ldx, const double* W, int ldw, bool trans,
                double alpha, double* Y, int ldy)
{
  if (m <= 0 || w <= 0) return;
  if (w > 64) throw std::runtime_error("trmm_panel: at most 64 columns");
  hipLaunchKernelGGL(k_trmm_panel, dim3((m + 63) / 64), dim3(256), 0, st, m, w, X, ldx, W, ldw, trans ? 1 : 0,
                     alpha, Y, ldy);
  HIP_CHECK(hipGetLastError());
}

}  // namespace mchip
