// FP64 kernels of the DF Fock build on gfx950.
//
//   k_slab_gemm_f64    grouped GEMM over B's pair slabs with gathered rows of a
//                      dense operand: HT(X,nu,j) = sum_mu B(X|mu nu) G(mu,j).
//                      Used for XMN,Mi->XiN (jkdf.cpp:361), mi,Xmn->Xin
//                      (z_full.cpp:202) and XMN,NL->XML (jkdf.cpp:566).
//   k_seg_gemm_nt_f64  split-K GEMM over a segmented contraction index,
//                      C(m,n) = sum_s sum_k A_s(m,k) B_s(n,k): the exchange
//                      SYRK XiM,XiN->MN (jkdf.cpp:417).
//   k_pair_gather / k_j_gamma / k_j_cols / k_pair_scatter
//                      HBM-streaming DF-J (jkdf.cpp:111-170).
//
// MFMA: v_mfma_f64_16x16x4_f64.  Operand layout per 64-lane wave:
//   A[i][k] from lane i + 16k, B[k][j] from lane j + 16k,
//   D[lane/16 + 4e][lane%16] in element e of the 4-double accumulator
//   (pinned by tests/test_gpu_mfma_layout.py).
#include "kernels.hpp"

namespace mchip {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// grouped slab GEMM
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_slab_gemm_f64(
    const double* __restrict__ bdata, const int64_t* __restrict__ slab_off,
    const int* __restrict__ ncol, const int* __restrict__ bsz, const int* __restrict__ boff,
    const int* __restrict__ colmap, const int* __restrict__ colmap_off,
    const Tile* __restrict__ tiles, int naux, int nbf,
    const double* __restrict__ G, int ldg, int ng,
    double* __restrict__ out, int ldo, double alpha)
{
  __shared__ double As[SG_BK][SG_BM + 1];
  __shared__ double Bs[SG_BK][SG_BN + 1];
  const Tile t = tiles[blockIdx.x];
  const int N = t.slab;
  const int bn = bsz[N];
  const int K = ncol[N];
  const int64_t rows = (int64_t)naux * bn;
  const double* A = bdata + slab_off[N];
  const int* cmap = colmap + colmap_off[N];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  const int a_row = tid >> 2, a_k = (tid & 3) * 4;
  const int b_k = tid >> 4, b_n = (tid & 15) * 4;
  const int64_t grow = (int64_t)t.m0 + a_row;
  const double* arow = A + grow * K;

  for (int k0 = 0; k0 < K; k0 += SG_BK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kk = k0 + a_k + q;
      As[a_k + q][a_row] = (grow < rows && kk < K) ? arow[kk] : 0.0;
    }
    {
      const int kk = k0 + b_k;
      const int mu = (kk < K) ? cmap[kk] : -1;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = t.n0 + b_n + q;
        Bs[b_k][b_n + q] = (mu >= 0 && j < ng) ? G[(int64_t)mu * ldg + j] : 0.0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kq = 0; kq < SG_BK; kq += 4) {
      const int kr = kq + (lane >> 4);
      const double a0 = As[kr][wm * 32 + (lane & 15)];
      const double a1 = As[kr][wm * 32 + 16 + (lane & 15)];
      const double b0 = Bs[kr][wn * 32 + (lane & 15)];
      const double b1 = Bs[kr][wn * 32 + 16 + (lane & 15)];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t r = (int64_t)t.m0 + wm * 32 + i * 16 + (lane >> 4) + 4 * e;
        const int c = t.n0 + wn * 32 + j * 16 + (lane & 15);
        if (r < rows && c < ng) {
          const int64_t x = r / bn;
          const int nu = (int)(r - x * bn);
          out[(x * nbf + boff[N] + nu) * (int64_t)ldo + c] = alpha * acc[i][j][e];
        }
      }
}

// ---------------------------------------------------------------------------
// segmented split-K NT GEMM: partial[chunk][m][n] = sum_{s in chunk} sum_k
//   A[s*sstride + m*lda + k] * B[s*sstride + n*lda + k]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_seg_gemm_nt_f64(
    const double* __restrict__ A, const double* __restrict__ B, int nseg, int64_t sstride,
    int lda, int klen, int n, const int2* __restrict__ tiles, int ntile, int segs_per_chunk,
    double* __restrict__ partial)
{
  __shared__ double As[SG_BK][SG_BM + 1];
  __shared__ double Bs[SG_BK][SG_BN + 1];
  const int2 tl = tiles[blockIdx.x % ntile];
  const int chunk = blockIdx.x / ntile;
  const int s0 = chunk * segs_per_chunk;
  const int s1 = min(nseg, s0 + segs_per_chunk);
  const int m0 = tl.x * SG_BM, n0 = tl.y * SG_BN;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  const int l_row = tid >> 2, l_k = (tid & 3) * 4;
  const int am = m0 + l_row, bnn = n0 + l_row;

  for (int s = s0; s < s1; ++s) {
    const double* As_g = A + (int64_t)s * sstride + (int64_t)am * lda;
    const double* Bs_g = B + (int64_t)s * sstride + (int64_t)bnn * lda;
    for (int k0 = 0; k0 < klen; k0 += SG_BK) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int kk = k0 + l_k + q;
        As[l_k + q][l_row] = (am < n && kk < klen) ? As_g[kk] : 0.0;
        Bs[l_k + q][l_row] = (bnn < n && kk < klen) ? Bs_g[kk] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int kq = 0; kq < SG_BK; kq += 4) {
        const int kr = kq + (lane >> 4);
        const double a0 = As[kr][wm * 32 + (lane & 15)];
        const double a1 = As[kr][wm * 32 + 16 + (lane & 15)];
        const double b0 = Bs[kr][wn * 32 + (lane & 15)];
        const double b1 = Bs[kr][wn * 32 + 16 + (lane & 15)];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
      }
      __syncthreads();
    }
  }
  double* P = partial + (int64_t)chunk * n * n;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = m0 + wm * 32 + i * 16 + (lane >> 4) + 4 * e;
        const int c = n0 + wn * 32 + j * 16 + (lane & 15);
        if (r < n && c < n) P[(int64_t)r * n + c] = acc[i][j][e];
      }
}

// out[r][c] = out[c][r] = alpha * sum_ch (P[ch][r][c] (+ P[ch][c][r] if sym2)), r <= c
__global__ void k_reduce_sym(const double* __restrict__ partial, int nchunk, int n, int sym2,
                             double alpha, double* __restrict__ out)
{
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx / n), c = (int)(idx % n);
  if (r > c) return;
  double v = 0.0;
  for (int ch = 0; ch < nchunk; ++ch) {
    const double* P = partial + (int64_t)ch * n * n;
    v += P[(int64_t)r * n + c];
    if (sym2) v += P[(int64_t)c * n + r];
  }
  v *= alpha;
  out[(int64_t)r * n + c] = v;
  out[(int64_t)c * n + r] = v;
}

// ---------------------------------------------------------------------------
// DF-J streaming kernels
// ---------------------------------------------------------------------------
// Pg[pg_off[N] + nu*ncol[N] + c] = P[colmap[N][c]*nbf + boff[N] + nu]  (grid.y = slab)
__global__ void k_pair_gather(const double* __restrict__ P, int nbf, const int* __restrict__ ncol,
                              const int* __restrict__ bsz, const int* __restrict__ boff,
                              const int* __restrict__ colmap, const int* __restrict__ colmap_off,
                              const int64_t* __restrict__ pg_off, double* __restrict__ Pg)
{
  const int N = blockIdx.y;
  const int K = ncol[N];
  const int R = bsz[N] * K;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= R) return;
  const int nu = e / K, c = e - nu * K;
  const int mu = colmap[colmap_off[N] + c];
  Pg[pg_off[N] + e] = P[(int64_t)mu * nbf + boff[N] + nu];
}

// J[mu][boff[N]+nu] = Jg[pg_off[N] + nu*ncol[N] + c]
__global__ void k_pair_scatter(const double* __restrict__ Jg, int nbf, const int* __restrict__ ncol,
                               const int* __restrict__ bsz, const int* __restrict__ boff,
                               const int* __restrict__ colmap, const int* __restrict__ colmap_off,
                               const int64_t* __restrict__ pg_off, double* __restrict__ J)
{
  const int N = blockIdx.y;
  const int K = ncol[N];
  const int R = bsz[N] * K;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= R) return;
  const int nu = e / K, c = e - nu * K;
  const int mu = colmap[colmap_off[N] + c];
  J[(int64_t)mu * nbf + boff[N] + nu] = Jg[pg_off[N] + e];
}

// gpart[N*naux + x] = sum_e slab_N[x][e] * Pg_N[e]; one wave per row, J_ROWS rows per block
__global__ __launch_bounds__(256) void k_j_gamma(
    const double* __restrict__ bdata, const int64_t* __restrict__ slab_off,
    const int* __restrict__ ncol, const int* __restrict__ bsz, const int64_t* __restrict__ pg_off,
    const double* __restrict__ Pg, int naux, double* __restrict__ gpart)
{
  const int N = blockIdx.y;
  const int R = bsz[N] * ncol[N];
  const double* slab = bdata + slab_off[N];
  const double* pg = Pg + pg_off[N];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int rr = wave; rr < J_ROWS; rr += 4) {
    const int x = blockIdx.x * J_ROWS + rr;
    if (x >= naux) break;
    const double* row = slab + (int64_t)x * R;
    double s = 0.0;
    for (int e = lane; e < R; e += 64) s += row[e] * pg[e];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) gpart[(int64_t)N * naux + x] = s;
  }
}

// g[x] = scale * sum_N gpart[N*naux + x]   (fixed order)
__global__ void k_reduce_slabs(const double* __restrict__ gpart, int nslab, int naux, double scale,
                               double* __restrict__ g)
{
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= naux) return;
  double s = 0.0;
  for (int N = 0; N < nslab; ++N) s += gpart[(int64_t)N * naux + x];
  g[x] = scale * s;
}

// Jpart[chunk][pg_off[N] + e] = sum_{x in chunk} gq[x] * slab_N[x][e]
__global__ void k_j_cols(const double* __restrict__ bdata, const int64_t* __restrict__ slab_off,
                         const int* __restrict__ ncol, const int* __restrict__ bsz,
                         const int64_t* __restrict__ pg_off, const double* __restrict__ gq,
                         int naux, int xchunk, int64_t pg_total, double* __restrict__ Jpart)
{
  const int N = blockIdx.y;
  const int chunk = blockIdx.z;
  const int R = bsz[N] * ncol[N];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= R) return;
  const double* slab = bdata + slab_off[N];
  const int x0 = chunk * xchunk, x1 = min(naux, x0 + xchunk);
  double s = 0.0;
  for (int x = x0; x < x1; ++x) s += gq[x] * slab[(int64_t)x * R + e];
  Jpart[(int64_t)chunk * pg_total + pg_off[N] + e] = s;
}

__global__ void k_reduce_chunks(const double* __restrict__ part, int nchunk, int64_t n,
                                double* __restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * n + i];
  out[i] = s;
}

__global__ void k_fill(double* __restrict__ p, int64_t n, double v)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_axpby(double* __restrict__ y, const double* __restrict__ x, int64_t n,
                        double a, double b)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

}  // namespace mchip
