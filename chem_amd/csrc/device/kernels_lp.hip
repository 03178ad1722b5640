// Reduced-precision kernels of the chosen DF-K form on gfx950.
//
//   k_to_f32             B~ (FP64) -> FP32 copy, once per geometry
//   k_slab_gemm_f32      W(X,nu,i) = sum_mu B~(X|mu nu) C(mu,i), FP32 MFMA
//                        (v_mfma_f32_16x16x4_f32), FP32 accumulation
//   k_seg_gemm_nt_f32    K partials = sum_{X,i} W(X,m,i) W(X,n,i), FP32 MFMA,
//                        chunk partial sums written as FP64 and reduced in FP64
//                        by k_reduce_sym
//
// FP32 MFMA lane layout (pinned by tests/test_gpu_mfma_layout.py):
//   A[i][k] from lane i + 16k, B[k][j] from lane j + 16k,
//   D[4*(lane/16) + e][lane%16] in element e.
#include "kernels.hpp"

namespace mchip {

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ void k_to_f32(const double* __restrict__ in, float* __restrict__ out, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

// row-major G64 (rows x ld64, first ncols) -> G32 (rows x ncols)
__global__ void k_rows_to_f32(const double* __restrict__ in, int ld64, int rows, int ncols,
                              float* __restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows * ncols) return;
  const int r = (int)(i / ncols), c = (int)(i % ncols);
  out[i] = (float)in[(int64_t)r * ld64 + c];
}

__global__ __launch_bounds__(256) void k_slab_gemm_f32(
    const float* __restrict__ bdata, const int64_t* __restrict__ slab_off,
    const int* __restrict__ ncol, const int* __restrict__ bsz, const int* __restrict__ boff,
    const int* __restrict__ colmap, const int* __restrict__ colmap_off,
    const Tile* __restrict__ tiles, int naux, int nbf,
    const float* __restrict__ G, int ldg, int ng,
    float* __restrict__ out, int ldo)
{
  __shared__ float As[SG_BK][SG_BM + 4];
  __shared__ float Bs[SG_BK][SG_BN + 4];
  const Tile t = tiles[blockIdx.x];
  const int N = t.slab;
  const int bn = bsz[N];
  const int K = ncol[N];
  const int64_t rows = (int64_t)naux * bn;
  const float* A = bdata + slab_off[N];
  const int* cmap = colmap + colmap_off[N];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const int a_row = tid >> 2, a_k = (tid & 3) * 4;
  const int b_k = tid >> 4, b_n = (tid & 15) * 4;
  const int64_t grow = (int64_t)t.m0 + a_row;
  const float* arow = A + grow * K;
  for (int k0 = 0; k0 < K; k0 += SG_BK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kk = k0 + a_k + q;
      As[a_k + q][a_row] = (grow < rows && kk < K) ? arow[kk] : 0.f;
    }
    {
      const int kk = k0 + b_k;
      const int mu = (kk < K) ? cmap[kk] : -1;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = t.n0 + b_n + q;
        Bs[b_k][b_n + q] = (mu >= 0 && j < ng) ? G[(int64_t)mu * ldg + j] : 0.f;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kq = 0; kq < SG_BK; kq += 4) {
      const int kr = kq + (lane >> 4);
      const float a0 = As[kr][wm * 32 + (lane & 15)];
      const float a1 = As[kr][wm * 32 + 16 + (lane & 15)];
      const float b0 = Bs[kr][wn * 32 + (lane & 15)];
      const float b1 = Bs[kr][wn * 32 + 16 + (lane & 15)];
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t r = (int64_t)t.m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + e;
        const int c = t.n0 + wn * 32 + j * 16 + (lane & 15);
        if (r < rows && c < ng) {
          const int64_t x = r / bn;
          const int nu = (int)(r - x * bn);
          out[(x * nbf + boff[N] + nu) * (int64_t)ldo + c] = acc[i][j][e];
        }
      }
}

__global__ __launch_bounds__(256) void k_seg_gemm_nt_f32(
    const float* __restrict__ A, const float* __restrict__ B, int nseg, int64_t sstride,
    int lda, int klen, int n, const int2* __restrict__ tiles, int ntile, int segs_per_chunk,
    double* __restrict__ partial)
{
  __shared__ float As[SG_BK][SG_BM + 4];
  __shared__ float Bs[SG_BK][SG_BN + 4];
  const int2 tl = tiles[blockIdx.x % ntile];
  const int chunk = blockIdx.x / ntile;
  const int s0 = chunk * segs_per_chunk;
  const int s1 = min(nseg, s0 + segs_per_chunk);
  const int m0 = tl.x * SG_BM, n0 = tl.y * SG_BN;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const int l_row = tid >> 2, l_k = (tid & 3) * 4;
  const int am = m0 + l_row, bnn = n0 + l_row;
  for (int s = s0; s < s1; ++s) {
    const float* As_g = A + (int64_t)s * sstride + (int64_t)am * lda;
    const float* Bs_g = B + (int64_t)s * sstride + (int64_t)bnn * lda;
    for (int k0 = 0; k0 < klen; k0 += SG_BK) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int kk = k0 + l_k + q;
        As[l_k + q][l_row] = (am < n && kk < klen) ? As_g[kk] : 0.f;
        Bs[l_k + q][l_row] = (bnn < n && kk < klen) ? Bs_g[kk] : 0.f;
      }
      __syncthreads();
#pragma unroll
      for (int kq = 0; kq < SG_BK; kq += 4) {
        const int kr = kq + (lane >> 4);
        const float a0 = As[kr][wm * 32 + (lane & 15)];
        const float a1 = As[kr][wm * 32 + 16 + (lane & 15)];
        const float b0 = Bs[kr][wn * 32 + (lane & 15)];
        const float b1 = Bs[kr][wn * 32 + 16 + (lane & 15)];
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
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
        const int r = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + e;
        const int c = n0 + wn * 32 + j * 16 + (lane & 15);
        if (r < n && c < n) P[(int64_t)r * n + c] = (double)acc[i][j][e];
      }
}

}  // namespace mchip
