// Exchange energy without the exchange matrix.
//
// The SCF energy needs K only through tr(K P) (E = tr(H P) + tr(F P),
// hfscf.cpp:278-296, F = H + J + K).  With P = C C^T (alpha, occupied C) and
// K = -sum_X W_X^T W_X, W_X = (L^-1 B)_X C (nbf x nocc per aux function X),
// tr(K P) = -sum_X || W_X^T C ||_F^2 = -sum_X || C^T (L^-1 B)_X C ||_F^2:
// the nocc x nocc matrices M_X = W_X^T C replace the nbf x nbf SYRK.
//
//   k_exchange_energy  one 128 x 128 tile of M_X per workgroup (upper tiles of
//                      the symmetric M_X, tile rows and columns over the occupied
//                      index), FP64 MFMA over the basis index, then the
//                      compensated sum of squares of the tile (off-diagonal
//                      tiles count twice) into partial[workgroup]
//   k_sum_partials     the fixed-order compensated sum of the partials
//   k_sumsq_pairs      compensated sum of squares of a dense array (the
//                      metric-applied M_X of an untransformed B), one (sum,
//                      error) pair per workgroup
#include "dd.hpp"
#include "gemm_mfma.hpp"
#include "kernels.hpp"

namespace mchip {

namespace {
constexpr int EN_BM = 128;
constexpr int EN_LD = EN_BM + 8;   // LDS rows [k][m]
}  // namespace

// W: [x][nu][i] (row stride ldw, nbf rows per aux function), C: row-major
// nbf x ldc (first nocc columns).  Work item (x, tile) = item index / ntile,
// item index % ntile; tiles[t] = (tile row, tile column), row <= column.
// The items of one aux function are neighbours and run on one XCD (grid a
// multiple of 8; padding items exit).
__global__ __launch_bounds__(256) void k_exchange_energy(const double* __restrict__ W, int ldw, int nbf, int nocc,
                                                         const double* __restrict__ C, int ldc, int naux,
                                                         const int2* __restrict__ tiles, int ntile,
                                                         double* __restrict__ partial)
{
  __shared__ double As[2][GM_BK * EN_LD];
  __shared__ double Bs[2][GM_BK * EN_LD];
  __shared__ double red[2][4];
  const int item = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (item >= naux * ntile) return;
  const int x = item / ntile;
  const int2 tl = tiles[item % ntile];
  const int m0 = tl.x * EN_BM, n0 = tl.y * EN_BM;
  const double* A = W + (int64_t)x * nbf * ldw;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * 64, wn0 = (wave & 1) * 64;
  Mfma<double>::v4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<double>::v4{0.0, 0.0, 0.0, 0.0};
  // staging: thread (column tid & 127, k rows (tid >> 7) + 2 j), 8 of each operand per step
  double ra[8], rb[8];
  const int cm = tid & 127;
  auto load = [&](int k0) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + (tid >> 7) + 2 * j;
      ra[j] = (m0 + cm < nocc && k < nbf) ? A[(int64_t)k * ldw + m0 + cm] : 0.0;
      rb[j] = (n0 + cm < nocc && k < nbf) ? C[(int64_t)k * ldc + n0 + cm] : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      As[buf][((tid >> 7) + 2 * j) * EN_LD + cm] = ra[j];
      Bs[buf][((tid >> 7) + 2 * j) * EN_LD + cm] = rb[j];
    }
  };
  const int nsteps = (nbf + GM_BK - 1) / GM_BK;
  load(0);
  store(0);
  __syncthreads();
  for (int s = 0; s < nsteps; ++s) {
    const int cur = s & 1;
    if (s + 1 < nsteps) load((s + 1) * GM_BK);
    mfma_step<double, 4, 4, EN_LD, EN_LD>(As[cur], Bs[cur], wm0, wn0, lane, acc);
    if (s + 1 < nsteps) store(cur ^ 1);
    __syncthreads();
  }
  // rows and columns past nocc hold zeros (their operands were loaded as 0)
  dsum q;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) dot2_add(q, acc[i][j][e], acc[i][j][e]);
  dsum_wave(q);
  if (lane == 0) {
    red[0][wave] = q.s;
    red[1][wave] = q.c;
  }
  __syncthreads();
  if (tid == 0) {
    dsum t;
    for (int w = 0; w < 4; ++w) {
      dsum o;
      o.s = red[0][w];
      o.c = red[1][w];
      dsum_add(t, o);
    }
    const double f = tl.x == tl.y ? 1.0 : 2.0;   // an upper off-diagonal tile stands for its mirror too
    partial[2 * (int64_t)item] = f * t.s;
    partial[2 * (int64_t)item + 1] = f * t.c;
  }
}

// partial[2 b], partial[2 b + 1] = (sum, error) of sum a[i]^2 over the
// grid-stride share of workgroup b (fixed order: the same bits on every run)
__global__ __launch_bounds__(256) void k_sumsq_pairs(const double* __restrict__ a, int64_t n,
                                                     double* __restrict__ partial)
{
  __shared__ double red[2][4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  dsum acc;
  for (int64_t i = (int64_t)blockIdx.x * 256 + tid; i < n; i += (int64_t)gridDim.x * 256) dot2_add(acc, a[i], a[i]);
  dsum_wave(acc);
  if (lane == 0) {
    red[0][wave] = acc.s;
    red[1][wave] = acc.c;
  }
  __syncthreads();
  if (tid == 0) {
    dsum t;
    for (int w = 0; w < 4; ++w) {
      dsum o;
      o.s = red[0][w];
      o.c = red[1][w];
      dsum_add(t, o);
    }
    partial[2 * blockIdx.x] = t.s;
    partial[2 * blockIdx.x + 1] = t.c;
  }
}

// out[0] = scale * sum of the n (sum, error) pairs in partial, in index order
// by one workgroup (the same bits on every run).
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partial, int64_t n, double scale,
                                                      double* __restrict__ out)
{
  __shared__ double red[2][4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  dsum a;
  for (int64_t i = tid; i < n; i += 256) {
    dsum o;
    o.s = partial[2 * i];
    o.c = partial[2 * i + 1];
    dsum_add(a, o);
  }
  dsum_wave(a);
  if (lane == 0) {
    red[0][wave] = a.s;
    red[1][wave] = a.c;
  }
  __syncthreads();
  if (tid == 0) {
    dsum t;
    for (int w = 0; w < 4; ++w) {
      dsum o;
      o.s = red[0][w];
      o.c = red[1][w];
      dsum_add(t, o);
    }
    out[0] = scale * dsum_value(t);
  }
}

}  // namespace mchip
