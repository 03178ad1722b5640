// FP64 exchange kernels, direct-to-LDS pipelined (gfx950):
//   k_pad_f64_slab      one FP64 slab (rows of ncol) -> rows padded to ldp (zeros past ncol)
//   k_gather_ct_f64     per-slab gathered, transposed occupied coefficients, padded rows
//   k_slab_gemm_f64_dl  W(X, mu, i) = sum_nu B~(X, mu nu) C(nu, i) per slab (K1)
//   k_seg_syrk_f64_dl   partial[chunk] = sum_{X in chunk} W_X W2_X^T (K3; W2 = W: upper tiles)
//
// Same padded slab layout and LDS row geometry as the FP32 kernels: a
// contraction step covers 8 FP64 = 64 bytes of a row, 4 chunks of 16 bytes.
// Lane (fr, fk) of a 16x16 fragment reads chunk fk of row fr, i.e. k = 2 fk + q
// for q = 0, 1, and feeds element q to the q-th v_mfma_f64_16x16x4_f64 (A[i][k]
// from lane i + 16 k): the two MFMAs together sum over all 8 k.  FP64 MFMA
// accumulation throughout; the accumulator layout is D[lane/16 + 4e][lane%16].
#include "dl_common.hpp"
#include "gemm_mfma.hpp"
#include "kernels.hpp"

namespace mchip {

typedef Mfma<double>::v4 f64v4;

__global__ void k_pad_f64_slab(const double* __restrict__ src, int N, const int* __restrict__ ncol,
                               const int* __restrict__ bsz, int naux, const int* __restrict__ ldp,
                               const int64_t* __restrict__ off, double* __restrict__ out)
{
  const int K = ncol[N], ld = ldp[N];
  const int64_t n = (int64_t)naux * bsz[N] * ld;
  double* o = out + off[N];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ld;
    const int c = (int)(i - r * ld);
    o[i] = c < K ? src[r * K + c] : 0.0;
  }
}

__global__ void k_gather_ct_f64(const double* __restrict__ C, int ldc, int ng, const int* __restrict__ ncol,
                                const int* __restrict__ ldp, const int64_t* __restrict__ ldp_pre,
                                const int* __restrict__ colmap, const int* __restrict__ colmap_off,
                                double* __restrict__ g)
{
  const int N = blockIdx.y;
  const int ld = ldp[N];
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)ng * ld) return;
  const int i = (int)(idx / ld), c = (int)(idx - (int64_t)(idx / ld) * ld);
  g[ldp_pre[N] * ng + idx] = c < ncol[N] ? C[(int64_t)colmap[colmap_off[N] + c] * ldc + i] : 0.0;
}

// One 8-element step on swizzled stage tiles (uint4 units, 4 per row).
template <int TM, int TN>
__device__ __forceinline__ void f64_step_swz(const uint4* A, const uint4* B, int wm0, int wn0, int lane,
                                             f64v4 (&acc)[TM][TN])
{
  const int fr = lane & 15, fk = lane >> 4;
  double2 a[TM], b[TN];
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int r = wm0 + 16 * i + fr;
    a[i] = __builtin_bit_cast(double2, A[r * 4 + swz(r, fk)]);
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int r = wn0 + 16 * j + fr;
    b[j] = __builtin_bit_cast(double2, B[r * 4 + swz(r, fk)]);
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc[i][j] = Mfma<double>::op(a[i].x, b[j].x, acc[i][j]);
      acc[i][j] = Mfma<double>::op(a[i].y, b[j].y, acc[i][j]);
    }
}

// Grouped slab GEMM in FP64 on the padded copy: W rows (X, boff[N] + nu) for
// the rows X*bsz + nu of slab N, columns [0, ldo), zero in [ng, ldo).  Tile
// table as for the BF16 kernel (m0 steps of BM, padded to a multiple of 8 with
// slab -1).
template <int WM, int WN, int TM, int TN, int S>
__global__ __launch_bounds__(64 * WM * WN) void k_slab_gemm_f64_dl(
    const double* __restrict__ data, const int64_t* __restrict__ off, const int* __restrict__ ldp,
    const int64_t* __restrict__ ldp_pre, const int* __restrict__ bsz, const int* __restrict__ boff,
    const Tile* __restrict__ tiles, int naux, int nbf, const double* __restrict__ g, int ng,
    double* __restrict__ out, int ldo)
{
  constexpr int NW = WM * WN;
  constexpr int BM = 16 * WM * TM, BN = 16 * WN * TN;
  // per-wave loads per stage, rounded down: a wave with one load more waits
  // for one stage-load more than it needs, never less
  constexpr int LPW = (BM / 16 + BN / 16) / NW;
  constexpr int STAGE = (BM + BN) * 4;   // uint4 per stage
  static_assert(S >= 2 && S <= 4, "pipeline depth");
  __shared__ uint4 lds0[STAGE], lds1[STAGE], lds2[STAGE], lds3[STAGE];
  const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  if (t.slab < 0) return;
  const int N = t.slab;
  const int bn = bsz[N];
  const int ld2 = ldp[N] >> 1;   // uint4 per row
  const int64_t rows = (int64_t)naux * bn;
  const uint4* A = reinterpret_cast<const uint4*>(data + off[N]);
  const uint4* G = reinterpret_cast<const uint4*>(g + ldp_pre[N] * ng);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WN) * (16 * TM), wn0 = (wave % WN) * (16 * TN);
  f64v4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f64v4{0.0, 0.0, 0.0, 0.0};
  const int nsteps = ld2 >> 2;
  const int64_t alast = rows - 1;
  const int blast = ng - 1;
  dl_pipeline<S, LPW>(
      nsteps,
      [&](int s, auto k) {
        issue_rows<BM, BN, NW>(pick_stage<S>(k, lds0, lds1, lds2, lds3), wave, lane, A, G, t.m0, alast, t.n0,
                               blast, ld2, ld2, s * 4);
      },
      [&](int, auto k) {
        const uint4* st = pick_stage<S>(k, lds0, lds1, lds2, lds3);
        f64_step_swz<TM, TN>(st, st + BM * 4, wm0, wn0, lane, acc);
      });
  const int fc = lane & 15, fr = lane >> 4;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t r = (int64_t)t.m0 + wm0 + 16 * i + fr + 4 * e;
      if (r >= rows) continue;
      const int64_t x = r / bn;
      const int nu = (int)(r - x * bn);
      const int64_t orow = (x * nbf + boff[N] + nu) * (int64_t)ldo;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int c = t.n0 + wn0 + 16 * j + fc;
        if (c < ldo) out[orow + c] = c < ng ? acc[i][j][e] : 0.0;
      }
    }
}

// Segmented SYRK in FP64 over W (rows of ldw, a multiple of 8, doubles; the
// columns past the occupied count are zero): partial[chunk](m, n) =
// sum_{X in chunk} sum_i W(X,m,i) W2(X,n,i) on the 128 x 128 tiles of the
// table (the upper ones when W2 = W, the exchange being symmetric; all of
// them for the two-sided exchange), 8 waves of 64 x 32.  All tiles of a
// chunk run on one XCD (grid padded to a multiple of 8; padding blocks exit).
template <int S>
__global__ __launch_bounds__(512) void k_seg_syrk_f64_dl(const double* __restrict__ W, const double* __restrict__ W2,
                                                         int nseg, int64_t sstride, int ldw, int n,
                                                         const int2* __restrict__ tiles, int ntile,
                                                         int segs_per_chunk, double* __restrict__ partial)
{
  constexpr int BM = 128, NW = 8, WN = 4, TM = 4, TN = 2;
  constexpr int LPW = (2 * BM / 16) / NW;
  constexpr int STAGE = 2 * BM * 4;
  static_assert(S >= 2 && S <= 4, "pipeline depth");
  __shared__ uint4 lds0[STAGE], lds1[STAGE], lds2[STAGE], lds3[STAGE];
  const int lid = xcd_tile(blockIdx.x, gridDim.x);
  const int chunk = lid / ntile;
  const int s0 = chunk * segs_per_chunk;
  if (s0 >= nseg) return;
  const int2 tl = tiles[lid % ntile];
  const int s1 = min(nseg, s0 + segs_per_chunk);
  const int m0 = tl.x * BM, n0 = tl.y * BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WN) * 64, wn0 = (wave % WN) * 32;
  f64v4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f64v4{0.0, 0.0, 0.0, 0.0};
  const int ksteps = ldw >> 3;
  const int nsteps = (s1 - s0) * ksteps;
  const int ld2 = ldw >> 1;
  const uint4* Wa = reinterpret_cast<const uint4*>(W);
  const uint4* Wb = reinterpret_cast<const uint4*>(W2);
  dl_pipeline<S, LPW>(
      nsteps,
      [&](int q, auto k) {
        const int sg = s0 + q / ksteps;
        const int kc0 = (q % ksteps) * 4;
        const int64_t so = (int64_t)sg * (sstride >> 1);
        issue_rows<BM, BM, NW>(pick_stage<S>(k, lds0, lds1, lds2, lds3), wave, lane, Wa + so, Wb + so, m0, n - 1,
                               n0, n - 1, ld2, ld2, kc0);
      },
      [&](int, auto k) {
        const uint4* st = pick_stage<S>(k, lds0, lds1, lds2, lds3);
        f64_step_swz<TM, TN>(st, st + BM * 4, wm0, wn0, lane, acc);
      });
  double* P = partial + (int64_t)chunk * n * n;
  const int fc = lane & 15, fr = lane >> 4;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = m0 + wm0 + 16 * i + fr + 4 * e;
        const int c = n0 + wn0 + 16 * j + fc;
        if (r < n && c < n) P[(int64_t)r * n + c] = acc[i][j][e];
      }
}

#define MCHIP_SLABF64DL_INST(WM, WN, TM, TN, S)                                                             \
  template __global__ void k_slab_gemm_f64_dl<WM, WN, TM, TN, S>(                                           \
      const double*, const int64_t*, const int*, const int64_t*, const int*, const int*, const Tile*, int, int, \
      const double*, int, double*, int);
MCHIP_SLABF64DL_INST(2, 2, 4, 2, 3)
MCHIP_SLABF64DL_INST(2, 4, 4, 2, 3)
MCHIP_SLABF64DL_INST(2, 2, 4, 5, 3)
MCHIP_SLABF64DL_INST(2, 4, 4, 5, 3)
MCHIP_SLABF64DL_INST(2, 4, 4, 5, 2)
#undef MCHIP_SLABF64DL_INST
template __global__ void k_seg_syrk_f64_dl<2>(const double*, const double*, int, int64_t, int, int, const int2*, int,
                                               int, double*);
template __global__ void k_seg_syrk_f64_dl<3>(const double*, const double*, int, int64_t, int, int, const int2*, int,
                                               int, double*);
template __global__ void k_seg_syrk_f64_dl<4>(const double*, const double*, int, int64_t, int, int, const int2*, int,
                                               int, double*);

}  // namespace mchip
