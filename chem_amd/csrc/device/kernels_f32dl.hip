// FP32 exchange on the transformed B~ = L^-1 B, direct-to-LDS pipelined (gfx950):
//   k_pad_f32_slab      one FP64 slab of B~ -> FP32 rows padded to ldp (zeros past ncol)
//   k_gather_ct_f32     per-slab gathered, transposed occupied coefficients in FP32
//   k_slab_gemm_f32_dl  W(X, mu, i) = sum_nu B~(X, mu nu) C(nu, i) per slab (K1)
//   k_seg_syrk_f32_dl   partial[chunk] = sum_{X in chunk} W_X W_X^T, upper tiles (K3)
// and, for the grouped SOS-MP2 graph:
//   k_transpose_pad_f32   b[X][nu][i] -> b^T[X][i][nu], rows padded to ldn
//   k_group_apply_f32_dl  M3: P_v,g b_g[X] for every (point, X), P_v as an FP32 pair
//   k_to_f32x2_rows       FP64 P_v -> FP32 hi/lo rows padded to ldn
//   k_pad_f32_rows, k_gather_ct_f32f, k_slab_gemm_f32_dl_f64, k_slab_out_gemm_f32_dl
//                         M2: b = B~ L on the padded FP32 copy, FP64 sums
//
// Operands use the padded slab layout of the BF16 copy: slab N has naux*bsz[N]
// rows of ldp[N] (multiple of 32) elements at element offset off[N]; the
// gathered coefficients of slab N are ng rows of ldp[N] at ldp_pre[N]*ng.
// A contraction step covers 16 FP32 = 64 bytes of a row, the same LDS row
// geometry (4 chunks of 16 bytes, XOR-swizzled) as one BF16 step of 32
// elements.  Lane (fr, fk) of a 16x16 MFMA fragment reads chunk fk of row fr,
// i.e. k = 4 fk + q for q = 0..3, and feeds element q to the q-th
// v_mfma_f32_16x16x4_f32: the four MFMAs together sum over all 16 k.
#include "dl_common.hpp"
#include "gemm_mfma.hpp"
#include "kernels.hpp"

namespace mchip {

typedef Mfma<float>::v4 f32v4;

__global__ void k_pad_f32_slab(const double* __restrict__ src, int N, const int* __restrict__ ncol,
                               const int* __restrict__ bsz, int naux, const int* __restrict__ ldp,
                               const int64_t* __restrict__ off, float* __restrict__ out)
{
  const int K = ncol[N], ld = ldp[N];
  const int64_t n = (int64_t)naux * bsz[N] * ld;
  float* o = out + off[N];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ld;
    const int c = (int)(i - r * ld);
    o[i] = c < K ? (float)src[r * K + c] : 0.f;
  }
}

__global__ void k_gather_ct_f32(const double* __restrict__ C, int ldc, int ng, const int* __restrict__ ncol,
                                const int* __restrict__ ldp, const int64_t* __restrict__ ldp_pre,
                                const int* __restrict__ colmap, const int* __restrict__ colmap_off,
                                float* __restrict__ g)
{
  const int N = blockIdx.y;
  const int ld = ldp[N];
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)ng * ld) return;
  const int i = (int)(idx / ld), c = (int)(idx - (int64_t)(idx / ld) * ld);
  g[ldp_pre[N] * ng + idx] = c < ncol[N] ? (float)C[(int64_t)colmap[colmap_off[N] + c] * ldc + i] : 0.f;
}

// One 16-element step on swizzled stage tiles (uint4 units, 4 per row).
template <int TM, int TN>
__device__ __forceinline__ void f32_step_swz(const uint4* A, const uint4* B, int wm0, int wn0, int lane,
                                             f32v4 (&acc)[TM][TN])
{
  const int fr = lane & 15, fk = lane >> 4;
  float4 a[TM], b[TN];
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int r = wm0 + 16 * i + fr;
    const uint4 v = A[r * 4 + swz(r, fk)];
    a[i] = make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int r = wn0 + 16 * j + fr;
    const uint4 v = B[r * 4 + swz(r, fk)];
    b[j] = make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc[i][j] = Mfma<float>::op(a[i].x, b[j].x, acc[i][j]);
      acc[i][j] = Mfma<float>::op(a[i].y, b[j].y, acc[i][j]);
      acc[i][j] = Mfma<float>::op(a[i].z, b[j].z, acc[i][j]);
      acc[i][j] = Mfma<float>::op(a[i].w, b[j].w, acc[i][j]);
    }
}

// Grouped slab GEMM in FP32: W rows (X, boff[N] + nu) for the rows X*bsz + nu
// of slab N, columns [0, ldo), zero in [ng, ldo).  Tile table as for the BF16
// kernel (m0 steps of BM, padded to a multiple of 8 with slab -1).  FP32
// accumulation over the slab's columns.
template <int WM, int WN, int TM, int TN, int S>
__global__ __launch_bounds__(64 * WM * WN) void k_slab_gemm_f32_dl(
    const float* __restrict__ data, const int64_t* __restrict__ off, const int* __restrict__ ldp,
    const int64_t* __restrict__ ldp_pre, const int* __restrict__ bsz, const int* __restrict__ boff,
    const Tile* __restrict__ tiles, int naux, int nbf, const float* __restrict__ g, int ng,
    float* __restrict__ out, int ldo)
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
  const int ld4 = ldp[N] >> 2;
  const int64_t rows = (int64_t)naux * bn;
  const uint4* A = reinterpret_cast<const uint4*>(data + off[N]);
  const uint4* G = reinterpret_cast<const uint4*>(g + ldp_pre[N] * ng);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WN) * (16 * TM), wn0 = (wave % WN) * (16 * TN);
  f32v4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32v4{0.f, 0.f, 0.f, 0.f};
  const int nsteps = ld4 >> 2;
  const int64_t alast = rows - 1;
  const int blast = ng - 1;
  dl_pipeline<S, LPW>(
      nsteps,
      [&](int s, auto k) {
        issue_rows<BM, BN, NW>(pick_stage<S>(k, lds0, lds1, lds2, lds3), wave, lane, A, G, t.m0, alast, t.n0,
                                    blast, ld4, ld4, s * 4);
      },
      [&](int, auto k) {
        const uint4* st = pick_stage<S>(k, lds0, lds1, lds2, lds3);
        f32_step_swz<TM, TN>(st, st + BM * 4, wm0, wn0, lane, acc);
      });
  const int fc = lane & 15, fr4 = 4 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t r = (int64_t)t.m0 + wm0 + 16 * i + fr4 + e;
      if (r >= rows) continue;
      const int64_t x = r / bn;
      const int nu = (int)(r - x * bn);
      const int64_t orow = (x * nbf + boff[N] + nu) * (int64_t)ldo;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int c = t.n0 + wn0 + 16 * j + fc;
        if (c < ldo) out[orow + c] = c < ng ? acc[i][j][e] : 0.f;
      }
    }
}

// Segmented SYRK in FP32 over W (rows of ldw, a multiple of 16, floats; the
// columns past the occupied count are zero): partial[chunk](m, n) =
// sum_{X in chunk} sum_i W(X,m,i) W(X,n,i) on the upper 128 x 128 tiles, 8
// waves of 64 x 32.  The FP32 MFMA accumulators are added into FP64 registers
// every 16 steps (64 MFMA accumulations) and restarted.  All tiles of a chunk
// run on one XCD (grid padded to a multiple of 8; padding blocks exit).
template <int S>
__global__ __launch_bounds__(512) void k_seg_syrk_f32_dl(const float* __restrict__ W, int nseg, int64_t sstride,
                                                         int ldw, int n, const int2* __restrict__ tiles, int ntile,
                                                         int segs_per_chunk, double* __restrict__ partial)
{
  constexpr int BM = 128, NW = 8, WN = 4, TM = 4, TN = 2, FLUSH = 16;
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
  f32v4 acc[TM][TN];
  double acc64[TM][TN][4];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc[i][j] = f32v4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int e = 0; e < 4; ++e) acc64[i][j][e] = 0.0;
    }
  const int ksteps = ldw >> 4;
  const int nsteps = (s1 - s0) * ksteps;
  const int ld4 = ldw >> 2;
  const uint4* W4 = reinterpret_cast<const uint4*>(W);
  dl_pipeline<S, LPW>(
      nsteps,
      [&](int q, auto k) {
        const int sg = s0 + q / ksteps;
        const int kc0 = (q % ksteps) * 4;
        const uint4* base = W4 + (int64_t)sg * (sstride >> 2);
        issue_rows<BM, BM, NW>(pick_stage<S>(k, lds0, lds1, lds2, lds3), wave, lane, base, base, m0, n - 1, n0,
                                    n - 1, ld4, ld4, kc0);
      },
      [&](int q, auto k) {
        const uint4* st = pick_stage<S>(k, lds0, lds1, lds2, lds3);
        f32_step_swz<TM, TN>(st, st + BM * 4, wm0, wn0, lane, acc);
        if ((q + 1) % FLUSH == 0) flush_f64<TM, TN>(acc, acc64);
      });
  flush_f64<TM, TN>(acc, acc64);
  double* P = partial + (int64_t)chunk * n * n;
  const int fc = lane & 15, fr4 = 4 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = m0 + wm0 + 16 * i + fr4 + e;
        const int c = n0 + wn0 + 16 * j + fc;
        if (r < n && c < n) P[(int64_t)r * n + c] = acc64[i][j][e];
      }
}

// Rows of an n x R block per aux function, transposed and zero-padded:
// out[X][c][k] = in[X][k][c] for k < n, 0 for n <= k < ldn (c < R).  32 x 32
// tiles through LDS (padded rows: no bank conflicts); grid (R/32, ldn/32, X).
__global__ __launch_bounds__(256) void k_transpose_pad_f32(const float* __restrict__ in, int n, int R, int ldn,
                                                           float* __restrict__ out)
{
  __shared__ float tile[32][33];
  const int64_t X = blockIdx.z;
  const int c0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  const float* src = in + X * n * R;
  float* dst = out + X * (int64_t)R * ldn;
#pragma unroll
  for (int r = ty; r < 32; r += 8) {
    const int k = k0 + r, c = c0 + tx;
    tile[r][tx] = (k < n && c < R) ? src[(int64_t)k * R + c] : 0.f;
  }
  __syncthreads();
#pragma unroll
  for (int r = ty; r < 32; r += 8) {
    const int c = c0 + r, k = k0 + tx;
    if (c < R && k < ldn) dst[(int64_t)c * ldn + k] = tile[tx][r];
  }
}

// M3 of the grouped SOS-MP2 contractions (the virtual pseudo-density applied
// to b for every X and Laplace point, z_full.cpp:224-240), direct-to-LDS:
// per point g, C_g = (Phi_g + Plo_g) B_g with B_g's columns the pairs (X, c),
// c < pwidth[g], column index X pwidth[g] + c:
//   out[X][m][pcol[g] + c] = sum_k P_g(m, k) bT[X][pcol[g] + c][k].
// Tiles (g, m0, column 0 of the tile, -) of 128 rows m and 128 columns (the
// columns of all aux functions run on, so a rank below 128 wastes no tile
// width).  The P rows (hi then lo plane) and the transposed b rows have ldn
// (multiple of 16) floats, zero past n; each 16-element step adds its FP32
// MFMA sums (16 products) into FP64 running sums.  Eight waves of 64 x 32.
template <int S>
__global__ __launch_bounds__(512, 2) void k_group_apply_f32_dl(const float* __restrict__ P, int npt,
                                                            const float* __restrict__ bT, float* __restrict__ out,
                                                            int n, int ldn, int R, int naux,
                                                            const int* __restrict__ pcol,
                                                            const int* __restrict__ pwidth,
                                                            const int4* __restrict__ tiles)
{
  constexpr int BM = 128, BN = 128, NW = 8, WN = 4, TM = 4, TN = 2;
  constexpr int QA = BM / 16, QB = BN / 16;
  constexpr int LPW = (QA + QB) / NW;
  static_assert((QA + QB) % NW == 0, "stage loads must split evenly over the waves");
  constexpr int STAGE = (BM + BN) * 4;
  static_assert(S >= 2 && S <= 4, "pipeline depth");
  __shared__ uint4 lds0[STAGE], lds1[STAGE], lds2[STAGE], lds3[STAGE];
  const int4 t = tiles[blockIdx.x];   // (point, m0, first column, -)
  const int g = t.x, w = pwidth[g], pc = pcol[g];
  const int ncols = naux * w;
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm0 = (wave / WN) * 64, wn0 = (wave % WN) * 32;
  // two sets of FP32 accumulators: step s sums into set s % 2 while the set of
  // step s - 1 (its MFMAs long done) is added into the FP64 sums, so no step
  // waits for its own MFMA results; the FP64 additions keep the step order
  static_assert(S % 2 == 0, "accumulator set = stage parity");
  f32v4 acc[2][TM][TN];
  double acc64[TM][TN][4];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc[0][i][j] = f32v4{0.f, 0.f, 0.f, 0.f};
      acc[1][i][j] = f32v4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int e = 0; e < 4; ++e) acc64[i][j][e] = 0.0;
    }
  const int ld4 = ldn >> 2;
  const int nh = ldn >> 4;   // steps per P plane
  static_assert(QA == NW && QB == NW, "one A row block and one B column block per wave and stage");
  const uint4* Phi = reinterpret_cast<const uint4*>(P + (int64_t)g * n * ldn);
  const int64_t lo_off = (int64_t)npt * n * ld4;   // the lo plane, in 16-byte units
  const uint4* B4 = reinterpret_cast<const uint4*>(bT);
  // per wave: A rows 16 wave + rr and B columns (X, c) 16 wave + rr (clamped
  // to the last valid), offsets in 16-byte units, fixed over the contraction
  const int rr = lane >> 2, pos = lane & 3;
  const int r = 16 * wave + rr;
  const int m = t.y + r < n - 1 ? t.y + r : n - 1;
  const int64_t offA = (int64_t)m * ld4 + swz(r, pos);
  const int col = t.z + r < ncols - 1 ? t.z + r : ncols - 1;
  const int Xc = col / w;
  const int64_t offB = ((int64_t)Xc * R + pc + (col - Xc * w)) * ld4 + swz(r, pos);
  dl_pipeline<S, LPW>(
      2 * nh,
      [&](int s, auto k) {
        uint4* st = pick_stage<S>(k, lds0, lds1, lds2, lds3);
        const bool lo = s >= nh;
        const int kc = (lo ? s - nh : s) * 4;
        lds_load16(Phi + (offA + (lo ? lo_off : 0) + kc), st + wave * 64);
        lds_load16(B4 + (offB + kc), st + BM * 4 + wave * 64);
      },
      [&](int, auto k) {
        constexpr int p = decltype(k)::value & 1;
        const uint4* st = pick_stage<S>(k, lds0, lds1, lds2, lds3);
        f32_step_swz<TM, TN>(st, st + BM * 4, wm0, wn0, lane, acc[p]);
        flush_f64<TM, TN>(acc[1 - p], acc64);   // step s - 1 (zero before the first step)
      });
  flush_f64<TM, TN>(acc[1], acc64);   // the last step: 2 nh - 1 is odd
  const int fc = lane & 15, fr4 = 4 * (lane >> 4);
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int col = t.z + wn0 + 16 * j + fc;
    if (col >= ncols) continue;
    const int X = col / w, c = col - X * w;
    float* o = out + (int64_t)X * n * R + pc + c;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = t.y + wm0 + 16 * i + fr4 + e;
        if (m < n) o[(int64_t)m * R] = (float)acc64[i][j][e];
      }
  }
}

// One slab of FP32 values in the slab layout (rows of ncol[N]) -> rows padded
// to ldp[N] with zeros, at element offset off[N].
__global__ void k_pad_f32_rows(const float* __restrict__ src, int N, const int* __restrict__ ncol,
                               const int* __restrict__ bsz, int naux, const int* __restrict__ ldp,
                               const int64_t* __restrict__ off, float* __restrict__ out)
{
  const int K = ncol[N], ld = ldp[N];
  const int64_t n = (int64_t)naux * bsz[N] * ld;
  float* o = out + off[N];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ld;
    const int c = (int)(i - r * ld);
    o[i] = c < K ? src[r * K + c] : 0.f;
  }
}

// Per-slab gathered, transposed FP32 coefficients from an FP32 n x ldc matrix:
// g[ldp_pre[N] ng + i ldp[N] + c] = C(colmap_N[c], i), zero for c >= ncol[N].
__global__ void k_gather_ct_f32f(const float* __restrict__ C, int ldc, int ng, const int* __restrict__ ncol,
                                 const int* __restrict__ ldp, const int64_t* __restrict__ ldp_pre,
                                 const int* __restrict__ colmap, const int* __restrict__ colmap_off,
                                 float* __restrict__ g)
{
  const int N = blockIdx.y;
  const int ld = ldp[N];
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)ng * ld) return;
  const int i = (int)(idx / ld), c = (int)(idx - (int64_t)(idx / ld) * ld);
  g[ldp_pre[N] * ng + idx] = c < ncol[N] ? C[(int64_t)colmap[colmap_off[N] + c] * ldc + i] : 0.f;
}

// The FP32 slab GEMM with FP64 sums (MP2 M2, b = B~ L): W rows (X, boff[N] +
// nu) for the rows X*bsz + nu of slab N, columns [0, ldo), zero in [ng, ldo).
// 128 x 128 tiles, eight waves of 64 x 32; each 16-element step's FP32 MFMA
// sums (16 products) are added into FP64 running sums, the step s - 1's set
// while step s accumulates into the other (double-buffered), in step order.
template <int S>
__global__ __launch_bounds__(512, 2) void k_slab_gemm_f32_dl_f64(
    const float* __restrict__ data, const int64_t* __restrict__ off, const int* __restrict__ ldp,
    const int64_t* __restrict__ ldp_pre, const int* __restrict__ bsz, const int* __restrict__ boff,
    const Tile* __restrict__ tiles, int naux, int nbf, const float* __restrict__ g, int ng,
    float* __restrict__ out, int ldo)
{
  constexpr int BM = 128, BN = 128, NW = 8, WN = 4, TM = 4, TN = 2;
  constexpr int QA = BM / 16, QB = BN / 16;
  static_assert(QA == NW && QB == NW, "one A and one B row block per wave and stage");
  static_assert(S % 2 == 0 && S >= 2 && S <= 4, "pipeline depth (even: accumulator set = stage parity)");
  constexpr int STAGE = (BM + BN) * 4;
  __shared__ uint4 lds0[STAGE], lds1[STAGE], lds2[STAGE], lds3[STAGE];
  const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  if (t.slab < 0) return;
  const int N = t.slab;
  const int bn = bsz[N];
  const int ld4 = ldp[N] >> 2;
  const int64_t rows = (int64_t)naux * bn;
  const uint4* A = reinterpret_cast<const uint4*>(data + off[N]);
  const uint4* G = reinterpret_cast<const uint4*>(g + ldp_pre[N] * ng);
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm0 = (wave / WN) * (16 * TM), wn0 = (wave % WN) * (16 * TN);
  f32v4 acc[2][TM][TN];
  double acc64[TM][TN][4];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc[0][i][j] = f32v4{0.f, 0.f, 0.f, 0.f};
      acc[1][i][j] = f32v4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int e = 0; e < 4; ++e) acc64[i][j][e] = 0.0;
    }
  const int nsteps = ld4 >> 2;
  // per wave: A row 16 wave + rr and B row 16 wave + rr (clamped), offsets in
  // 16-byte units, fixed over the contraction
  const int rr = lane >> 2, pos = lane & 3;
  const int r = 16 * wave + rr;
  const int64_t ar = (int64_t)t.m0 + r < rows - 1 ? (int64_t)t.m0 + r : rows - 1;
  const int br = t.n0 + r < ng - 1 ? t.n0 + r : ng - 1;
  const int64_t offA = ar * ld4 + swz(r, pos), offB = (int64_t)br * ld4 + swz(r, pos);
  dl_pipeline<S, 2>(
      nsteps,
      [&](int s, auto k) {
        uint4* st = pick_stage<S>(k, lds0, lds1, lds2, lds3);
        lds_load16(A + (offA + s * 4), st + wave * 64);
        lds_load16(G + (offB + s * 4), st + BM * 4 + wave * 64);
      },
      [&](int, auto k) {
        constexpr int p = decltype(k)::value & 1;
        const uint4* st = pick_stage<S>(k, lds0, lds1, lds2, lds3);
        f32_step_swz<TM, TN>(st, st + BM * 4, wm0, wn0, lane, acc[p]);
        flush_f64<TM, TN>(acc[1 - p], acc64);   // step s - 1 (zero before the first step)
      });
  // the last step's set (constant indices: a run-time index would put the
  // accumulators in scratch memory)
  if (nsteps > 0) {
    if ((nsteps - 1) & 1)
      flush_f64<TM, TN>(acc[1], acc64);
    else
      flush_f64<TM, TN>(acc[0], acc64);
  }
  const int fc = lane & 15, fr4 = 4 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t rw = (int64_t)t.m0 + wm0 + 16 * i + fr4 + e;
      if (rw >= rows) continue;
      const int64_t x = rw / bn;
      const int nu = (int)(rw - x * bn);
      const int64_t orow = (x * nbf + boff[N] + nu) * (int64_t)ldo;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int c = t.n0 + wn0 + 16 * j + fc;
        if (c < ldo) out[orow + c] = c < ng ? (float)acc64[i][j][e] : 0.f;
      }
    }
}

template __global__ void k_slab_gemm_f32_dl_f64<2>(const float*, const int64_t*, const int*, const int64_t*,
                                                   const int*, const int*, const Tile*, int, int, const float*, int,
                                                   float*, int);
template __global__ void k_slab_gemm_f32_dl_f64<4>(const float*, const int64_t*, const int*, const int64_t*,
                                                   const int*, const int*, const Tile*, int, int, const float*, int,
                                                   float*, int);

// MP2 M4 on direct-to-LDS pipelines, grouped over the Laplace points:
// out_g[slab_off[N] + r nc + c] = sum_{i < pwidth[g]} A[(x nbf + boff[N] + nu) ldr
// + pcol[g] + i] G[colmap_N[c] ldr + pcol[g] + i] for the rows r = x bn + nu
// and columns c < nc of slab N (b3 = L_o b2 on B's pair pattern).  Both
// operands are contraction-contiguous rows of ldr floats; ldr, pcol[g] and
// pwidth[g] are multiples of 16 (whole 64-byte steps).  Tile = (slab,
// m0, n0, point g in pad): 128 x 128, eight waves of 64 x 32, one A and one B
// row block per wave and stage; each step's FP32 MFMA sums (16 products) go
// into FP64 running sums, double-buffered as in the M2 kernel.
template <int S>
__global__ __launch_bounds__(512, 2) void k_slab_out_gemm_f32_dl(
    const float* __restrict__ A, int ldr, const int64_t* __restrict__ slab_off, const int* __restrict__ ncol,
    const int* __restrict__ bsz, const int* __restrict__ boff, const int* __restrict__ colmap,
    const int* __restrict__ colmap_off, const Tile* __restrict__ tiles, int naux, int nbf,
    const float* __restrict__ G, float* __restrict__ out, const int* __restrict__ pcol,
    const int* __restrict__ pwidth, int64_t out_stride)
{
  constexpr int BM = 128, BN = 128, NW = 8, WN = 4, TM = 4, TN = 2;
  constexpr int QA = BM / 16, QB = BN / 16;
  static_assert(QA == NW && QB == NW, "one A and one B row block per wave and stage");
  static_assert(S % 2 == 0 && S >= 2 && S <= 4, "pipeline depth (even: accumulator set = stage parity)");
  constexpr int STAGE = (BM + BN) * 4;
  __shared__ uint4 lds0[STAGE], lds1[STAGE], lds2[STAGE], lds3[STAGE];
  const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  if (t.slab < 0) return;
  const int N = t.slab, g = t.pad;
  const int bn = bsz[N], nc = ncol[N];
  const int64_t rows = (int64_t)naux * bn;
  const int ld4 = ldr >> 2, pc4 = pcol[g] >> 2;
  const uint4* A4 = reinterpret_cast<const uint4*>(A);
  const uint4* G4 = reinterpret_cast<const uint4*>(G);
  const int* cmap = colmap + colmap_off[N];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm0 = (wave / WN) * (16 * TM), wn0 = (wave % WN) * (16 * TN);
  f32v4 acc[2][TM][TN];
  double acc64[TM][TN][4];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc[0][i][j] = f32v4{0.f, 0.f, 0.f, 0.f};
      acc[1][i][j] = f32v4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int e = 0; e < 4; ++e) acc64[i][j][e] = 0.0;
    }
  const int nsteps = pwidth[g] >> 4;
  // per wave: A row 16 wave + rr (a row (x, nu) of b2) and B row 16 wave + rr
  // (the L_o row of column c), both clamped into range, offsets in 16-byte
  // units, fixed over the contraction
  const int rr = lane >> 2, pos = lane & 3;
  const int r = 16 * wave + rr;
  const int64_t ra = (int64_t)t.m0 + r < rows - 1 ? (int64_t)t.m0 + r : rows - 1;
  const int64_t x = ra / bn;
  const int64_t arow = x * nbf + boff[N] + (ra - x * bn);
  const int cb = t.n0 + r < nc - 1 ? t.n0 + r : nc - 1;
  const int64_t offA = arow * ld4 + pc4 + swz(r, pos), offB = (int64_t)cmap[cb] * ld4 + pc4 + swz(r, pos);
  dl_pipeline<S, 2>(
      nsteps,
      [&](int s, auto k) {
        uint4* st = pick_stage<S>(k, lds0, lds1, lds2, lds3);
        lds_load16(A4 + (offA + s * 4), st + wave * 64);
        lds_load16(G4 + (offB + s * 4), st + BM * 4 + wave * 64);
      },
      [&](int, auto k) {
        constexpr int p = decltype(k)::value & 1;
        const uint4* st = pick_stage<S>(k, lds0, lds1, lds2, lds3);
        f32_step_swz<TM, TN>(st, st + BM * 4, wm0, wn0, lane, acc[p]);
        flush_f64<TM, TN>(acc[1 - p], acc64);   // step s - 1 (zero before the first step)
      });
  if (nsteps > 0) {
    if ((nsteps - 1) & 1)
      flush_f64<TM, TN>(acc[1], acc64);
    else
      flush_f64<TM, TN>(acc[0], acc64);
  }
  float* o = out + (int64_t)g * out_stride + slab_off[N];
  const int fc = lane & 15, fr4 = 4 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t rw = (int64_t)t.m0 + wm0 + 16 * i + fr4 + e;
      if (rw >= rows) continue;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int c = t.n0 + wn0 + 16 * j + fc;
        if (c < nc) o[rw * nc + c] = (float)acc64[i][j][e];
      }
    }
}

template __global__ void k_slab_out_gemm_f32_dl<2>(const float*, int, const int64_t*, const int*, const int*,
                                                   const int*, const int*, const int*, const Tile*, int, int,
                                                   const float*, float*, const int*, const int*, int64_t);
template __global__ void k_slab_out_gemm_f32_dl<4>(const float*, int, const int64_t*, const int*, const int*,
                                                   const int*, const int*, const int*, const Tile*, int, int,
                                                   const float*, float*, const int*, const int*, int64_t);

// FP64 n x n -> FP32 hi/lo pair, rows padded with zeros to ldn.
__global__ void k_to_f32x2_rows(const double* __restrict__ in, int n, int ldn, float* __restrict__ hi,
                                float* __restrict__ lo)
{
  const int64_t total = (int64_t)n * ldn;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ldn;
    const int c = (int)(i - r * ldn);
    float h = 0.f, l = 0.f;
    if (c < n) {
      const double v = in[r * n + c];
      h = (float)v;
      l = (float)(v - (double)h);
    }
    hi[i] = h;
    lo[i] = l;
  }
}

template __global__ void k_group_apply_f32_dl<2>(const float*, int, const float*, float*, int, int, int, int,
                                                 const int*, const int*, const int4*);
template __global__ void k_group_apply_f32_dl<4>(const float*, int, const float*, float*, int, int, int, int,
                                                 const int*, const int*, const int4*);

#define MCHIP_SLABF32DL_INST(WM, WN, TM, TN, S)                                                            \
  template __global__ void k_slab_gemm_f32_dl<WM, WN, TM, TN, S>(                                          \
      const float*, const int64_t*, const int*, const int64_t*, const int*, const int*, const Tile*, int, int, \
      const float*, int, float*, int);
MCHIP_SLABF32DL_INST(2, 2, 4, 2, 3)
MCHIP_SLABF32DL_INST(2, 4, 4, 2, 3)
MCHIP_SLABF32DL_INST(2, 2, 4, 5, 3)
MCHIP_SLABF32DL_INST(2, 4, 4, 5, 3)
MCHIP_SLABF32DL_INST(2, 4, 6, 5, 3)
MCHIP_SLABF32DL_INST(2, 4, 4, 5, 2)
#undef MCHIP_SLABF32DL_INST
template __global__ void k_seg_syrk_f32_dl<2>(const float*, int, int64_t, int, int, const int2*, int, int, double*);
template __global__ void k_seg_syrk_f32_dl<3>(const float*, int, int64_t, int, int, const int2*, int, int, double*);
template __global__ void k_seg_syrk_f32_dl<4>(const float*, int, int64_t, int, int, const int2*, int, int, double*);

}  // namespace mchip
