// BF16 split-precision exchange kernels (v_mfma_f32_16x16x32_bf16, gfx950).
//
// Every FP32/FP64 operand x is carried as hi = bf16(x), lo = bf16(x - hi)
// (16 significant bits), and a product is formed as lo*hi + hi*lo + hi*hi
// with FP32 accumulation (the lo*lo term, 2^-16 relative, is dropped).
// Chunk partials (FP32 sums over at most 32 aux functions) are reduced in FP64.
//
//   k_split_slabs / k_split_slab   B~ (FP64 pair slabs) -> hi/lo slabs whose rows
//                         are padded to ldp[N] = roundup(ncol[N], 32) (zeros)
//   k_gather_ct           per-slab gathered, transposed occupied coefficients
//                         G_N[i][c] = C(colmap_N[c], i) as hi/lo, row length ldp[N]
//   k_slab_gemm_bf16x3    W(X,nu,i) = sum_c B~_N(X nu, c) G_N(i, c) -> FP32 W,
//                         register-staged (XMN,Mi->XiN of jkdf.cpp:361 on B~)
//   k_slab_gemm_bf16x3_dl / _dla   the same, direct-to-LDS pipelined (one
//                         pipeline for both operands / separate ones), W as
//                         bf16 hi/lo
//   k_seg_syrk_bf16x3     partial[chunk](m,n) = sum_{X in chunk} sum_i W(X,m,i) W(X,n,i)
//                         (XiM,XiN->MN of jkdf.cpp:417), W split to hi/lo on load
//   k_seg_syrk_bf16x3_dl  the same on bf16 hi/lo W, direct-to-LDS pipelined
//
// Both GEMMs read their operands with the contraction index contiguous, 16
// bytes (8 bf16) per lane, which is exactly the MFMA operand layout:
//   A[l%16][8*(l/16)+q], B[8*(l/16)+q][l%16], D[4*(l/16)+e][l%16]
// (pinned by tests/test_gpu_mfma_layout.py).  In the register-staged kernels
// LDS rows hold 32 bf16 plus 8 of padding (80 bytes); the direct-to-LDS ones
// use 64-byte rows with XOR-swizzled 16-byte chunks (dl_common.hpp).
#include "dl_common.hpp"
#include "kernels.hpp"

namespace mchip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int LROW = 5;   // LDS row stride in 16-byte units (32 bf16 + 8 padding)

__device__ __forceinline__ f32x4 mfma_bf16(const uint4& a, const uint4& b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

__device__ __forceinline__ uint16_t bf16_bits(float x)
{
  return __builtin_bit_cast(uint16_t, (__bf16)x);
}

__device__ __forceinline__ float bf16_value(uint16_t b)
{
  return (float)__builtin_bit_cast(__bf16, b);
}

// hi/lo split of an FP64 value (hi from the FP32 rounding, lo from the exact remainder)
__device__ __forceinline__ void split64(double x, uint16_t& hi, uint16_t& lo)
{
  hi = bf16_bits((float)x);
  lo = bf16_bits((float)(x - (double)bf16_value(hi)));
}

__device__ __forceinline__ void split32(float x, uint16_t& hi, uint16_t& lo)
{
  hi = bf16_bits(x);
  lo = bf16_bits(x - bf16_value(hi));
}

__global__ void k_split_slabs(const double* __restrict__ data64, const int64_t* __restrict__ slab_off,
                              const int* __restrict__ ncol, const int* __restrict__ bsz, int naux,
                              const int* __restrict__ ldp, const int64_t* __restrict__ off16,
                              uint16_t* __restrict__ hi, uint16_t* __restrict__ lo)
{
  const int N = blockIdx.y;
  const int K = ncol[N], ld = ldp[N];
  const int64_t n = (int64_t)naux * bsz[N] * ld;
  const double* src = data64 + slab_off[N];
  uint16_t* h = hi + off16[N];
  uint16_t* l = lo + off16[N];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ld;
    const int c = (int)(i - r * ld);
    uint16_t a = 0, b = 0;
    if (c < K) split64(src[r * K + c], a, b);
    h[i] = a;
    l[i] = b;
  }
}

// Copies of L^-1 B from the canonical half of a symmetric tensor
// (B(X|mu nu) = B(X|nu mu) and every kept pair (M, N) has its transpose kept).
// T holds the transformed canonical columns c < ce of slab N, rows (x, nu) of
// length ce.  The values go into the padded copies (element offsets off[],
// row length ldp[]; any of out64, out32, hi/lo may be null):
//   k_tilde_canon   slab N's columns c < ce, and zeros in its padding
//                   columns [ncol, ldp) (columns [ce, ncol) are the mirrors
//                   written for the slabs of their blocks)
//   k_tilde_mirror  the columns of each block M < N (segment s: block M at
//                   column col of slab N, width bM) into slab M's columns for
//                   block N (from column q): element (x, mu, nu) of the
//                   segment goes to row (x, mu), column q + nu of slab M
//                   (nu fastest, so the stores run along a row)
__device__ __forceinline__ void tilde_store(double v, int64_t j, double* out64, float* out32, uint16_t* hi,
                                            uint16_t* lo)
{
  if (out64) out64[j] = v;
  if (out32) out32[j] = (float)v;
  if (hi) {
    uint16_t a, b;
    split64(v, a, b);
    hi[j] = a;
    lo[j] = b;
  }
}

__global__ void k_tilde_canon(const double* __restrict__ T, int N, int ce, const int* __restrict__ ncol,
                              const int* __restrict__ bsz, int naux, const int* __restrict__ ldp,
                              const int64_t* __restrict__ off, double* __restrict__ out64,
                              float* __restrict__ out32, uint16_t* __restrict__ hi, uint16_t* __restrict__ lo)
{
  const int K = ncol[N], ld = ldp[N];
  const int64_t n = (int64_t)naux * bsz[N] * ld, o = off[N];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ld;
    const int c = (int)(i - r * ld);
    if (c >= ce && c < K) continue;
    tilde_store(c < ce ? T[r * ce + c] : 0.0, o + i, out64, out32, hi, lo);
  }
}

__global__ void k_tilde_mirror(const double* __restrict__ T, int N, int ce, const int4* __restrict__ seg,
                               const int* __restrict__ bsz, int naux, const int* __restrict__ ldp,
                               const int64_t* __restrict__ off, double* __restrict__ out64,
                               float* __restrict__ out32, uint16_t* __restrict__ hi, uint16_t* __restrict__ lo)
{
  const int4 sg = seg[blockIdx.y];   // (M, column of block M in slab N, column q of block N in slab M, unused)
  const int M = sg.x, bM = bsz[M], bN = bsz[N], ldM = ldp[M];
  const int64_t n = (int64_t)naux * bM * bN, oM = off[M];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = i / bN;
    const int nu = (int)(i - t * bN);
    const int64_t x = t / bM;
    const int mu = (int)(t - x * bM);
    const double v = T[(x * bN + nu) * ce + sg.y + mu];
    tilde_store(v, oM + (x * bM + mu) * ldM + sg.z + nu, out64, out32, hi, lo);
  }
}

// One slab N whose FP64 values are at src (same row layout as in data64).
__global__ void k_split_slab(const double* __restrict__ src, int N, const int* __restrict__ ncol,
                             const int* __restrict__ bsz, int naux, const int* __restrict__ ldp,
                             const int64_t* __restrict__ off16, uint16_t* __restrict__ hi,
                             uint16_t* __restrict__ lo)
{
  const int K = ncol[N], ld = ldp[N];
  const int64_t n = (int64_t)naux * bsz[N] * ld;
  uint16_t* h = hi + off16[N];
  uint16_t* l = lo + off16[N];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ld;
    const int c = (int)(i - r * ld);
    uint16_t a = 0, b = 0;
    if (c < K) split64(src[r * K + c], a, b);
    h[i] = a;
    l[i] = b;
  }
}

__global__ void k_gather_ct(const double* __restrict__ C, int ldc, int ng, const int* __restrict__ ncol,
                            const int* __restrict__ ldp, const int64_t* __restrict__ ldp_pre,
                            const int* __restrict__ colmap, const int* __restrict__ colmap_off,
                            uint16_t* __restrict__ ghi, uint16_t* __restrict__ glo)
{
  const int N = blockIdx.y;
  const int ld = ldp[N];
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)ng * ld) return;
  const int i = (int)(idx / ld), c = (int)(idx - (int64_t)(idx / ld) * ld);
  uint16_t a = 0, b = 0;
  if (c < ncol[N]) split64(C[(int64_t)colmap[colmap_off[N] + c] * ldc + i], a, b);
  const int64_t o = ldp_pre[N] * ng + idx;
  ghi[o] = a;
  glo[o] = b;
}

// One BK = 32 step on LDS tiles: 3 MFMAs per (i, j) sub-tile.
template <int TM, int TN>
__device__ __forceinline__ void bf16x3_step(const uint4* sAh, const uint4* sAl, const uint4* sBh,
                                            const uint4* sBl, int wm0, int wn0, int lane,
                                            f32x4 (&acc)[TM][TN])
{
  const int fr = lane & 15, fk = lane >> 4;
  uint4 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    ah[i] = sAh[(wm0 + 16 * i + fr) * LROW + fk];
    al[i] = sAl[(wm0 + 16 * i + fr) * LROW + fk];
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    bh[j] = sBh[(wn0 + 16 * j + fr) * LROW + fk];
    bl[j] = sBl[(wn0 + 16 * j + fr) * LROW + fk];
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc[i][j] = mfma_bf16(al[i], bh[j], acc[i][j]);
      acc[i][j] = mfma_bf16(ah[i], bl[j], acc[i][j]);
      acc[i][j] = mfma_bf16(ah[i], bh[j], acc[i][j]);
    }
}

// Grouped slab GEMM.  Tile (slab, m0, n0); slab < 0 marks padding of the
// tile table to a multiple of 8.  Rows r = X*bsz[N] + nu of slab N, columns
// n = occupied index i < ng, contraction over the slab's padded columns.
// WM x WN waves, each owning TM x TN MFMA tiles of 16 x 16: the workgroup tile
// is BM = 16*WM*TM rows by BN = 16*WN*TN columns.  B~ (operand A) streams from
// HBM once per column tile, so wide column tiles (all occupied orbitals in one
// tile) keep the kernel off the HBM roof.
// NBUF = 2: double-buffered LDS, one barrier per step.  NBUF = 1: one LDS
// buffer (two barriers per step) so that two workgroups share a CU and one's
// prologue and epilogue overlap the other's MFMA loop.
template <int WM, int WN, int TM, int TN, int NBUF>
__global__ __launch_bounds__(64 * WM * WN, NBUF == 1 ? 4 : 1) void k_slab_gemm_bf16x3(
    const uint16_t* __restrict__ bhi, const uint16_t* __restrict__ blo, const int64_t* __restrict__ off16,
    const int* __restrict__ ldp, const int64_t* __restrict__ ldp_pre, const int* __restrict__ bsz,
    const int* __restrict__ boff, const Tile* __restrict__ tiles, int naux, int nbf,
    const uint16_t* __restrict__ ghi, const uint16_t* __restrict__ glo, int ng, float* __restrict__ out,
    int ldo)
{
  constexpr int NT = 64 * WM * WN;
  constexpr int BM = 16 * WM * TM, BN = 16 * WN * TN;
  constexpr int CA = (BM * 4 + NT - 1) / NT, CB = (BN * 4 + NT - 1) / NT;   // 16-byte chunks per thread
  __shared__ uint4 sAh[NBUF][BM * LROW], sAl[NBUF][BM * LROW], sBh[NBUF][BN * LROW], sBl[NBUF][BN * LROW];
  const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  if (t.slab < 0) return;
  const int N = t.slab;
  const int bn = bsz[N];
  const int ld8 = ldp[N] >> 3;
  const int64_t rows = (int64_t)naux * bn;
  const uint4* Ah = reinterpret_cast<const uint4*>(bhi + off16[N]);
  const uint4* Al = reinterpret_cast<const uint4*>(blo + off16[N]);
  const uint4* Gh = reinterpret_cast<const uint4*>(ghi + ldp_pre[N] * ng);
  const uint4* Gl = reinterpret_cast<const uint4*>(glo + ldp_pre[N] * ng);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WN) * (16 * TM), wn0 = (wave % WN) * (16 * TN);
  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const uint4 zero = {0u, 0u, 0u, 0u};
  uint4 rah[CA], ral[CA], rbh[CB], rbl[CB];
  auto load = [&](int s) {
    const int kc0 = s * 4;
#pragma unroll
    for (int j = 0; j < CA; ++j) {
      const int q = tid + NT * j;
      const int64_t r = (int64_t)t.m0 + (q >> 2);
      const int64_t o = r * ld8 + kc0 + (q & 3);
      const bool ok = q < BM * 4 && r < rows;
      rah[j] = ok ? Ah[o] : zero;
      ral[j] = ok ? Al[o] : zero;
    }
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      const int q = tid + NT * j;
      const int c = t.n0 + (q >> 2);
      const int64_t o = (int64_t)c * ld8 + kc0 + (q & 3);
      const bool ok = q < BN * 4 && c < ng;
      rbh[j] = ok ? Gh[o] : zero;
      rbl[j] = ok ? Gl[o] : zero;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < CA; ++j) {
      const int q = tid + NT * j;
      if (q < BM * 4) {
        sAh[buf][(q >> 2) * LROW + (q & 3)] = rah[j];
        sAl[buf][(q >> 2) * LROW + (q & 3)] = ral[j];
      }
    }
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      const int q = tid + NT * j;
      if (q < BN * 4) {
        sBh[buf][(q >> 2) * LROW + (q & 3)] = rbh[j];
        sBl[buf][(q >> 2) * LROW + (q & 3)] = rbl[j];
      }
    }
  };
  const int nsteps = ld8 >> 2;
  if constexpr (NBUF == 2) {
    if (nsteps > 0) {
      load(0);
      store(0);
    }
    __syncthreads();
    for (int s = 0; s < nsteps; ++s) {
      const int cur = s & 1;
      if (s + 1 < nsteps) load(s + 1);
      bf16x3_step<TM, TN>(sAh[cur], sAl[cur], sBh[cur], sBl[cur], wm0, wn0, lane, acc);
      if (s + 1 < nsteps) store(cur ^ 1);
      __syncthreads();
    }
  }
  else {
    if (nsteps > 0) load(0);
    for (int s = 0; s < nsteps; ++s) {
      if (s > 0) __syncthreads();   // every wave is done reading the previous step
      store(0);
      __syncthreads();
      if (s + 1 < nsteps) load(s + 1);
      bf16x3_step<TM, TN>(sAh[0], sAl[0], sBh[0], sBl[0], wm0, wn0, lane, acc);
    }
  }
  const int fc = lane & 15, fr4 = 4 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t r = (int64_t)t.m0 + wm0 + 16 * i + fr4 + e;
      if (r >= rows) continue;
      const int64_t x = r / bn;
      const int nu = (int)(r - x * bn);
      float* orow = out + (x * nbf + boff[N] + nu) * (int64_t)ldo;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int c = t.n0 + wn0 + 16 * j + fc;
        if (c < ldo) orow[c] = acc[i][j][e];
      }
    }
}

// Segmented SYRK on FP32 W (segment stride sstride, rows of length ldw, a
// multiple of 32; columns past the occupied count are zero).  tiles[] lists
// the upper 128 x 128 tiles; workgroup b handles tile b % ntile of chunk
// b / ntile (segments [chunk*per, chunk*per + per)).  Each wave owns 64 x 64
// of the tile.  The FP32 accumulation runs over one chunk only (the driver
// keeps chunks short: 3 MFMA accumulations per 32 occupied columns per aux
// function) and the chunks are summed in FP64.
template <int BM>
__global__ __launch_bounds__(256) void k_seg_syrk_bf16x3(
    const float* __restrict__ W, int nseg, int64_t sstride, int ldw, int n, const int2* __restrict__ tiles,
    int ntile, int segs_per_chunk, float* __restrict__ partial)
{
  constexpr int TM = BM / 32;
  constexpr int CA = BM * 8 / 256;   // float4 chunks per thread per operand
  __shared__ uint4 sAh[2][BM * LROW], sAl[2][BM * LROW], sBh[2][BM * LROW], sBl[2][BM * LROW];
  const int2 tl = tiles[blockIdx.x % ntile];
  const int chunk = blockIdx.x / ntile;
  const int s0 = chunk * segs_per_chunk;
  const int s1 = min(nseg, s0 + segs_per_chunk);
  const int m0 = tl.x * BM, n0 = tl.y * BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * (BM / 2), wn0 = (wave & 1) * (BM / 2);
  f32x4 acc[TM][TM];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TM; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int ksteps = ldw >> 5;
  const int nsteps = (s1 > s0) ? (s1 - s0) * ksteps : 0;
  const float4 fz = {0.f, 0.f, 0.f, 0.f};
  float4 ra[CA], rb[CA];
  auto load = [&](int q) {
    const int s = s0 + q / ksteps;
    const int k0 = (q % ksteps) * 32;
    const float* Wg = W + (int64_t)s * sstride + k0;
#pragma unroll
    for (int j = 0; j < CA; ++j) {
      const int c = tid + 256 * j;
      const int row = c >> 3, kq = (c & 7) * 4;
      const int m = m0 + row, nn = n0 + row;
      ra[j] = (m < n) ? *reinterpret_cast<const float4*>(Wg + (int64_t)m * ldw + kq) : fz;
      rb[j] = (nn < n) ? *reinterpret_cast<const float4*>(Wg + (int64_t)nn * ldw + kq) : fz;
    }
  };
  auto put = [&](uint4* dh, uint4* dl, int c, const float4& v) {
    uint16_t h0, h1, h2, h3, l0, l1, l2, l3;
    split32(v.x, h0, l0);
    split32(v.y, h1, l1);
    split32(v.z, h2, l2);
    split32(v.w, h3, l3);
    const int row = c >> 3, kq = c & 7;
    uint2* ph = reinterpret_cast<uint2*>(dh + row * LROW) + kq;
    uint2* pl = reinterpret_cast<uint2*>(dl + row * LROW) + kq;
    *ph = uint2{(uint32_t)h0 | ((uint32_t)h1 << 16), (uint32_t)h2 | ((uint32_t)h3 << 16)};
    *pl = uint2{(uint32_t)l0 | ((uint32_t)l1 << 16), (uint32_t)l2 | ((uint32_t)l3 << 16)};
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < CA; ++j) {
      const int c = tid + 256 * j;
      put(sAh[buf], sAl[buf], c, ra[j]);
      put(sBh[buf], sBl[buf], c, rb[j]);
    }
  };
  if (nsteps > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int q = 0; q < nsteps; ++q) {
    const int cur = q & 1;
    if (q + 1 < nsteps) load(q + 1);
    bf16x3_step<TM, TM>(sAh[cur], sAl[cur], sBh[cur], sBl[cur], wm0, wn0, lane, acc);
    if (q + 1 < nsteps) store(cur ^ 1);
    __syncthreads();
  }
  float* P = partial + (int64_t)chunk * n * n;   // the FP32 chunk sum, exactly
  const int fc = lane & 15, fr4 = 4 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TM; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = m0 + wm0 + 16 * i + fr4 + e;
        const int c = n0 + wn0 + 16 * j + fc;
        if (r < n && c < n) P[(int64_t)r * n + c] = acc[i][j][e];
      }
}


template __global__ void k_seg_syrk_bf16x3<128>(const float*, int, int64_t, int, int, const int2*, int,
                                                 int, float*);

// ---------------------------------------------------------------------------
// Direct-to-LDS pipelined variants (global_load_lds_dwordx4, gfx950).
//
// Operand tiles go from HBM straight into LDS, S stages deep, with no
// register staging: a wave-wide load writes 64 x 16 bytes contiguously (lane
// i at +16 i), i.e. 16 rows of 32 bf16.  Lane i of the load for rows
// [16q, 16q + 16) fetches row r = 16q + i/4, 16-byte chunk c = swz(r, i%4),
// so that a row's chunk c sits at position swz(r, c) (an XOR, its own
// inverse): the 16 rows an MFMA fragment read touches (one chunk each) then
// hit distinct bank groups (dl_common.hpp).  Rows past the operand's end are clamped onto its last row
// (finite data whose products are never stored, or stored as zero).
// Every wave issues the same number of loads per stage, so a stage is
// complete once vmcnt drops to (stages still in flight) x (loads per wave).
// ---------------------------------------------------------------------------

// One BK = 32 step on swizzled stage tiles (uint4 units, 4 per row).
template <int TM, int TN>
__device__ __forceinline__ void bf16x3_step_swz(const uint4* Ah, const uint4* Al, const uint4* Bh,
                                                const uint4* Bl, int wm0, int wn0, int lane,
                                                f32x4 (&acc)[TM][TN])
{
  const int fr = lane & 15, fk = lane >> 4;
  uint4 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int r = wm0 + 16 * i + fr;
    ah[i] = Ah[r * 4 + swz(r, fk)];
    al[i] = Al[r * 4 + swz(r, fk)];
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int r = wn0 + 16 * j + fr;
    bh[j] = Bh[r * 4 + swz(r, fk)];
    bl[j] = Bl[r * 4 + swz(r, fk)];
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc[i][j] = mfma_bf16(al[i], bh[j], acc[i][j]);
      acc[i][j] = mfma_bf16(ah[i], bl[j], acc[i][j]);
      acc[i][j] = mfma_bf16(ah[i], bh[j], acc[i][j]);
    }
}

// The same step, software-pipelined: every B fragment is read first, then the
// A fragments of row i + 1 are read while the 3 TN MFMAs of row i run (two
// A fragment registers sets), so that LDS latency overlaps the MFMAs.  The
// order is pinned with scheduling-group barriers (the scheduler otherwise
// reads fragments just before their first MFMA and waits for each).  Each
// accumulator receives its three products in the same order as in
// bf16x3_step_swz, so the results are the same bits.
template <int TM, int TN>
__device__ __forceinline__ void bf16x3_step_pipe(const uint4* Ah, const uint4* Al, const uint4* Bh,
                                                 const uint4* Bl, int wm0, int wn0, int lane,
                                                 f32x4 (&acc)[TM][TN])
{
  const int fr = lane & 15, fk = lane >> 4;
  uint4 bh[TN], bl[TN], ah[2], al[2];
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int r = wn0 + 16 * j + fr;
    bh[j] = Bh[r * 4 + swz(r, fk)];
    bl[j] = Bl[r * 4 + swz(r, fk)];
  }
  {
    const int r = wm0 + fr;
    ah[0] = Ah[r * 4 + swz(r, fk)];
    al[0] = Al[r * 4 + swz(r, fk)];
  }
  __builtin_amdgcn_sched_group_barrier(0x100, 2 * TN + 2, 0);   // DS reads: B fragments, A row 0
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    if (i + 1 < TM) {
      const int r = wm0 + 16 * (i + 1) + fr;
      ah[(i + 1) & 1] = Ah[r * 4 + swz(r, fk)];
      al[(i + 1) & 1] = Al[r * 4 + swz(r, fk)];
      __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);       // DS reads: A row i + 1
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc[i][j] = mfma_bf16(al[i & 1], bh[j], acc[i][j]);
      acc[i][j] = mfma_bf16(ah[i & 1], bl[j], acc[i][j]);
      acc[i][j] = mfma_bf16(ah[i & 1], bh[j], acc[i][j]);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, 3 * TN, 0);    // MFMAs of row i
  }
}

// The software-pipelined step with the next stages' global loads spread over
// its first SPAN rows: issue(u) starts wave-load u (u < NL), and row i < SPAN
// issues loads [NL i / SPAN, NL (i + 1) / SPAN) between its A-fragment reads
// and its MFMAs.
// Issued in one burst at the step boundary, a workgroup's stage loads queue
// in the address unit (64 bytes per clock) while no MFMA runs; spread over
// the rows they are processed under the MFMAs.  Same products, same order.
template <int TM, int TN, int NL, int SPAN = TM, class Issue>
__device__ __forceinline__ void bf16x3_step_pipe_io(const uint4* Ah, const uint4* Al, const uint4* Bh,
                                                    const uint4* Bl, int wm0, int wn0, int lane,
                                                    f32x4 (&acc)[TM][TN], Issue&& issue)
{
  const int fr = lane & 15, fk = lane >> 4;
  uint4 bh[TN], bl[TN], ah[2], al[2];
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int r = wn0 + 16 * j + fr;
    bh[j] = Bh[r * 4 + swz(r, fk)];
    bl[j] = Bl[r * 4 + swz(r, fk)];
  }
  {
    const int r = wm0 + fr;
    ah[0] = Ah[r * 4 + swz(r, fk)];
    al[0] = Al[r * 4 + swz(r, fk)];
  }
  __builtin_amdgcn_sched_group_barrier(0x100, 2 * TN + 2, 0);   // DS reads: B fragments, A row 0
  static_for<TM>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    constexpr int u0 = i < SPAN ? NL * i / SPAN : NL, u1 = i < SPAN ? NL * (i + 1) / SPAN : NL;
    if constexpr (i + 1 < TM) {
      const int r = wm0 + 16 * (i + 1) + fr;
      ah[(i + 1) & 1] = Ah[r * 4 + swz(r, fk)];
      al[(i + 1) & 1] = Al[r * 4 + swz(r, fk)];
      __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);       // DS reads: A row i + 1
    }
    if constexpr (u1 > u0) {
      // global -> LDS loads, one group each: they keep their program order
      // (the caller's closing vmcnt count assumes which loads were issued last)
      static_for<u1 - u0>([&](auto uu) {
        issue(std::integral_constant<int, u0 + decltype(uu)::value>{});
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
      });
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc[i][j] = mfma_bf16(al[i & 1], bh[j], acc[i][j]);
      acc[i][j] = mfma_bf16(ah[i & 1], bl[j], acc[i][j]);
      acc[i][j] = mfma_bf16(ah[i & 1], bh[j], acc[i][j]);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, 3 * TN, 0);    // MFMAs of row i
  });
}

// Issue the loads of one stage: the wave-loads j = wave + NW*u (u < LPW) over
// the list [A hi rows/16][A lo][B hi][B lo].
template <int BM, int BN, int NW>
__device__ __forceinline__ void issue_stage(uint4* st, int wave, int lane, const uint4* Ah, const uint4* Al,
                                            const uint4* Bh, const uint4* Bl, int64_t a0, int64_t alast,
                                            int b0, int blast, int lda8, int ldb8, int kc0)
{
  constexpr int QA = BM / 16, QB = BN / 16, NJ = 2 * QA + 2 * QB;
  static_assert(NJ % NW == 0, "stage loads must split evenly over the waves");
  const int rr = lane >> 2, pos = lane & 3;
#pragma unroll
  for (int u = 0; u < NJ / NW; ++u) {
    const int j = wave + NW * u;
    if (j < 2 * QA) {
      const bool lo = j >= QA;
      const int q = lo ? j - QA : j;
      const int r = 16 * q + rr;
      const int64_t gr = a0 + r < alast ? a0 + r : alast;
      const uint4* src = (lo ? Al : Ah) + gr * lda8 + kc0 + swz(r, pos);
      lds_load16(src, st + (lo ? BM * 4 : 0) + q * 64);
    }
    else {
      const int jb = j - 2 * QA;
      const bool lo = jb >= QB;
      const int q = lo ? jb - QB : jb;
      const int r = 16 * q + rr;
      const int64_t gr = b0 + r < blast ? b0 + r : blast;
      const uint4* src = (lo ? Bl : Bh) + gr * ldb8 + kc0 + swz(r, pos);
      lds_load16(src, st + 8 * BM + (lo ? BN * 4 : 0) + q * 64);
    }
  }
}

// Quad lane exchange: lane i of each group of 4 lanes receives v from lane
// i ^ T (DPP quad_perm, no LDS traffic).
template <int T>
__device__ __forceinline__ uint32_t quad_xor(uint32_t v)
{
  constexpr int perm = (0 ^ T) | ((1 ^ T) << 2) | ((2 ^ T) << 4) | ((3 ^ T) << 6);
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, perm, 0xF, 0xF, false);
}

__device__ __forceinline__ uint32_t sel4(const uint32_t (&v)[4], int k)
{
  return k == 0 ? v[0] : k == 1 ? v[1] : k == 2 ? v[2] : v[3];
}

// W rows (X, boff + nu) for the tile's rows X*bn + nu of the slab, columns
// [n0, n0 + BN) below ldo, zero at and past ng: FP32 (out) or bf16 hi/lo
// (whi, wlo).  Accumulator layout of the 16x16 MFMA: D[4*(lane/16) + e][lane%16].
// bf16 output: the 4 x 4 blocks (rows 4g..4g+3 of lane group g, columns of a
// lane quad) are transposed across the quad with DPP, hi and lo packed into
// one 32-bit word, so that each lane stores 4 consecutive columns of one row
// as one 8-byte store per plane (ldo is a multiple of 32).
template <int TM, int TN>
__device__ __forceinline__ void store_w_tile(const f32x4 (&acc)[TM][TN], const Tile& t, int wm0, int wn0, int lane,
                                             int64_t rows, int bn, int nbf, int boffN, int ng, float* out,
                                             uint16_t* whi, uint16_t* wlo, int ldo)
{
  const int fc = lane & 15, g = lane >> 4;
  if (whi) {
    const int cq = fc & 3, c4 = fc & ~3;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int r = t.m0 + wm0 + 16 * i + 4 * g + cq;   // the row this lane stores after the transpose
      const bool rok = r < rows;
      const int x = r / bn;
      const int nu = r - x * bn;
      const int64_t orow = ((int64_t)x * nbf + boffN + nu) * (int64_t)ldo;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int col = t.n0 + wn0 + 16 * j + fc;   // the column this lane holds before the transpose
        uint32_t p[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          uint16_t h, l;
          split32(col < ng ? acc[i][j][e] : 0.f, h, l);
          p[e] = (uint32_t)h | ((uint32_t)l << 16);
        }
        // o[k] = element (row 4g + cq, column c4 + k): lane cq ^ T holds it in p[cq]
        uint32_t o[4];
        o[0] = o[1] = o[2] = o[3] = 0u;
        const uint32_t r1 = quad_xor<1>(sel4(p, cq ^ 1));
        const uint32_t r2 = quad_xor<2>(sel4(p, cq ^ 2));
        const uint32_t r3 = quad_xor<3>(sel4(p, cq ^ 3));
        const uint32_t own = sel4(p, cq);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int d = k ^ cq;
          o[k] = d == 0 ? own : d == 1 ? r1 : d == 2 ? r2 : r3;
        }
        const int c0 = t.n0 + wn0 + 16 * j + c4;
        if (rok && c0 < ldo) {
          uint2 hv, lv;
          hv.x = (o[0] & 0xFFFFu) | (o[1] << 16);
          hv.y = (o[2] & 0xFFFFu) | (o[3] << 16);
          lv.x = (o[0] >> 16) | (o[1] & 0xFFFF0000u);
          lv.y = (o[2] >> 16) | (o[3] & 0xFFFF0000u);
          *reinterpret_cast<uint2*>(whi + orow + c0) = hv;
          *reinterpret_cast<uint2*>(wlo + orow + c0) = lv;
        }
      }
    }
    return;
  }
  const int fr4 = 4 * g;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t r = (int64_t)t.m0 + wm0 + 16 * i + fr4 + e;
      if (r >= rows) continue;
      const int64_t x = r / bn;
      const int nu = (int)(r - x * bn);
      const int64_t orow = (x * nbf + boffN + nu) * (int64_t)ldo;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int c = t.n0 + wn0 + 16 * j + fc;
        if (c >= ldo) continue;
        out[orow + c] = c < ng ? acc[i][j][e] : 0.f;
      }
    }
}

// Grouped slab GEMM, direct-to-LDS pipelined.  Same operands and tile table
// as k_slab_gemm_bf16x3 (m0 steps of BM).  W columns in [ng, ldo) are
// written as zero.  split_out: W is written as bf16 hi/lo (whi, wlo, same
// layout) instead of FP32 (out).
// Four-wave shapes of at most 20 MFMA tiles per wave fit two workgroups per
// CU (LDS and registers): ask the compiler for two waves per SIMD.
template <int WM, int WN, int TM, int TN>
constexpr int dl_min_blocks()
{
  return (WM * WN == 4 && TM * TN <= 20) ? 2 : 1;
}

template <int WM, int WN, int TM, int TN, int S>
__global__ __launch_bounds__(64 * WM * WN, (dl_min_blocks<WM, WN, TM, TN>())) void k_slab_gemm_bf16x3_dl(
    const uint16_t* __restrict__ bhi, const uint16_t* __restrict__ blo, const int64_t* __restrict__ off16,
    const int* __restrict__ ldp, const int64_t* __restrict__ ldp_pre, const int* __restrict__ bsz,
    const int* __restrict__ boff, const Tile* __restrict__ tiles, int naux, int nbf,
    const uint16_t* __restrict__ ghi, const uint16_t* __restrict__ glo, int ng, float* __restrict__ out,
    uint16_t* __restrict__ whi, uint16_t* __restrict__ wlo, int ldo)
{
  constexpr int NW = WM * WN;
  constexpr int BM = 16 * WM * TM, BN = 16 * WN * TN;
  constexpr int LPW = (2 * BM / 16 + 2 * BN / 16) / NW;
  constexpr int STAGE = (BM + BN) * 8;   // uint4 per stage
  static_assert(S >= 2 && S <= 4, "pipeline depth");
  __shared__ uint4 lds0[STAGE], lds1[STAGE], lds2[STAGE], lds3[STAGE];
  const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  if (t.slab < 0) return;
  const int N = t.slab;
  const int bn = bsz[N];
  const int ld8 = ldp[N] >> 3;
  const int64_t rows = (int64_t)naux * bn;
  const uint4* Ah = reinterpret_cast<const uint4*>(bhi + off16[N]);
  const uint4* Al = reinterpret_cast<const uint4*>(blo + off16[N]);
  const uint4* Gh = reinterpret_cast<const uint4*>(ghi + ldp_pre[N] * ng);
  const uint4* Gl = reinterpret_cast<const uint4*>(glo + ldp_pre[N] * ng);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WN) * (16 * TM), wn0 = (wave % WN) * (16 * TN);
  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int nsteps = ld8 >> 2;
  const int64_t alast = rows - 1;
  const int blast = ng - 1;
  dl_pipeline<S, LPW>(
      nsteps,
      [&](int s, auto k) {
        issue_stage<BM, BN, NW>(pick_stage<S>(k, lds0, lds1, lds2, lds3), wave, lane, Ah, Al, Gh, Gl, t.m0, alast,
                                t.n0, blast, ld8, ld8, s * 4);
      },
      [&](int, auto k) {
        const uint4* st = pick_stage<S>(k, lds0, lds1, lds2, lds3);
        bf16x3_step_swz<TM, TN>(st, st + BM * 4, st + BM * 8, st + BM * 8 + BN * 4, wm0, wn0, lane, acc);
      });
  store_w_tile<TM, TN>(acc, t, wm0, wn0, lane, rows, bn, nbf, boff[N], ng, out, whi, wlo, ldo);
}

// Grouped slab GEMM with separate pipelines for the two operands: the B~ rows
// (streamed once from HBM) are loaded SA - 1 steps ahead of use, the gathered
// coefficients (re-read by every row tile of a slab, served from L2) one.  At
// BN = 320 a stage of both would take 56 KB and allow only double buffering;
// split, the B~ stream gets SA - 1 steps of latency cover in the same LDS.
// Every stage is an LDS array of its own and the step loop is unrolled over a
// whole cycle of stages, so a step waits only for its own stage's loads.
// Same operands, tile table and output as k_slab_gemm_bf16x3_dl (bf16 hi/lo W).
// Zero rows: 256 rows of 64 bytes (hi or lo plane of a row tile's k-step),
// the source of the B~ stage of steps past a tile's last (read, never written).
__device__ uint4 g_zero_rows[1024] = {};

template <int WM, int WN, int TM, int TN, int SA>
__device__ __forceinline__ void slab_bf16x3_dla_tile(
    const Tile& t, const uint16_t* __restrict__ bhi, const uint16_t* __restrict__ blo,
    const int64_t* __restrict__ off16, const int* __restrict__ ldp, const int64_t* __restrict__ ldp_pre,
    const int* __restrict__ bsz, const int* __restrict__ boff, int naux, int nbf, const uint16_t* __restrict__ ghi,
    const uint16_t* __restrict__ glo, int ng, uint16_t* __restrict__ whi, uint16_t* __restrict__ wlo, int ldo)
{
  static_assert(SA >= 2 && SA <= 4, "B~ pipeline depth");
  constexpr int NW = WM * WN;
  constexpr int BM = 16 * WM * TM, BN = 16 * WN * TN;
  constexpr int QA = BM / 16, QB = BN / 16;
  static_assert((2 * QA) % NW == 0 && (2 * QB) % NW == 0, "stage loads must split evenly over the waves");
  constexpr int LA = 2 * QA / NW, LB = 2 * QB / NW;   // wave-loads per wave per stage
  constexpr int SGA = BM * 8, SGB = BN * 8;           // uint4 per stage: hi rows, then lo rows
  constexpr int CYC = SA % 2 == 0 ? SA : 2 * SA;      // steps until both stage cycles repeat
  __shared__ uint4 sa0[SGA], sa1[SGA], sa2[SGA], sa3[SGA];
  __shared__ uint4 sb0[SGB], sb1[SGB];
  auto stage_a = [&](auto k) -> uint4* {
    constexpr int i = decltype(k)::value % SA;
    if constexpr (i == 0) return sa0;
    else if constexpr (i == 1) return sa1;
    else if constexpr (i == 2) return sa2;
    else return sa3;
  };
  auto stage_b = [&](auto k) -> uint4* {
    if constexpr (decltype(k)::value % 2 == 0) return sb0;
    else return sb1;
  };
  const int N = t.slab;
  const int bn = bsz[N];
  const int ld8 = ldp[N] >> 3;
  const int64_t rows = (int64_t)naux * bn;
  const uint4* Ah = reinterpret_cast<const uint4*>(bhi + off16[N]);
  const uint4* Al = reinterpret_cast<const uint4*>(blo + off16[N]);
  const uint4* Gh = reinterpret_cast<const uint4*>(ghi + ldp_pre[N] * ng);
  const uint4* Gl = reinterpret_cast<const uint4*>(glo + ldp_pre[N] * ng);
  // the wave index as a scalar: the operand bases chosen by it stay in SGPRs
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm0 = (wave / WN) * (16 * TM), wn0 = (wave % WN) * (16 * TN);
  const int rr = lane >> 2, pos = lane & 3;
  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int nsteps = ld8 >> 2;
  const int64_t alast = rows - 1;
  const int blast = ng - 1;
  // wave-loads j = wave + NW*u over [hi rows/16][lo rows/16] of one operand
  // per wave-load: a 32-bit row offset (in 16-byte units) from a wave-uniform
  // base (slabs and gathered coefficients stay below 2^32 bytes), so that the
  // loads take the scalar-base form and the row addresses one VGPR each
  uint32_t offa[LA], offb[LB];
#pragma unroll
  for (int u = 0; u < LA; ++u) {
    const int j = wave + NW * u;
    const int q = j >= QA ? j - QA : j;
    const int r = 16 * q + rr;
    const int64_t gr = t.m0 + r < alast ? t.m0 + r : alast;
    offa[u] = (uint32_t)(gr * ld8 + swz(r, pos));
  }
#pragma unroll
  for (int u = 0; u < LB; ++u) {
    const int j = wave + NW * u;
    const int q = j >= QB ? j - QB : j;
    const int r = 16 * q + rr;
    const int gr = t.n0 + r < blast ? t.n0 + r : blast;
    offb[u] = (uint32_t)((int64_t)gr * ld8 + swz(r, pos));
  }
  auto issue_a = [&](int s, uint4* st) {
#pragma unroll
    for (int u = 0; u < LA; ++u) {
      const int j = wave + NW * u;
      const bool lo = j >= QA;
      const int q = lo ? j - QA : j;
      lds_load16((lo ? Al : Ah) + (offa[u] + (uint32_t)(s * 4)), st + (lo ? BM * 4 : 0) + q * 64);
    }
  };
  auto issue_b = [&](int s, uint4* st) {
#pragma unroll
    for (int u = 0; u < LB; ++u) {
      const int j = wave + NW * u;
      const bool lo = j >= QB;
      const int q = lo ? j - QB : j;
      lds_load16((lo ? Gl : Gh) + (offb[u] + (uint32_t)(s * 4)), st + (lo ? BN * 4 : 0) + q * 64);
    }
  };
  // Prologue: A(0..SA-2) and B(0), all landed.  One barrier per step.  Step
  // s: barrier (every wave's loads for step s have landed, every wave is done
  // with step s - 1's stages), then the MFMAs of step s with the loads of
  // B(s+1) and then A(s+SA-1) spread over its rows (into the stages step s - 1
  // used), then vmcnt(LA) (SA >= 3): B(s+1) and everything issued before it,
  // A(s+1) among them, have landed.  A(s+1) was issued one step earlier than
  // B(s+1): the B~ stream (HBM) gets about SA - 1 steps of latency cover, the
  // coefficients (L2) about one.  SA == 2: A(s+1) is issued after B(s+1) in
  // step s and must land with it (vmcnt(0)).
  // No step has a branch around its loads or MFMAs: the compiler's count of
  // the direct-to-LDS writes then stays exact (after a branch around a load
  // it waits for all of them, vmcnt(0), before the next LDS read).  The loop
  // runs whole stage cycles; the steps past the last (fewer than a cycle)
  // load their B~ stage from zero rows, so their products are exact zeros
  // and the sums are unchanged, and their coefficient stage from the last
  // step's columns (in bounds).
  static_assert(BM <= 256, "zero rows cover 256 rows");
  constexpr int WAIT = SA >= 3 ? LA : 0;
  uint32_t offz[LA];
#pragma unroll
  for (int u = 0; u < LA; ++u) {
    const int j = wave + NW * u;
    const int q = j >= QA ? j - QA : j;
    const int r = 16 * q + rr;
    offz[u] = (uint32_t)(r * 4 + swz(r, pos));
  }
  if (nsteps > 0) {
    const uint4* zr = reinterpret_cast<const uint4*>(g_zero_rows);
    static_for<SA - 1>([&](auto k) {   // A(k), zero rows for a step past the last
      uint4* st = stage_a(k);
      const bool z = k.value >= nsteps;
#pragma unroll
      for (int u = 0; u < LA; ++u) {
        const int j = wave + NW * u;
        const bool lo = j >= QA;
        const int q = lo ? j - QA : j;
        lds_load16(z ? zr + offz[u] : (lo ? Al : Ah) + (offa[u] + (uint32_t)(k.value * 4)),
                   st + (lo ? BM * 4 : 0) + q * 64);
      }
    });
    issue_b(0, stage_b(std::integral_constant<int, 0>{}));
    wait_vm<0>();
    // step s: MFMAs on stages (sa, sb), loads of B(s + 1) into nb and of
    // A(s + SA - 1) into na
    auto step = [&](uint4* nb, uint4* na, const uint4* sa, const uint4* sb, int s) __attribute__((always_inline)) {
        const int sb1 = s + 1 < nsteps ? s + 1 : nsteps - 1;
        const int sa1 = s + SA - 1;
        const bool za = sa1 >= nsteps;   // uniform: zero rows for a step past the last
        auto issue = [&](auto uu) {
          constexpr int u = decltype(uu)::value;
          if constexpr (u < LB) {
            const int j = wave + NW * u;
            const bool lo = j >= QB;
            const int q = lo ? j - QB : j;
            lds_load16((lo ? Gl : Gh) + (offb[u] + (uint32_t)(sb1 * 4)), nb + (lo ? BN * 4 : 0) + q * 64);
          }
          else {
            constexpr int v = u - LB;
            const int j = wave + NW * v;
            const bool lo = j >= QA;
            const int q = lo ? j - QA : j;
            const uint4* src = za ? zr + offz[v] : (lo ? Al : Ah) + (offa[v] + (uint32_t)(sa1 * 4));
            lds_load16(src, na + (lo ? BM * 4 : 0) + q * 64);
          }
        };
        block_sync();
        bf16x3_step_pipe_io<TM, TN, LA + LB>(sa, sa + BM * 4, sb, sb + BN * 4, wm0, wn0, lane, acc, issue);
        wait_vm<WAIT>();
    };
    // whole stage cycles with no branch (stages fixed at compile time), then
    // the last nsteps mod CYC steps in a loop with the stages chosen at run
    // time (the compiler may wait for every load there; the steps past the
    // last are not run at all)
    const int nfull = nsteps - nsteps % CYC;
    for (int s0 = 0; s0 < nfull; s0 += CYC)
      static_for<CYC>([&](auto k) {
        step(stage_b(std::integral_constant<int, k.value + 1>{}),
             stage_a(std::integral_constant<int, k.value + SA - 1>{}), stage_a(k), stage_b(k), s0 + k.value);
      });
    auto sas = [&](int i) -> uint4* {
      return i == 0   ? stage_a(std::integral_constant<int, 0>{})
             : i == 1 ? stage_a(std::integral_constant<int, 1>{})
             : i == 2 ? stage_a(std::integral_constant<int, 2>{})
                      : stage_a(std::integral_constant<int, 3>{});
    };
    for (int s = nfull; s < nsteps; ++s) {
      const int k = s - nfull;
      step((k & 1) ? sb0 : sb1, sas((k + SA - 1) % SA), sas(k % SA), (k & 1) ? sb1 : sb0, s);
    }
    wait_vm<0>();   // no load into this workgroup's LDS outlives it
  }
  store_w_tile<TM, TN>(acc, t, wm0, wn0, lane, rows, bn, nbf, boff[N], ng, nullptr, whi, wlo, ldo);
}

template <int WM, int WN, int TM, int TN, int SA>
__global__ __launch_bounds__(64 * WM * WN, (dl_min_blocks<WM, WN, TM, TN>())) void k_slab_gemm_bf16x3_dla(
    const uint16_t* __restrict__ bhi, const uint16_t* __restrict__ blo, const int64_t* __restrict__ off16,
    const int* __restrict__ ldp, const int64_t* __restrict__ ldp_pre, const int* __restrict__ bsz,
    const int* __restrict__ boff, const Tile* __restrict__ tiles, int naux, int nbf,
    const uint16_t* __restrict__ ghi, const uint16_t* __restrict__ glo, int ng, uint16_t* __restrict__ whi,
    uint16_t* __restrict__ wlo, int ldo)
{
  const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  if (t.slab < 0) return;
  slab_bf16x3_dla_tile<WM, WN, TM, TN, SA>(t, bhi, blo, off16, ldp, ldp_pre, bsz, boff, naux, nbf, ghi, glo, ng,
                                           whi, wlo, ldo);
}

// The same slab GEMM over the tiles of several exchange builds at once (one
// launch for the live requests of the request loop): tile.pad is the build's
// index in reqs, whose operands and W it reads and writes.
template <int WM, int WN, int TM, int TN, int SA>
__global__ __launch_bounds__(64 * WM * WN, (dl_min_blocks<WM, WN, TM, TN>())) void k_slab_gemm_bf16x3_dla_group(
    const ExGroupReq* __restrict__ reqs, const Tile* __restrict__ tiles)
{
  const Tile t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  if (t.slab < 0) return;
  const ExGroupReq& r = reqs[__builtin_amdgcn_readfirstlane(t.pad)];   // uniform: scalar loads
  slab_bf16x3_dla_tile<WM, WN, TM, TN, SA>(t, r.bhi, r.blo, r.off16, r.ldp, r.ldp_pre, r.bsz, r.boff, r.naux, r.nbf,
                                           r.ghi, r.glo, r.ng, r.whi, r.wlo, r.ldw);
}

// Segmented SYRK over bf16 hi/lo W, direct-to-LDS pipelined: partial[chunk](m,n)
// = sum_{X in chunk} sum_i W(X,m,i) W(X,n,i) on the upper BM x BM tiles.
// NW waves in two rows of BM/2 rows each: BM = 128: 4 waves of 64 x 64;
// BM = 256: 8 waves of 128 x 64 (each W row block is fetched by half as many
// tiles).  (4 waves of 128 x 128 at BM = 256, one wave per SIMD with half the
// LDS fragment reads per MFMA, measured the same: the step is bound by the
// operand fetches into LDS.)  Rows of W have ldw (multiple of 32) bf16.
// One output tile tl of one aux chunk of the SYRK.
template <int S, int BM, int NW, bool DIAG1 = true>
__device__ __forceinline__ void syrk_bf16x3_dl_tile(const uint16_t* __restrict__ whi,
                                                    const uint16_t* __restrict__ wlo, int nseg, int64_t sstride,
                                                    int ldw, int n, int chunk, int2 tl, int segs_per_chunk,
                                                    float* __restrict__ partial)
{
  static_assert(BM == 128 || BM == 256, "SYRK tile");
  static_assert(NW == 4 || NW == BM / 32, "SYRK wave layout");
  constexpr int WN = NW / 2;
  constexpr int TM = BM / 32, TN = BM / WN / 16;
  constexpr int LPW = (4 * BM / 16) / NW;
  constexpr int STAGE = 2 * BM * 8;
  static_assert(S >= 2 && S <= 4, "pipeline depth");
  __shared__ uint4 lds0[STAGE], lds1[STAGE], lds2[STAGE], lds3[STAGE];
  const int s0 = chunk * segs_per_chunk;
  if (s0 >= nseg) return;
  const int s1 = min(nseg, s0 + segs_per_chunk);
  const int m0 = tl.x * BM, n0 = tl.y * BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave / WN) * (BM / 2), wn0 = (wave % WN) * (16 * TN);
  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int ksteps = ldw >> 5;
  const int nsteps = (s1 > s0) ? (s1 - s0) * ksteps : 0;
  const int ld8 = ldw >> 3;
  const uint4* Wh = reinterpret_cast<const uint4*>(whi);
  const uint4* Wl = reinterpret_cast<const uint4*>(wlo);
  // a diagonal tile (m0 == n0) has the same W rows on both sides: its stages
  // hold them once, with half the loads per wave (the waits count them)
  // One barrier per step.  Step q: barrier (every wave's loads for step q have
  // landed, every wave is done with step q - 1's stage), then the MFMAs of
  // step q with the loads of step q + S - 1 (into the stage step q - 1 used)
  // spread over the first half of its rows, then vmcnt((S - 2) LP): step q + 1
  // has landed.  No step branches around its loads (the compiler's count of
  // direct-to-LDS writes stays exact): whole stage cycles run, and the steps
  // past the last load their m-side rows from zero rows (exact zero products)
  // and their n-side rows from the last step (in bounds).
  const uint4* zr = reinterpret_cast<const uint4*>(g_zero_rows);
  const int uwave = __builtin_amdgcn_readfirstlane(wave);
  auto run = [&](auto diag) {
    constexpr bool D = decltype(diag)::value;
    constexpr int QA = BM / 16, QB = D ? 0 : BM / 16;
    static_assert((2 * QA + 2 * QB) % NW == 0, "stage loads must split evenly over the waves");
    constexpr int LP = (2 * QA + 2 * QB) / NW;   // wave-loads per wave per stage
    const int rr = lane >> 2, pos = lane & 3;
    uint32_t roff[LP], zoff[LP];
#pragma unroll
    for (int u = 0; u < LP; ++u) {
      const int j = uwave + NW * u;
      const bool sa = j < 2 * QA;
      const int jj = sa ? j : j - 2 * QA;
      const int q = jj % (sa ? QA : QB > 0 ? QB : 1);
      const int r = 16 * q + rr;
      const int g = (sa ? m0 : n0) + r;
      roff[u] = (uint32_t)((int64_t)(g < n - 1 ? g : n - 1) * ld8 + swz(r, pos));
      zoff[u] = (uint32_t)(r * 4 + swz(r, pos));
    }
    // wave-load u of step qs into stage st
    auto issue_u = [&](int qs, uint4* st, int u) {
      const int j = uwave + NW * u;
      const bool sa = j < 2 * QA;
      const int jj = sa ? j : j - 2 * QA;
      const int qa = sa ? QA : QB;
      const bool lo = jj >= qa;
      const int q = lo ? jj - qa : jj;
      const bool z = sa && qs >= nsteps;
      const int qc = qs < nsteps ? qs : nsteps - 1;
      const int sg = s0 + qc / ksteps;
      const int64_t base = (int64_t)sg * (sstride >> 3) + (qc % ksteps) * 4;
      const uint4* src = z ? zr + zoff[u] : (lo ? Wl : Wh) + (base + roff[u]);
      lds_load16(src, st + (sa ? 0 : BM * 8) + (lo ? BM * 4 : 0) + q * 64);
    };
    if (nsteps > 0) {
      static_for<S - 1>([&](auto k) {
        uint4* st = pick_stage<S>(k, lds0, lds1, lds2, lds3);
#pragma unroll
        for (int u = 0; u < LP; ++u) issue_u(k.value, st, u);
      });
      wait_vm<0>();
      for (int q0 = 0; q0 < nsteps; q0 += S) {
        static_for<S>([&](auto k) {
          const int q = q0 + k.value;
          uint4* ns = pick_stage<S>(std::integral_constant<int, k.value + S - 1>{}, lds0, lds1, lds2, lds3);
          auto issue = [&](auto uu) { issue_u(q + S - 1, ns, decltype(uu)::value); };
          block_sync();
          const uint4* st = pick_stage<S>(k, lds0, lds1, lds2, lds3);
          const uint4* sb = D ? st : st + BM * 8;
          bf16x3_step_pipe_io<TM, TN, LP, (TM + 1) / 2>(st, st + BM * 4, sb, sb + BM * 4, wm0, wn0, lane, acc,
                                                         issue);
          wait_vm<(S - 2) * LP>();
        });
      }
      wait_vm<0>();   // no load into this workgroup's LDS outlives it
    }
  };
  // DIAG1 = false: diagonal tiles load both sides like the others, so that
  // every tile of a chunk takes the same time (its tiles share W rows in L2
  // only while they stay in step)
  if (DIAG1 && m0 == n0)
    run(std::true_type{});
  else
    run(std::false_type{});
  float* P = partial + (int64_t)chunk * n * n;   // the FP32 chunk sum, exactly
  const int fc = lane & 15, fr4 = 4 * (lane >> 4);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = m0 + wm0 + 16 * i + fr4 + e;
        const int c = n0 + wn0 + 16 * j + fc;
        if (r < n && c < n) P[(int64_t)r * n + c] = acc[i][j][e];
      }
}

template <int S, int BM, int NW, bool DIAG1>
__global__ __launch_bounds__(64 * NW) void k_seg_syrk_bf16x3_dl(
    const uint16_t* __restrict__ whi, const uint16_t* __restrict__ wlo, int nseg, int64_t sstride, int ldw,
    int n, const int2* __restrict__ tiles, int ntile, int segs_per_chunk, float* __restrict__ partial)
{
  // all tiles of a chunk run on one XCD, dispatched together: the W rows of
  // the chunk come from HBM about once per XCD and are re-read from its L2
  // (grid padded to a multiple of 8; padding blocks exit)
  const int lid = xcd_tile(blockIdx.x, gridDim.x);
  syrk_bf16x3_dl_tile<S, BM, NW, DIAG1>(whi, wlo, nseg, sstride, ldw, n, lid / ntile, tiles[lid % ntile],
                                        segs_per_chunk, partial);
}

// The same SYRK over the work items of several exchange builds at once: item
// (request, chunk, tile row, tile column), request < 0 for padding.
template <int S, int BM, int NW>
__global__ __launch_bounds__(64 * NW) void k_seg_syrk_bf16x3_dl_group(const ExGroupReq* __restrict__ reqs,
                                                                      const int4* __restrict__ items)
{
  const int4 it = items[xcd_tile(blockIdx.x, gridDim.x)];
  if (it.x < 0) return;
  const ExGroupReq& r = reqs[__builtin_amdgcn_readfirstlane(it.x)];   // uniform: scalar loads
  syrk_bf16x3_dl_tile<S, BM, NW>(r.whi, r.wlo, r.naux, (int64_t)r.nbf * r.ldw, r.ldw, r.nbf, it.y, int2{it.z, it.w},
                                 r.spc, r.partial);
}

#define MCHIP_SLAB16DL_INST(WM, WN, TM, TN, S)                                                         \
  template __global__ void k_slab_gemm_bf16x3_dl<WM, WN, TM, TN, S>(                                  \
      const uint16_t*, const uint16_t*, const int64_t*, const int*, const int64_t*, const int*,         \
      const int*, const Tile*, int, int, const uint16_t*, const uint16_t*, int, float*, uint16_t*,      \
      uint16_t*, int);
MCHIP_SLAB16DL_INST(2, 4, 4, 2, 4)
MCHIP_SLAB16DL_INST(2, 4, 4, 3, 3)
MCHIP_SLAB16DL_INST(2, 4, 4, 5, 2)
MCHIP_SLAB16DL_INST(2, 2, 8, 5, 3)
MCHIP_SLAB16DL_INST(2, 2, 4, 5, 2)
#undef MCHIP_SLAB16DL_INST
#define MCHIP_SLAB16DLA_INST(WM, WN, TM, TN, SA)                                                       \
  template __global__ void k_slab_gemm_bf16x3_dla<WM, WN, TM, TN, SA>(                                \
      const uint16_t*, const uint16_t*, const int64_t*, const int*, const int64_t*, const int*,         \
      const int*, const Tile*, int, int, const uint16_t*, const uint16_t*, int, uint16_t*, uint16_t*, int);
MCHIP_SLAB16DLA_INST(2, 4, 4, 5, 4)
MCHIP_SLAB16DLA_INST(2, 4, 4, 5, 3)
MCHIP_SLAB16DLA_INST(2, 4, 4, 3, 4)
MCHIP_SLAB16DLA_INST(2, 2, 4, 5, 2)
MCHIP_SLAB16DLA_INST(2, 4, 6, 5, 3)
MCHIP_SLAB16DLA_INST(2, 4, 6, 5, 2)
#undef MCHIP_SLAB16DLA_INST
#define MCHIP_SYRK16DL_INST(S, BM, NW)                                                                 \
  template __global__ void k_seg_syrk_bf16x3_dl<S, BM, NW, true>(const uint16_t*, const uint16_t*, int, int64_t,   \
                                                           int, int, const int2*, int, int, float*);
MCHIP_SYRK16DL_INST(2, 128, 4)
MCHIP_SYRK16DL_INST(3, 128, 4)
MCHIP_SYRK16DL_INST(4, 128, 4)
MCHIP_SYRK16DL_INST(2, 256, 8)
#undef MCHIP_SYRK16DL_INST
template __global__ void k_seg_syrk_bf16x3_dl<2, 256, 8, false>(const uint16_t*, const uint16_t*, int, int64_t, int,
                                                                int, const int2*, int, int, float*);
template __global__ void k_slab_gemm_bf16x3_dla_group<2, 4, 4, 5, 3>(const ExGroupReq*, const Tile*);
template __global__ void k_slab_gemm_bf16x3_dla_group<2, 2, 4, 5, 2>(const ExGroupReq*, const Tile*);
template __global__ void k_seg_syrk_bf16x3_dl_group<2, 256, 8>(const ExGroupReq*, const int4*);
template __global__ void k_seg_syrk_bf16x3_dl_group<2, 128, 4>(const ExGroupReq*, const int4*);

#define MCHIP_SLAB16_INST(WM, WN, TM, TN, NB)                                                          \
  template __global__ void k_slab_gemm_bf16x3<WM, WN, TM, TN, NB>(                                     \
      const uint16_t*, const uint16_t*, const int64_t*, const int*, const int64_t*, const int*,         \
      const int*, const Tile*, int, int, const uint16_t*, const uint16_t*, int, float*, int);
MCHIP_SLAB16_INST(2, 2, 4, 2, 2)
MCHIP_SLAB16_INST(2, 4, 4, 2, 2)
#undef MCHIP_SLAB16_INST

}  // namespace mchip
