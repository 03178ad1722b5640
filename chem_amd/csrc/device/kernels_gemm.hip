// Contraction kernels of the DF Fock build (FP64 and FP32 MFMA).
//
//   k_slab_gemm<T,BM,BN>  grouped GEMM over B's pair slabs with gathered rows
//                         of a dense operand: out(X,nu,j) = sum_mu B(X|mu nu) G(mu,j).
//                         XMN,Mi->XiN (jkdf.cpp:361) and mi,Xmn->Xin
//                         (z_full.cpp:202).  One tile table entry per block.
// The exchange SYRK XiM,XiN->MN (jkdf.cpp:417) of every precision is in the
// direct-to-LDS kernel files.
#include "gemm_mfma.hpp"
#include "kernels.hpp"

namespace mchip {

template <class T, int BM, int BN, int FL>
__global__ __launch_bounds__(256) void k_slab_gemm(
    const T* __restrict__ bdata, const int64_t* __restrict__ slab_off, const int* __restrict__ ncol,
    const int* __restrict__ bsz, const int* __restrict__ boff, const int* __restrict__ colmap,
    const int* __restrict__ colmap_off, const Tile* __restrict__ tiles, int naux, int nbf,
    const T* __restrict__ G, int ldg, int ng, T* __restrict__ out, int ldo)
{
  constexpr int TM = BM / 32, TN = BN / 32;
  constexpr int LDA = BM + 16, LDB = BN + 16;
  constexpr int RA = BM / 16, RB = BN / 16;
  __shared__ T As[2][GM_BK * LDA];
  __shared__ T Bs[2][GM_BK * LDB];
  const Tile t = tiles[blockIdx.x];
  const int N = t.slab;
  const int bn = bsz[N];
  const int K = ncol[N];
  const int64_t rows = (int64_t)naux * bn;
  const T* A = bdata + slab_off[N];
  const int* cmap = colmap + colmap_off[N];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * (BM / 2), wn0 = (wave & 1) * (BN / 2);
  typename Mfma<T>::v4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = typename Mfma<T>::v4{0, 0, 0, 0};
  static_assert(FL == 0 || sizeof(T) == 4, "FP64 flushing applies to the FP32 kernel");
  double acc64[TM][TN][4] = {};
  const int ak = tid & 15, ar = tid >> 4;
  const int bk = tid >> 4, bc = tid & 15;
  T ra[RA], rb[RB];
  auto load = [&](int k0) {
    const int kk = k0 + ak;
#pragma unroll
    for (int j = 0; j < RA; ++j) {
      const int64_t r = (int64_t)t.m0 + ar + 16 * j;
      ra[j] = (r < rows && kk < K) ? A[r * K + kk] : T(0);
    }
    const int kb = k0 + bk;
    const int mu = (kb < K) ? cmap[kb] : -1;
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      const int c = t.n0 + bc + 16 * j;
      rb[j] = (mu >= 0 && c < ng) ? G[(int64_t)mu * ldg + c] : T(0);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < RA; ++j) As[buf][ak * LDA + ar + 16 * j] = ra[j];
#pragma unroll
    for (int j = 0; j < RB; ++j) Bs[buf][bk * LDB + bc + 16 * j] = rb[j];
  };
  const int nsteps = (K + GM_BK - 1) / GM_BK;
  if (nsteps > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int s = 0; s < nsteps; ++s) {
    const int cur = s & 1;
    if (s + 1 < nsteps) load((s + 1) * GM_BK);
    mfma_step<T, TM, TN, LDA, LDB>(As[cur], Bs[cur], wm0, wn0, lane, acc);
    if constexpr (FL > 0) flush_f64<TM, TN>(acc, acc64);
    if (s + 1 < nsteps) store(cur ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t r = (int64_t)t.m0 + wm0 + 16 * i + Mfma<T>::row(lane, e);
        const int c = t.n0 + wn0 + 16 * j + (lane & 15);
        if (r < rows && c < ng) {
          const int64_t x = r / bn;
          const int nu = (int)(r - x * bn);
          if constexpr (FL > 0)
            out[(x * nbf + boff[N] + nu) * (int64_t)ldo + c] = (T)acc64[i][j][e];
          else
            out[(x * nbf + boff[N] + nu) * (int64_t)ldo + c] = acc[i][j][e];
        }
      }
}


// explicit instantiations
template __global__ void k_slab_gemm<double, 128, 128>(const double*, const int64_t*, const int*, const int*,
    const int*, const int*, const int*, const Tile*, int, int, const double*, int, int, double*, int);
template __global__ void k_slab_gemm<double, 128, 64>(const double*, const int64_t*, const int*, const int*,
    const int*, const int*, const int*, const Tile*, int, int, const double*, int, int, double*, int);
template __global__ void k_slab_gemm<float, 128, 128>(const float*, const int64_t*, const int*, const int*,
    const int*, const int*, const int*, const Tile*, int, int, const float*, int, int, float*, int);
template __global__ void k_slab_gemm<float, 128, 64>(const float*, const int64_t*, const int*, const int*,
    const int*, const int*, const int*, const Tile*, int, int, const float*, int, int, float*, int);
template __global__ void k_slab_gemm<float, 128, 128, 1>(const float*, const int64_t*, const int*, const int*,
    const int*, const int*, const int*, const Tile*, int, int, const float*, int, int, float*, int);
template __global__ void k_slab_gemm<float, 128, 64, 1>(const float*, const int64_t*, const int*, const int*,
    const int*, const int*, const int*, const Tile*, int, int, const float*, int, int, float*, int);

}  // namespace mchip

namespace mchip {

// out_slab[N](r, c) = sum_k A(row(r), k) G(colmap_N[c], k)
//   rows r = (X, nu in N) read from A at (X*nbf + boff[N] + nu)*lda, k < klen
//   columns c over slab N's significant mu, G row-major with leading dim ldg
// b3(X|M N) = sum_i L(M,i) b2(X,i,N) on B's pair pattern (z_full.cpp:255-291).
template <class T, int BM, int BN>
__global__ __launch_bounds__(256) void k_slab_out_gemm(
    const T* __restrict__ A, int lda, int klen, const int64_t* __restrict__ slab_off,
    const int* __restrict__ ncol, const int* __restrict__ bsz, const int* __restrict__ boff,
    const int* __restrict__ colmap, const int* __restrict__ colmap_off, const Tile* __restrict__ tiles,
    int naux, int nbf, const T* __restrict__ G, int ldg, T* __restrict__ out, const int* __restrict__ pcol,
    const int* __restrict__ pwidth, int64_t out_stride)
{
  // grouped use (pcol != null): tile.pad is the group; its contraction columns
  // are [pcol, pcol + pwidth) of A and G and its output block starts at
  // out + group * out_stride
  if (pcol) {
    const int g = tiles[blockIdx.x].pad;
    A += pcol[g];
    G += pcol[g];
    klen = pwidth[g];
    out += (int64_t)g * out_stride;
  }
  constexpr int TM = BM / 32, TN = BN / 32;
  constexpr int LDA = BM + 16, LDB = BN + 16;
  constexpr int RA = BM / 16, RB = BN / 16;
  __shared__ T As[2][GM_BK * LDA];
  __shared__ T Bs[2][GM_BK * LDB];
  const Tile t = tiles[blockIdx.x];
  const int N = t.slab;
  const int bn = bsz[N];
  const int nc = ncol[N];
  const int64_t rows = (int64_t)naux * bn;
  const int* cmap = colmap + colmap_off[N];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * (BM / 2), wn0 = (wave & 1) * (BN / 2);
  typename Mfma<T>::v4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = typename Mfma<T>::v4{0, 0, 0, 0};
  constexpr bool F64 = sizeof(T) == 4;   // FP32: FP64 running sums, flushed every step
  double acc64[TM][TN][4] = {};
  const int lk = tid & 15, lr = tid >> 4;
  int64_t arow[RA];
  int grow[RB];
#pragma unroll
  for (int j = 0; j < RA; ++j) {
    const int64_t r = (int64_t)t.m0 + lr + 16 * j;
    if (r < rows) {
      const int64_t x = r / bn;
      arow[j] = (x * nbf + boff[N] + (r - x * bn)) * (int64_t)lda;
    }
    else {
      arow[j] = -1;
    }
  }
#pragma unroll
  for (int j = 0; j < RB; ++j) {
    const int c = t.n0 + lr + 16 * j;
    grow[j] = (c < nc) ? cmap[c] : -1;
  }
  T ra[RA], rb[RB];
  auto load = [&](int k0) {
    const int kk = k0 + lk;
#pragma unroll
    for (int j = 0; j < RA; ++j) ra[j] = (arow[j] >= 0 && kk < klen) ? A[arow[j] + kk] : T(0);
#pragma unroll
    for (int j = 0; j < RB; ++j) rb[j] = (grow[j] >= 0 && kk < klen) ? G[(int64_t)grow[j] * ldg + kk] : T(0);
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < RA; ++j) As[buf][lk * LDA + lr + 16 * j] = ra[j];
#pragma unroll
    for (int j = 0; j < RB; ++j) Bs[buf][lk * LDB + lr + 16 * j] = rb[j];
  };
  const int nsteps = (klen + GM_BK - 1) / GM_BK;
  if (nsteps > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int s = 0; s < nsteps; ++s) {
    const int cur = s & 1;
    if (s + 1 < nsteps) load((s + 1) * GM_BK);
    mfma_step<T, TM, TN, LDA, LDB>(As[cur], Bs[cur], wm0, wn0, lane, acc);
    if constexpr (F64) flush_f64<TM, TN>(acc, acc64);
    if (s + 1 < nsteps) store(cur ^ 1);
    __syncthreads();
  }
  T* o = out + slab_off[N];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t r = (int64_t)t.m0 + wm0 + 16 * i + Mfma<T>::row(lane, e);
        const int c = t.n0 + wn0 + 16 * j + (lane & 15);
        if (r < rows && c < nc) {
          if constexpr (F64)
            o[r * nc + c] = (T)acc64[i][j][e];
          else
            o[r * nc + c] = acc[i][j][e];
        }
      }
}

// partial[chunk](m,n) = sum_{slabs s in chunk} sum_{k < len_s} A(off_s + m*len_s + k) B(off_s + n*len_s + k)
// Z_XY = sum_{mu nu} b3(X|mu nu) B(Y|mu nu) over the pair slabs (z_full.cpp:373-377).
template <class T, int BM, int BN>
__global__ __launch_bounds__(256) void k_slabseg_gemm(
    const T* __restrict__ A, const T* __restrict__ B, const int64_t* __restrict__ seg_off,
    const int64_t* __restrict__ seg_len, const int* __restrict__ chunk_seg, int n,
    const int2* __restrict__ tiles, int ntile, double* __restrict__ partial)
{
  constexpr int TM = BM / 32, TN = BN / 32;
  constexpr int LDA = BM + 16, LDB = BN + 16;
  constexpr int RA = BM / 16, RB = BN / 16;
  __shared__ T As[2][GM_BK * LDA];
  __shared__ T Bs[2][GM_BK * LDB];
  const int2 tl = tiles[blockIdx.x % ntile];
  const int chunk = blockIdx.x / ntile;
  const int sb = chunk_seg[chunk], se = chunk_seg[chunk + 1];
  const int m0 = tl.x * BM, n0 = tl.y * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * (BM / 2), wn0 = (wave & 1) * (BN / 2);
  typename Mfma<T>::v4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = typename Mfma<T>::v4{0, 0, 0, 0};
  const int lk = tid & 15, lr = tid >> 4;
  T ra[RA], rb[RB];
  int seg = sb;
  int64_t k0 = 0;
  // skip empty segments
  while (seg < se && seg_len[seg] == 0) ++seg;
  auto load = [&](int sg, int64_t kb) {
    const int64_t len = seg_len[sg];
    const int64_t kk = kb + lk;
    const T* Ab = A + seg_off[sg];
    const T* Bb = B + seg_off[sg];
#pragma unroll
    for (int j = 0; j < RA; ++j) {
      const int m = m0 + lr + 16 * j;
      ra[j] = (m < n && kk < len) ? Ab[(int64_t)m * len + kk] : T(0);
    }
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      const int nn = n0 + lr + 16 * j;
      rb[j] = (nn < n && kk < len) ? Bb[(int64_t)nn * len + kk] : T(0);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < RA; ++j) As[buf][lk * LDA + lr + 16 * j] = ra[j];
#pragma unroll
    for (int j = 0; j < RB; ++j) Bs[buf][lk * LDB + lr + 16 * j] = rb[j];
  };
  auto advance = [&](int& sg, int64_t& kb) {
    kb += GM_BK;
    if (kb >= seg_len[sg]) {
      kb = 0;
      ++sg;
      while (sg < se && seg_len[sg] == 0) ++sg;
    }
  };
  bool have = seg < se;
  if (have) {
    load(seg, k0);
    store(0);
  }
  __syncthreads();
  int cur = 0;
  while (have) {
    int nseg = seg;
    int64_t nk = k0;
    advance(nseg, nk);
    const bool more = nseg < se;
    if (more) load(nseg, nk);
    mfma_step<T, TM, TN, LDA, LDB>(As[cur], Bs[cur], wm0, wn0, lane, acc);
    if (more) store(cur ^ 1);
    __syncthreads();
    cur ^= 1;
    seg = nseg;
    k0 = nk;
    have = more;
  }
  double* P = partial + (int64_t)chunk * n * n;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = m0 + wm0 + 16 * i + Mfma<T>::row(lane, e);
        const int c = n0 + wn0 + 16 * j + (lane & 15);
        if (r < n && c < n) P[(int64_t)r * n + c] = (double)acc[i][j][e];
      }
}

// Grouped slab-segment SYRK in FP32 with FP64-flushed accumulators (8 waves,
// 64 x 32 each, flush every 4 steps = 64 products):
//   partial[g][chunk](m,n) = sum_{slabs s in chunk} sum_k A_g(off_s + m*len_s + k) B(off_s + n*len_s + k)
// with A_g = A + g * a_stride.  Z_g(X,Y) = sum b3_g(X|mu nu) B(Y|mu nu) for all
// Laplace points g in one launch (z_full.cpp:373-377).
__global__ __launch_bounds__(512) void k_slabseg_f32(const float* __restrict__ A, int64_t a_stride,
                                                     const float* __restrict__ B, const int64_t* __restrict__ seg_off,
                                                     const int64_t* __restrict__ seg_len,
                                                     const int* __restrict__ chunk_seg, int nchunk, int n,
                                                     const int2* __restrict__ tiles, int ntile,
                                                     double* __restrict__ partial)
{
  constexpr int BM = 128, TM = 4, TN = 2, LDA = BM + 16, RA = BM / 32, FLUSH = 4;
  __shared__ float As[2][GM_BK * LDA];
  __shared__ float Bs[2][GM_BK * LDA];
  const int2 tl = tiles[blockIdx.x % ntile];
  const int rest = blockIdx.x / ntile;
  const int chunk = rest % nchunk, g = rest / nchunk;
  const float* Ag = A + (int64_t)g * a_stride;
  const int sb = chunk_seg[chunk], se = chunk_seg[chunk + 1];
  const int m0 = tl.x * BM, n0 = tl.y * BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 2) * 64, wn0 = (wave & 3) * 32;
  Mfma<float>::v4 acc[TM][TN];
  double acc64[TM][TN][4];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc[i][j] = Mfma<float>::v4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int e = 0; e < 4; ++e) acc64[i][j][e] = 0.0;
    }
  const int lk = tid & 15, lr = tid >> 4;
  float ra[RA], rb[RA];
  auto load = [&](int sg, int64_t kb) {
    const int64_t len = seg_len[sg];
    const int64_t kk = kb + lk;
    const float* Ab = Ag + seg_off[sg];
    const float* Bb = B + seg_off[sg];
#pragma unroll
    for (int j = 0; j < RA; ++j) {
      const int m = m0 + lr + 32 * j, nn = n0 + lr + 32 * j;
      ra[j] = (m < n && kk < len) ? Ab[(int64_t)m * len + kk] : 0.f;
      rb[j] = (nn < n && kk < len) ? Bb[(int64_t)nn * len + kk] : 0.f;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < RA; ++j) {
      As[buf][lk * LDA + lr + 32 * j] = ra[j];
      Bs[buf][lk * LDA + lr + 32 * j] = rb[j];
    }
  };
  auto advance = [&](int& sg, int64_t& kb) {
    kb += GM_BK;
    if (kb >= seg_len[sg]) {
      kb = 0;
      ++sg;
      while (sg < se && seg_len[sg] == 0) ++sg;
    }
  };
  auto flush = [&]() {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
#pragma unroll
        for (int e = 0; e < 4; ++e) acc64[i][j][e] += (double)acc[i][j][e];
        acc[i][j] = Mfma<float>::v4{0.f, 0.f, 0.f, 0.f};
      }
  };
  int seg = sb;
  int64_t k0 = 0;
  while (seg < se && seg_len[seg] == 0) ++seg;
  bool have = seg < se;
  if (have) {
    load(seg, k0);
    store(0);
  }
  __syncthreads();
  int cur = 0, step = 0;
  while (have) {
    int nseg = seg;
    int64_t nk = k0;
    advance(nseg, nk);
    const bool more = nseg < se;
    if (more) load(nseg, nk);
    mfma_step<float, TM, TN, LDA, LDA>(As[cur], Bs[cur], wm0, wn0, lane, acc);
    if (more) store(cur ^ 1);
    if (++step % FLUSH == 0) flush();
    __syncthreads();
    cur ^= 1;
    seg = nseg;
    k0 = nk;
    have = more;
  }
  flush();
  double* Pp = partial + ((int64_t)g * nchunk + chunk) * n * n;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = m0 + wm0 + 16 * i + Mfma<float>::row(lane, e);
        const int c = n0 + wn0 + 16 * j + (lane & 15);
        if (r < n && c < n) Pp[(int64_t)r * n + c] = acc64[i][j][e];
      }
}


template __global__ void k_slab_out_gemm<double, 128, 128>(const double*, int, int, const int64_t*,
    const int*, const int*, const int*, const int*, const int*, const Tile*, int, int, const double*, int,
    double*, const int*, const int*, int64_t);
template __global__ void k_slab_out_gemm<float, 128, 128>(const float*, int, int, const int64_t*,
    const int*, const int*, const int*, const int*, const int*, const Tile*, int, int, const float*, int,
    float*, const int*, const int*, int64_t);
template __global__ void k_slabseg_gemm<double, 128, 128>(const double*, const double*, const int64_t*,
    const int64_t*, const int*, int, const int2*, int, double*);

}  // namespace mchip
