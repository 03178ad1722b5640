// FP64 kernels of the DF Fock build on gfx950.
//
//   k_reduce_sym       FP64 reduction of split-K partials (FP64 or FP32) into a symmetric K
//   k_pair_gather / k_j_gamma / k_j_cols / k_pair_scatter
//                      HBM-streaming DF-J (jkdf.cpp:111-170); every sum of
//                      J (gamma = B.P, V^-1 gamma, J = B^T gamma~) is
//                      compensated (dd.hpp), so J is as accurate as a
//                      double-FP64 evaluation rounded once.
// The MFMA contractions live in kernels_gemm.hip.
#include "dd.hpp"
#include "kernels.hpp"

namespace mchip {

// out[r][c] = out[c][r] = alpha * sum_ch (P[ch][r][c] (+ P[ch][c][r] if sym2)), r <= c
template <class T>
__global__ void k_reduce_sym(const T* __restrict__ partial, int nchunk, int n, int sym2,
                             double alpha, double* __restrict__ out)
{
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int r = (int)(idx / n), c = (int)(idx % n);
  if (r > c) return;
  double v = 0.0;
  for (int ch = 0; ch < nchunk; ++ch) {
    const T* P = partial + (int64_t)ch * n * n;
    v += (double)P[(int64_t)r * n + c];
    if (sym2) v += (double)P[(int64_t)c * n + r];
  }
  v *= alpha;
  out[(int64_t)r * n + c] = v;
  out[(int64_t)c * n + r] = v;
}
template __global__ void k_reduce_sym<double>(const double*, int, int, int, double, double*);

// out[i] = alpha * sum_ch partial[ch][i], i < n2 (fixed order; no symmetrisation)
__global__ void k_reduce_full(const double* __restrict__ partial, int nchunk, int64_t n2, double alpha,
                              double* __restrict__ out)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x) {
    double v = 0.0;
    for (int ch = 0; ch < nchunk; ++ch) v += partial[(int64_t)ch * n2 + i];
    out[i] = alpha * v;
  }
}
template __global__ void k_reduce_sym<float>(const float*, int, int, int, double, double*);

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

// J from the canonical half: J[mu][boff[N]+nu] for the columns c < cdn[N].y,
// and the transpose J[boff[N]+nu][mu] for the columns of blocks M < N
__global__ void k_pair_scatter_sym(const double* __restrict__ Jg, int nbf, const int* __restrict__ ncol,
                                   const int* __restrict__ bsz, const int* __restrict__ boff,
                                   const int2* __restrict__ cdn, const int* __restrict__ colmap,
                                   const int* __restrict__ colmap_off, const int64_t* __restrict__ pg_off,
                                   double* __restrict__ J)
{
  const int N = blockIdx.y;
  const int bn = bsz[N], nc = ncol[N], cd = cdn[N].x, ce = cdn[N].y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= bn * ce) return;
  const int nu = q / ce, c = q - nu * ce;
  const int mu = colmap[colmap_off[N] + c];
  const double v = Jg[pg_off[N] + nu * nc + c];
  J[(int64_t)mu * nbf + boff[N] + nu] = v;
  if (c < cd) J[(int64_t)(boff[N] + nu) * nbf + mu] = v;
}

// gpart[N*naux + x] = sum_e slab_N[x][e] * Pg_N[e] as a compensated pair (sum
// in gpart, rounding errors in gpart_c); one wave per row, J_ROWS rows per block
__global__ __launch_bounds__(256) void k_j_gamma(
    const double* __restrict__ bdata, const int64_t* __restrict__ slab_off,
    const int* __restrict__ ncol, const int* __restrict__ bsz, const int64_t* __restrict__ pg_off,
    const double* __restrict__ Pg, int naux, double* __restrict__ gpart, double* __restrict__ gpart_c)
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
    dsum s;
    for (int e = lane; e < R; e += 64) dot2_add(s, row[e], pg[e]);
    dsum_wave(s);
    if (lane == 0) {
      gpart[(int64_t)N * naux + x] = s.s;
      gpart_c[(int64_t)N * naux + x] = s.c;
    }
  }
}

// The same gamma sums over the canonical half of the slabs: B(X|mu nu) =
// B(X|nu mu) and P symmetric, so slab N contributes its columns of blocks
// M < N twice and of its own block M = N once (cdn[N] = (first column of
// block N, first column past it); columns are in block order), and the
// columns of blocks M > N, the transposes of other slabs' entries, are not
// read.  One wave per aux row, J_ROWS rows per block.
__global__ __launch_bounds__(256) void k_j_gamma_sym(
    const double* __restrict__ bdata, const int64_t* __restrict__ slab_off, const int* __restrict__ ncol,
    const int* __restrict__ bsz, const int2* __restrict__ cdn, const int64_t* __restrict__ pg_off,
    const double* __restrict__ Pg, int naux, double* __restrict__ gpart, double* __restrict__ gpart_c)
{
  const int N = blockIdx.y;
  const int bn = bsz[N], nc = ncol[N], cd = cdn[N].x, ce = cdn[N].y;
  const double* slab = bdata + slab_off[N];
  const double* pg = Pg + pg_off[N];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int rr = wave; rr < J_ROWS; rr += 4) {
    const int x = blockIdx.x * J_ROWS + rr;
    if (x >= naux) break;
    dsum s;
    for (int nu = 0; nu < bn; ++nu) {
      const double* row = slab + ((int64_t)x * bn + nu) * nc;
      const double* pr = pg + (int64_t)nu * nc;
      for (int c = lane; c < ce; c += 64) dot2_add(s, c < cd ? 2.0 * row[c] : row[c], pr[c]);
    }
    dsum_wave(s);
    if (lane == 0) {
      gpart[(int64_t)N * naux + x] = s.s;
      gpart_c[(int64_t)N * naux + x] = s.c;
    }
  }
}

// g[x] = scale * sum_N (gpart + gpart_c)[N*naux + x]  (compensated, fixed order;
// scale is a power of two)
__global__ void k_reduce_slabs(const double* __restrict__ gpart, const double* __restrict__ gpart_c, int nslab,
                               int naux, double scale, double* __restrict__ g)
{
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= naux) return;
  dsum s;
  for (int N = 0; N < nslab; ++N) {
    sum2_add(s, gpart[(int64_t)N * naux + x]);
    s.c += gpart_c[(int64_t)N * naux + x];
  }
  g[x] = scale * dsum_value(s);
}

// y = A x for a symmetric n x n A (row r of A is column r), compensated; one
// wave per row
__global__ __launch_bounds__(256) void k_symv_dot2(const double* __restrict__ A, const double* __restrict__ x, int n,
                                                   double* __restrict__ y)
{
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const double* row = A + (int64_t)r * n;
  dsum s;
  for (int e = lane; e < n; e += 64) dot2_add(s, row[e], x[e]);
  dsum_wave(s);
  if (lane == 0) y[r] = dsum_value(s);
}

// Jpart[chunk][pg_off[N] + e] = sum_{x in chunk} gq[x] * slab_N[x][e] as a
// compensated pair (sums in Jpart, errors in Jpart_c)
__global__ void k_j_cols(const double* __restrict__ bdata, const int64_t* __restrict__ slab_off,
                         const int* __restrict__ ncol, const int* __restrict__ bsz,
                         const int64_t* __restrict__ pg_off, const double* __restrict__ gq,
                         int naux, int xchunk, int64_t pg_total, double* __restrict__ Jpart,
                         double* __restrict__ Jpart_c)
{
  const int N = blockIdx.y;
  const int chunk = blockIdx.z;
  const int R = bsz[N] * ncol[N];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= R) return;
  const double* slab = bdata + slab_off[N];
  const int x0 = chunk * xchunk, x1 = min(naux, x0 + xchunk);
  dsum s;
  for (int x = x0; x < x1; ++x) dot2_add(s, gq[x], slab[(int64_t)x * R + e]);
  Jpart[(int64_t)chunk * pg_total + pg_off[N] + e] = s.s;
  Jpart_c[(int64_t)chunk * pg_total + pg_off[N] + e] = s.c;
}

// k_j_cols over the canonical half: the columns c < cdn[N].y of slab N only
// (thread q = nu * ce + c); the other entries of Jpart are left as they were
__global__ void k_j_cols_sym(const double* __restrict__ bdata, const int64_t* __restrict__ slab_off,
                             const int* __restrict__ ncol, const int* __restrict__ bsz, const int2* __restrict__ cdn,
                             const int64_t* __restrict__ pg_off, const double* __restrict__ gq, int naux,
                             int xchunk, int64_t pg_total, double* __restrict__ Jpart, double* __restrict__ Jpart_c)
{
  const int N = blockIdx.y;
  const int chunk = blockIdx.z;
  const int bn = bsz[N], nc = ncol[N], ce = cdn[N].y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= bn * ce) return;
  const int nu = q / ce, c = q - nu * ce;
  const int e = nu * nc + c;
  const int64_t R = (int64_t)bn * nc;
  const double* slab = bdata + slab_off[N];
  const int x0 = chunk * xchunk, x1 = min(naux, x0 + xchunk);
  dsum s;
  for (int x = x0; x < x1; ++x) dot2_add(s, gq[x], slab[(int64_t)x * R + e]);
  Jpart[(int64_t)chunk * pg_total + pg_off[N] + e] = s.s;
  Jpart_c[(int64_t)chunk * pg_total + pg_off[N] + e] = s.c;
}

// out[i] = sum_chunk (part + part_c)[chunk][i], compensated, fixed order
__global__ void k_reduce_chunks_dd(const double* __restrict__ part, const double* __restrict__ part_c, int nchunk,
                                   int64_t n, double* __restrict__ out)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    dsum s;
    for (int c = 0; c < nchunk; ++c) {
      sum2_add(s, part[(int64_t)c * n + i]);
      s.c += part_c[(int64_t)c * n + i];
    }
    out[i] = dsum_value(s);
  }
}

// Element-wise kernels below are grid-stride loops: launch them with
// gs_blocks(n), which keeps gridDim.x * blockDim.x below 2^32 for any n.
__global__ void k_reduce_chunks(const double* __restrict__ part, int nchunk, int64_t n,
                                double* __restrict__ out)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * n + i];
    out[i] = s;
  }
}

__global__ void k_fill(double* __restrict__ p, int64_t n, double v)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_axpby(double* __restrict__ y, const double* __restrict__ x, int64_t n,
                        double a, double b)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}

}  // namespace mchip
