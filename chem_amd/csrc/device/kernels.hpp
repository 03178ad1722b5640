// Kernel declarations shared by the device drivers.
#pragma once

#include <algorithm>
#include <stdexcept>

#include "dev.hpp"

namespace mchip {

constexpr int SG_BM = 128;    // rows per slab-GEMM / SYRK tile
constexpr int J_ROWS = 16;    // aux rows per block in k_j_gamma

// FL = 1 (FP32 only): FP64 running sums, flushed after every 16-product step.
template <class T, int BM, int BN, int FL = 0>
__global__ void k_slab_gemm(const T* bdata, const int64_t* slab_off, const int* ncol, const int* bsz,
                            const int* boff, const int* colmap, const int* colmap_off,
                            const Tile* tiles, int naux, int nbf, const T* G, int ldg, int ng, T* out,
                            int ldo);
template <class T, int BM, int BN>
__global__ void k_slab_out_gemm(const T* A, int lda, int klen, const int64_t* slab_off, const int* ncol,
                                const int* bsz, const int* boff, const int* colmap,
                                const int* colmap_off, const Tile* tiles, int naux, int nbf,
                                const T* G, int ldg, T* out, const int* pcol, const int* pwidth,
                                int64_t out_stride);
__global__ void k_transpose_pad_f32(const float* in, int n, int R, int ldn, float* out);
template <int S>
__global__ void k_group_apply_f32_dl(const float* P, int npt, const float* bT, float* out, int n, int ldn, int R,
                                     int naux, const int* pcol, const int* pwidth, const int4* tiles);
__global__ void k_to_f32x2_rows(const double* in, int n, int ldn, float* hi, float* lo);
__global__ void k_pad_f32_rows(const float* src, int N, const int* ncol, const int* bsz, int naux, const int* ldp,
                               const int64_t* off, float* out);
__global__ void k_gather_ct_f32f(const float* C, int ldc, int ng, const int* ncol, const int* ldp,
                                 const int64_t* ldp_pre, const int* colmap, const int* colmap_off, float* g);
template <int S>
__global__ void k_slab_out_gemm_f32_dl(const float* A, int ldr, const int64_t* slab_off, const int* ncol,
                                       const int* bsz, const int* boff, const int* colmap, const int* colmap_off,
                                       const Tile* tiles, int naux, int nbf, const float* G, float* out,
                                       const int* pcol, const int* pwidth, int64_t out_stride);
template <int S>
__global__ void k_slab_gemm_f32_dl_f64(const float* data, const int64_t* off, const int* ldp, const int64_t* ldp_pre,
                                       const int* bsz, const int* boff, const Tile* tiles, int naux, int nbf,
                                       const float* g, int ng, float* out, int ldo);
__global__ void k_slabseg_f32(const float* A, int64_t a_stride, const float* B, const int64_t* seg_off,
                              const int64_t* seg_len, const int* chunk_seg, int nchunk, int n,
                              const int2* tiles, int ntile, double* partial);
template <class T, int BM, int BN>
__global__ void k_slabseg_gemm(const T* A, const T* B, const int64_t* seg_off, const int64_t* seg_len,
                               const int* chunk_seg, int n, const int2* tiles, int ntile,
                               double* partial);
__global__ void k_zero_upper_cm(double* A, int n);
template <class T>
__global__ void k_reduce_sym(const T* partial, int nchunk, int n, int sym2, double alpha,
                             double* out);
__global__ void k_pair_gather(const double* P, int nbf, const int* ncol, const int* bsz,
                              const int* boff, const int* colmap, const int* colmap_off,
                              const int64_t* pg_off, double* Pg);
__global__ void k_pair_scatter(const double* Jg, int nbf, const int* ncol, const int* bsz,
                               const int* boff, const int* colmap, const int* colmap_off,
                               const int64_t* pg_off, double* J);
__global__ void k_j_gamma_sym(const double* bdata, const int64_t* slab_off, const int* ncol, const int* bsz,
                              const int2* cdn, const int64_t* pg_off, const double* Pg, int naux, double* gpart,
                              double* gpart_c);
__global__ void k_j_cols_sym(const double* bdata, const int64_t* slab_off, const int* ncol, const int* bsz,
                             const int2* cdn, const int64_t* pg_off, const double* gq, int naux, int xchunk,
                             int64_t pg_total, double* Jpart, double* Jpart_c);
__global__ void k_pair_scatter_sym(const double* Jg, int nbf, const int* ncol, const int* bsz, const int* boff,
                                   const int2* cdn, const int* colmap, const int* colmap_off, const int64_t* pg_off,
                                   double* J);
__global__ void k_j_gamma(const double* bdata, const int64_t* slab_off, const int* ncol,
                          const int* bsz, const int64_t* pg_off, const double* Pg, int naux,
                          double* gpart, double* gpart_c);
__global__ void k_reduce_slabs(const double* gpart, const double* gpart_c, int nslab, int naux, double scale,
                               double* g);
__global__ void k_symv_dot2(const double* A, const double* x, int n, double* y);
__global__ void k_j_cols(const double* bdata, const int64_t* slab_off, const int* ncol,
                         const int* bsz, const int64_t* pg_off, const double* gq, int naux,
                         int xchunk, int64_t pg_total, double* Jpart, double* Jpart_c);
__global__ void k_reduce_chunks(const double* part, int nchunk, int64_t n, double* out);
__global__ void k_reduce_chunks_dd(const double* part, const double* part_c, int nchunk, int64_t n, double* out);
__global__ void k_fill(double* p, int64_t n, double v);
__global__ void k_axpby(double* y, const double* x, int64_t n, double a, double b);

__global__ void k_to_f32(const double* in, float* out, int64_t n);
__global__ void k_to_f32x2(const double* in, float* hi, float* lo, int64_t n);

// BF16 split-precision exchange (kernels_bf16.hip)
__global__ void k_split_slabs(const double* data64, const int64_t* slab_off, const int* ncol,
                              const int* bsz, int naux, const int* ldp, const int64_t* off16,
                              uint16_t* hi, uint16_t* lo);
template <int WM, int WN, int TM, int TN, int S>
__global__ void k_slab_gemm_bf16x3_dl(const uint16_t* bhi, const uint16_t* blo, const int64_t* off16,
                                      const int* ldp, const int64_t* ldp_pre, const int* bsz,
                                      const int* boff, const Tile* tiles, int naux, int nbf,
                                      const uint16_t* ghi, const uint16_t* glo, int ng, float* out,
                                      uint16_t* whi, uint16_t* wlo, int ldo);
template <int WM, int WN, int TM, int TN, int SA>
__global__ void k_slab_gemm_bf16x3_dla(const uint16_t* bhi, const uint16_t* blo, const int64_t* off16,
                                       const int* ldp, const int64_t* ldp_pre, const int* bsz,
                                       const int* boff, const Tile* tiles, int naux, int nbf,
                                       const uint16_t* ghi, const uint16_t* glo, int ng, uint16_t* whi,
                                       uint16_t* wlo, int ldo);
__global__ void k_exchange_energy(const double* W, int ldw, int nbf, int nocc, const double* C, int ldc, int naux,
                                  const int2* tiles, int ntile, double* partial);
__global__ void k_sum_partials(const double* partial, int64_t n, double scale, double* out);
__global__ void k_sumsq_pairs(const double* a, int64_t n, double* partial);
template <int WM, int WN, int TM, int TN, int SA>
__global__ void k_slab_gemm_bf16x3_dla_group(const ExGroupReq* reqs, const Tile* tiles);
template <int S, int BM, int NW>
__global__ void k_seg_syrk_bf16x3_dl_group(const ExGroupReq* reqs, const int4* items);
template <int S, int BM, int NW, bool DIAG1>
__global__ void k_seg_syrk_bf16x3_dl(const uint16_t* whi, const uint16_t* wlo, int nseg, int64_t sstride,
                                     int ldw, int n, const int2* tiles, int ntile, int segs_per_chunk,
                                     float* partial);
__global__ void k_tilde_canon(const double* T, int N, int ce, const int* ncol, const int* bsz, int naux,
                              const int* ldp, const int64_t* off, double* out64, float* out32, uint16_t* hi,
                              uint16_t* lo);
__global__ void k_tilde_mirror(const double* T, int N, int ce, const int4* seg, const int* bsz, int naux,
                               const int* ldp, const int64_t* off, double* out64, float* out32, uint16_t* hi,
                               uint16_t* lo);
__global__ void k_split_slab(const double* src, int N, const int* ncol, const int* bsz, int naux,
                             const int* ldp, const int64_t* off16, uint16_t* hi, uint16_t* lo);
__global__ void k_gather_ct(const double* C, int ldc, int ng, const int* ncol, const int* ldp,
                            const int64_t* ldp_pre, const int* colmap, const int* colmap_off,
                            uint16_t* ghi, uint16_t* glo);
template <int WM, int WN, int TM, int TN, int NBUF>
__global__ void k_slab_gemm_bf16x3(const uint16_t* bhi, const uint16_t* blo, const int64_t* off16,
                                   const int* ldp, const int64_t* ldp_pre, const int* bsz,
                                   const int* boff, const Tile* tiles, int naux, int nbf,
                                   const uint16_t* ghi, const uint16_t* glo, int ng, float* out,
                                   int ldo);
template <int BM>
__global__ void k_seg_syrk_bf16x3(const float* W, int nseg, int64_t sstride, int ldw, int n,
                                  const int2* tiles, int ntile, int segs_per_chunk, float* partial);

// FP32 exchange on the padded copy of B~ (kernels_f32dl.hip)
__global__ void k_pad_f32_slab(const double* src, int N, const int* ncol, const int* bsz, int naux,
                               const int* ldp, const int64_t* off, float* out);
__global__ void k_gather_ct_f32(const double* C, int ldc, int ng, const int* ncol, const int* ldp,
                                const int64_t* ldp_pre, const int* colmap, const int* colmap_off, float* g);
template <int WM, int WN, int TM, int TN, int S>
__global__ void k_slab_gemm_f32_dl(const float* data, const int64_t* off, const int* ldp, const int64_t* ldp_pre,
                                   const int* bsz, const int* boff, const Tile* tiles, int naux, int nbf,
                                   const float* g, int ng, float* out, int ldo);

// FP64 exchange on padded operands (kernels_f64dl.hip)
__global__ void k_pad_f64_slab(const double* src, int N, const int* ncol, const int* bsz, int naux,
                               const int* ldp, const int64_t* off, double* out);
__global__ void k_gather_ct_f64(const double* C, int ldc, int ng, const int* ncol, const int* ldp,
                                const int64_t* ldp_pre, const int* colmap, const int* colmap_off, double* g);
template <int WM, int WN, int TM, int TN, int S>
__global__ void k_slab_gemm_f64_dl(const double* data, const int64_t* off, const int* ldp, const int64_t* ldp_pre,
                                   const int* bsz, const int* boff, const Tile* tiles, int naux, int nbf,
                                   const double* g, int ng, double* out, int ldo);
template <int S>
__global__ void k_seg_syrk_f64_dl(const double* W, const double* W2, int nseg, int64_t sstride, int ldw, int n,
                                  const int2* tiles, int ntile, int segs_per_chunk, double* partial);
__global__ void k_reduce_full(const double* partial, int nchunk, int64_t n2, double alpha, double* out);
template <int S>
__global__ void k_seg_syrk_f32_dl(const float* W, int nseg, int64_t sstride, int ldw, int n, const int2* tiles,
                                  int ntile, int segs_per_chunk, double* partial);

// Dense FP64 GEMM on MFMA (kernels_dgemm.hip): C = alpha op(A) op(B) + beta C,
// C row-major (ldc); A(m,k) = ak ? A[m*lda + k] : A[k*lda + m]; B(k,n) =
// bk ? B[n*ldb + k] : B[k*ldb + n].  tri: zeros of triangular operands whose
// contraction range a tile skips (DG_TRI_*).  Asynchronous on st.
constexpr int DG_TRI_A_LOWER = 1;   // A(m,k) = 0 for k > m
constexpr int DG_TRI_A_UPPER = 2;   // A(m,k) = 0 for k < m
constexpr int DG_TRI_B_LOWER = 4;   // B(k,n) = 0 for k < n
constexpr int DG_TRI_B_UPPER = 8;   // B(k,n) = 0 for k > n
void dgemm_rm(hipStream_t st, bool ak, bool bk, int64_t M, int64_t N, int K, const double* A, int64_t lda,
              const double* B, int64_t ldb, double* C, int64_t ldc, double alpha, double beta, int tri);
// Y = alpha X Wop for a column-major m x w panel X (w <= 64) and a lower-
// triangular w x w W (column-major, ldw); Wop = W^T if trans; Y may be X.
void trmm_panel(hipStream_t st, int m, int w, const double* X, int ldx, const double* W, int ldw, bool trans,
                double alpha, double* Y, int ldy);

// Blocks of a one-thread-per-element launch.  A dispatch holds at most 2^32 - 1
// work-items per grid dimension; larger element counts must use a grid-stride
// kernel launched with gs_blocks.
inline unsigned nblk(int64_t n, int bs)
{
  const int64_t b = (n + bs - 1) / bs;
  if (b * bs >= ((int64_t)1 << 32)) throw std::runtime_error("1-D launch exceeds 2^32 work-items");
  return (unsigned)b;
}

// Blocks of a grid-stride launch (bounded, at least 1).
inline unsigned gs_blocks(int64_t n, int bs = 256)
{
  const int64_t b = (n + bs - 1) / bs;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 65536));
}

}  // namespace mchip
