// Kernel declarations shared by the device drivers.
#pragma once

#include "dev.hpp"

namespace mchip {

constexpr int SG_BM = 64, SG_BN = 64, SG_BK = 16;
constexpr int J_ROWS = 16;   // aux rows per block in k_j_gamma

__global__ void k_slab_gemm_f64(const double* bdata, const int64_t* slab_off, const int* ncol,
                                const int* bsz, const int* boff, const int* colmap,
                                const int* colmap_off, const Tile* tiles, int naux, int nbf,
                                const double* G, int ldg, int ng, double* out, int ldo,
                                double alpha);
__global__ void k_seg_gemm_nt_f64(const double* A, const double* B, int nseg, int64_t sstride,
                                  int lda, int klen, int n, const int2* tiles, int ntile,
                                  int segs_per_chunk, double* partial);
__global__ void k_reduce_sym(const double* partial, int nchunk, int n, int sym2, double alpha,
                             double* out);
__global__ void k_pair_gather(const double* P, int nbf, const int* ncol, const int* bsz,
                              const int* boff, const int* colmap, const int* colmap_off,
                              const int64_t* pg_off, double* Pg);
__global__ void k_pair_scatter(const double* Jg, int nbf, const int* ncol, const int* bsz,
                               const int* boff, const int* colmap, const int* colmap_off,
                               const int64_t* pg_off, double* J);
__global__ void k_j_gamma(const double* bdata, const int64_t* slab_off, const int* ncol,
                          const int* bsz, const int64_t* pg_off, const double* Pg, int naux,
                          double* gpart);
__global__ void k_reduce_slabs(const double* gpart, int nslab, int naux, double scale, double* g);
__global__ void k_j_cols(const double* bdata, const int64_t* slab_off, const int* ncol,
                         const int* bsz, const int64_t* pg_off, const double* gq, int naux,
                         int xchunk, int64_t pg_total, double* Jpart);
__global__ void k_reduce_chunks(const double* part, int nchunk, int64_t n, double* out);
__global__ void k_fill(double* p, int64_t n, double v);
__global__ void k_axpby(double* y, const double* x, int64_t n, double a, double b);

__global__ void k_to_f32(const double* in, float* out, int64_t n);
__global__ void k_rows_to_f32(const double* in, int ld64, int rows, int ncols, float* out);
__global__ void k_slab_gemm_f32(const float* bdata, const int64_t* slab_off, const int* ncol,
                                const int* bsz, const int* boff, const int* colmap,
                                const int* colmap_off, const Tile* tiles, int naux, int nbf,
                                const float* G, int ldg, int ng, float* out, int ldo);
__global__ void k_seg_gemm_nt_f32(const float* A, const float* B, int nseg, int64_t sstride,
                                  int lda, int klen, int n, const int2* tiles, int ntile,
                                  int segs_per_chunk, double* partial);

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace mchip
