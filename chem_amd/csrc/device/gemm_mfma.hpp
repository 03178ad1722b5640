// MFMA GEMM building blocks shared by the DF-K / MP2 contraction kernels.
//
// Tiles: BM x BN per 256-thread workgroup, 2 x 2 waves, each wave owns a
// (BM/2) x (BN/2) sub-tile made of 16x16 MFMA tiles; BK = 16 along the
// contraction.  Global loads of step s+1 are issued into registers before the
// MFMAs of step s and written to the other LDS buffer afterwards (register
// prefetch, double-buffered LDS, one barrier per step).
//
// MFMA lane layouts (pinned by tests/test_gpu_mfma_layout.py):
//   FP64 v_mfma_f64_16x16x4_f64: A[i][k] lane i+16k, B[k][j] lane j+16k,
//        D[lane/16 + 4e][lane%16]
//   FP32 v_mfma_f32_16x16x4_f32: same operands, D[4*(lane/16) + e][lane%16]
#pragma once

#include <hip/hip_runtime.h>

namespace mchip {

template <class T>
struct Mfma;

template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(double a, double b, v4 c)
  {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int e) { return (lane >> 4) + 4 * e; }
};

template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(float a, float b, v4 c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int e) { return 4 * (lane >> 4) + e; }
};

constexpr int GM_BK = 16;

// Adds FP32 MFMA accumulators into FP64 running sums and clears them: called
// after every BK step, no FP32 sum runs over more than GM_BK products.
template <int TM, int TN>
__device__ __forceinline__ void flush_f64(Mfma<float>::v4 (&acc)[TM][TN], double (&acc64)[TM][TN][4])
{
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
#pragma unroll
      for (int e = 0; e < 4; ++e) acc64[i][j][e] += (double)acc[i][j][e];
      acc[i][j] = Mfma<float>::v4{0.f, 0.f, 0.f, 0.f};
    }
}

// One BK step of MFMAs on LDS tiles As[k][m], Bs[k][n] (leading dims LDA, LDB).
template <class T, int TM, int TN, int LDA, int LDB>
__device__ __forceinline__ void mfma_step(const T* As, const T* Bs, int wm0, int wn0, int lane,
                                          typename Mfma<T>::v4 (&acc)[TM][TN])
{
#pragma unroll
  for (int kq = 0; kq < GM_BK; kq += 4) {
    const int kr = kq + (lane >> 4);
    T a[TM], b[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) a[i] = As[kr * LDA + wm0 + i * 16 + (lane & 15)];
#pragma unroll
    for (int j = 0; j < TN; ++j) b[j] = Bs[kr * LDB + wn0 + j * 16 + (lane & 15)];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<T>::op(a[i], b[j], acc[i][j]);
  }
}

}  // namespace mchip
