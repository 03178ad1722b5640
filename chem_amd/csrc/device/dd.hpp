// Error-free transformations for compensated FP64 sums and dot products
// (Ogita, Rump and Oishi, "Accurate sum and dot product", SIAM J. Sci.
// Comput. 26 (2005): TwoSum after Knuth, TwoProduct by one fused multiply-add).
//
// A running sum is kept as an unevaluated pair (s, c): s the FP64 sum of the
// rounded terms, c the sum of every rounding error made on the way.  s + c
// evaluated once at the end is as accurate as if the sum had been carried in
// twice the working precision, then rounded (Dot2 / Sum2 of that paper).
//
// Contraction of a * b + s into one fma would destroy the error terms, so the
// helpers switch floating-point contraction off and spell out the one fma the
// product error needs.
#pragma once

#include <hip/hip_runtime.h>

namespace mchip {

struct dsum {
  double s = 0.0, c = 0.0;
};

// s + e == a + b exactly, s = fl(a + b)
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
#pragma clang fp contract(off)
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// acc += a * b with both rounding errors kept
__device__ __forceinline__ void dot2_add(dsum& acc, double a, double b)
{
#pragma clang fp contract(off)
  const double p = a * b;
  const double pe = __fma_rn(a, b, -p);
  double s, e;
  two_sum(acc.s, p, s, e);
  acc.s = s;
  acc.c += e + pe;
}

// acc += x (an FP64 term)
__device__ __forceinline__ void sum2_add(dsum& acc, double x)
{
#pragma clang fp contract(off)
  double s, e;
  two_sum(acc.s, x, s, e);
  acc.s = s;
  acc.c += e;
}

// acc += other (pairs combined: their sums by TwoSum, the error terms added)
__device__ __forceinline__ void dsum_add(dsum& acc, const dsum& o)
{
#pragma clang fp contract(off)
  double s, e;
  two_sum(acc.s, o.s, s, e);
  acc.s = s;
  acc.c += o.c + e;
}

__device__ __forceinline__ double dsum_value(const dsum& a)
{
#pragma clang fp contract(off)
  return a.s + a.c;
}

// Sum of a pair over the 64 lanes of a wave (butterfly; every lane ends with
// the total).
__device__ __forceinline__ void dsum_wave(dsum& a)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    dsum o;
    o.s = __shfl_xor(a.s, off, 64);
    o.c = __shfl_xor(a.c, off, 64);
    dsum_add(a, o);
  }
}

}  // namespace mchip
