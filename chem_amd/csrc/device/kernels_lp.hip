// Reduced-precision kernels of the chosen DF-K form on gfx950.
//
//   k_to_f32             B~ (FP64) -> FP32 copy, once per geometry
//   k_to_f32x2           FP64 -> FP32 pair hi + lo (x = hi + lo to ~2^-48)
// The FP32 MFMA contractions of the exchange are in kernels_f32dl.hip, those
// of MP2 are the FP32 instantiations in kernels_gemm.hip.
#include "kernels.hpp"

namespace mchip {

// grid-stride (launch with gs_blocks(n))
__global__ void k_to_f32(const double* __restrict__ in, float* __restrict__ out, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}

// hi = fl32(x), lo = fl32(x - hi): the pair carries x to about 48 bits, so a
// product with an FP32 operand loses no more than that operand's rounding.
__global__ void k_to_f32x2(const double* __restrict__ in, float* __restrict__ hi, float* __restrict__ lo,
                           int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double x = in[i];
    const float h = (float)x;
    hi[i] = h;
    lo[i] = (float)(x - (double)h);
  }
}

}  // namespace mchip
